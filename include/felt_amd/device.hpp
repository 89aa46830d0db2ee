// device.hpp — device functors for Felt::Surface<D,L>::update, for HIP translation units
// (hipcc --offload-arch=gfx950 -ffp-contract=off).
//
//   #include <felt_amd/device.hpp>
//   Felt::Surface<3, 3> s(size, part);
//   s.update(Felt::device_fn([] __device__ (const Felt::VecDi<3>& p,
//                                           const Felt::device::Grid<3>& g) {
//       return 0.2f * g.curv(p);
//   }));
//
// The functor is compiled into k_device_speed below, in the caller's translation unit, and runs
// once per zero-layer voxel on the surface's stream between felt_surface_update_begin and
// felt_surface_update_commit (Surface.hpp:254-296, 310-386): it sees the grid as it was before
// the step, as the reference's functors do, and its result is the voxel's delta.  Nothing is
// copied to the host.  The reference's three call shapes are accepted, tried in this order
// (Surface.hpp:1754-1802): fn(pos, grid), fn(pos), fn(); the result must be a float.
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <type_traits>
#include <utility>

#include "Surface.hpp"
#include "device_grid.hpp"

namespace Felt {
namespace detail {

template <int N>
struct rank : rank<N - 1> {};
template <>
struct rank<0> {};

template <int D, class F>
__device__ auto call_device(F& fn, const VecDi<D>& p, const device::Grid<D>& g, rank<2>)
    -> decltype(fn(p, g)) {
    return fn(p, g);
}
template <int D, class F>
__device__ auto call_device(F& fn, const VecDi<D>& p, const device::Grid<D>&, rank<1>)
    -> decltype(fn(p)) {
    return fn(p);
}
template <int D, class F>
__device__ auto call_device(F& fn, const VecDi<D>&, const device::Grid<D>&, rank<0>)
    -> decltype(fn()) {
    return fn();
}

// delta[i] = fn(position of voxels[i]) for span[0] <= i < span[1] (the span lives on the device,
// so the kernel strides over whatever the step listed).
template <int D, class F>
__global__ void __launch_bounds__(256)
k_device_speed(device::Grid<D> g, const uint32_t* voxels, const int32_t* span, float* delta,
               F fn) {
    const int lo = span[0], hi = span[1];
    for (int i = lo + (int)(blockIdx.x * blockDim.x + threadIdx.x); i < hi;
         i += (int)(gridDim.x * blockDim.x)) {
        const VecDi<D> p = g.position(voxels[i]);
        const auto r = call_device<D>(fn, p, g, rank<2>{});
        static_assert(std::is_same<std::decay_t<decltype(r)>, Distance>::value,
                      "a speed functor must return Felt::Distance (float)");
        delta[i] = r;
    }
}

template <int D, class F>
struct DeviceSpeed {
    // Launches fn over the job's zero layer on the job's stream; an empty string on success.
    static std::string launch(const felt_zero_job& job, F& fn) {
        const uint64_t n = job.capacity ? job.capacity : 1;
        const unsigned blocks = (unsigned)(n / 256 + 1 < 4096 ? n / 256 + 1 : 4096);
        hipLaunchKernelGGL((k_device_speed<D, F>), dim3(blocks), dim3(256), 0,
                           (hipStream_t)job.stream, device::Grid<D>{job.grid}, job.voxels,
                           job.span, job.delta, fn);
        const hipError_t e = hipGetLastError();
        return e == hipSuccess ? std::string() : std::string(hipGetErrorString(e));
    }
};

}  // namespace detail
}  // namespace Felt
