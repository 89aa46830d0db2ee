// felt_stencil.hip — the isogrid's numeric API on the device, and the speed functor queried at
// chosen positions.
//
// felt_surface_stencil* evaluate the grid operators a speed functor calls (NumericMixin.hpp:
// 72-500) through Felt::device::Grid, the same header-only view a caller's device functor is
// compiled against; felt_surface_speed_eval runs the update kernel's own per-voxel speed code
// (speed_voxel) at listed positions.
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../include/felt_amd/device_grid.hpp"
#include "felt_kernels.h"
#include "felt_surface.h"

using namespace felt;

namespace {

constexpr int TPB = 256;

template <int D, class T>
__global__ void k_stencil(Felt::device::Grid<D> grid, int op, int n, const T* pos, float* out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    std::array<T, D> p{};
    for (int a = 0; a < D; a++) p[a] = pos[(size_t)i * D + a];
    Felt::VecDf<D> v{};
    switch (op) {
    case FELT_STENCIL_GET:
        out[i] = grid.get(p);
        return;
    case FELT_STENCIL_CURV:
        out[i] = grid.curv(p);
        return;
    case FELT_STENCIL_DIVERGENCE:
        out[i] = grid.divergence(p);
        return;
    case FELT_STENCIL_GRAD:
        v = grid.grad(p);
        break;
    case FELT_STENCIL_GRAD_E:
        v = grid.gradE(p);
        break;
    case FELT_STENCIL_GRAD_F:
        v = grid.gradF(p);
        break;
    case FELT_STENCIL_GRAD_B:
        v = grid.gradB(p);
        break;
    default:
        v = grid.gradC(p);
        break;
    }
    for (int a = 0; a < D; a++) out[(size_t)i * D + a] = v[a];
}

// Voxel ids of member m's grid positions (3 ints each); positions in inactive partitions count
// in *bad.
template <int D>
__global__ void k_pos_voxels(Geo G, Pool P, int m, const int* gpos, int n, uint32_t* out,
                             int* bad) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int q[3] = {gpos[i * 3], gpos[i * 3 + 1], gpos[i * 3 + 2]};
    int p, l;
    locate<D>(G, P, m, q, p, l);
    const int s = P.page[p];
    if (s < 0) {
        atomicAdd(bad, 1);
        out[i] = 0;
        return;
    }
    out[i] = (uint32_t)s * (uint32_t)G.pv + (uint32_t)l;
}

template <int D>
__global__ void k_speed_eval(Geo G, Pool P, const uint32_t* vox, int n, const Speed* sp,
                             float* out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = speed_voxel<D>(G, P, vox[i], sp);
}

int blocks(size_t n) { return (int)((n + TPB - 1) / TPB); }

#define HIPX(expr)                                                                              \
    do {                                                                                       \
        hipError_t e_ = (expr);                                                                \
        if (e_ != hipSuccess)                                                                  \
            return felt_fail(FELT_E_DEVICE, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

}  // namespace

void felt_fill_device_grid(const felt_surface* s, int m, felt_device_grid* out) {
    const Geo& G = s->G;
    const Mem& M = s->mems[m];
    felt_device_grid d{};
    d.D = G.D;
    d.L = G.L;
    for (int a = 0; a < 3; a++) {
        d.size[a] = M.size[a];
        d.offset[a] = M.offset[a];
        d.part[a] = G.part[a];
        d.part_shift[a] = G.fpart[a].sh;
        d.cstride[a] = M.cstr[a];
        d.tile_shift[a] = G.tsh[a];
        d.tile_stride[a] = G.tstr[a];
        d.in_stride[a] = G.wstr[a];
        d.tiles[a] = G.part[a] >> G.tsh[a];
    }
    d.pbase = M.pbase;
    d.tile_vshift = G.tvsh;
    d.pv = (uint32_t)G.pv;
    d.dx = 1.0f;
    d.page = s->P.page;
    d.pbg = s->P.pbg;
    d.phi = s->P.phi;
    d.slot_org = reinterpret_cast<const int32_t*>(s->P.slot_org);
    *out = d;
}

// op at n world positions of member m: integer positions (ipos) or real ones (fpos).
int felt_stencil_impl(felt_surface* s, int m, int op, float dx, size_t n, const int32_t* ipos,
                      const float* fpos, float* out) {
    if (op < FELT_STENCIL_GET || op > FELT_STENCIL_GRAD_C)
        return felt_fail(FELT_E_ARG, "unknown stencil operator");
    if (!(dx > 0.0f) || dx != dx) return felt_fail(FELT_E_ARG, "dx must be positive");
    if (n == 0) return FELT_OK;
    if (n > (size_t)0x7fffffff / 4) return felt_fail(FELT_E_ARG, "too many positions");
    const int D = s->G.D;
    const size_t width = op >= FELT_STENCIL_GRAD ? (size_t)D : 1;
    int rc;
    if ((rc = s->tmp_f.ensure(n * D + n * width + 1))) return rc;
    float* d_out = s->tmp_f.ptr + n * D;
    felt_device_grid dg;
    felt_fill_device_grid(s, m, &dg);
    dg.dx = dx;
    if (fpos) {
        HIPX(hipMemcpyAsync(s->tmp_f.ptr, fpos, n * D * sizeof(float), hipMemcpyHostToDevice, s->st));
        if (D == 2)
            hipLaunchKernelGGL((k_stencil<2, float>), dim3(blocks(n)), dim3(TPB), 0, s->st,
                               Felt::device::Grid<2>{dg}, op, (int)n, s->tmp_f.ptr, d_out);
        else
            hipLaunchKernelGGL((k_stencil<3, float>), dim3(blocks(n)), dim3(TPB), 0, s->st,
                               Felt::device::Grid<3>{dg}, op, (int)n, s->tmp_f.ptr, d_out);
    } else {
        if ((rc = s->tmp_i.ensure(n * D + 1))) return rc;
        HIPX(hipMemcpyAsync(s->tmp_i.ptr, ipos, n * D * sizeof(int), hipMemcpyHostToDevice, s->st));
        if (D == 2)
            hipLaunchKernelGGL((k_stencil<2, int>), dim3(blocks(n)), dim3(TPB), 0, s->st,
                               Felt::device::Grid<2>{dg}, op, (int)n, s->tmp_i.ptr, d_out);
        else
            hipLaunchKernelGGL((k_stencil<3, int>), dim3(blocks(n)), dim3(TPB), 0, s->st,
                               Felt::device::Grid<3>{dg}, op, (int)n, s->tmp_i.ptr, d_out);
    }
    HIPX(hipGetLastError());
    HIPX(hipMemcpyAsync(out, d_out, n * width * sizeof(float), hipMemcpyDeviceToHost, s->st));
    HIPX(hipStreamSynchronize(s->st));
    return FELT_OK;
}

// The speed functor at member m's grid positions gpos (3 ints each, inside the grid).
int felt_speed_eval_impl(felt_surface* s, int m, const felt_speed* fn, size_t n,
                         const int32_t* gpos, float* out) {
    if (n == 0) return FELT_OK;
    if (n > (size_t)0x7fffffff / 4) return felt_fail(FELT_E_ARG, "too many positions");
    int rc;
    if ((rc = s->tmp_i.ensure(n * 3 + 1))) return rc;
    if ((rc = s->tmp_u.ensure(n + 1))) return rc;
    if ((rc = s->tmp_f.ensure(n + 1))) return rc;
    // the functor of member m; the other members' descriptors are left as they are
    s->h_speed[m] = Speed{fn->kind, fn->a, fn->b, fn->c, (unsigned long long)fn->seed, fn->cell};
    s->speed_valid = false;  // d_speed no longer holds the last global update's functors
    HIPX(hipMemcpyAsync(s->d_speed, s->h_speed, s->G.nmem * sizeof(Speed), hipMemcpyHostToDevice,
                        s->st));
    HIPX(hipMemcpyAsync(s->tmp_i.ptr, gpos, n * 3 * sizeof(int), hipMemcpyHostToDevice, s->st));
    HIPX(hipMemsetAsync(&s->ctr->misc, 0, sizeof(int), s->st));
    const Geo& G = s->G;
    if (G.D == 2)
        hipLaunchKernelGGL(k_pos_voxels<2>, dim3(blocks(n)), dim3(TPB), 0, s->st, G, s->P, m,
                           s->tmp_i.ptr, (int)n, s->tmp_u.ptr, &s->ctr->misc);
    else
        hipLaunchKernelGGL(k_pos_voxels<3>, dim3(blocks(n)), dim3(TPB), 0, s->st, G, s->P, m,
                           s->tmp_i.ptr, (int)n, s->tmp_u.ptr, &s->ctr->misc);
    int bad = 0;
    HIPX(hipMemcpyAsync(&bad, &s->ctr->misc, sizeof(int), hipMemcpyDeviceToHost, s->st));
    HIPX(hipStreamSynchronize(s->st));
    if (bad) return felt_fail(FELT_E_ARG, "speed_eval: position in an inactive partition");
    if (G.D == 2)
        hipLaunchKernelGGL(k_speed_eval<2>, dim3(blocks(n)), dim3(TPB), 0, s->st, G, s->P,
                           s->tmp_u.ptr, (int)n, s->d_speed, s->tmp_f.ptr);
    else
        hipLaunchKernelGGL(k_speed_eval<3>, dim3(blocks(n)), dim3(TPB), 0, s->st, G, s->P,
                           s->tmp_u.ptr, (int)n, s->d_speed, s->tmp_f.ptr);
    HIPX(hipGetLastError());
    HIPX(hipMemcpyAsync(out, s->tmp_f.ptr, n * sizeof(float), hipMemcpyDeviceToHost, s->st));
    HIPX(hipStreamSynchronize(s->st));
    return FELT_OK;
}
