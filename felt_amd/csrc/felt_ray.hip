// felt_ray.hip — Surface::ray on the device: one thread per ray (Surface.hpp:468-616, 1450-1598).
//
// A ray first tries the partition holding its origin, then every partition whose face it
// crosses, nearest first.  The reference gathers those crossings into a list, sorts it by squared
// distance from the origin and drops consecutive repeats of a partition; here a thread instead
// re-walks the crossing planes and takes the smallest (distance, crossing order) after the last
// one it visited, which is the same order as the reference's sort with ties kept in crossing
// order, and needs no per-ray list.  Within a partition the ray samples the field every half
// voxel; at the first zero-layer sample whose outward normal faces the ray it walks down the
// normal onto the zero curve (at most 100 steps).  All arithmetic is fp32 in the reference's
// operation order (no contraction), so hits agree with the fp32 reference.
#include <hip/hip_runtime.h>

#include <cfloat>
#include <cmath>
#include <mutex>
#include <string>
#include <vector>

#include "felt_surface.h"

using namespace felt;

namespace {

constexpr int TPB = 256;
constexpr float kTiny = 0.00001f;  // Surface.hpp:67
constexpr int kConverge = 100;      // MAX_CONVERGE_STEPS, Surface.hpp:1486

// Member geometry in registers: grid offset / size, children grid and pool partition base.
struct RayMem {
    int off[3], size[3], csize[3], cstr[3];
    int pbase;
};

template <int D>
__device__ __forceinline__ float sumv(const float v[3]) {
    return D == 3 ? v[0] + (v[1] + v[2]) : v[0] + v[1];  // Eigen's unrolled reduction order
}

template <int D>
__device__ __forceinline__ float dotv(const float a[3], const float b[3]) {
    float p[3] = {0.f, 0.f, 0.f};
#pragma unroll
    for (int i = 0; i < D; i++) p[i] = a[i] * b[i];
    return sumv<D>(p);
}

// Real-valued Util::inside over the member's grid (Util.hpp:94-106); NaN is outside.
template <int D>
__device__ __forceinline__ bool in_grid(const RayMem& M, const float p[3]) {
#pragma unroll
    for (int a = 0; a < D; a++)
        if (!(p[a] < (float)(M.off[a] + M.size[a]) && p[a] >= (float)M.off[a])) return false;
    return true;
}

// N-linear interpolation of the partitioned isogrid (NumericMixin.hpp:403-481): corners are read
// through get(), so positions outside the grid or in inactive partitions give the background.
template <int D>
__device__ float interp(const Geo& G, const Pool& P, int m, const RayMem& M, const float pos[3]) {
    int fl[3] = {0, 0, 0};
    float t[3] = {0.f, 0.f, 0.f};
#pragma unroll
    for (int a = 0; a < D; a++) {
        const float f = floorf(pos[a]);
        fl[a] = (int)f - M.off[a];
        t[a] = pos[a] - f;
    }
    // the 2^D corners in levels: partition and leaf, then page, then value
    constexpr int NC = 1 << D;
    float v[NC];
    int part[NC], leaf[NC], pg[NC];
    bool in[NC];
#pragma unroll
    for (int i = 0; i < NC; i++) {
        int lc[3] = {0, 0, 0};
        int p = M.pbase;
        bool inside = true;
#pragma unroll
        for (int a = 0; a < D; a++) {
            const int q = fl[a] + ((i >> a) & 1);
            if (q < 0 || q >= M.size[a]) inside = false;
            const int qa = (int)fdiv((uint32_t)q, G.fpart[a]);
            lc[a] = q - qa * G.part[a];
            p += qa * M.cstr[a];
        }
        in[i] = inside;
        part[i] = p;
        leaf[i] = leaf_of<D>(G, lc);
    }
#pragma unroll
    for (int i = 0; i < NC; i++) pg[i] = in[i] ? P.page[part[i]] : -1;
#pragma unroll
    for (int i = 0; i < NC; i++) {
        float x = (float)(G.L + 1);
        if (in[i]) x = pg[i] < 0 ? P.pbg[part[i]] : P.phi[(size_t)pg[i] * G.pv + leaf[i]];
        v[i] = x;
    }
    int n = 1 << D;
#pragma unroll
    for (int axis = 0; axis < D; axis++) {
        const float w = t[axis];
#pragma unroll
        for (int i = 0; i < (1 << (D - 1)); i++)
            if (i < n / 2) v[i] = w * v[2 * i + 1] + (1.0f - w) * v[2 * i];
        n /= 2;
    }
    return v[0];
}

// Safe gradient at a real position (NumericMixin.hpp:188-231): the probe is stepped -1, +2, -1
// in fp32 along each axis in turn, as the reference does.
template <int D>
__device__ void grad(const Geo& G, const Pool& P, int m, const RayMem& M, const float pos[3],
                     float gv[3]) {
    float pt[3] = {pos[0], pos[1], pos[2]};
    const float centre = interp<D>(G, P, m, M, pos);
#pragma unroll
    for (int a = 0; a < D; a++) {
        float back = centre, fwd = centre;
        int order = 0;
        pt[a] -= 1.0f;
        if (in_grid<D>(M, pt)) {
            back = interp<D>(G, P, m, M, pt);
            order++;
        }
        pt[a] += 2.0f;
        if (in_grid<D>(M, pt)) {
            fwd = interp<D>(G, P, m, M, pt);
            order++;
        }
        pt[a] -= 1.0f;
        gv[a] = order != 0 ? (fwd - back) / (float)order : 0.0f;
    }
}

template <int D>
__device__ __forceinline__ void normalize(float n[3]) {
    float p[3] = {0.f, 0.f, 0.f};
#pragma unroll
    for (int i = 0; i < D; i++) p[i] = n[i] * n[i];
    const float z = sumv<D>(p);
    if (z > 0.0f) {
        const float r = sqrtf(z);
#pragma unroll
        for (int i = 0; i < D; i++) n[i] = n[i] / r;
    }
}

// layer_id(interp(pos)) == 0, i.e. iround(v + FLT_EPSILON) == 0 (Surface.hpp:1677-1697).
__device__ __forceinline__ bool zero_layer(float v) {
    return __builtin_roundf(v + kEps) == 0.0f;
}

// Ray through one partition (Surface.hpp:1450-1527).  `lim` bounds the half-voxel samples by the
// most a unit direction can take inside one partition.
template <int D>
__device__ bool ray_child(const Geo& G, const Pool& P, int m, const RayMem& M, const float org[3],
                          const float dir[3], const int c[3], int lim, float hit[3]) {
    int lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
#pragma unroll
    for (int a = 0; a < D; a++) {
        lo[a] = c[a] * G.part[a] + M.off[a];
        hi[a] = lo[a] + G.part[a];
    }
    float ps[3] = {org[0], org[1], org[2]};
    float t = 0.0f;
    for (int it = 0; it < lim; it++) {
        bool in = true;
#pragma unroll
        for (int a = 0; a < D; a++)
            if (!(ps[a] < (float)hi[a] && ps[a] >= (float)lo[a])) in = false;
        if (!in) break;
        if (zero_layer(interp<D>(G, P, m, M, ps))) {
            float n[3] = {0.f, 0.f, 0.f};
            grad<D>(G, P, m, M, ps, n);
            normalize<D>(n);
            if (dotv<D>(n, dir) < 0.0f) {
                for (int step = 0; step < kConverge; step++) {
                    const float dist = interp<D>(G, P, m, M, ps);
#pragma unroll
                    for (int a = 0; a < D; a++) ps[a] = ps[a] - n[a] * dist;
                    if (!in_grid<D>(M, ps)) return false;
                    if (fabsf(dist) <= kTiny || dotv<D>(n, dir) >= 0.0f) break;
                    grad<D>(G, P, m, M, ps, n);
                    normalize<D>(n);
                }
#pragma unroll
                for (int a = 0; a < D; a++) hit[a] = ps[a];
                return true;
            }
        }
        t += 0.5f;
#pragma unroll
        for (int a = 0; a < D; a++) ps[a] = org[a] + dir[a] * t;
    }
    return false;
}

// round_to_next (Surface.hpp:1567-1583): the next partition face along `dim` in direction sd.
__device__ __forceinline__ float round_to_next(const Geo& G, const RayMem& M, int dim, float sd,
                                               float pos) {
    float v = (pos - (float)M.off[dim]) / (float)G.part[dim];
    v = sd == -1.0f ? floorf(v) : ceilf(v);
    v *= (float)G.part[dim];
    v += (float)M.off[dim];
    return v;
}

// One ray_check_track_child (Surface.hpp:1529-1565) against the plane x[dim] = plane: returns
// false when the crossing lies past the grid in the ray's direction (the reference compares the
// upper side with the grid size, not offset + size), else true; `on` is set when the crossing
// is inside the grid, with the crossing point in pi.
template <int D>
__device__ __forceinline__ bool cross(const RayMem& M, const float o[3], const float dir[3],
                                      int dim, float sd, float plane, float pi[3], bool& on) {
    const float nd = -sd;
    const float d = plane * sd;
    const float t = -(d + nd * o[dim]) / (nd * dir[dim]);
#pragma unroll
    for (int i = 0; i < D; i++) pi[i] = (o[i] + dir[i] * t) + dir[i] * kTiny;
    on = false;
#pragma unroll
    for (int i = 0; i < D; i++)
        if ((dir[i] > 0.0f && pi[i] > (float)M.size[i]) ||
            (dir[i] < 0.0f && pi[i] < (float)M.off[i]))
            return false;
    on = in_grid<D>(M, pi);
    return true;
}

// Walks the planes of ray(): per axis with a non-zero direction, the first partition face ahead
// of the origin (or the grid's near face), the rounded face after it, then every face in turn
// until a crossing lies past the grid (Surface.hpp:490-569).  f(order, pi) is called for each
// crossing inside the grid, `order` counting them in the reference's insertion order.  `cap`
// bounds the faces per axis above what any accepted ray reaches.
template <int D, class F>
__device__ void walk_planes(const Geo& G, const RayMem& M, const float o[3], const float dir[3],
                            F&& f) {
    int order = 0;
    float pi[3];
    bool on;
    for (int dim = 0; dim < D; dim++) {
        const float sd = (float)((0.0f < dir[dim]) - (dir[dim] < 0.0f));
        if (sd == 0.0f) continue;
        float ppd = round_to_next(G, M, dim, sd, o[dim]);
        float plane = ppd;
        if (!(ppd >= (float)M.off[dim] && ppd < (float)(M.off[dim] + M.size[dim]))) {
            float gd;
            if (sd == -1.0f) {
                gd = (float)(M.off[dim] + M.size[dim]);
                if (ppd < gd) continue;
            } else {
                gd = (float)M.off[dim];
                if (ppd > gd) continue;
            }
            plane = gd;
        }
        cross<D>(M, o, dir, dim, sd, plane, pi, on);
        if (on) f(order++, pi);
        ppd = round_to_next(G, M, dim, sd, plane);
        if (ppd != plane) {
            plane = ppd;
            if (!cross<D>(M, o, dir, dim, sd, plane, pi, on)) continue;
            if (on) f(order++, pi);
        }
        const float cs = (float)G.part[dim];
        const int cap = 2 * M.csize[dim] + 8;
        for (int k = 0; k < cap; k++) {
            plane += sd * cs;
            if (!cross<D>(M, o, dir, dim, sd, plane, pi, on)) break;
            if (on) f(order++, pi);
        }
    }
}

// Partition of an in-grid crossing and whether it tracks layer 0 or +-1 (its lists non-empty).
template <int D>
__device__ __forceinline__ int tracked_part(const Geo& G, const Pool& P, const RayMem& M,
                                            const float pi[3], int c[3]) {
    int p = M.pbase;
#pragma unroll
    for (int a = 0; a < D; a++) {
        const int q = (int)floorf(pi[a]) - M.off[a];
        c[a] = (int)fdiv((uint32_t)q, G.fpart[a]);
        p += c[a] * M.cstr[a];
    }
    const int* cnt = P.pcnt + (size_t)p * G.NL;
    int n = 0;
    for (int k = max(0, G.L - 1); k <= min(G.NL - 1, G.L + 1); k++) n += cnt[k];
    return n > 0 ? p : -1;
}

template <int D>
__global__ void __launch_bounds__(TPB)
k_ray(Geo G, Pool P, const int* members, int m0, int n, const float* origins, const float* dirs,
      float* out) {
    const int i = blockIdx.x * TPB + threadIdx.x;
    if (i >= n) return;
    const int m = members ? members[i] : m0;
    RayMem M;
    {
        const Mem& src = P.mem[m];
#pragma unroll
        for (int a = 0; a < 3; a++) {
            M.off[a] = src.offset[a];
            M.size[a] = src.size[a];
            M.csize[a] = src.csize[a];
            M.cstr[a] = src.cstr[a];
        }
        M.pbase = src.pbase;
    }
    float o[3] = {0.f, 0.f, 0.f}, dir[3] = {0.f, 0.f, 0.f}, hit[3] = {0.f, 0.f, 0.f};
    float dmax = 0.0f;
    int pmax = 1;
#pragma unroll
    for (int a = 0; a < D; a++) {
        o[a] = origins[(size_t)i * D + a];
        dir[a] = dirs[(size_t)i * D + a];
        dmax = fmaxf(dmax, fabsf(dir[a]));
        pmax = max(pmax, G.part[a]);
    }
    // Samples one partition can hold along a direction whose largest component is dmax (a unit
    // direction has dmax >= 1/sqrt(D)); capped for directions that are not.
    const float lf = 2.0f * (float)pmax / dmax;
    const int lim = lf < 4096.0f ? (int)ceilf(lf) + 4 : 4100;
    bool found = false;

    // The partition of the origin first, chosen by truncating the origin (Surface.hpp:473-484).
    if (in_grid<D>(M, o)) {
        int c[3] = {0, 0, 0};
#pragma unroll
        for (int a = 0; a < D; a++)
            c[a] = (int)fdiv((uint32_t)((int)o[a] - M.off[a]), G.fpart[a]);
        found = ray_child<D>(G, P, m, M, o, dir, c, lim, hit);
    }

    // Then the tracked partitions crossed, by (squared distance, crossing order), skipping a
    // partition equal to the one visited just before (Surface.hpp:578-611).
    float cur_key = -1.0f;
    int cur_ord = -1, prev_p = -1;
    while (!found) {
        float best_key = INFINITY, best_pi[3] = {0.f, 0.f, 0.f};
        int best_ord = 0x7fffffff, best_p = -1, best_c[3] = {0, 0, 0};
        walk_planes<D>(G, M, o, dir, [&](int ord, const float pi[3]) {
            float dv[3] = {0.f, 0.f, 0.f};
#pragma unroll
            for (int a = 0; a < D; a++) dv[a] = pi[a] - o[a];
            const float key = dotv<D>(dv, dv);
            const bool after = key > cur_key || (key == cur_key && ord > cur_ord);
            const bool before = key < best_key || (key == best_key && ord < best_ord);
            if (!after || !before) return;
            int c[3] = {0, 0, 0};
            const int p = tracked_part<D>(G, P, M, pi, c);
            if (p < 0) return;
            best_key = key;
            best_ord = ord;
            best_p = p;
#pragma unroll
            for (int a = 0; a < D; a++) {
                best_pi[a] = pi[a];
                best_c[a] = c[a];
            }
        });
        if (best_p < 0) break;
        cur_key = best_key;
        cur_ord = best_ord;
        if (best_p == prev_p) continue;
        prev_p = best_p;
        found = ray_child<D>(G, P, m, M, best_pi, dir, best_c, lim, hit);
    }
#pragma unroll
    for (int a = 0; a < D; a++) out[(size_t)i * D + a] = found ? hit[a] : FLT_MAX;
}

#define HIPX(expr)                                                                              \
    do {                                                                                       \
        hipError_t e_ = (expr);                                                                \
        if (e_ != hipSuccess)                                                                  \
            return felt_fail(FELT_E_DEVICE, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

// Rays the device path accepts: finite, origins within 2^20 of the grid centre, unit directions
// (Surface::ray takes a normal).  These keep every loop of the kernel within its bounds.
int check_rays(int D, size_t n, const float* origins, const float* dirs) {
    for (size_t i = 0; i < n; i++) {
        double len2 = 0.0;
        for (int a = 0; a < D; a++) {
            const float o = origins[i * D + a], d = dirs[i * D + a];
            if (!std::isfinite(o) || !std::isfinite(d))
                return felt_fail(FELT_E_ARG, "ray " + std::to_string(i) + " is not finite");
            if (std::fabs(o) > 1048576.0f)
                return felt_fail(FELT_E_ARG, "ray " + std::to_string(i) + " origin beyond 2^20");
            len2 += (double)d * d;
        }
        if (len2 < 0.98 || len2 > 1.02)
            return felt_fail(FELT_E_ARG, "ray " + std::to_string(i) + " direction is not a unit vector");
    }
    return FELT_OK;
}

}  // namespace

int felt_ray_device(felt_surface* s, const int* d_members, int m0, size_t n, const float* d_org,
                    const float* d_dir, float* d_out) {
    if (n == 0) return FELT_OK;
    const int blocks = (int)((n + TPB - 1) / TPB);
    if (s->G.D == 2)
        hipLaunchKernelGGL(k_ray<2>, dim3(blocks), dim3(TPB), 0, s->st, s->G, s->P, d_members, m0,
                           (int)n, d_org, d_dir, d_out);
    else
        hipLaunchKernelGGL(k_ray<3>, dim3(blocks), dim3(TPB), 0, s->st, s->G, s->P, d_members, m0,
                           (int)n, d_org, d_dir, d_out);
    HIPX(hipGetLastError());
    return FELT_OK;
}

int felt_ray_impl(felt_surface* s, const int32_t* members, int m0, size_t n, const float* origins,
                  const float* dirs, float* out) {
    int rc;
    if (n == 0) return FELT_OK;
    if (n > (size_t)0x7fffffff / 4) return felt_fail(FELT_E_ARG, "too many rays");
    const int D = s->G.D;
    if ((rc = check_rays(D, n, origins, dirs))) return rc;
    if (members)
        for (size_t i = 0; i < n; i++)
            if (members[i] < 0 || members[i] >= s->G.nmem)
                return felt_fail(FELT_E_ARG, "ray " + std::to_string(i) + " member out of range");
    const size_t nf = n * D;
    if ((rc = s->tmp_f.ensure(3 * nf))) return rc;
    float* d_org = s->tmp_f.ptr;
    float* d_dir = d_org + nf;
    float* d_out = d_dir + nf;
    HIPX(hipMemcpyAsync(d_org, origins, nf * sizeof(float), hipMemcpyHostToDevice, s->st));
    HIPX(hipMemcpyAsync(d_dir, dirs, nf * sizeof(float), hipMemcpyHostToDevice, s->st));
    int* d_mem = nullptr;
    if (members) {
        if ((rc = s->tmp_i.ensure(n))) return rc;
        d_mem = s->tmp_i.ptr;
        HIPX(hipMemcpyAsync(d_mem, members, n * sizeof(int), hipMemcpyHostToDevice, s->st));
    }
    if ((rc = felt_ray_device(s, d_mem, m0, n, d_org, d_dir, d_out))) return rc;
    HIPX(hipMemcpyAsync(out, d_out, nf * sizeof(float), hipMemcpyDeviceToHost, s->st));
    HIPX(hipStreamSynchronize(s->st));
    return FELT_OK;
}

int felt_time_ray_impl(felt_surface* s, size_t n, const float* origins, const float* dirs, int reps,
                       double* us) {
    int rc;
    if (n == 0 || reps < 1) return felt_fail(FELT_E_ARG, "need rays and at least one repetition");
    if (n > (size_t)0x7fffffff / 4) return felt_fail(FELT_E_ARG, "too many rays");
    if ((rc = check_rays(s->G.D, n, origins, dirs))) return rc;
    const size_t nf = n * s->G.D;
    if ((rc = s->tmp_f.ensure(3 * nf))) return rc;
    float* d_org = s->tmp_f.ptr;
    float* d_dir = d_org + nf;
    float* d_out = d_dir + nf;
    HIPX(hipMemcpyAsync(d_org, origins, nf * sizeof(float), hipMemcpyHostToDevice, s->st));
    HIPX(hipMemcpyAsync(d_dir, dirs, nf * sizeof(float), hipMemcpyHostToDevice, s->st));
    if ((rc = felt_ray_device(s, nullptr, 0, n, d_org, d_dir, d_out))) return rc;  // warm-up
    hipEvent_t a = nullptr, b = nullptr;
    float ms = 0.0f;
    hipError_t e = hipEventCreate(&a);
    if (e == hipSuccess) e = hipEventCreate(&b);
    if (e == hipSuccess) e = hipEventRecord(a, s->st);
    for (int r = 0; r < reps && e == hipSuccess; r++) {
        if ((rc = felt_ray_device(s, nullptr, 0, n, d_org, d_dir, d_out))) break;
    }
    if (e == hipSuccess && !rc) e = hipEventRecord(b, s->st);
    if (e == hipSuccess && !rc) e = hipEventSynchronize(b);
    if (e == hipSuccess && !rc) e = hipEventElapsedTime(&ms, a, b);
    if (a) (void)hipEventDestroy(a);
    if (b) (void)hipEventDestroy(b);
    if (rc) return rc;
    if (e != hipSuccess) return felt_fail(FELT_E_DEVICE, std::string("ray timing: ") + hipGetErrorString(e));
    *us = (double)ms * 1e3 / reps;
    return FELT_OK;
}

// ---- rays without device allocation -------------------------------------------------------
// felt_surface_ray_noalloc stages its rays through this array instead of device buffers the
// library would allocate: origins, directions and hits of at most kNoallocRays rays, one block of
// kNoallocRays * 3 floats each (2.25 MiB).  Every device holds its own copy until the process
// ends; host code reaches it only through the address hipGetSymbolAddress returns with that
// device current, never through the array's name.
constexpr size_t kNoallocRays = 65536;
constexpr size_t kNoallocBlock = kNoallocRays * 3;
static __device__ __align__(256) float scratch[3 * kNoallocBlock];

namespace {

std::mutex g_scratch_mu;            // every call on a device shares its array: one call at a time
std::vector<float*> g_scratch_ptr;  // device address of `scratch` per device, resolved once

int scratch_on(int dev, float** out) {
    if ((size_t)dev >= g_scratch_ptr.size()) g_scratch_ptr.resize((size_t)dev + 1, nullptr);
    if (!g_scratch_ptr[dev]) {
        int cur = 0;
        HIPX(hipGetDevice(&cur));
        if (cur != dev) HIPX(hipSetDevice(dev));
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (cur != dev) (void)hipSetDevice(cur);
        if (e != hipSuccess || !p)
            return felt_fail(FELT_E_DEVICE,
                             std::string("hipGetSymbolAddress: ") + hipGetErrorString(e));
        g_scratch_ptr[dev] = static_cast<float*>(p);
    }
    *out = g_scratch_ptr[dev];
    return FELT_OK;
}

}  // namespace

int felt_ray_noalloc_impl(felt_surface* s, size_t n, const float* origins, const float* dirs,
                          float* out) {
    if (n == 0) return FELT_OK;
    if (n > kNoallocRays)
        return felt_fail(FELT_E_ARG, "felt_surface_ray_noalloc takes at most 65536 rays per call");
    const int D = s->G.D;
    int rc;
    if ((rc = check_rays(D, n, origins, dirs))) return rc;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    float* base = nullptr;
    if ((rc = scratch_on(s->device, &base))) return rc;
    const size_t nf = n * D;
    float* d_org = base;
    float* d_dir = base + kNoallocBlock;
    float* d_out = base + 2 * kNoallocBlock;
    HIPX(hipMemcpyAsync(d_org, origins, nf * sizeof(float), hipMemcpyHostToDevice, s->st));
    HIPX(hipMemcpyAsync(d_dir, dirs, nf * sizeof(float), hipMemcpyHostToDevice, s->st));
    if ((rc = felt_ray_device(s, nullptr, 0, n, d_org, d_dir, d_out))) return rc;
    HIPX(hipMemcpyAsync(out, d_out, nf * sizeof(float), hipMemcpyDeviceToHost, s->st));
    HIPX(hipStreamSynchronize(s->st));
    return FELT_OK;
}
