// felt_kernels.h — HIP kernels of the sparse-field update (gfx950).  Each kernel restates one
// phase of Surface<D,L>::update (Surface.hpp:254-1440) over compacted voxel lists.
#pragma once

#include "felt_dev.h"
#include "../../include/felt_amd.h"

namespace felt {

// Device counters of one surface.
// Every list length a kernel needs lives here, so launches never wait on the host.  A span
// {lo, hi} is a half-open range of list positions.
struct Ctr {
    int free_top;        // free-slot stack height
    int sc_n;            // status-change records this round
    int exp_n;           // of which expanding (from +-L to +-(L-1))
    int lu_span[2][2];   // {0, entries} of the two lookups (affected / affected buffer)
    int flag;            // a non-zero-layer status change happened this round
    int need_n;          // partitions to activate
    int nan_events;
    int act_n;           // active partitions
    int z_span[2];       // {0, zero-layer delta entries} of a bounded update
    int band_span[2];    // {0, band entries}
    int err;             // 1: brick pool exhausted; 2: layer counts and lists disagree;
                         // 3: step refused, its worst-case activations exceed the free bricks
    int misc;
    int one;             // constant 1: the live flag of a round that always runs
    int live[8];         // round r of an unrolled step runs while live[r] != 0
    int band_total;      // tracked voxels, refreshed at the end of an unrolled step
    int front[20];       // breadth-first frontier bounds of calc_affected
    int reserve;         // bricks the step may activate at most (k_reserve_count)
    int dead_n;          // partitions deactivated since the neighbour records were last fixed
    unsigned stamp;      // this step's stamp in the reserve count's partition stamps
    // A bounded step's affected set sorted by layer for its first round: the entries of layer
    // k - L (k != L) are seg0[k] .. seg0[k + 1]; seg0_cnt and seg0_cur count and place them.
    int seg0[18];
    int seg0_cnt[17];
    int seg0_cur[17];
};

// Expansion records (expand_narrow_band, Surface.hpp:1249-1355): the status changes out of the
// outermost layers, kept for the expansion after the flush.
struct SC {
    uint32_t* xv;   // the expanding records (outermost layer to the next one in): voxel
    int8_t* xside;  // and side
};

// A lookup: the reference's AffectedLookupGrid as a tagged voxel list (tag = layer).
struct Lu {
    uint32_t* v;
    int8_t* tag;
    float* d;    // distance scratch aligned with v
    int8_t* tk;  // layer of the new value, aligned with v
};

struct Speed {
    int kind;
    float a, b, c;
    unsigned long long seed;
    int cell;
};

__device__ __forceinline__ unsigned long long splitmix64(unsigned long long x) {
    x += 0x9E3779B97F4A7C15ULL;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ULL;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBULL;
    return x ^ (x >> 31);
}

// Value noise: lattice values U[-1,1) from SplitMix64 of the lattice point, N-linear
// interpolation reducing axis 0 first, fp32 without contraction.
template <int D>
__device__ __forceinline__ float value_noise(const int w[3], unsigned long long seed, int cell) {
    int base[3] = {0, 0, 0};
    float t[3] = {0.f, 0.f, 0.f};
#pragma unroll
    for (int a = 0; a < D; a++) {
        const int q = w[a] >= 0 ? w[a] / cell : -((-w[a] + cell - 1) / cell);
        base[a] = q;
        t[a] = (float)(w[a] - q * cell) / (float)cell;
    }
    const unsigned long long mul[3] = {0x9E3779B97F4A7C15ULL, 0xC2B2AE3D27D4EB4FULL,
                                       0x165667B19E3779F9ULL};
    // corner i's key, sum over a of (base[a] + bit a of i) * mul[a] mod 2^64: the base's
    // products once, plus the multipliers of the corner's bits
    unsigned long long key0 = 0;
#pragma unroll
    for (int a = 0; a < D; a++) key0 += (unsigned long long)(long long)base[a] * mul[a];
    float val[8];
#pragma unroll
    for (int i = 0; i < (1 << D); i++) {
        unsigned long long key = key0;
#pragma unroll
        for (int a = 0; a < D; a++)
            if ((i >> a) & 1) key += mul[a];
        const unsigned long long h = splitmix64(seed ^ key);
        const float u = (float)(h >> 40) * (1.0f / 16777216.0f);
        val[i] = u * 2.0f - 1.0f;
    }
    int n = 1 << D;
#pragma unroll
    for (int axis = 0; axis < D; axis++) {
        const float wt = t[axis];
        for (int i = 0; i < n / 2; i++) val[i] = wt * val[2 * i + 1] + (1.0f - wt) * val[2 * i];
        n /= 2;
    }
    return val[0];
}

// The speed functor sp at a voxel whose curvature is curv (when the functor uses it) and whose
// world position is w (for the noise).
template <int D>
__device__ __forceinline__ float speed_from(const Speed& sp, float curv, const int w[3]) {
    switch (sp.kind) {
    case FELT_SPEED_CONSTANT:
        return sp.a;
    case FELT_SPEED_CURVATURE:
        return sp.b * curv;
    case FELT_SPEED_CONST_CURV:
        return sp.a + sp.b * curv;
    default: {
        const float n = value_noise<D>(w, sp.seed, sp.cell);
        float r = sp.a * (1.0f + sp.c * n) + sp.b * curv;
        if (r < -1.0f) r = -1.0f;
        else if (r > 1.0f) r = 1.0f;
        return r;
    }
    }
}

// The speed functor of voxel v's member (sp[member]) at v: the curvature stencil is gathered in
// levels (gather) and reduced in the reference's order (curvature_of).  The update kernel and the
// speed query both call this.
template <int D>
__device__ __forceinline__ float speed_voxel(const Geo& G, const Pool& P, uint32_t v,
                                             const Speed* sp) {
    VoxCtx c;
    vox_ctx<D>(G, P, v, c);
    const Speed s = sp[c.m];
    float curv = 0.0f;
    if (s.kind != FELT_SPEED_CONSTANT) {
        float val[Stencil<D>::K];
        const int (*off)[3] = D == 3 ? kStencil3 : kStencil2;
        // partition edges of 1 (a voxel on both sides of its brick) take the general gather
        if (G.part[0] > 1 && (D < 2 || G.part[1] > 1) && (D < 3 || G.part[2] > 1))
            gather_sep<D, Stencil<D>::K>(G, P, c, off, val);
        else
            gather<D, Stencil<D>::K>(G, P, c, off, val);
        curv = curvature_of<D>(P.phi[c.v], val);
    }
    const Mem& M = P.mem[c.m];
    int w[3] = {0, 0, 0};
    for (int a = 0; a < D; a++) w[a] = c.g[a] + M.offset[a];
    return speed_from<D>(s, curv, w);
}

}  // namespace felt
