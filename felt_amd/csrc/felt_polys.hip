// felt_polys.hip — per-partition marching cubes / squares on the device (gfx950).
//
// Replaces Felt::Polys<S> and Impl::Poly::Single (Polys.hpp:29-201, Poly.hpp:17-368).  For each
// partition to polygonise, one workgroup visits the band voxels of its brick; the cube of voxel
// x spans x..x+1 on every axis (the reference's pos_calc = x - SpxGridPosOffset with its corner
// table, PolyMixin.hpp:380-409).  A crossed edge yields one vertex per partition, as the
// reference's per-partition lookup does (Poly.hpp:296-346); here the vertex belongs to the
// band cube of lowest leaf index among the cubes of the partition that share the edge, so
// vertex indices come from two block-wide exclusive scans and no atomics.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "felt_polys.h"

using namespace felt;

namespace {

constexpr int TPB = 256;

#include "mc_tables.inc"

__constant__ unsigned long long c_mc_tri[256];
__constant__ unsigned int c_ms_seg[16];

// Edge table in the reference's corner convention: offset of the edge's low end from pos_calc
// and its axis (PolyMixin.hpp:271-276, 390-406).
__constant__ int c_edge3[12][4] = {{0, 0, 0, 0},  {1, 0, -1, 2}, {0, 0, -1, 0}, {0, 0, -1, 2},
                                   {0, 1, 0, 0},  {1, 1, -1, 2}, {0, 1, -1, 0}, {0, 1, -1, 2},
                                   {0, 0, 0, 1},  {1, 0, 0, 1},  {1, 0, -1, 1}, {0, 0, -1, 1}};
__constant__ int c_corner3[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 0, -1}, {0, 0, -1},
                                    {0, 1, 0}, {1, 1, 0}, {1, 1, -1}, {0, 1, -1}};
__constant__ int c_edge2[4][3] = {{0, 0, 0}, {1, 0, 1}, {0, 1, 0}, {0, 0, 1}};
__constant__ int c_corner2[4][2] = {{0, 0}, {1, 0}, {1, 1}, {0, 1}};
// For each axis, the edges along it (3D: 4, 2D: 2).
__constant__ int c_axis_edges3[3][4] = {{0, 2, 4, 6}, {8, 9, 10, 11}, {1, 3, 5, 7}};
__constant__ int c_axis_edges2[2][2] = {{0, 2}, {1, 3}};

template <int D>
struct McT;
template <>
struct McT<3> {
    static constexpr int nedge = 12, ncorner = 8, per_axis = 4;
    __device__ static int edge(int e, int a) { return c_edge3[e][a]; }
    __device__ static int corner(int c, int a) { return c_corner3[c][a]; }
    __device__ static int axis_edge(int axis, int i) { return c_axis_edges3[axis][i]; }
    // pos_calc = pos - SpxGridPosOffset, SpxGridPosOffset = (0,0,-1)
    __device__ static int calc_off(int a) { return a == 2 ? 1 : 0; }
    __device__ static unsigned edges_of(unsigned m) {
        const unsigned long long t = c_mc_tri[m];
        const int n = (int)(t >> 60) * 3;
        unsigned em = 0;
        for (int i = 0; i < n; i++) em |= 1u << ((t >> (4 * i)) & 0xf);
        return em;
    }
    __device__ static int nsimplex(unsigned m) { return (int)(c_mc_tri[m] >> 60); }
    __device__ static int simplex_edge(unsigned m, int i, int j) {
        return (int)((c_mc_tri[m] >> (4 * (i * 3 + j))) & 0xf);
    }
};
template <>
struct McT<2> {
    static constexpr int nedge = 4, ncorner = 4, per_axis = 2;
    __device__ static int edge(int e, int a) { return c_edge2[e][a]; }
    __device__ static int corner(int c, int a) { return c_corner2[c][a]; }
    __device__ static int axis_edge(int axis, int i) { return c_axis_edges2[axis][i]; }
    __device__ static int calc_off(int) { return 0; }
    __device__ static unsigned edges_of(unsigned m) {
        const unsigned t = c_ms_seg[m];
        const int n = (int)(t >> 28) * 2;
        unsigned em = 0;
        for (int i = 0; i < n; i++) em |= 1u << ((t >> (4 * i)) & 0xf);
        return em;
    }
    __device__ static int nsimplex(unsigned m) { return (int)(c_ms_seg[m] >> 28); }
    __device__ static int simplex_edge(unsigned m, int i, int j) {
        return (int)((c_ms_seg[m] >> (4 * (i * 2 + j))) & 0xf);
    }
};

// The member a marched voxel belongs to (grid size, world offset) and the brick being marched
// (slot, grid position of its first voxel): reads inside that brick skip the page table.
struct MemV {
    int m;
    int size[3];
    int offset[3];
    int slot;
    int org[3];
    int pbase;
    int cstr[3];
};

__device__ __forceinline__ MemV mem_view(const Pool& P, int slot) {
    const int4 o = P.slot_org[slot];
    const int m = o.w & kMemberMask;
    const Mem& M = P.mem[m];
    return MemV{m,    {M.size[0], M.size[1], M.size[2]}, {M.offset[0], M.offset[1], M.offset[2]},
                slot, {o.x, o.y, o.z},                   M.pbase,
                {M.cstr[0], M.cstr[1], M.cstr[2]}};
}

// get(pos) at K grid positions, loaded in levels: positions in the marched brick are read
// directly; the others need their partition's page first.  All K loads of a level are issued
// before any is used.  Background L+1 outside the grid, the partition background when inactive.
template <int D, int K>
__device__ __forceinline__ void get_many(const Geo& G, const Pool& P, const MemV& V,
                                         const int (*q)[3], float* out) {
    uint32_t idx[K];
    int part[K];
    bool in[K], same[K];
#pragma unroll
    for (int j = 0; j < K; j++) {
        bool inside = true, mine = true;
        int lc[3] = {0, 0, 0}, lq[3] = {0, 0, 0};
        int p = V.pbase;
#pragma unroll
        for (int a = 0; a < D; a++) {
            if (q[j][a] < 0 || q[j][a] >= V.size[a]) inside = false;
            lc[a] = q[j][a] - V.org[a];
            if (lc[a] < 0 || lc[a] >= G.part[a]) mine = false;
            const int qa = (int)fdiv((uint32_t)q[j][a], G.fpart[a]);
            lq[a] = q[j][a] - qa * G.part[a];
            p += qa * V.cstr[a];
        }
        in[j] = inside;
        same[j] = inside && mine;
        idx[j] = same[j] ? (uint32_t)V.slot * (uint32_t)G.pv + (uint32_t)leaf_of<D>(G, lc)
                         : (uint32_t)leaf_of<D>(G, lq);
        part[j] = p;
    }
    // every page entry and value is loaded, with no branch between the loads: a position
    // outside the grid reads the brick's first voxel (and its member's first page entry) and
    // is replaced by L+1 afterwards
    int pg[K];
#pragma unroll
    for (int j = 0; j < K; j++) pg[j] = P.page[(in[j] && !same[j]) ? part[j] : V.pbase];
    float raw[K];
#pragma unroll
    for (int j = 0; j < K; j++) {
        const float* src = P.phi + (size_t)V.slot * G.pv;
        if (same[j]) src = P.phi + idx[j];
        else if (in[j])
            src = pg[j] < 0 ? P.pbg + part[j] : P.phi + ((size_t)pg[j] * G.pv + idx[j]);
        raw[j] = *src;
    }
#pragma unroll
    for (int j = 0; j < K; j++) out[j] = in[j] ? raw[j] : (float)(G.L + 1);
}

// get(pos) at grid position q of the member.
template <int D>
__device__ __forceinline__ float get_v(const Geo& G, const Pool& P, const MemV& V, const int q[3]) {
    int lc[3] = {0, 0, 0};
    bool in = true;
    for (int a = 0; a < D; a++) {
        lc[a] = q[a] - V.org[a];
        if (lc[a] < 0 || lc[a] >= G.part[a] || q[a] >= V.size[a]) in = false;
    }
    if (in) return P.phi[(size_t)V.slot * G.pv + leaf_of<D>(G, lc)];
    return get_g<D>(G, P, V.m, V.size, q);
}

// Corner mask of the cube based at grid position g (bit i: phi(corner i) > 0, Poly.hpp:356-368).
template <int D>
__device__ unsigned cube_mask(const Geo& G, const Pool& P, const MemV& V, const int g[3]) {
    constexpr int NC = McT<D>::ncorner;
    int q[NC][3];
#pragma unroll
    for (int c = 0; c < NC; c++) {
        q[c][0] = q[c][1] = q[c][2] = 0;
#pragma unroll
        for (int a = 0; a < D; a++) q[c][a] = g[a] + McT<D>::calc_off(a) + McT<D>::corner(c, a);
    }
    float v[NC];
    get_many<D, NC>(G, P, V, q, v);
    unsigned m = 0;
#pragma unroll
    for (int c = 0; c < NC; c++) m |= (unsigned)(v[c] > 0.0f) << c;
    return m;
}

// Per-axis view of the cube based at grid position g, whose base voxel is at brick-local lc of the
// marched brick: the leaf terms (leaf_of is a sum of per-axis terms) of lc - 1, lc and lc + 1
// (the last wrapped into the next brick), whether lc is the brick's last coordinate, and whether
// g - 1 / g + 1 stay in the brick / the grid.
struct CubeAxes {
    int Fm[3], F0[3], F1[3];
    bool hi[3], okm[3], okp[3], inp[3];
};

template <int D>
__device__ __forceinline__ void cube_axes(const Geo& G, const MemV& V, const int lc[3],
                                          const int g[3], CubeAxes& A) {
#pragma unroll
    for (int a = 0; a < D; a++) {
        const int l = lc[a];
        A.hi[a] = l == G.part[a] - 1;
        A.F0[a] = leaf_term(G, a, l);
        A.F1[a] = leaf_term(G, a, A.hi[a] ? 0 : l + 1);
        A.Fm[a] = leaf_term(G, a, l > 0 ? l - 1 : 0);
        A.inp[a] = g[a] + 1 < V.size[a];
        A.okm[a] = l > 0;
        A.okp[a] = !A.hi[a] && A.inp[a];
    }
}

// Corner c's offset from the cube's base voxel (calc_off + corner, each component 0 or 1) as
// bits, bit a for axis a: c_corner3 shifted by pos_calc's (0, 0, 1), and c_corner2.
__device__ constexpr unsigned kCornerUp3[8] = {4, 5, 1, 0, 6, 7, 3, 2};
__device__ constexpr unsigned kCornerUp2[4] = {0, 1, 3, 2};

// cube_mask for the cube whose base voxel is in the marched brick (partition p): a corner's
// leaf is the sum of its axes' terms, and the bricks its corners can fall in (the marched one
// and those across the + faces it borders, their edges and corner) are looked up once each.
template <int D>
__device__ unsigned cube_mask_at(const Geo& G, const Pool& P, const MemV& V, int p,
                                 const CubeAxes& A) {
    constexpr int NC = McT<D>::ncorner, NB = 1 << D;
    unsigned hm = 0;  // axes along which the + corners leave the brick
#pragma unroll
    for (int a = 0; a < D; a++) hm |= (unsigned)A.hi[a] << a;
    // brick k (bit a: across axis a): partition p + the crossed axes' strides; one outside the
    // grid reads the marched brick's page and is never used (its corners are outside too)
    const float* src[NB];
    bool inact[NB];
    src[0] = P.phi + (uint32_t)V.slot * (uint32_t)G.pv;
    inact[0] = false;
#pragma unroll
    for (int k = 1; k < NB; k++) {
        bool ok = true;
        int pk = p;
#pragma unroll
        for (int a = 0; a < D; a++)
            if ((k >> a) & 1) {
                ok = ok && A.hi[a] && A.inp[a];
                pk += V.cstr[a];
            }
        const int pg = P.page[ok ? pk : p];
        src[k] = pg >= 0 ? P.phi + (uint32_t)pg * (uint32_t)G.pv : P.pbg + pk;
        inact[k] = pg < 0;
    }
    unsigned m = 0;
#pragma unroll
    for (int c = 0; c < NC; c++) {
        const unsigned cm = D == 3 ? kCornerUp3[c] : kCornerUp2[c];
        int leaf = 0;
        bool inside = true;
#pragma unroll
        for (int a = 0; a < D; a++) {
            const bool up = (cm >> a) & 1u;
            leaf += up ? A.F1[a] : A.F0[a];
            inside = inside && (!up || A.inp[a]);
        }
        const unsigned kb = cm & hm;  // the corner's brick
        const float* sp = src[0];
        bool ia = false;
#pragma unroll
        for (unsigned k = 1; k < (unsigned)NB; k++) {
            if ((k & ~cm) != 0u) continue;  // not a subset of the corner's axes
            sp = kb == k ? src[k] : sp;
            ia = kb == k ? inact[k] : ia;
        }
        const float v = *(inside ? (ia ? sp : sp + leaf) : src[0]);
        m |= (unsigned)(inside ? v > 0.0f : (float)(G.L + 1) > 0.0f) << c;
    }
    return m;
}

// band_mask_around for a cube base at brick-local lc of brick slot s, from its axes' terms.
template <int D>
__device__ __forceinline__ unsigned band_mask_at(const Geo& G, const Pool& P, int s,
                                                 const CubeAxes& A) {
    constexpr int NK = D == 3 ? 27 : 9;
    const size_t base = (size_t)s * G.pv;
    unsigned m = 0;
#pragma unroll
    for (int k = 0; k < NK; k++) {
        const int d[3] = {D == 3 ? k / 9 - 1 : k / 3 - 1, D == 3 ? (k / 3) % 3 - 1 : k % 3 - 1,
                          D == 3 ? k % 3 - 1 : 0};
        if (d[0] != 0 && d[1] != 0 && (D == 2 || d[2] != 0)) continue;  // shares no edge
        int leaf = 0;
        bool ok = true;
#pragma unroll
        for (int a = 0; a < D; a++) {
            leaf += d[a] < 0 ? A.Fm[a] : (d[a] > 0 ? A.F1[a] : A.F0[a]);
            ok = ok && (d[a] < 0 ? A.okm[a] : (d[a] > 0 ? A.okp[a] : true));
        }
        const int8_t lay = P.lay[ok ? base + (size_t)leaf : base];
        if (ok && lay != NONE) m |= 1u << k;
    }
    return m;
}

// Band bits of the cubes at b + d, d in {-1, 0, 1}^D with at least one zero component (every cube
// that can share an edge with the cube at b): bit (d0+1)*9 + (d1+1)*3 + (d2+1) is set when that
// cube's base voxel lies in the marched brick and in the grid and is in a layer list.  The layer
// bytes are loaded together, with no branch between the loads (a cube outside reads the brick's
// first voxel and is masked).
template <int D>
__device__ __forceinline__ unsigned band_mask_around(const Geo& G, const Pool& P, const MemV& V,
                                                     int s, const int org[3], const int b[3]) {
    constexpr int NK = D == 3 ? 27 : 9;
    const size_t base = (size_t)s * G.pv;
    unsigned m = 0;
#pragma unroll
    for (int k = 0; k < NK; k++) {
        const int d[3] = {D == 3 ? k / 9 - 1 : k / 3 - 1, D == 3 ? (k / 3) % 3 - 1 : k % 3 - 1,
                          D == 3 ? k % 3 - 1 : 0};
        if (d[0] != 0 && d[1] != 0 && (D == 2 || d[2] != 0)) continue;  // shares no edge
        int lc[3] = {0, 0, 0};
        bool ok = true;
#pragma unroll
        for (int a = 0; a < D; a++) {
            lc[a] = b[a] + d[a] - org[a];
            if (lc[a] < 0 || lc[a] >= G.part[a] || b[a] + d[a] >= V.size[a]) ok = false;
        }
        const int8_t lay = P.lay[ok ? base + leaf_of<D>(G, lc) : base];
        if (ok && lay != NONE) m |= 1u << k;
    }
    return m;
}

// Owner of the vertex on edge e of the cube at base b: the lowest-leaf band cube of the partition
// among those sharing the edge (the cubes at b + edge(e) - edge(f) for the edges f along e's
// axis), found from the band bits of the cubes around b (band_mask_around).  Returns the owner's
// row-major leaf PosIdx, and which of the sharers (McT::axis_edge order) it is in cand.
template <int D>
__device__ __forceinline__ int edge_owner_m(const Geo& G, unsigned mask, const int org[3],
                                            const int b[3], int e, int* cand) {
    const int axis = McT<D>::edge(e, D);
    int best = -1, best_i = 0;
    for (int i = 0; i < McT<D>::per_axis; i++) {
        const int f = McT<D>::axis_edge(axis, i);
        int d[3] = {0, 0, 0};
        for (int a = 0; a < D; a++) d[a] = McT<D>::edge(e, a) - McT<D>::edge(f, a);
        const int k = D == 3 ? (d[0] + 1) * 9 + (d[1] + 1) * 3 + (d[2] + 1)
                             : (d[0] + 1) * 3 + (d[1] + 1);
        if (!((mask >> k) & 1u)) continue;
        int lc[3] = {0, 0, 0};
        for (int a = 0; a < D; a++) lc[a] = b[a] + d[a] - org[a];
        const int l = rm_of<D>(G, lc);
        if (best < 0 || l < best) {
            best = l;
            best_i = i;
        }
    }
    *cand = best_i;
    return best;
}

template <int D>
__device__ bool inside_wf(const MemV& V, const float w[3]) {
    for (int a = 0; a < D; a++) {
        if (w[a] >= (float)(V.offset[a] + V.size[a])) return false;
        if (w[a] < (float)V.offset[a]) return false;
    }
    return true;
}

// Safe gradient at an integer grid position.
template <int D>
__device__ void grad_i(const Geo& G, const Pool& P, const MemV& V, const int g[3], float out[3]) {
    int pt[3] = {g[0], g[1], g[2]};
    const float centre = get_v<D>(G, P, V, g);
    for (int a = 0; a < D; a++) {
        float back = centre, fwd = centre;
        int order = 0;
        pt[a] -= 1;
        if (inside<D>(V.size, pt)) {
            back = get_v<D>(G, P, V, pt);
            order++;
        }
        pt[a] += 2;
        if (inside<D>(V.size, pt)) {
            fwd = get_v<D>(G, P, V, pt);
            order++;
        }
        pt[a] -= 1;
        out[a] = order != 0 ? (fwd - back) / (float)order : 0.0f;
        out[a] = out[a] / 1.0f;
    }
}

// The reference's safe gradient (NumericMixin.hpp:188-231) of the N-linear interpolation
// (NumericMixin.hpp:403-481) at a vertex interpolated on an edge along axis ax: the probe is
// stepped -1, +2, -1 along each axis in fp32 and the value at each probe interpolated.  The
// vertex's other coordinates are whole numbers, and so are those of every probe (whole numbers
// +-1 are exact in fp32), so at a probe every axis but ax has weight 0 and the value is
// t*hi + (1-t)*lo between the two grid values along ax, t from the probe's own ax coordinate.
// Only those 2 values per probe are read, 14 in all, in one level-split gather.  The terms the
// reference multiplies by weight 0 are finite, so the result is the same number; only the sign
// of an exact zero can differ.
template <int D>
__device__ void grad_edge(const Geo& G, const Pool& P, const MemV& V, const float w[3], int ax,
                          float out[3]) {
    constexpr int NP = 1 + 2 * D;
    float pr[NP][3];
    float pt[3] = {w[0], w[1], w[2]};
    pr[0][0] = pt[0], pr[0][1] = pt[1], pr[0][2] = pt[2];
#pragma unroll
    for (int a = 0; a < D; a++) {
        pt[a] -= 1.0f;
        pr[1 + 2 * a][0] = pt[0], pr[1 + 2 * a][1] = pt[1], pr[1 + 2 * a][2] = pt[2];
        pt[a] += 2.0f;
        pr[2 + 2 * a][0] = pt[0], pr[2 + 2 * a][1] = pt[1], pr[2 + 2 * a][2] = pt[2];
        pt[a] -= 1.0f;
    }
    int q[2 * NP][3];
    float t[NP];
#pragma unroll
    for (int j = 0; j < NP; j++) {
        float x = 0.f;
        q[2 * j][0] = q[2 * j][1] = q[2 * j][2] = 0;
        for (int a = 0; a < D; a++) {
            const float f = floorf(pr[j][a]);
            if (a == ax) x = pr[j][a] - f;
            q[2 * j][a] = (int)f - V.offset[a];
        }
        t[j] = x;
#pragma unroll
        for (int a = 0; a < 3; a++) q[2 * j + 1][a] = q[2 * j][a] + (a == ax ? 1 : 0);
    }
    float v[2 * NP];
    get_many<D, 2 * NP>(G, P, V, q, v);
    float val[NP];
#pragma unroll
    for (int j = 0; j < NP; j++) val[j] = t[j] * v[2 * j + 1] + (1.0f - t[j]) * v[2 * j];
#pragma unroll
    for (int a = 0; a < D; a++) {
        float back = val[0], fwd = val[0];
        int order = 0;
        if (inside_wf<D>(V, pr[1 + 2 * a])) {
            back = val[1 + 2 * a];
            order++;
        }
        if (inside_wf<D>(V, pr[2 + 2 * a])) {
            fwd = val[2 + 2 * a];
            order++;
        }
        out[a] = order != 0 ? (fwd - back) / (float)order : 0.0f;
        out[a] = out[a] / 1.0f;
    }
}

__device__ void normalize3(float n[3]) {
    const float sq = n[0] * n[0] + (n[1] * n[1] + n[2] * n[2]);
    if (sq > 0.0f) {
        const float r = __builtin_sqrtf(sq);
        n[0] = n[0] / r;
        n[1] = n[1] / r;
        n[2] = n[2] / r;
    }
}

// Vertex on the edge from grid position ga along axis (Poly.hpp:296-346; PolyMixin.hpp:142-148).
template <int D>
__device__ void make_vertex(const Geo& G, const Pool& P, const MemV& V, const int ga[3], int axis,
                            float* out) {
    int gb[3] = {ga[0], ga[1], ga[2]};
    gb[axis] += 1;
    const float va = get_v<D>(G, P, V, ga), vb = get_v<D>(G, P, V, gb);
    float pos[3] = {0.f, 0.f, 0.f}, n[3] = {0.f, 0.f, 0.f};
    if (fabsf(va) <= kEps || fabsf(vb) <= kEps) {
        const int* e = fabsf(va) <= kEps ? ga : gb;
        for (int a = 0; a < D; a++) pos[a] = (float)(e[a] + V.offset[a]);
        if (D == 3) grad_i<D>(G, P, V, e, n);
    } else {
        float mu;
        if (fabsf(va - vb) <= kEps) mu = 0.5f;
        else mu = va / (va - vb);
        for (int a = 0; a < D; a++) {
            const float fa = (float)(ga[a] + V.offset[a]), fb = (float)(gb[a] + V.offset[a]);
            pos[a] = fa + (fb - fa) * mu;
        }
        if (D == 3) grad_edge<D>(G, P, V, pos, axis, n);
    }
    if (D == 3) {
        normalize3(n);
        for (int a = 0; a < 3; a++) {
            out[a] = pos[a];
            out[3 + a] = n[a];
        }
    } else {
        out[0] = pos[0];
        out[1] = pos[1];
    }
}

// ---- data-parallel march: one thread per band voxel of the partitions being marched ----------

constexpr int BT = 1024;
constexpr int ITEMS = 4;

__device__ __forceinline__ int block_excl_scan(int val, int* total) {
    __shared__ int wsum[BT / 64];
    __shared__ int wpre[BT / 64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    int x = val;
    for (int o = 1; o < 64; o <<= 1) {
        const int y = __shfl_up(x, o);
        if (lane >= o) x += y;
    }
    if (lane == 63) wsum[wave] = x;
    __syncthreads();
    if (wave == 0) {
        const int nw = BT / 64;
        const int w = lane < nw ? wsum[lane] : 0;
        int z = w;
        for (int o = 1; o < 64; o <<= 1) {
            const int y = __shfl_up(z, o);
            if (lane >= o) z += y;
        }
        if (lane < nw) wpre[lane] = z - w;
        if (lane == nw - 1) wsum[0] = z;
    }
    __syncthreads();
    const int res = wpre[wave] + x - val;
    *total = wsum[0];
    __syncthreads();
    return res;
}

// Band voxels per marched partition: the sum of its layer counts (a voxel is in a layer list
// exactly when its layer byte is not NONE).
__global__ void k_mb_count(Geo G, Pool P, const int* parts, int nb, int* cnt) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= nb) return;
    const int p = parts[j];
    int c = 0;
    if (P.page[p] >= 0)
        for (int k = 0; k < G.NL; k++) c += P.pcnt[p * G.NL + k];
    cnt[j] = c;
}

// Single block: start[j] = exclusive prefix of cnt, start[nb] = total.
__global__ void k_mb_scan(const int* cnt, int nb, int* start) {
    __shared__ int carry;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (int b0 = 0; b0 < nb; b0 += BT) {
        const int b = b0 + threadIdx.x;
        int tot;
        const int ex = block_excl_scan(b < nb ? cnt[b] : 0, &tot);
        if (b < nb) start[b] = carry + ex;
        __syncthreads();
        if (threadIdx.x == 0) carry += tot;
        __syncthreads();
    }
    if (threadIdx.x == 0) start[nb] = carry;
}

// The band voxels of each marched partition in its brick's storage order, from the partition's
// first item: each wavefront takes a segment of consecutive leaves, counts its band voxels by
// ballot, the segment counts are exchanged once through LDS, and a second pass writes them.
template <int D>
__global__ void k_mb_scatter(Geo G, Pool P, const int* parts, const int* start, int* item_j,
                             int* item_l) {
    __shared__ int wsum[BT / 64];
    const int s = P.page[parts[blockIdx.x]];
    if (s < 0) return;
    const int lane = __lane_id(), w = threadIdx.x >> 6, nw = blockDim.x >> 6;
    const int seg = (((G.pv + nw - 1) / nw) + 63) & ~63;
    const int l0 = w * seg, l1 = min(l0 + seg, G.pv);
    const int8_t* lay = P.lay + (size_t)s * G.pv;
    int c = 0;
    for (int b = l0; b < l1; b += 64) {
        const int l = b + lane;
        c += __popcll(__ballot(l < l1 && lay[l] != NONE));
    }
    if (lane == 0) wsum[w] = c;
    __syncthreads();
    int at = start[blockIdx.x];
    const int end = start[blockIdx.x + 1];  // from the layer counts: nothing is written past it
    for (int k = 0; k < w; k++) at += wsum[k];
    for (int b = l0; b < l1; b += 64) {
        const int l = b + lane;
        const bool in = l < l1 && lay[l] != NONE;
        const unsigned long long m = __ballot(in);
        const int e = at + __popcll(m & ((1ull << lane) - 1ull));
        if (in && e < end) {
            int lc[3] = {0, 0, 0};
            leaf_coords<D>(G, (uint32_t)l, lc);
            item_j[e] = blockIdx.x;
            item_l[e] = rm_of<D>(G, lc);  // items and their scratch are keyed by leaf PosIdx
        }
        at += __popcll(m);
    }
    // counts above the layer bytes (never, while the two agree) leave slots that must still
    // name a voxel of this brick for the passes that read them
    if (threadIdx.x == 0) {
        int total = 0;
        for (int k = 0; k < nw; k++) total += wsum[k];
        for (int e = start[blockIdx.x] + total; e < end; e++) {
            item_j[e] = blockIdx.x;
            item_l[e] = 0;
        }
    }
}

// Cube mask, owned edges and vertex / simplex counts of each band voxel (the scan that follows
// lists the voxels with simplices, so that the simplex pass runs on them alone).
template <int D>
__global__ void k_m_count(Geo G, Pool P, const int* parts, const int* item_j, const int* item_l,
                          int M, uint8_t* cmask, uint16_t* own, uint32_t* ocand, int* nv,
                          int* nt) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < M) {
        const int j = item_j[i], l = item_l[i];
        const int p = parts[j];
        const int s = P.page[p];
        const MemV V = mem_view(P, s);
        const int* org = V.org;
        int g[3] = {0, 0, 0}, lc[3] = {0, 0, 0};
        rm_coords<D>(G, (uint32_t)l, lc);
        for (int a = 0; a < D; a++) g[a] = lc[a] + org[a];
        CubeAxes A;
        cube_axes<D>(G, V, lc, g, A);
        const unsigned m = cube_mask_at<D>(G, P, V, p, A);
        const unsigned em = McT<D>::edges_of(m);
        const unsigned bm = em ? band_mask_at<D>(G, P, s, A) : 0u;
        unsigned om = 0, oc = 0;
        for (int e = 0; e < McT<D>::nedge; e++) {
            if (!((em >> e) & 1)) continue;
            int ci;
            if (edge_owner_m<D>(G, bm, org, g, e, &ci) == l) om |= 1u << e;
            oc |= (unsigned)ci << (2 * e);
        }
        cmask[(size_t)j * G.pv + l] = (uint8_t)m;
        own[(size_t)j * G.pv + l] = (uint16_t)om;
        ocand[(size_t)j * G.pv + l] = oc;
        nv[i] = __popc(om);
        nt[i] = McT<D>::nsimplex(m);  // an owned edge is a crossed edge, so it has simplices
    }
}

// The owned edges of the busy voxels as item * 16 + edge, in vertex order.
__global__ void k_m_vlist(const int* work, int nwork, const uint16_t* own, const int* item_j,
                          const int* item_l, int pv, const int* vb, int* vwork) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nwork) return;
    const int i = work[t];
    unsigned om = own[(size_t)item_j[i] * pv + item_l[i]];
    for (int at = vb[i]; om; at++, om &= om - 1u) vwork[at] = i * 16 + (__ffs(om) - 1);
}

// Device-wide exclusive scans of the vertex counts a, the simplex counts b and the busy flags
// (b > 0), in three kernels (block sums, top-level scan, apply) and without atomics, so every
// output is in item order: the vertex and simplex bases, and the list of busy items.
__global__ void k_s3_blocks(const int* a, const int* b, int n, int* bs) {
    const int base = blockIdx.x * BT * ITEMS + threadIdx.x;
    int sa = 0, sb = 0, sc = 0;
    for (int k = 0; k < ITEMS; k++) {
        const int i = base + k * BT;
        if (i < n) {
            sa += a[i];
            sb += b[i];
            sc += b[i] > 0;
        }
    }
    int ta, tb, tc;
    block_excl_scan(sa, &ta);
    block_excl_scan(sb, &tb);
    block_excl_scan(sc, &tc);
    if (threadIdx.x == 0) {
        bs[3 * blockIdx.x] = ta;
        bs[3 * blockIdx.x + 1] = tb;
        bs[3 * blockIdx.x + 2] = tc;
    }
}

__global__ void k_s3_top(int* bs, int nblk, int* tot) {
    __shared__ int carry[3];
    if (threadIdx.x < 3) carry[threadIdx.x] = 0;
    __syncthreads();
    for (int b0 = 0; b0 < nblk; b0 += BT) {
        const int b = b0 + threadIdx.x;
        int t[3], e[3];
        for (int c = 0; c < 3; c++) {
            const int x = b < nblk ? bs[3 * b + c] : 0;
            e[c] = block_excl_scan(x, &t[c]);
        }
        if (b < nblk)
            for (int c = 0; c < 3; c++) bs[3 * b + c] = carry[c] + e[c];
        __syncthreads();
        if (threadIdx.x == 0)
            for (int c = 0; c < 3; c++) carry[c] += t[c];
        __syncthreads();
    }
    if (threadIdx.x < 3) tot[threadIdx.x] = carry[threadIdx.x];
}

// Each thread scans ITEMS consecutive entries so that the output order is the input order.
__global__ void k_s3_apply(const int* a, const int* b, int n, const int* bs, int* ea, int* eb,
                           int* work) {
    const int i0 = (blockIdx.x * BT + threadIdx.x) * ITEMS;
    int sa = 0, sb = 0, sc = 0;
    for (int k = 0; k < ITEMS; k++)
        if (i0 + k < n) {
            sa += a[i0 + k];
            sb += b[i0 + k];
            sc += b[i0 + k] > 0;
        }
    int ta, tb, tc;
    int ra = block_excl_scan(sa, &ta) + bs[3 * blockIdx.x];
    int rb = block_excl_scan(sb, &tb) + bs[3 * blockIdx.x + 1];
    int rc = block_excl_scan(sc, &tc) + bs[3 * blockIdx.x + 2];
    for (int k = 0; k < ITEMS; k++)
        if (i0 + k < n) {
            const int va = a[i0 + k], vb = b[i0 + k];
            ea[i0 + k] = ra;
            eb[i0 + k] = rb;
            if (vb > 0) work[rc++] = i0 + k;
            ra += va;
            rb += vb;
        }
}

// Vertex base of each band voxel, by leaf, for the owner lookups of the simplex pass.
__global__ void k_m_leafbase(const int* item_j, const int* item_l, int M, int pv, const int* vb,
                             int* vleaf) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < M) vleaf[(size_t)item_j[i] * pv + item_l[i]] = vb[i];
}

// First vertex / simplex of each marched partition; entry nb holds the totals.
__global__ void k_m_partstart(const int* start, int nb, const int* vb, const int* tb, int M,
                              const int* tot, int* vst, int* tst) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j > nb) return;
    const int it = start[j];
    vst[j] = it < M ? vb[it] : tot[0];
    tst[j] = it < M ? tb[it] : tot[1];
}

// One vertex per owned edge (Poly.hpp:296-346): its cube's vertex base + the edge's rank among
// the cube's owned edges numbers it within its partition, whose vertices start at vdst[j].
template <int D>
__global__ void __launch_bounds__(TPB) k_m_vert(Geo G, Pool P, const int* parts,
                                                const int* item_j, const int* item_l,
                                                const int* vwork, int nvw, const uint16_t* own,
                                                const int* vb, const int* vst,
                                                const long long* vdst, float* varena) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nvw) return;
    const int x = vwork[t];
    const int i = x >> 4, e = x & 15;
    const int j = item_j[i], l = item_l[i];
    const MemV V = mem_view(P, P.page[parts[j]]);
    int ga[3] = {0, 0, 0};
    rm_coords<D>(G, (uint32_t)l, ga);
    for (int a = 0; a < D; a++) ga[a] += V.org[a] + McT<D>::calc_off(a) + McT<D>::edge(e, a);
    const unsigned om = own[(size_t)j * G.pv + l];
    const int rank = __popc(om & ((1u << e) - 1u));
    const int vw = D == 3 ? 6 : 2;
    const long long at = vdst[j] + (vb[i] + rank - vst[j]);
    make_vertex<D>(G, P, V, ga, McT<D>::edge(e, D), varena + (size_t)at * vw);
}

// Simplices of every band cube (Poly.hpp:198-287), vertex indices local to the partition: a
// simplex edge's vertex belongs to the owner the count pass recorded for it.  Partition j's
// simplices start at tdst[j].
template <int D>
__global__ void __launch_bounds__(TPB) k_m_tri(Geo G, const int* item_j, const int* item_l,
                                               const int* work, int nwork, const uint8_t* cmask,
                                               const uint16_t* own, const uint32_t* ocand,
                                               const int* tb, const int* vleaf, const int* vst,
                                               const int* tst, const long long* tdst,
                                               uint32_t* tarena) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nwork) return;
    const int i = work[t];
    const int j = item_j[i], l = item_l[i];
    const size_t jb = (size_t)j * G.pv;
    const unsigned m = cmask[jb + l];
    const int ns = McT<D>::nsimplex(m);
    int lc0[3] = {0, 0, 0};
    rm_coords<D>(G, (uint32_t)l, lc0);
    const int base = vst[j];
    uint32_t* tout = tarena + (size_t)(tdst[j] + (tb[i] - tst[j])) * D;
    const unsigned oc = ocand[jb + l];
    // every simplex vertex's owner first, then all their loads together (no branch between
    // them: the slots past the cube's simplices read the cube itself), then the stores
    constexpr int NV = (D == 3 ? 5 : 2) * D;  // at most 5 triangles (3D) or 2 segments (2D)
    int ow[NV], eo[NV];
#pragma unroll
    for (int q = 0; q < NV; q++) {
        const int t = q / D, k = q % D;
        ow[q] = l;
        eo[q] = 0;
        if (t < ns) {
            const int e = McT<D>::simplex_edge(m, t, k);
            // sharer ci of the edge owns its vertex, as its edge eo
            const int axis = McT<D>::edge(e, D);
            eo[q] = McT<D>::axis_edge(axis, (int)((oc >> (2 * e)) & 3u));
            int lc[3] = {0, 0, 0};
            for (int a = 0; a < D; a++)
                lc[a] = lc0[a] + McT<D>::edge(e, a) - McT<D>::edge(eo[q], a);
            ow[q] = rm_of<D>(G, lc);
        }
    }
    int vl[NV];
    unsigned om[NV];
#pragma unroll
    for (int q = 0; q < NV; q++) {
        vl[q] = vleaf[jb + ow[q]];
        om[q] = own[jb + ow[q]];
    }
#pragma unroll
    for (int q = 0; q < NV; q++)
        if (q / D < ns)
            tout[q] = (uint32_t)(vl[q] + __popc(om[q] & ((1u << eo[q]) - 1u)) - base);
}

// Copies live mesh segments into a fresh arena when the old one is compacted.
__global__ void k_copy_segments(const long long* src, const long long* dst, const long long* len,
                                int n, const float* from, float* into) {
    for (int j = blockIdx.x; j < n; j += gridDim.x)
        for (long long i = threadIdx.x; i < len[j]; i += blockDim.x) into[dst[j] + i] = from[src[j] + i];
}
__global__ void k_copy_segments_u(const long long* src, const long long* dst,
                                  const long long* len, int n, const uint32_t* from,
                                  uint32_t* into) {
    for (int j = blockIdx.x; j < n; j += gridDim.x)
        for (long long i = threadIdx.x; i < len[j]; i += blockDim.x) into[dst[j] + i] = from[src[j] + i];
}

// Partitions whose bits meet `want`, as (partition, bits) pairs: bit 0 in
// delta(0)|delta(2L), bit 1 in status_change(0)|status_change(2L), bit 2 tracked by the
// isogrid in list 0 or 2L.
__global__ void k_part_bits(Geo G, Pool P, const unsigned* dmask, const unsigned* smask,
                            unsigned want, int* out, int* cnt) {
    for (int b0 = blockIdx.x * blockDim.x; b0 < G.np; b0 += gridDim.x * blockDim.x) {
        const int p = b0 + threadIdx.x;
        unsigned b = 0;
        if (p < G.np) {
            const unsigned outer = 1u | (1u << (G.NL - 1));
            if (dmask[p] & outer) b |= 1;
            if (smask[p] & outer) b |= 2;
            if (P.pcnt[p * G.NL] > 0 || P.pcnt[p * G.NL + G.NL - 1] > 0) b |= 4;
        }
        const bool take = (b & want) != 0;
        const int at = wave_append(cnt, take);
        if (take) {
            out[2 * at] = p;
            out[2 * at + 1] = (int)b;
        }
    }
}

// Whether each listed partition is active in the isogrid.
__global__ void k_gather_active(Pool P, const int* parts, int n, int* out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = P.page[parts[i]] >= 0;
}

}  // namespace

#define HIPX(expr)                                                                              \
    do {                                                                                       \
        hipError_t e_ = (expr);                                                                \
        if (e_ != hipSuccess)                                                                  \
            return felt_fail(FELT_E_DEVICE, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

#define DISPATCHP(D_, KERNEL, GRID, BLOCK, ...)                                                \
    do {                                                                                       \
        if ((D_) == 2) hipLaunchKernelGGL(KERNEL<2>, GRID, BLOCK, 0, st, __VA_ARGS__);          \
        else hipLaunchKernelGGL(KERNEL<3>, GRID, BLOCK, 0, st, __VA_ARGS__);                    \
    } while (0)

void felt_mc_table_host(uint64_t* out) {
    for (int i = 0; i < 256; i++) out[i] = kMcTri[i];
}

void felt_ms_table_host(uint32_t* out) {
    for (int i = 0; i < 16; i++) out[i] = kMsSeg[i];
}

int felt_polys_upload_tables() {
    static bool done = false;
    if (done) return FELT_OK;
    HIPX(hipMemcpyToSymbol(HIP_SYMBOL(c_mc_tri), kMcTri, sizeof(kMcTri)));
    HIPX(hipMemcpyToSymbol(HIP_SYMBOL(c_ms_seg), kMsSeg, sizeof(kMsSeg)));
    done = true;
    return FELT_OK;
}

void Changes::track(int64_t p) {
    if (look[p] != -1) return;
    look[p] = (int64_t)list.size();
    list.push_back(p);
}
void Changes::untrack(int64_t p) {
    const int64_t at = look[p];
    if (at == -1) return;
    const int64_t last = (int64_t)list.size() - 1;
    if (at < last) {
        const int64_t m = list[last];
        list[at] = m;
        look[m] = at;
    }
    look[p] = -1;
    list.pop_back();
}
void Changes::reset() {
    for (int64_t p : list) look[p] = -1;
    list.clear();
}

int felt_polys_init(felt_polys* q) {
    const Geo& G = q->s->G;
    q->kids.assign(G.np, PolyKid{});
    q->a.look.assign(G.np, -1);
    q->alive.look.assign(G.np, -1);
    q->b.look.assign(G.np, -1);
    q->pending = &q->a;
    q->done = &q->b;
    q->vw = G.D == 3 ? 6 : 2;
    return felt_polys_upload_tables();
}

void felt_polys_release(felt_polys* q) {
    q->varena.release();
    q->tarena.release();
    q->varena2.release();
    q->tarena2.release();
    q->dst.release();
    q->hdst.release();
    q->hbuf.release();
    q->cmask.release();
    q->own.release();
    q->ocand.release();
    q->vwork.release();
    q->vbase.release();
    q->tbase.release();
    q->tot.release();
    q->dparts.release();
    q->dseg.release();
    q->bits.release();
    q->plist.release();
    q->pstart.release();
    q->vst.release();
    q->item_j.release();
    q->item_l.release();
    q->nv.release();
    q->nt.release();
    q->vleaf.release();
    q->work.release();
}

// Partitions whose notify bits meet `want`, sorted by partition index.
static int part_bits(felt_polys* q, unsigned want, std::vector<std::pair<int, int>>& out) {
    felt_surface* s = q->s;
    const Geo& G = s->G;
    hipStream_t st = s->st;
    int rc;
    if ((rc = q->plist.ensure(2 * (size_t)G.np + 2))) return rc;
    int* cnt = q->plist.ptr + 2 * (size_t)G.np;
    HIPX(hipMemsetAsync(cnt, 0, sizeof(int), st));
    const int grid = (int)std::min<int64_t>(4096, ((int64_t)G.np + TPB - 1) / TPB);
    hipLaunchKernelGGL(k_part_bits, dim3(grid), dim3(TPB), 0, st, G, s->P, s->dmask, s->smask,
                       want, q->plist.ptr, cnt);
    int n = 0;
    HIPX(hipMemcpyAsync(&n, cnt, sizeof(int), hipMemcpyDeviceToHost, st));
    HIPX(hipStreamSynchronize(st));
    std::vector<int> raw(2 * (size_t)n);
    if (n) {
        HIPX(hipMemcpyAsync(raw.data(), q->plist.ptr, raw.size() * sizeof(int),
                            hipMemcpyDeviceToHost, st));
        HIPX(hipStreamSynchronize(st));
    }
    out.resize(n);
    for (int i = 0; i < n; i++) out[i] = {raw[2 * i], raw[2 * i + 1]};
    std::sort(out.begin(), out.end());
    return FELT_OK;
}

// notify (Polys.hpp:92-133).
int felt_polys_notify_impl(felt_polys* q) {
    std::vector<std::pair<int, int>> l;
    if (int rc = part_bits(q, 3u, l)) return rc;
    for (const auto& e : l) {
        const int p = e.first, bits = e.second;
        if (bits & 1) {
            const bool act = q->kids[p].active || (bits & 4);
            if (act) q->pending->track(p);
            else q->pending->untrack(p);
        }
    }
    for (const auto& e : l)
        if ((e.second & 2) && q->kids[e.first].active) q->pending->track(e.first);
    return FELT_OK;
}

// invalidate (Polys.hpp:171-191): the isogrid's outermost-layer partitions.
int felt_polys_invalidate_impl(felt_polys* q) {
    std::vector<std::pair<int, int>> l;
    if (int rc = part_bits(q, 4u, l)) return rc;
    q->pending->reset();
    for (const auto& e : l) q->pending->track(e.first);
    return FELT_OK;
}

// Segment capacity for a mesh of n entries: a re-marched partition whose mesh grows a little
// still fits in place.
static long long seg_cap(long long n) { return n ? n + n / 4 + 8 : 0; }

// Compacts the arenas into the spare ones: every live segment is copied with a fresh capacity,
// children whose mesh count is zero (about to be re-marched) give up their segment, and at least
// extra_v / extra_t entries are left free after the live ones.  The spare arenas only grow, so
// a compaction in steady state allocates nothing.
static int arena_compact(felt_polys* q, long long extra_v, long long extra_t) {
    hipStream_t st = q->s->st;
    const int D = q->s->G.D, vw = q->vw;
    long long live_v = 0, live_t = 0;
    size_t n = 0;
    for (int64_t p : q->alive.list) {
        const PolyKid& k = q->kids[p];
        live_v += seg_cap(k.nv);
        live_t += seg_cap(k.nt);
        n += k.nv || k.nt ? 1 : 0;
    }
    const long long cap_v = std::max<long long>(4096, 2 * (live_v + extra_v));
    const long long cap_t = std::max<long long>(4096, 2 * (live_t + extra_t));
    int rc;
    if ((rc = q->varena2.ensure((size_t)cap_v * vw))) return rc;
    if ((rc = q->tarena2.ensure((size_t)cap_t * D))) return rc;
    // copy plan: (source, destination, length) of the vertex then the simplex segments
    if ((rc = q->hdst.ensure(6 * n + 6))) return rc;
    if ((rc = q->dst.ensure(6 * n + 6))) return rc;
    long long* h = q->hdst.ptr;
    long long cv = 0, ct = 0;
    size_t j = 0;
    for (int64_t p : q->alive.list) {
        PolyKid& k = q->kids[p];
        if (!k.nv && !k.nt) {
            k.voff = k.toff = k.vcap = k.tcap = 0;
            continue;
        }
        h[j] = k.voff * vw;
        h[n + j] = cv * vw;
        h[2 * n + j] = k.nv * vw;
        h[3 * n + j] = k.toff * D;
        h[4 * n + j] = ct * D;
        h[5 * n + j] = k.nt * D;
        k.voff = cv;
        k.toff = ct;
        k.vcap = seg_cap(k.nv);
        k.tcap = seg_cap(k.nt);
        cv += k.vcap;
        ct += k.tcap;
        j++;
    }
    if (n) {
        HIPX(hipMemcpyAsync(q->dst.ptr, h, 6 * n * sizeof(long long), hipMemcpyHostToDevice, st));
        const int grid = (int)std::min<size_t>(n, 65535);
        const long long* d = q->dst.ptr;
        hipLaunchKernelGGL(k_copy_segments, dim3(grid), dim3(TPB), 0, st, d, d + n, d + 2 * n,
                           (int)n, q->varena.ptr, q->varena2.ptr);
        hipLaunchKernelGGL(k_copy_segments_u, dim3(grid), dim3(TPB), 0, st, d + 3 * n, d + 4 * n,
                           d + 5 * n, (int)n, q->tarena.ptr, q->tarena2.ptr);
        HIPX(hipGetLastError());
        HIPX(hipStreamSynchronize(st));  // the plan's pinned host copy may be reused after this
    }
    std::swap(q->varena, q->varena2);
    std::swap(q->tarena, q->tarena2);
    q->vtop = cv;
    q->ttop = ct;
    return FELT_OK;
}

// Partitions marched per batch: the batch's per-voxel scratch (corner masks, owners, owner
// candidates, vertex bases: 11 B per voxel) stays under 768 MB.
static size_t march_batch(const Geo& G) {
    const size_t per = (size_t)G.pv * (1 + 2 + 4 + 4);
    return std::max<size_t>(1, std::min<size_t>(16384, (size_t)(768u << 20) / per));
}

// Allocates up front what a march of a full batch needs and a spare arena as large as the
// arena, so that marches in steady state do not allocate (a server calls this at start).
int felt_polys_reserve(felt_polys* q) {
    const Geo& G = q->s->G;
    const size_t nb = std::min<size_t>(march_batch(G), (size_t)G.np);
    const size_t vox = nb * (size_t)G.pv;
    const size_t items = vox / 2 + 1;  // band voxels of the batch's partitions (most are fewer)
    int rc;
    for (DBuf<int>* d : {&q->dparts, &q->pstart})
        if ((rc = d->ensure(nb + 1))) return rc;
    if ((rc = q->vst.ensure(2 * (nb + 1) + 2))) return rc;
    for (DBuf<int>* d : {&q->item_j, &q->item_l, &q->nv, &q->nt, &q->vbase, &q->tbase, &q->work})
        if ((rc = d->ensure(items))) return rc;
    if ((rc = q->cmask.ensure(vox))) return rc;
    if ((rc = q->own.ensure(vox))) return rc;
    if ((rc = q->ocand.ensure(vox))) return rc;
    if ((rc = q->vleaf.ensure(vox))) return rc;
    if ((rc = q->dst.ensure(2 * nb + 2))) return rc;
    if ((rc = q->hdst.ensure(2 * nb + 2))) return rc;
    if ((rc = q->varena2.ensure(q->varena.cap))) return rc;
    if ((rc = q->tarena2.ensure(q->tarena.cap))) return rc;
    return FELT_OK;
}

// march (Polys.hpp:138-165).
int felt_polys_march_impl(felt_polys* q) {
    felt_surface* s = q->s;
    const Geo& G = s->G;
    hipStream_t st = s->st;
    int rc;
    const std::vector<int64_t>& pend = q->pending->list;
    const size_t npend = pend.size();
    std::vector<int> pl(pend.begin(), pend.end()), active(npend);
    if (npend) {
        if ((rc = q->plist.ensure(2 * npend + 2))) return rc;
        HIPX(hipMemcpyAsync(q->plist.ptr, pl.data(), npend * sizeof(int), hipMemcpyHostToDevice,
                            st));
        hipLaunchKernelGGL(k_gather_active, dim3((unsigned)((npend + TPB - 1) / TPB)), dim3(TPB),
                           0, st, s->P, q->plist.ptr, (int)npend, q->plist.ptr + npend);
        HIPX(hipMemcpyAsync(active.data(), q->plist.ptr + npend, npend * sizeof(int),
                            hipMemcpyDeviceToHost, st));
        HIPX(hipStreamSynchronize(st));
    }
    std::vector<int> todo;
    for (size_t i = 0; i < npend; i++) {
        const int64_t p = pend[i];
        PolyKid& k = q->kids[p];
        if (active[i]) {
            k.active = true;
            q->alive.track(p);
            k.nv = k.nt = 0;
            todo.push_back((int)p);
        } else if (k.active) {
            k = PolyKid{};
            q->alive.untrack(p);
        }
    }
    const size_t batch = march_batch(G);
    const int bt = std::min(BT, std::max(64, ((G.pv + 63) / 64) * 64));
    q->last_spx = 0;
    for (size_t b0 = 0; b0 < todo.size(); b0 += batch) {
        const int nb = (int)std::min(batch, todo.size() - b0);
        if ((rc = q->dparts.ensure(nb))) return rc;
        if ((rc = q->pstart.ensure(nb + 1))) return rc;
        if ((rc = q->vst.ensure(2 * (nb + 1) + 2))) return rc;
        HIPX(hipMemcpyAsync(q->dparts.ptr, todo.data() + b0, nb * sizeof(int),
                            hipMemcpyHostToDevice, st));
        hipLaunchKernelGGL(k_mb_count, dim3((nb + TPB - 1) / TPB), dim3(TPB), 0, st, G, s->P,
                           q->dparts.ptr, nb, q->pstart.ptr);
        hipLaunchKernelGGL(k_mb_scan, dim3(1), dim3(BT), 0, st, q->pstart.ptr, nb, q->pstart.ptr);
        int M = 0;
        HIPX(hipMemcpyAsync(&M, q->pstart.ptr + nb, sizeof(int), hipMemcpyDeviceToHost, st));
        HIPX(hipStreamSynchronize(st));
        if (M == 0) continue;
        const int nblk = (M + BT * ITEMS - 1) / (BT * ITEMS);
        if ((rc = q->item_j.ensure(M))) return rc;
        if ((rc = q->item_l.ensure(M))) return rc;
        if ((rc = q->nv.ensure(M))) return rc;
        if ((rc = q->nt.ensure(M))) return rc;
        if ((rc = q->vbase.ensure(M))) return rc;
        if ((rc = q->tbase.ensure(M))) return rc;
        if ((rc = q->tot.ensure(3 * (size_t)nblk + 3))) return rc;
        if ((rc = q->work.ensure(M))) return rc;
        if ((rc = q->cmask.ensure((size_t)nb * G.pv))) return rc;
        if ((rc = q->own.ensure((size_t)nb * G.pv))) return rc;
        if ((rc = q->ocand.ensure((size_t)nb * G.pv))) return rc;
        if ((rc = q->vleaf.ensure((size_t)nb * G.pv))) return rc;
        DISPATCHP(G.D, k_mb_scatter, dim3(nb), dim3(bt), G, s->P, q->dparts.ptr, q->pstart.ptr,
                  q->item_j.ptr, q->item_l.ptr);
        int* bs = q->tot.ptr;
        int* totals = q->tot.ptr + 3 * nblk;  // vertices, simplices, busy items
        int* nwork = totals + 2;
        prof_begin(s, PK_MARCH_COUNT, M);
        DISPATCHP(G.D, k_m_count, dim3((M + TPB - 1) / TPB), dim3(TPB), G, s->P, q->dparts.ptr,
                  q->item_j.ptr, q->item_l.ptr, M, q->cmask.ptr, q->own.ptr, q->ocand.ptr,
                  q->nv.ptr, q->nt.ptr);
        prof_end(s);
        hipLaunchKernelGGL(k_s3_blocks, dim3(nblk), dim3(BT), 0, st, q->nv.ptr, q->nt.ptr, M, bs);
        hipLaunchKernelGGL(k_s3_top, dim3(1), dim3(BT), 0, st, bs, nblk, totals);
        hipLaunchKernelGGL(k_s3_apply, dim3(nblk), dim3(BT), 0, st, q->nv.ptr, q->nt.ptr, M, bs,
                           q->vbase.ptr, q->tbase.ptr, q->work.ptr);
        hipLaunchKernelGGL(k_m_leafbase, dim3((M + TPB - 1) / TPB), dim3(TPB), 0, st, q->item_j.ptr,
                           q->item_l.ptr, M, G.pv, q->vbase.ptr, q->vleaf.ptr);
        int* vst = q->vst.ptr;
        int* tst = q->vst.ptr + (nb + 1);
        hipLaunchKernelGGL(k_m_partstart, dim3((nb + 1 + TPB - 1) / TPB), dim3(TPB), 0, st,
                           q->pstart.ptr, nb, q->vbase.ptr, q->tbase.ptr, M, totals, vst, tst);
        std::vector<int> hv(2 * (nb + 1));
        int W = 0;
        HIPX(hipMemcpyAsync(hv.data(), vst, 2 * (nb + 1) * sizeof(int), hipMemcpyDeviceToHost, st));
        HIPX(hipMemcpyAsync(&W, nwork, sizeof(int), hipMemcpyDeviceToHost, st));
        HIPX(hipStreamSynchronize(st));
        const long long NV = hv[nb], NT = hv[(nb + 1) + nb];
        // Destinations: a partition whose new mesh fits its segment is rewritten in place, the
        // others get a new segment with slack at the bump pointers (after a compaction when the
        // arenas are full).
        auto place = [&](bool dry, long long* need_v, long long* need_t) {
            *need_v = *need_t = 0;
            for (int j = 0; j < nb; j++) {
                PolyKid& k = q->kids[todo[b0 + j]];
                const long long nv = hv[j + 1] - hv[j];
                const long long nt = hv[(nb + 1) + j + 1] - hv[(nb + 1) + j];
                const bool fits = nv <= k.vcap && nt <= k.tcap;
                if (!fits) {
                    *need_v += seg_cap(nv);
                    *need_t += seg_cap(nt);
                }
                if (dry) continue;
                if (!fits) {
                    k.vcap = seg_cap(nv);
                    k.tcap = seg_cap(nt);
                    k.voff = q->vtop;
                    k.toff = q->ttop;
                    q->vtop += k.vcap;
                    q->ttop += k.tcap;
                }
                k.nv = nv;
                k.nt = nt;
                q->hdst.ptr[j] = k.voff;
                q->hdst.ptr[nb + j] = k.toff;
            }
        };
        long long need_v, need_t;
        place(true, &need_v, &need_t);
        if (q->vtop + need_v > (long long)(q->varena.cap / q->vw) ||
            q->ttop + need_t > (long long)(q->tarena.cap / G.D)) {
            long long all_v = 0, all_t = 0;
            for (int j = 0; j < nb; j++) {
                all_v += seg_cap(hv[j + 1] - hv[j]);
                all_t += seg_cap(hv[(nb + 1) + j + 1] - hv[(nb + 1) + j]);
            }
            if ((rc = arena_compact(q, all_v, all_t))) return rc;
        }
        if ((rc = q->hdst.ensure(2 * (size_t)nb + 2))) return rc;
        if ((rc = q->dst.ensure(2 * (size_t)nb + 2))) return rc;
        place(false, &need_v, &need_t);
        HIPX(hipMemcpyAsync(q->dst.ptr, q->hdst.ptr, 2 * (size_t)nb * sizeof(long long),
                            hipMemcpyHostToDevice, st));
        q->last_spx += NT;
        if (NV > 0 && (rc = q->vwork.ensure((size_t)NV))) return rc;
        prof_begin(s, PK_MARCH_EMIT, W);
        if (W > 0) {
            const int nvw = (int)NV;
            hipLaunchKernelGGL(k_m_vlist, dim3((W + TPB - 1) / TPB), dim3(TPB), 0, st,
                               q->work.ptr, W, q->own.ptr, q->item_j.ptr, q->item_l.ptr, G.pv,
                               q->vbase.ptr, q->vwork.ptr);
            if (nvw > 0)
                DISPATCHP(G.D, k_m_vert, dim3((nvw + TPB - 1) / TPB), dim3(TPB), G, s->P,
                          q->dparts.ptr, q->item_j.ptr, q->item_l.ptr, q->vwork.ptr, nvw,
                          q->own.ptr, q->vbase.ptr, vst, q->dst.ptr, q->varena.ptr);
            DISPATCHP(G.D, k_m_tri, dim3((W + TPB - 1) / TPB), dim3(TPB), G, q->item_j.ptr,
                      q->item_l.ptr, q->work.ptr, W, q->cmask.ptr, q->own.ptr, q->ocand.ptr,
                      q->tbase.ptr, q->vleaf.ptr, vst, tst, q->dst.ptr + nb, q->tarena.ptr);
        }
        prof_end(s);
    }
    HIPX(hipStreamSynchronize(st));
    std::swap(q->pending, q->done);
    q->pending->reset();
    return FELT_OK;
}

int felt_polys_child_mesh_impl(felt_polys* q, uint64_t p, float* vtx, size_t vcap, uint32_t* spx,
                               size_t scap) {
    const PolyKid& k = q->kids[p];
    const int D = q->s->G.D;
    if (vcap < (size_t)k.nv * q->vw || scap < (size_t)k.nt * D)
        return felt_fail(FELT_E_CAPACITY, "mesh buffers too small");
    if (k.nv)
        HIPX(hipMemcpyAsync(vtx, q->varena.ptr + k.voff * q->vw, k.nv * q->vw * sizeof(float),
                            hipMemcpyDeviceToHost, q->s->st));
    if (k.nt)
        HIPX(hipMemcpyAsync(spx, q->tarena.ptr + k.toff * D, k.nt * D * sizeof(uint32_t),
                            hipMemcpyDeviceToHost, q->s->st));
    HIPX(hipStreamSynchronize(q->s->st));
    return FELT_OK;
}

uint64_t felt_polys_device_bytes(const felt_polys* q) {
    uint64_t b = (q->varena.cap + q->varena2.cap) * sizeof(float) +
                 (q->tarena.cap + q->tarena2.cap) * sizeof(uint32_t) + q->dst.cap * sizeof(long long);
    b += q->cmask.cap + q->own.cap * sizeof(uint16_t) + q->ocand.cap * sizeof(uint32_t);
    b += q->bits.cap + q->dseg.cap * sizeof(long long);
    for (const DBuf<int>* d : {&q->vbase, &q->tbase, &q->tot, &q->dparts, &q->pstart, &q->vst,
                               &q->item_j, &q->item_l, &q->nv, &q->nt, &q->vleaf, &q->work,
                               &q->vwork, &q->plist})
        b += d->cap * sizeof(int);
    return b;
}
