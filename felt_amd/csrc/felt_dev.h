// felt_dev.h — device-side data layout and stencil helpers shared by the surface and polygon
// kernels (gfx950).
//
// One handle owns a pool of one or more member surfaces of the same dimension, layer count and
// partition size (a single Felt::Surface is a pool of one).  Their partitions share one index
// space: member m owns pool partitions [pbase_m, pbase_m + np_m), row-major over its children
// grid.  Layout in HBM:
//   page[np]        partition -> brick slot, or -1 when the partition is inactive
//   pbg[np]         value every voxel of an inactive partition reads (the reference's child
//                   background, GridMixin.hpp:587-621; PartitionedMixin.hpp:502-516)
//   pcnt[np*NL]     voxels per partition per layer list; a partition is in the reference's
//                   parent list k exactly when pcnt[p*NL+k] > 0 (PartitionedMixin.hpp:489-558)
//   part_mem[np]    partition -> member surface
//   slot_part[cap]  brick slot -> partition, -1 when free
//   slot_org[cap]   brick slot -> grid coordinates of its partition's first voxel, and member
//                   (bits 0-29) with bit 30 set when the brick touches a face of its grid
//   nbr[2*cap]      brick slot -> its face-neighbour record (the six neighbour bricks, or which
//                   neighbour partitions are inactive or outside the grid; its in-grid extent)
//   phi[cap*pv]     signed distances, one brick of pv = prod(part) voxels per slot, stored in
//                   tiles (below)
//   lay[cap*pv]     layer list of each voxel (-L..L) or NONE: replaces the reference's per-child
//                   lookup grid and per-layer leaf lists (LookupMixin.hpp:225-387)
//   mem[nmem]       member geometry
// A voxel is named by v = slot*pv + leaf.  Grid coordinates g are world coordinates minus the
// member's grid offset, so 0 <= g < size; partition coordinates are g / part.
//
// Brick storage order.  A brick is cut into tiles of tl[0] x tl[1] x tl[2] voxels (4x4x2 in 3D,
// 4x8 in 2D, each edge the largest power of two up to that which divides the partition edge);
// tiles are row-major in the brick and voxels row-major in a tile, so one 128-byte line of phi
// holds a 4x4x2 block.  The band is a thin sheet: a line then carries several band voxels of a
// layer and most of their face neighbours, where a row-major line (32 voxels along one axis) is
// crossed by the sheet about once.  Partition edges that are not multiples of 2 give 1-voxel tile
// edges, i.e. the reference's row-major order (Util.hpp:23-40).  The reference's leaf PosIdx (row-
// major in the partition) is kept wherever it is observable: mesh ordering and the on-disk format.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace felt {

constexpr int8_t NONE = -128;
constexpr float kEps = 1.1920928955078125e-07f;  // FLT_EPSILON

// Division by a run-time invariant divisor with one high multiply (Granlund & Montgomery
// 1994; Warren, Hacker's Delight 10-8): s = ceil(log2 d), m = floor(2^32 (2^s - d) / d) + 1,
// q = (t + ((n - t) >> 1)) >> (s - 1) with t = mulhi(n, m).  Exact for every 32-bit n.
// A power-of-two divisor (sh >= 0) divides by a shift.
struct FastDiv {
    uint32_t d, m;
    int s;
    int sh;
};

inline FastDiv make_fastdiv(uint32_t d) {
    FastDiv f{d, 0u, 0, -1};
    if (d != 0 && (d & (d - 1)) == 0) {
        while ((1u << (f.sh + 1)) != d) f.sh++;
        f.sh++;
    }
    if (d <= 1) return f;
    int s = 0;
    while ((1ull << s) < d) s++;
    f.s = s;
    f.m = (uint32_t)((((1ull << 32) * ((1ull << s) - d)) / d) + 1);
    return f;
}

__device__ __forceinline__ uint32_t fdiv(uint32_t n, const FastDiv& f) {
    if (f.sh >= 0) return n >> f.sh;
    if (f.d <= 1) return n;
    const uint32_t t = __umulhi(n, f.m);
    return (t + ((n - t) >> 1)) >> (f.s - 1);
}

// Geometry shared by every member of a pool.
struct Geo {
    int D, L, NL;
    int part[3];
    int lstr[3];  // row-major leaf strides inside a partition (the reference's leaf PosIdx)
    int tl[3], tsh[3], tmask[3];  // tile edge, its log2, edge - 1
    int wstr[3];  // stride of a step inside a tile
    int tstr[3];  // stride of a step to the next tile
    int tvsh;     // log2 of the tile volume
    FastDiv fnt[3];  // tiles per axis
    int pv;
    int np;       // partitions of all members
    int nmem;     // member surfaces
    FastDiv fpart[3], fpv;
};

// One member surface: its grid and children grid (Surface.hpp:163-173; PartitionedMixin.hpp:
// 131-141).  The first two 16-byte rows are read per voxel by the stencil kernels.
struct alignas(16) Mem {
    int size[3];
    int pbase;      // its first pool partition
    int cstr[3];    // partition strides in its children grid
    int np;         // its partitions
    int csize[3];
    int offset[3];  // world position of grid position 0: -size/2 (Surface.hpp:1716-1719)
    FastDiv fcsize[3];
};

struct Pool {
    int* page;
    float* pbg;
    int* pcnt;
    int* slot_part;
    float* phi;
    int8_t* lay;
    int4* slot_org;
    const Mem* mem;
    const int* part_mem;
    int4* nbr;  // two per slot: the brick's face-neighbour record (below)
};

// Face-neighbour record of an active brick, two int4 per slot, kept in step with the page table
// by k_nbr_fix after every activation and deactivation.  Field j = 2a + h names the neighbour
// partition one step along axis a (h = 0: -1, h = 1: +1): its brick slot when active,
// kNbrOut when it lies outside the member's children grid, else kNbrInactive - p for pool
// partition p (whose voxels read pbg[p]).  Fields 0-3 are the first int4, 4-5 the x and y of the
// second; its z holds the in-grid extent of the brick along axes 0 and 1 (16 bits each) and its
// w along axis 2: min(part, size - origin), less than part only for a brick the grid cuts.  A
// distance gather then needs no page-table, brick-origin or member read.
constexpr int kNbrOut = -1;
constexpr int kNbrInactive = -2;

// slot_org.w: the member and whether the brick touches a face of its member's grid.
constexpr int kMemberMask = 0x3fffffff;
constexpr int kEdgeBrick = 1 << 30;
constexpr int kFar = 0x7fffffff;  // a grid size no neighbour of an inner brick reaches

// A voxel of the band being processed: id, member, coordinates in its brick and in its grid,
// and its grid's size.  Every face neighbour of a voxel of a brick that touches no grid face is
// inside the grid, so for those bricks `size` is left at kFar and the member's geometry is read
// only when a neighbour lies in another brick.
struct Vx {
    uint32_t v;
    int m;
    int loc[3];
    int g[3];
    int size[3];
};

__device__ __forceinline__ int4 mem_row0(const Pool& P, int m) {
    return *reinterpret_cast<const int4*>(&P.mem[m].size[0]);
}
__device__ __forceinline__ int4 mem_row1(const Pool& P, int m) {
    return *reinterpret_cast<const int4*>(&P.mem[m].cstr[0]);
}

// iround(v + FLT_EPSILON), exact round-half-away-from-zero (Surface.hpp:1694-1697).  NaN has
// no reference value; it maps outside the band and is counted.
__device__ __forceinline__ int layer_of(float v, int L, int* nan_ctr) {
    const float t = v + kEps;
    if (t != t) {
        atomicAdd(nan_ctr, 1);
        return L + 1;
    }
    return (int)__builtin_roundf(t);
}

__device__ __forceinline__ bool in_band(int k, int L) { return k >= -L && k <= L; }

// Storage leaf of brick-local coordinates lc.
template <int D>
__host__ __device__ __forceinline__ int leaf_of(const Geo& G, const int lc[3]) {
    int l = 0;
#pragma unroll
    for (int a = 0; a < D; a++) l += (lc[a] >> G.tsh[a]) * G.tstr[a] + (lc[a] & G.tmask[a]) * G.wstr[a];
    return l;
}

// Brick-local coordinates of storage leaf l.
template <int D>
__device__ __forceinline__ void leaf_coords(const Geo& G, uint32_t l, int lc[3]) {
    uint32_t w = l & ((1u << G.tvsh) - 1u), t = l >> G.tvsh;
#pragma unroll
    for (int a = D - 1; a >= 0; a--) {
        const int in = (int)(w & (uint32_t)G.tmask[a]);
        w >>= G.tsh[a];
        const uint32_t q = fdiv(t, G.fnt[a]);
        lc[a] = ((int)(t - q * (uint32_t)(G.part[a] >> G.tsh[a])) << G.tsh[a]) | in;
        t = q;
    }
}

// Brick-local coordinates of the reference's row-major leaf PosIdx l.
template <int D>
__device__ __forceinline__ void rm_coords(const Geo& G, uint32_t l, int lc[3]) {
#pragma unroll
    for (int a = D - 1; a >= 0; a--) {
        const uint32_t q = fdiv(l, G.fpart[a]);
        lc[a] = (int)(l - q * (uint32_t)G.part[a]);
        l = q;
    }
}

template <int D>
__host__ __device__ __forceinline__ int rm_of(const Geo& G, const int lc[3]) {
    int l = 0;
#pragma unroll
    for (int a = 0; a < D; a++) l += lc[a] * G.lstr[a];
    return l;
}

// Storage offset of a step d (components in {-1,0,1}) from brick-local coordinates lc that
// stays in the brick.
template <int D>
__device__ __forceinline__ int step_offset(const Geo& G, const int lc[3], const int d[3]) {
    int off = 0;
#pragma unroll
    for (int a = 0; a < D; a++) {
        if (d[a] == 0) continue;
        const int in = (lc[a] & G.tmask[a]) + d[a];
        off += (in >= 0 && in <= G.tmask[a]) ? d[a] * G.wstr[a]
                                              : d[a] * (G.tstr[a] - G.tmask[a] * G.wstr[a]);
    }
    return off;
}

template <int D>
__device__ __forceinline__ void decode(const Geo& G, const Pool& P, uint32_t v, Vx& x) {
    const uint32_t slot = fdiv(v, G.fpv);
    x.v = v;
#pragma unroll
    for (int a = D; a < 3; a++) {
        x.loc[a] = 0;
        x.g[a] = 0;
        x.size[a] = 1;
    }
    leaf_coords<D>(G, v - slot * (uint32_t)G.pv, x.loc);
    const int4 o = P.slot_org[slot];
    x.m = o.w & kMemberMask;
    int4 r0 = make_int4(kFar, kFar, kFar, 0);
    if (o.w & kEdgeBrick) r0 = mem_row0(P, x.m);
    x.g[0] = o.x + x.loc[0];
    x.size[0] = r0.x;
    if (D > 1) {
        x.g[1] = o.y + x.loc[1];
        x.size[1] = r0.y;
    }
    if (D > 2) {
        x.g[2] = o.z + x.loc[2];
        x.size[2] = r0.z;
    }
}

// Member-local partition coordinates of local partition index p.
template <int D>
__device__ __forceinline__ void part_coords(const Mem& M, int p, int pc[3]) {
    uint32_t r = (uint32_t)p;
#pragma unroll
    for (int a = D - 1; a >= 0; a--) {
        const uint32_t q = fdiv(r, M.fcsize[a]);
        pc[a] = (int)(r - q * (uint32_t)M.csize[a]);
        r = q;
    }
}

template <int D>
__device__ __forceinline__ bool inside(const int size[3], const int q[3]) {
#pragma unroll
    for (int a = 0; a < D; a++)
        if (q[a] < 0 || q[a] >= size[a]) return false;
    return true;
}

// Pool partition and leaf of an in-grid position of member m.
template <int D>
__device__ __forceinline__ void locate(const Geo& G, const Pool& P, int m, const int q[3], int& p,
                                       int& l) {
    const int4 r0 = mem_row0(P, m), r1 = mem_row1(P, m);
    const int cs[3] = {r1.x, r1.y, r1.z};
    p = r0.w;
    int lc[3] = {0, 0, 0};
#pragma unroll
    for (int a = 0; a < D; a++) {
        const int qa = (int)fdiv((uint32_t)q[a], G.fpart[a]);
        p += qa * cs[a];
        lc[a] = q[a] - qa * G.part[a];
    }
    l = leaf_of<D>(G, lc);
}

template <int D>
__device__ __forceinline__ int part_of(const Geo& G, const Pool& P, int m, const int q[3]) {
    int p, l;
    locate<D>(G, P, m, q, p, l);
    return p;
}

// get(pos) of member m's partitioned isogrid (PartitionedMixin.hpp:594-601): L+1 outside the
// grid, the partition background when inactive.
template <int D>
__device__ __forceinline__ float get_g(const Geo& G, const Pool& P, int m, const int size[3],
                                       const int q[3]) {
    if (!inside<D>(size, q)) return (float)(G.L + 1);
    int p, l;
    locate<D>(G, P, m, q, p, l);
    const int s = P.page[p];
    if (s < 0) return P.pbg[p];
    return P.phi[(size_t)s * G.pv + l];
}

// Value at x.g + d (d components in {-1,0,1}); reads the voxel's own brick directly when the
// neighbour stays inside it and inside the grid.
template <int D>
__device__ __forceinline__ float nb(const Geo& G, const Pool& P, const Vx& x, int d0, int d1,
                                    int d2) {
    const int d[3] = {d0, d1, d2};
    bool same = true;
    int q[3];
#pragma unroll
    for (int a = 0; a < D; a++) {
        const int la = x.loc[a] + d[a];
        q[a] = x.g[a] + d[a];
        if (la < 0 || la >= G.part[a] || q[a] >= x.size[a]) same = false;
    }
    if (same) return P.phi[(int64_t)x.v + step_offset<D>(G, x.loc, d)];
    return get_g<D>(G, P, x.m, x.size, q);
}

// distance(pos, dist, dir) (Surface.hpp:1387-1440): signed minimum over in-grid face
// neighbours in the direction of the zero curve, plus one step.
template <int D>
__device__ __forceinline__ float distance(const Geo& G, const Pool& P, const Vx& x, float self,
                                          float dir) {
    float dist = self * dir;
#pragma unroll
    for (int a = 0; a < D; a++) {
#pragma unroll
        for (int s = -1; s <= 1; s += 2) {
            const int qa = x.g[a] + s;
            if (qa < 0 || qa >= x.size[a]) continue;
            const float dn = nb<D>(G, P, x, a == 0 ? s : 0, a == 1 ? s : 0, a == 2 ? s : 0);
            if (dn * dir < dist) dist = dn * dir;
        }
    }
    const float dn = dist * dir;
    return dn + dir;
}

// distance() of one band voxel through its brick's face-neighbour record, with the voxel's own
// value given: the record and the voxel's in-brick neighbours are loaded together, then the
// values of the neighbours in other bricks (or their partitions' backgrounds), so the gather is
// two dependent levels after the list entry.  Same result as distance().
template <int D>
__device__ __forceinline__ float distance_rec_self(const Geo& G, const Pool& P, uint32_t v,
                                                   float self, float dir) {
    constexpr int NN = 2 * D;
    const uint32_t slot = fdiv(v, G.fpv);
    int loc[3] = {0, 0, 0};
    leaf_coords<D>(G, v - slot * (uint32_t)G.pv, loc);
    const int4 ra = P.nbr[2 * (size_t)slot], rb = P.nbr[2 * (size_t)slot + 1];
    const int rec[6] = {ra.x, ra.y, ra.z, ra.w, rb.x, rb.y};
    const int ext[3] = {(int)((unsigned)rb.z & 0xffffu), (int)((unsigned)rb.z >> 16), rb.w};
    // per neighbour: whether it is in the grid, and where its value is -- a voxel of this brick
    // or of the neighbour brick, or the background of an inactive neighbour partition.  Every
    // neighbour is loaded (one outside the grid reads this voxel and is masked afterwards), so
    // the loads are issued together with no branch between them.
    bool in[NN];
    const float* src[NN];
#pragma unroll
    for (int a = 0; a < D; a++) {
#pragma unroll
        for (int h = 0; h < 2; h++) {
            const int sd = h ? 1 : -1, j = 2 * a + h;
            const int la = loc[a] + sd;
            const bool cross = la < 0 || la >= G.part[a];
            int d[3] = {0, 0, 0};
            d[a] = sd;
            int lc[3] = {loc[0], loc[1], loc[2]};
            lc[a] = sd < 0 ? G.part[a] - 1 : 0;
            const int t = rec[j];
            const uint32_t own = v + (uint32_t)step_offset<D>(G, loc, d);
            const uint32_t other = (uint32_t)max(t, 0) * (uint32_t)G.pv + (uint32_t)leaf_of<D>(G, lc);
            in[j] = cross ? t != kNbrOut : la < ext[a];
            src[j] = !cross ? P.phi + own
                            : (t >= 0 ? P.phi + other : (t < kNbrOut ? P.pbg + (kNbrInactive - t)
                                                                    : P.phi + v));
        }
    }
    float nv[NN];
#pragma unroll
    for (int j = 0; j < NN; j++) nv[j] = *src[j];
    float dist = self * dir;
#pragma unroll
    for (int j = 0; j < NN; j++)
        if (in[j] && nv[j] * dir < dist) dist = nv[j] * dir;
    const float dn = dist * dir;
    return dn + dir;
}

// The same, reading the voxel's own value alongside the record.
template <int D>
__device__ __forceinline__ float distance_rec(const Geo& G, const Pool& P, uint32_t v, float dir) {
    return distance_rec_self<D>(G, P, v, P.phi[v], dir);
}

// A band voxel with what neighbour gathers need: coordinates in its brick and grid, its grid's
// size, its member's partition strides and its own pool partition.
struct VoxCtx {
    uint32_t v;
    int m, own;
    int loc[3], g[3], sz[3], cs[3];
};

template <int D>
__device__ __forceinline__ void vox_ctx(const Geo& G, const Pool& P, uint32_t v, VoxCtx& c) {
    const uint32_t slot = fdiv(v, G.fpv);
    const int4 o = P.slot_org[slot];
    c.v = v;
    c.m = o.w & kMemberMask;
    const int4 r0 = mem_row0(P, c.m), r1 = mem_row1(P, c.m);
    c.loc[0] = c.loc[1] = c.loc[2] = 0;
    leaf_coords<D>(G, v - slot * (uint32_t)G.pv, c.loc);
    const int org[3] = {o.x, o.y, o.z};
    const int sz[3] = {r0.x, r0.y, r0.z}, cs[3] = {r1.x, r1.y, r1.z};
    c.own = r0.w;
#pragma unroll
    for (int a = 0; a < 3; a++) {
        c.g[a] = org[a] + c.loc[a];
        c.sz[a] = sz[a];
        c.cs[a] = cs[a];
        if (a < D) c.own += (int)fdiv((uint32_t)org[a], G.fpart[a]) * cs[a];
    }
}

// get() (background L+1 outside the grid) at K offsets off[j] (components in {-1,0,1}) of a
// voxel, gathered in levels: every neighbour's brick or partition first, then the pages of the
// neighbours in other bricks, then all values, so the K loads of a level are in flight together.
template <int D, int K>
__device__ __forceinline__ void gather(const Geo& G, const Pool& P, const VoxCtx& c,
                                       const int (*off)[3], float* val, bool* in_grid = nullptr) {
    int idx[K], part[K];
    bool in[K];
#pragma unroll
    for (int j = 0; j < K; j++) {
        bool inside = true, cross = false;
        int lc[3] = {c.loc[0], c.loc[1], c.loc[2]};
        int np = c.own;
#pragma unroll
        for (int a = 0; a < D; a++) {
            const int q = c.g[a] + off[j][a];
            if (q < 0 || q >= c.sz[a]) inside = false;
            lc[a] += off[j][a];
            if (lc[a] < 0) {
                lc[a] += G.part[a];
                np -= c.cs[a];
                cross = true;
            } else if (lc[a] >= G.part[a]) {
                lc[a] -= G.part[a];
                np += c.cs[a];
                cross = true;
            }
        }
        in[j] = inside;
        if (cross) {
            idx[j] = leaf_of<D>(G, lc);
            part[j] = np;
        } else {
            idx[j] = (int)c.v + step_offset<D>(G, c.loc, off[j]);
            part[j] = -1;
        }
    }
    // every page entry and value is loaded, with no branch between the loads: a neighbour
    // outside the grid reads the voxel itself (and its own partition's page) and is replaced
    // by L+1 afterwards
    int pg[K];
#pragma unroll
    for (int j = 0; j < K; j++) pg[j] = P.page[(in[j] && part[j] >= 0) ? part[j] : c.own];
    float raw[K];
#pragma unroll
    for (int j = 0; j < K; j++) {
        const float* src = P.phi + c.v;
        if (in[j])
            src = part[j] < 0 ? P.phi + (uint32_t)idx[j]
                              : (pg[j] < 0 ? P.pbg + part[j]
                                           : P.phi + ((size_t)pg[j] * G.pv + idx[j]));
        raw[j] = *src;
    }
#pragma unroll
    for (int j = 0; j < K; j++) {
        val[j] = in[j] ? raw[j] : (float)(G.L + 1);
        if (in_grid) in_grid[j] = in[j];
    }
}

// Storage leaf term of brick-local coordinate x along axis a: the leaf of coordinates lc is the
// sum of its axes' terms (leaf_of).  Both products are below pv <= 2^24, so 24-bit multiplies
// are exact.
__device__ __forceinline__ int leaf_term(const Geo& G, int a, int x) {
    return (int)(__umul24((uint32_t)(x >> G.tsh[a]), (uint32_t)G.tstr[a]) +
                 __umul24((uint32_t)(x & G.tmask[a]), (uint32_t)G.wstr[a]));
}

// gather<D, K>(G, P, c, off, val) for offsets whose components are compile-time constants, with
// the work shared between the K neighbours: per axis, the leaf terms of the coordinate and its
// two steps (wrapped into the neighbouring brick), whether each step leaves the brick or the
// grid; per brick that a neighbour can fall in (the voxel's own, the one across each axis the
// voxel borders, the one across each pair of those), one page lookup.  A neighbour's value is
// then read from its brick's base plus the sum of its axes' leaf terms, with no branch.
template <int D, int K>
__device__ __forceinline__ void gather_sep(const Geo& G, const Pool& P, const VoxCtx& c,
                                           const int (*off)[3], float* val) {
    int F[3][3];     // leaf term of the coordinate stepped by -1, 0, +1
    bool inn[3][3];  // the stepped coordinate is inside the grid
    int dir[3];      // the side of the brick the voxel borders along the axis (0: neither; every
                     // partition edge is at least 2, so a voxel borders at most one side)
    bool bord[3];    // the step across that side stays inside the grid
    int leaf = 0;
#pragma unroll
    for (int a = 0; a < D; a++) {
        const int l = c.loc[a], pa = G.part[a];
        const bool lo = l == 0, hi = l == pa - 1;
        F[a][0] = leaf_term(G, a, lo ? pa - 1 : l - 1);
        F[a][1] = leaf_term(G, a, l);
        F[a][2] = leaf_term(G, a, hi ? 0 : l + 1);
        leaf += F[a][1];
        inn[a][0] = c.g[a] > 0;
        inn[a][1] = true;
        inn[a][2] = c.g[a] + 1 < c.sz[a];
        dir[a] = lo ? -1 : (hi ? 1 : 0);
        bord[a] = lo ? inn[a][0] : (hi && inn[a][2]);
    }
    const uint32_t own_base = c.v - (uint32_t)leaf;
    // bricks: [0] own; [1 + a] across axis a; [1 + D + pair(a, b)] across axes a and b
    constexpr int NB = D == 3 ? 7 : 4;
    int bp[NB];  // partition index (valid when the brick is inside the grid)
    bool bin[NB];
    bp[0] = c.own;
    bin[0] = true;
#pragma unroll
    for (int a = 0; a < D; a++) {
        bp[1 + a] = c.own + dir[a] * c.cs[a];
        bin[1 + a] = bord[a];
    }
#pragma unroll
    for (int a = 0; a < D; a++)
#pragma unroll
        for (int b = a + 1; b < D; b++) {
            const int k = 1 + D + (D == 2 ? 0 : (a == 0 ? b - 1 : 2));
            bp[k] = c.own + dir[a] * c.cs[a] + dir[b] * c.cs[b];
            bin[k] = bin[1 + a] && bin[1 + b];
        }
    int pg[NB];
    pg[0] = 0;
#pragma unroll
    for (int k = 1; k < NB; k++) pg[k] = P.page[bin[k] ? bp[k] : c.own];
    // brick k's values: phi + its base (active) or its background (inactive); a brick outside
    // the grid reads the voxel's own page and is never used, as its neighbours are outside too
    const float* src_b[NB];
    src_b[0] = P.phi + own_base;
#pragma unroll
    for (int k = 1; k < NB; k++)
        src_b[k] = pg[k] >= 0 ? P.phi + (uint32_t)pg[k] * (uint32_t)G.pv : P.pbg + bp[k];
#pragma unroll
    for (int j = 0; j < K; j++) {
        int l = 0, a0 = -1, a1 = -1;  // a0, a1: the axes the offset steps along (at most two)
        bool inside = true;
#pragma unroll
        for (int a = 0; a < D; a++) {
            const int o = off[j][a];
            l += F[a][o + 1];
            inside = inside && inn[a][o + 1];
            if (o != 0) {
                if (a0 < 0) a0 = a;
                else a1 = a;
            }
        }
        // the brick: across the axes along which the step leaves the voxel's brick
        const float* src = src_b[0];
        bool inactive = false;
        if (a0 >= 0) {
            const bool c0 = off[j][a0] == dir[a0];
            src = c0 ? src_b[1 + a0] : src;
            inactive = c0 && pg[1 + a0] < 0;
            if (a1 >= 0) {
                const bool c1 = off[j][a1] == dir[a1];
                const int kp = 1 + D + (D == 2 ? 0 : (a0 == 0 ? a1 - 1 : 2));
                src = c1 ? (c0 ? src_b[kp] : src_b[1 + a1]) : src;
                inactive = c1 ? (c0 ? pg[kp] < 0 : pg[1 + a1] < 0) : inactive;
            }
        }
        // an inactive brick's background is one value; a neighbour outside the grid reads the
        // voxel itself and is replaced by L+1
        const float raw = *(inside ? (inactive ? src : src + l) : P.phi + c.v);
        val[j] = inside ? raw : (float)(G.L + 1);
    }
}

// The curvature stencil: 2D face neighbours (face(a, s) at 2a + (s > 0)), then for each axis
// pair a < b the four edge neighbours (a, b) signs (-,-), (-,+), (+,-), (+,+).
template <int D>
struct Stencil;
template <>
struct Stencil<2> {
    static constexpr int K = 8;
};
template <>
struct Stencil<3> {
    static constexpr int K = 18;
};
__device__ constexpr int kStencil3[18][3] = {
    {-1, 0, 0}, {1, 0, 0}, {0, -1, 0}, {0, 1, 0}, {0, 0, -1}, {0, 0, 1},
    {-1, -1, 0}, {-1, 1, 0}, {1, -1, 0}, {1, 1, 0},      // x,y
    {-1, 0, -1}, {-1, 0, 1}, {1, 0, -1}, {1, 0, 1},      // x,z
    {0, -1, -1}, {0, -1, 1}, {0, 1, -1}, {0, 1, 1}};     // y,z
__device__ constexpr int kStencil2[8][3] = {
    {-1, 0, 0}, {1, 0, 0}, {0, -1, 0}, {0, 1, 0},
    {-1, -1, 0}, {-1, 1, 0}, {1, -1, 0}, {1, 1, 0}};

template <int D>
__device__ __forceinline__ int st_face(int a, int s) { return 2 * a + (s > 0); }
// Edge neighbour with sign sa along axis a and sb along axis b (a != b).
template <int D>
__device__ __forceinline__ int st_edge(int a, int sa, int b, int sb) {
    const int p = a < b ? a : b, q = a < b ? b : a;
    const int sp = a < b ? sa : sb, sq = a < b ? sb : sa;
    const int pair = D == 2 ? 0 : (p == 0 ? (q == 1 ? 0 : 1) : 2);
    return 2 * D + 4 * pair + 2 * (sp > 0) + (sq > 0);
}

// Mean curvature by difference of normals (NumericMixin.hpp:72-156) from the gathered stencil,
// same operation order as the reference with no contraction: forward normals, backward normals,
// then (dn0 + (dn1 + dn2)) / 2 in 3D or (dn0 + dn1) / 2 in 2D.
template <int D>
__device__ __forceinline__ float curvature_of(float c, const float* v) {
    float nf[3], nbk[3];
#pragma unroll
    for (int a = 0; a < D; a++) {
        const float va = v[st_face<D>(a, 1)] - c;
        float sq = 0.0f;
#pragma unroll
        for (int b = 0; b < D; b++) {
            if (b == a) continue;
            const float vl = v[st_edge<D>(a, 1, b, -1)];
            const float vh = v[st_edge<D>(a, 1, b, 1)];
            const float t = (vh - vl) / 2.0f;
            sq += t * t;
        }
        nf[a] = va / __builtin_sqrtf(va * va + sq);
    }
#pragma unroll
    for (int a = 0; a < D; a++) {
        const float va = c - v[st_face<D>(a, -1)];
        float sq = 0.0f;
#pragma unroll
        for (int b = 0; b < D; b++) {
            if (b == a) continue;
            const float vl = v[st_edge<D>(a, -1, b, -1)];
            const float vh = v[st_edge<D>(a, -1, b, 1)];
            const float t = (vh - vl) / 2.0f;
            sq += t * t;
        }
        nbk[a] = va / __builtin_sqrtf(va * va + sq);
    }
    float dn[3];
#pragma unroll
    for (int a = 0; a < D; a++) dn[a] = nf[a] - nbk[a];
    float s;
    if (D == 3) s = dn[0] + (dn[1] + dn[D - 1]);
    else s = dn[0] + dn[1];
    return s / 2.0f;
}

// Number of lanes below this one whose bit is set in the wavefront mask m.
__device__ __forceinline__ int lanes_below(unsigned long long m) {
    return (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32),
                                          __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

// Wave-aggregated append: one atomic per wavefront.  Every lane of the wave must call it.
__device__ __forceinline__ int wave_append(int* counter, bool pred) {
    const unsigned long long m = __ballot(pred);
    const int lane = __lane_id();
    if (m == 0ull) return -1;
    const int leader = __ffsll((long long)m) - 1;
    int base = 0;
    if (lane == leader) base = atomicAdd(counter, __popcll(m));
    base = __shfl(base, leader);
    const unsigned long long below = lane == 0 ? 0ull : (m & ((~0ull) >> (64 - lane)));
    return pred ? base + __popcll(below) : -1;
}

// Wave-aggregated append of k entries per lane: one atomic per wavefront; returns the lane's
// first index.  Every lane of the wave must call it.
__device__ __forceinline__ int wave_append_n(int* counter, int k) {
    const int lane = __lane_id();
    int x = k;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int y = __shfl_up(x, o);
        if (lane >= o) x += y;
    }
    const int total = __shfl(x, 63);
    int base = 0;
    if (lane == 63 && total) base = atomicAdd(counter, total);
    base = __shfl(base, 63);
    return base + x - k;
}

// Sets a 0/1 mark byte to 1 with one 32-bit atomic OR on its word; true when it was 0.
__device__ __forceinline__ bool claim_mark(uint8_t* addr) {
    const uintptr_t a = (uintptr_t)addr;
    unsigned int* w = (unsigned int*)(a & ~(uintptr_t)3);
    const int sh = (int)(a & 3) * 8;
    return ((atomicOr(w, 1u << sh) >> sh) & 0xffu) == 0u;
}

// Compare-and-swap of one byte through its aligned 32-bit word.
__device__ __forceinline__ bool cas_byte(int8_t* addr, int8_t expect, int8_t desired) {
    const uintptr_t a = (uintptr_t)addr;
    unsigned int* w = (unsigned int*)(a & ~(uintptr_t)3);
    const int sh = (int)(a & 3) * 8;
    unsigned int old = __hip_atomic_load(w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    while (true) {
        if ((int8_t)((old >> sh) & 0xffu) != expect) return false;
        const unsigned int nw = (old & ~(0xffu << sh)) | ((unsigned int)(uint8_t)desired << sh);
        const unsigned int prev = atomicCAS(w, old, nw);
        if (prev == old) return true;
        old = prev;
    }
}

}  // namespace felt
