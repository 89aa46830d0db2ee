// felt_step_kernels.h — kernels of one update step (included by felt_surface.hip only).
//
// Each round of Surface::update_end[_global] (Surface.hpp:404-430, 739-762) runs over one
// "processed list" of voxels (the band, the affected set or the last round's buffer): every
// entry carries its voxel id v, its layer at the start of the round (tag) and the layer its new
// value falls in (tk).  Status changes are the entries with tk != tag; they are compacted into
// the status-change records and the next round's buffer by count / scan / scatter passes, so no
// list append contends on a global atomic.  Per-partition counters are updated with one atomic
// per (wavefront, partition, layer).
//
// Every kernel reads its list bounds from device memory (a span {lo, hi}) and loops with a
// grid stride, and every kernel of a round takes the round's live flag, so a whole step can be
// launched without reading anything back; rounds that turn out not to be needed exit at once.
#pragma once

#include "felt_kernels.h"

namespace felt {
namespace step {

constexpr int BT = 1024;  // threads of the scan / compaction blocks
constexpr int ITEMS = 16;  // consecutive list entries per thread in the change compaction

// Exclusive scan of one int per thread over a block of BT threads; returns the block total.
__device__ __forceinline__ int block_excl_scan(int val, int* total) {
    __shared__ int wsum[BT / 64];
    __shared__ int wpre[BT / 64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    int x = val;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int y = __shfl_up(x, o);
        if (lane >= o) x += y;
    }
    if (lane == 63) wsum[wave] = x;
    __syncthreads();
    if (wave == 0) {
        const int nw = BT / 64;
        int w = lane < nw ? wsum[lane] : 0;
        int z = w;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const int y = __shfl_up(z, o);
            if (lane >= o) z += y;
        }
        if (lane < nw) wpre[lane] = z - w;
        if (lane == nw - 1) wsum[0] = z;  // block total (wsum[0] no longer needed)
    }
    __syncthreads();
    const int res = wpre[wave] + x - val;
    *total = wsum[0];
    __syncthreads();
    return res;
}

// Exclusive scans of three ints per thread over a block of BT threads, with one set of barriers
// for the three; tot[c] is the block total of column c.
__device__ __forceinline__ void block_excl_scan3(const int v[3], int ex[3], int tot[3]) {
    constexpr int NW = BT / 64;
    __shared__ int wsum3[3][NW];
    __shared__ int wpre3[3][NW + 1];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    int x[3] = {v[0], v[1], v[2]};
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const int y = __shfl_up(x[c], o);
            if (lane >= o) x[c] += y;
        }
    }
    if (lane == 63)
#pragma unroll
        for (int c = 0; c < 3; c++) wsum3[c][wave] = x[c];
    __syncthreads();
    if (wave == 0) {
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const int w = lane < NW ? wsum3[c][lane] : 0;
            int z = w;
#pragma unroll
            for (int o = 1; o < NW; o <<= 1) {
                const int y = __shfl_up(z, o);
                if (lane >= o) z += y;
            }
            if (lane < NW) wpre3[c][lane] = z - w;
            if (lane == NW - 1) wpre3[c][NW] = z;
        }
    }
    __syncthreads();
#pragma unroll
    for (int c = 0; c < 3; c++) {
        ex[c] = wpre3[c][wave] + x[c] - v[c];
        tot[c] = wpre3[c][NW];
    }
    __syncthreads();  // the shared arrays are free again
}

// Adds +1 to pcnt[p][inc] and -1 to pcnt[p][dec] (-1 = none) for every active lane, with one
// atomic per distinct (partition, layer) of the wavefront.  All lanes must call it.
__device__ __forceinline__ void agg_pcnt(int* pcnt, int NL, int p, int dec, int inc) {
    const int lane = __lane_id();
    unsigned long long todo = __ballot(p >= 0);
    while (todo) {
        const int leader = __ffsll((long long)todo) - 1;
        const int p0 = __shfl(p, leader);
        const bool mine = p >= 0 && p == p0;
        const unsigned long long mm = __ballot(mine);
        for (int k = 0; k < NL; k++) {
            const int d = __popcll(__ballot(mine && inc == k)) - __popcll(__ballot(mine && dec == k));
            if (lane == leader && d != 0) atomicAdd(&pcnt[p0 * NL + k], d);
        }
        todo &= ~mm;
    }
}

// ---- step and round control -----------------------------------------------------------------

// ---- brick reservation ------------------------------------------------------------------
// Every voxel that expands the band in a step (Surface.hpp:1249-1355) was tracked when the step
// started: the first round processes the band (or, in a bounded step, the affected set), and
// each later round only the voxels whose status the previous round changed; voxels the
// expansion adds are tracked directly and are never processed again in the step.  So the
// partitions a step can activate are inactive partitions face-adjacent to a partition holding
// band voxels at the start (bounded step: holding affected voxels), and a partition activated in
// the step is never deactivated in it; reactivating one the step deactivated reuses the brick it
// freed.  Counting those partitions before anything is written lets a step that might exhaust the
// brick pool be refused with the surface unchanged.

// Inactive partitions with a face neighbour marked in `src` (nullptr: any active partition).
template <int D>
__global__ void k_reserve_count(Geo G, Pool P, const unsigned* src, Ctr* c) {
    int n = 0;
    for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < G.np; p += gridDim.x * blockDim.x) {
        if (P.page[p] >= 0) continue;
        const Mem& M = P.mem[P.part_mem[p]];
        int pc[3] = {0, 0, 0};
        part_coords<D>(M, p - M.pbase, pc);
        bool near = false;
        for (int a = 0; a < D && !near; a++)
            for (int sd = -1; sd <= 1 && !near; sd += 2) {
                const int q = pc[a] + sd;
                if (q < 0 || q >= M.csize[a]) continue;
                const int pn = p + sd * M.cstr[a];
                near = src ? src[pn] != 0u : P.page[pn] >= 0;
            }
        n += near ? 1 : 0;
    }
    for (int o = 32; o > 0; o >>= 1) n += __shfl_down(n, o);
    if ((threadIdx.x & 63) == 0 && n) atomicAdd(&c->reserve, n);
}

// The same over the partitions of members ids[blockIdx.y] only.
template <int D>
__global__ void k_reserve_count_members(Geo G, Pool P, const unsigned* src, const int* ids,
                                        Ctr* c) {
    const Mem& M = P.mem[ids[blockIdx.y]];
    int n = 0;
    for (int lp = blockIdx.x * blockDim.x + threadIdx.x; lp < M.np; lp += gridDim.x * blockDim.x) {
        const int p = M.pbase + lp;
        if (P.page[p] >= 0) continue;
        int pc[3] = {0, 0, 0};
        part_coords<D>(M, lp, pc);
        bool near = false;
        for (int a = 0; a < D && !near; a++)
            for (int sd = -1; sd <= 1 && !near; sd += 2) {
                const int q = pc[a] + sd;
                if (q < 0 || q >= M.csize[a]) continue;
                near = src[p + sd * M.cstr[a]] != 0u;
            }
        n += near ? 1 : 0;
    }
    for (int o = 32; o > 0; o >>= 1) n += __shfl_down(n, o);
    if ((threadIdx.x & 63) == 0 && n) atomicAdd(&c->reserve, n);
}

// The count for a global step (any active partition marks its inactive face neighbours), from
// the other side: each active brick's face-neighbour record names its inactive neighbour
// partitions, which are stamped with the step's stamp and counted by the brick that stamps each
// first.  A partition past the grid's end (the +1 rule) is recorded as outside and never
// counted: it holds no grid voxel, so no step can activate it.
template <int D>
__global__ void k_reserve_count_active(Geo G, Pool P, int cap, unsigned* stamp, Ctr* c) {
    const unsigned st = c->stamp;
    int n = 0;
    for (int s = blockIdx.x * blockDim.x + threadIdx.x; s < cap; s += gridDim.x * blockDim.x) {
        if (P.slot_part[s] < 0) continue;
        const int4 ra = P.nbr[2 * (size_t)s], rb = P.nbr[2 * (size_t)s + 1];
        const int rec[6] = {ra.x, ra.y, ra.z, ra.w, rb.x, rb.y};
#pragma unroll
        for (int j = 0; j < 2 * D; j++)
            if (rec[j] <= kNbrInactive) n += atomicExch(&stamp[kNbrInactive - rec[j]], st) != st;
    }
    for (int o = 32; o > 0; o >>= 1) n += __shfl_down(n, o);
    if ((threadIdx.x & 63) == 0 && n) atomicAdd(&c->reserve, n);
}

// A global step whose reserve exceeds the free bricks does nothing: its lists are emptied and
// its rounds switched off before any value is written.
__global__ void k_reserve_decide(Ctr* c, int* bbase, int NL) {
    if (c->reserve <= c->free_top) return;
    c->err = 3;
    c->one = 0;
    for (int r = 0; r < 8; r++) c->live[r] = 0;
    for (int k = 0; k <= NL; k++) bbase[k] = 0;
    c->band_span[0] = c->band_span[1] = 0;
}

// Start of a step: counters of the previous step cleared, first round live.
__device__ __forceinline__ void step_reset(Ctr* c) {
    // a fresh stamp per step for k_reserve_count_active (0 is the stamps' initial value)
    c->stamp = c->stamp == 0xffffffffu ? 1u : c->stamp + 1u;
    c->sc_n = 0;
    c->flag = 0;
    c->lu_span[0][0] = c->lu_span[0][1] = 0;
    c->lu_span[1][0] = c->lu_span[1][1] = 0;
    c->z_span[0] = c->z_span[1] = 0;
    c->band_span[0] = c->band_span[1] = 0;
    c->need_n = 0;
    c->act_n = 0;
    c->one = 1;
    c->live[0] = 1;
    for (int r = 1; r < 8; r++) c->live[r] = 0;
    c->reserve = 0;
}

// update_start in one launch: k_step_reset's work by the first thread, and the partitions' delta
// and status-change masks cleared.
__global__ void k_step_start(Ctr* c, unsigned* dmask, unsigned* smask, int np) {
    if (blockIdx.x == 0 && threadIdx.x == 0) step_reset(c);
    for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < np; p += gridDim.x * blockDim.x) {
        dmask[p] = 0u;
        smask[p] = 0u;
    }
}

__global__ void k_step_reset(Ctr* c) { step_reset(c); }

// Round r runs only if round r-1 ran and changed an outer layer (Surface.hpp:752-760); its
// status-change state starts empty (m_grid_status_change.reset).
__global__ void k_round_gate(Ctr* c, int r) {
    const int live = c->live[r - 1] && c->flag;
    c->live[r] = live;
    if (live) {
        c->flag = 0;
        c->sc_n = 0;
        c->need_n = 0;
    }
}

__global__ void k_fill_u32_live(unsigned* p, int n, unsigned v, const int* live) {
    if (!*live) return;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = v;
}

// Empties a lookup (plookup->reset, Surface.hpp:424, 758).
__global__ void k_lu_clear(Ctr* c, int lu, const int* live) {
    if (*live) c->lu_span[lu][1] = 0;
}

// Between unrolled rounds r and r + 1: round r's lookup lu emptied (when round r ran), then
// round r + 1's gate.
__global__ void k_round_next(Ctr* c, int lu, int r) {
    if (c->live[r]) c->lu_span[lu][1] = 0;
    const int live = c->live[r] && c->flag;
    c->live[r + 1] = live;
    if (live) {
        c->flag = 0;
        c->sc_n = 0;
        c->need_n = 0;
    }
}

// After the last unrolled round r: its lookup lu emptied (when it ran), and the band total
// cleared for k_band_total2.
__global__ void k_round_end(Ctr* c, int lu, int r) {
    if (c->live[r]) c->lu_span[lu][1] = 0;
    c->band_total = 0;
}


// ---- band compaction ---------------------------------------------------------------------
// The band list is layer-major: segment k holds the layer k-L voxels, partitions in index
// order within a segment.  Partition p holds pcnt[p*NL + k] voxels of layer k, so its offset
// in segment k is the segment start + the layer-k counts of the partitions before it.  The scan
// runs partition-major, one thread per partition and PB partitions per block, over NL+1
// columns: the NL layer counts and a flag for partitions holding any band voxel, whose scan
// places them in the list of partitions the scatter visits.

constexpr int PB = 256;  // partitions per scan block

// Column c (c < NL: layer count, c == NL: the band flag) of partition p.
__device__ __forceinline__ int bscan_col(const int* pcnt, int np, int NL, int p, int c,
                                         int flag) {
    if (p >= np) return 0;
    return c < NL ? pcnt[p * NL + c] : (c == NL ? flag : 0);
}

__device__ __forceinline__ int bscan_flag(const int* pcnt, int np, int NL, int p) {
    if (p >= np) return 0;
    int any = 0;
    for (int k = 0; k < NL; k++) any |= pcnt[p * NL + k];
    return any != 0;
}

// Exclusive scan over the block of 8 columns packed as 32-bit pairs (v[i]: columns 2i, 2i+1);
// tot gets the block totals.
__device__ __forceinline__ void bscan_block8(unsigned long long v[4], unsigned long long tot[4]) {
    __shared__ unsigned long long wsum[PB / 64][4];
    const int lane = __lane_id(), w = threadIdx.x >> 6;
    unsigned long long x[4];
#pragma unroll
    for (int i = 0; i < 4; i++) x[i] = v[i];
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const unsigned long long y = __shfl_up(x[i], o);
            if (lane >= o) x[i] += y;
        }
    }
    if (lane == 63) {
#pragma unroll
        for (int i = 0; i < 4; i++) wsum[w][i] = x[i];
    }
    __syncthreads();
    unsigned long long pre[4] = {0, 0, 0, 0}, t[4] = {0, 0, 0, 0};
#pragma unroll
    for (int k = 0; k < PB / 64; k++) {
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const unsigned long long a = wsum[k][i];
            if (k < w) pre[i] += a;
            t[i] += a;
        }
    }
#pragma unroll
    for (int i = 0; i < 4; i++) {
        v[i] = x[i] - v[i] + pre[i];
        tot[i] = t[i];
    }
    __syncthreads();  // wsum is free again
}

__device__ __forceinline__ int col32(const unsigned long long v[4], int j) {
    return (int)(uint32_t)(v[j >> 1] >> (32 * (j & 1)));
}

// Pass 1: per block the totals of its partitions' columns, bs[b*(NL+1) + c].
__global__ void __launch_bounds__(PB) k_bscan_blocks(const int* pcnt, int np, int NL, int* bs) {
    const int p = blockIdx.x * PB + threadIdx.x;
    const int flag = bscan_flag(pcnt, np, NL, p);
    for (int g = 0; g <= NL; g += 8) {
        unsigned long long v[4], tot[4];
#pragma unroll
        for (int i = 0; i < 4; i++)
            v[i] = (unsigned long long)(uint32_t)bscan_col(pcnt, np, NL, p, g + 2 * i, flag) |
                   ((unsigned long long)(uint32_t)bscan_col(pcnt, np, NL, p, g + 2 * i + 1, flag)
                    << 32);
        bscan_block8(v, tot);
        if (threadIdx.x < 8 && g + threadIdx.x <= NL)
            bs[(size_t)blockIdx.x * (NL + 1) + g + threadIdx.x] = col32(tot, threadIdx.x);
    }
}

// Pass 2 (single block): the layer columns scanned in layer-major order (segment k's blocks
// start where segment k-1's end), the flag column on its own.  base[k] is segment k's start,
// base[NL] the band size; the flag total is the number of partitions in the visit list.
__global__ void k_bscan_top(int* bs, int nb, int NL, int* base, Ctr* c) {
    __shared__ int carry;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (int k = 0; k <= NL; k++) {
        if (k == NL) {
            __syncthreads();
            if (threadIdx.x == 0) {
                base[NL] = carry;
                c->band_span[0] = 0;
                c->band_span[1] = carry;
                carry = 0;
            }
            __syncthreads();
        } else if (threadIdx.x == 0) {
            base[k] = carry;
        }
        for (int b0 = 0; b0 < nb; b0 += BT) {
            const int b = b0 + threadIdx.x;
            int tot;
            const size_t at = (size_t)b * (NL + 1) + k;
            const int ex = block_excl_scan(b < nb ? bs[at] : 0, &tot);
            if (b < nb) bs[at] = carry + ex;
            __syncthreads();
            if (threadIdx.x == 0) carry += tot;
            __syncthreads();
        }
    }
    if (threadIdx.x == 0) c->act_n = carry;
}

// Pass 3: each partition's offset in every segment, and the visit list.
__global__ void __launch_bounds__(PB) k_bscan_apply(const int* pcnt, int np, int NL,
                                                    const int* bs, int* offsets, int* visit) {
    const int p = blockIdx.x * PB + threadIdx.x;
    const int flag = bscan_flag(pcnt, np, NL, p);
    const int* bb = bs + (size_t)blockIdx.x * (NL + 1);
    for (int g = 0; g <= NL; g += 8) {
        unsigned long long v[4], tot[4];
#pragma unroll
        for (int i = 0; i < 4; i++)
            v[i] = (unsigned long long)(uint32_t)bscan_col(pcnt, np, NL, p, g + 2 * i, flag) |
                   ((unsigned long long)(uint32_t)bscan_col(pcnt, np, NL, p, g + 2 * i + 1, flag)
                    << 32);
        bscan_block8(v, tot);
        if (p >= np) continue;
#pragma unroll
        for (int j = 0; j < 8; j++) {
            const int c = g + j;
            if (c < NL) offsets[p * NL + c] = bb[c] + col32(v, j);
            else if (c == NL && flag) visit[bb[NL] + col32(v, j)] = p;
        }
    }
}

// Pass 4: band voxels written into their layer segments, the entries of each partition in its
// brick's storage order.  A block holds kScatterBlock wavefronts; a partition gets wpp of them
// (1, 2, 4 or 8, fewer for small bricks so a block serves several partitions).  In each round a
// wavefront reads one span of 64*VPL voxels of its brick, VPL consecutive layer bytes per lane
// (VPL = 64 for bricks of more than 512 voxels, else 8), counts
// each lane's voxels per layer in 16-bit fields (layers 0-3 and 4-7 in two 64-bit words) and
// scans the counts over the wave; the wavefronts of a partition then exchange their span totals
// through LDS.  A span's entries go to segment start (from the scan of partition counts) + the
// totals of earlier rounds and earlier spans + their position within the span; they are first
// ordered by layer in LDS so that the stores to each segment are contiguous across lanes.
// More than 8 layers take a per-layer ballot path in the partition's first wavefront.  A
// partition whose layers and counts disagree writes no entry beyond its counts and is reported
// through err.
constexpr int kScatterBlock = 8;  // wavefronts per block

// The four bytes of x as the four 16-bit fields of a 64-bit word.
__device__ __forceinline__ unsigned long long spread16(uint32_t x) {
    return (unsigned long long)(x & 0xffu) | ((unsigned long long)(x & 0xff00u) << 8) |
           ((unsigned long long)(x & 0xff0000u) << 16) |
           ((unsigned long long)(x & 0xff000000u) << 24);
}

__device__ __forceinline__ int field16(unsigned long long lo, unsigned long long hi, int f) {
    return (int)(((f < 4 ? lo : hi) >> (16 * (f & 3))) & 0xffffull);
}

// The spans of partition p (an entry of the visit list, or -1) for the wavefronts wv - wi ..
// wv - wi + wpp - 1 of the block; rr counts the block's rounds, whose parity picks the span_tot
// buffer.  Returns whether this lane found counts and layers disagreeing.
template <int VPL>
__device__ __forceinline__ bool scatter_partition(const Geo& G, const Pool& P, int p, int wpp,
                                                  int wi, int wv, int lane, int& rr,
                                                  const int* offsets, uint32_t* band,
                                                  int8_t* tag,
                                                  unsigned long long (*span_tot)[kScatterBlock][2],
                                                  uint16_t* stage) {
    const int s = p >= 0 ? P.page[p] : -1;
    const bool act = s >= 0;
    const int NL = G.NL;
    // lane k < NL: next free entry and end of layer k's entries of this partition
    int nxt = 0, end = 0;
    if (act && lane < NL) {
        nxt = offsets[p * NL + lane];
        end = nxt + P.pcnt[p * NL + lane];
    }
    const uint32_t v0 = act ? (uint32_t)s * (uint32_t)G.pv : 0u;
    const int8_t* lay = P.lay + (size_t)v0;
    bool bad = false;
    if (NL > 8) {
        if (act && wi == 0) {
            for (int i0 = 0; i0 < G.pv; i0 += 64) {
                const int i = i0 + lane;
                const int f = i < G.pv && lay[i] != NONE ? lay[i] + G.L : -1;
                for (int k = 0; k < NL; k++) {
                    const unsigned long long m = __ballot(f == k);
                    if (!m) continue;
                    const int b = __shfl(nxt, k), e = __shfl(end, k);
                    if (f == k) {
                        const int at = b + __popcll(m & ((1ull << lane) - 1ull));
                        if (at < e) {
                            band[at] = v0 + (uint32_t)i;
                            tag[at] = (int8_t)(k - G.L);
                        } else {
                            bad = true;
                        }
                    }
                    if (lane == k) nxt += __popcll(m);
                }
            }
        }
    } else {
        constexpr int SPAN = 64 * VPL;        // voxels a wavefront reads per round
        constexpr int CB = VPL < 16 ? VPL : 16;  // bytes per load
        constexpr int NW = VPL / 4;              // words of layer bytes per lane
        const int rounds = (G.pv + SPAN * wpp - 1) / (SPAN * wpp);
        for (int r = 0; r < rounds; r++, rr++) {
            // this lane's VPL consecutive voxels of the wavefront's span
            const int s0 = (r * wpp + wi) * SPAN;
            const int i0 = s0 + VPL * lane;
            uint32_t q[NW];
#pragma unroll
            for (int c = 0; c < VPL / CB; c++) {
                const int ic = i0 + CB * c;
                uint32_t* qc = q + c * (CB / 4);
#pragma unroll
                for (int k = 0; k < CB / 4; k++) qc[k] = 0x80808080u;
                if (!act || ic >= G.pv) continue;
                if ((G.pv & (CB - 1)) == 0) {
                    if constexpr (CB == 16) {
                        const uint4 w = *reinterpret_cast<const uint4*>(lay + ic);
                        qc[0] = w.x;
                        qc[1] = w.y;
                        qc[2] = w.z;
                        qc[3] = w.w;
                    } else {
                        const uint2 w = *reinterpret_cast<const uint2*>(lay + ic);
                        qc[0] = w.x;
                        qc[1] = w.y;
                    }
                } else {
                    for (int j = 0; j < CB; j++) {
                        const uint32_t byte =
                            ic + j < G.pv ? (uint32_t)(uint8_t)lay[ic + j] : 0x80u;
                        qc[j >> 2] = (qc[j >> 2] & ~(0xffu << (8 * (j & 3)))) |
                                     (byte << (8 * (j & 3)));
                    }
                }
            }
            // Each layer byte becomes its field f = k + L (0..NL-1, NL <= 7), and a voxel outside
            // the band field 7, four bytes at a time: adding L byte-wise takes the band's bytes
            // to 0..2L and NONE (0x80) to 0x80 + L, whose top bit marks it.  q then holds 8f per
            // byte.  The counts go into 8-bit fields (at most VPL <= 64 per field) with no
            // branch, then into 16-bit fields for the scan (layers 0-3 and 4-7 in two words).
            const uint32_t Lrep = (uint32_t)G.L * 0x01010101u;
#pragma unroll
            for (int w = 0; w < NW; w++) {
                const uint32_t x = q[w];
                const uint32_t u = ((x & 0x7f7f7f7fu) + Lrep) ^ (x & 0x80808080u);
                const uint32_t out = (u >> 7) & 0x01010101u;
                q[w] = ((u & 0x07070707u) | ((out << 3) - out)) << 3;
            }
            unsigned long long c8 = 0;
#pragma unroll
            for (int j = 0; j < VPL; j++) c8 += 1ull << ((q[j >> 2] >> (8 * (j & 3))) & 0xffu);
            const unsigned long long n0 = spread16((uint32_t)c8),
                                     n1 = spread16((uint32_t)(c8 >> 32) & 0x00ffffffu);
            unsigned long long x0 = n0, x1 = n1;
#pragma unroll
            for (int o = 1; o < 64; o <<= 1) {
                const unsigned long long y0 = __shfl_up(x0, o), y1 = __shfl_up(x1, o);
                if (lane >= o) {
                    x0 += y0;
                    x1 += y1;
                }
            }
            const unsigned long long w0 = __shfl(x0, 63), w1 = __shfl(x1, 63);
            // the partition's spans of this round: totals before this wavefront's and in all
            unsigned long long pre0 = 0, pre1 = 0, all0 = w0, all1 = w1;
            if (wpp > 1) {
                if (lane == 0) {
                    span_tot[rr & 1][wv][0] = w0;
                    span_tot[rr & 1][wv][1] = w1;
                }
                __syncthreads();
                all0 = all1 = 0;
                for (int k = 0; k < wpp; k++) {
                    const unsigned long long a0 = span_tot[rr & 1][wv - wi + k][0];
                    const unsigned long long a1 = span_tot[rr & 1][wv - wi + k][1];
                    if (k < wi) {
                        pre0 += a0;
                        pre1 += a1;
                    }
                    all0 += a0;
                    all1 += a1;
                }
            }
            // The round's entries of layer f fill [nxt_f, nxt_f + all_f): one bound check per
            // layer, and a partition whose counts disagree with its layers writes nothing.
            const bool skip = __ballot(lane < NL && nxt + field16(all0, all1, lane) > end) != 0ull;
            bad = bad || skip;
            // The span's entries ordered by layer, then storage order: layer f starts at
            // st_f = sum of the span's counts of layers below f, packed like the counts, so
            // entry = field f of (entries before this lane's + st) + running count.
            const unsigned long long rep = 0x0001000100010001ull;
            const unsigned long long lo_tot =
                ((w0 & 0xffffull) + ((w0 >> 16) & 0xffffull) + ((w0 >> 32) & 0xffffull) +
                 (w0 >> 48));
            const unsigned long long st0 = (w0 << 16) + (w0 << 32) + (w0 << 48);
            const unsigned long long st1 = lo_tot * rep + (w1 << 16) + (w1 << 32) + (w1 << 48);
            const int total = (int)(lo_tot + (w1 & 0xffffull) + ((w1 >> 16) & 0xffffull) +
                                    ((w1 >> 32) & 0xffffull) + (w1 >> 48));
            // the field bytes are read again below rather than kept unpacked across the scan
#pragma unroll
            for (int w = 0; w < NW; w++) asm volatile("" : "+v"(q[w]));
            // r0, r1: each field's next position in the stage.  Field 7 (outside the band)
            // starts at total, so those voxels are written past the span's entries, into the
            // stage's VPL spare slots, with no branch.
            unsigned long long r0 = x0 - n0 + st0, r1 = x1 - n1 + st1;
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
#pragma unroll
            for (int j = 0; j < VPL; j++) {
                const uint32_t sh = ((q[j >> 2] >> (8 * (j & 3))) & 0xffu) << 1;  // 16 f
                const bool hi = sh >= 64u;
                const unsigned long long r = hi ? r1 : r0;
                const int e = (int)((r >> (sh & 63u)) & 0xffffull);
                const unsigned long long rn = r + (1ull << (sh & 63u));
                r0 = hi ? r0 : rn;
                r1 = hi ? rn : r1;
                stage[e] = (uint16_t)(VPL * lane + j);
            }
            __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
            __builtin_amdgcn_wave_barrier();
            // copy out: consecutive lanes write consecutive entries of a layer segment
            const int first = nxt + field16(pre0, pre1, lane) - field16(st0, st1, lane);
            for (int e0 = 0; e0 < total; e0 += 64) {
                const int e = e0 + lane;
                int f = 0;
#pragma unroll
                for (int g = 1; g < 8; g++) f += (g < NL && field16(st0, st1, g) <= e) ? 1 : 0;
                const int at = __shfl(first, f) + e;
                if (e < total && !skip) {
                    band[at] = v0 + (uint32_t)(s0 + stage[e]);
                    tag[at] = (int8_t)(f - G.L);
                }
            }
            __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
            __builtin_amdgcn_wave_barrier();
            if (lane < NL) nxt += field16(all0, all1, lane);
        }
    }
    return bad;
}

template <int D, int VPL>
__global__ void __launch_bounds__(64 * kScatterBlock)
k_band_scatter(Geo G, Pool P, int wpp, const int* visit, const int* offsets, uint32_t* band,
               int8_t* tag, Ctr* ctr) {
    __shared__ unsigned long long span_tot[2][kScatterBlock][2];
    // a span's entries, staged, and VPL spare slots for the voxels outside the band
    __shared__ uint16_t span_voxel[kScatterBlock][64 * VPL + VPL];
    const int lane = __lane_id();
    const int wv = threadIdx.x >> 6;
    const int wi = wv & (wpp - 1);
    const int ppb = kScatterBlock / wpp;
    const int nvisit = ctr->act_n;
    const int ntask = (nvisit + ppb - 1) / ppb;
    int rr = 0;
    bool bad = false;
    for (int t = blockIdx.x; t < ntask; t += gridDim.x) {
        const int i = t * ppb + wv / wpp;
        const int p = i < nvisit ? visit[i] : -1;
        bad = scatter_partition<VPL>(G, P, p, wpp, wi, wv, lane, rr, offsets, band, tag,
                                     span_tot, span_voxel[wv]) || bad;
    }
    if (__ballot(bad) != 0ull && lane == 0) ctr->err = 2;
}

// ---- per-round processing --------------------------------------------------------------------

#define FELT_SPAN_LOOP(i, rng)                                                                 \
    for (int i = (rng)[0] + blockIdx.x * blockDim.x + threadIdx.x, i##_hi = (rng)[1]; i < i##_hi; \
         i += gridDim.x * blockDim.x)

// The same loop with the span cut into eight contiguous parts, one per XCD: the hardware deals
// workgroups to the eight XCDs in turn (block b runs on XCD b mod 8), so block b works in part
// b mod 8, and entries next to each other in the list -- the same brick, or neighbouring bricks
// -- are processed on one XCD and share its L2, instead of each brick's lines being fetched
// into several XCDs' L2s.  Blocks past the last multiple of eight have no part; a grid of fewer
// than eight blocks works on the whole span as one part.
#define FELT_XCD_SPAN_LOOP(i, rng)                                                             \
    for (int i##_k = gridDim.x >= 8u ? 8 : 1, i##_x = (int)blockIdx.x % i##_k,                 \
             i##_l = (int)blockIdx.x / i##_k, i##_per = (int)gridDim.x / i##_k,                \
             i##_n = (rng)[1] - (rng)[0],                                                      \
             i = (rng)[0] + (int)((long long)i##_n * i##_x / i##_k) + i##_l * (int)blockDim.x + \
                 (int)threadIdx.x,                                                            \
             i##_hi = (rng)[0] + (int)((long long)i##_n * (i##_x + 1) / i##_k);                \
         i##_l < i##_per && i < i##_hi; i += i##_per * (int)blockDim.x)

__global__ void k_tk_init(const int8_t* tag, int8_t* tk, const int* rng, const int* live) {
    if (!*live) return;
    FELT_SPAN_LOOP(i, rng) tk[i] = tag[i];
}

__device__ __forceinline__ int8_t clamp_k(int k) { return (int8_t)(k < -100 ? -100 : (k > 100 ? 100 : k)); }

template <int D>
__global__ void k_speed(Geo G, Pool P, const uint32_t* list, const int* rng, const Speed* sp,
                        float* out) {
    FELT_SPAN_LOOP(i, rng) out[i] = speed_voxel<D>(G, P, list[i], sp);
}

// update_zero_layer (Surface.hpp:920-979): phi += delta; tk = new layer.  tk is indexed like
// list, from the span's start.
__global__ void k_apply(Geo G, Pool P, const uint32_t* list, const float* delta, const int* rng,
                        int8_t* tk, int* nan_ctr) {
    FELT_SPAN_LOOP(i, rng) {
        const uint32_t v = list[i];
        const float nv = P.phi[v] + delta[i];
        P.phi[v] = nv;
        tk[i] = clamp_k(layer_of(nv, G.L, nan_ctr));
    }
}

// update_distance pass 1 (Surface.hpp:1029-1073) over the entries of one layer: d = the
// distance from the layer's neighbours.  Tagged lists hold every layer and are filtered by tag.
// The pass is bound by its gathers (entry -> brick record and own value -> neighbours in other
// bricks), not by arithmetic.
template <int D, bool Tagged>
__device__ __forceinline__ void dist1_body(const Geo& G, const Pool& P, const uint32_t* list,
                                           const int8_t* tag, const int* rng, int layer,
                                           float* out) {
    const float side = layer < 0 ? -1.0f : 1.0f;
    if (Tagged) {
        // a lookup's entries of one layer cluster where the previous round's list held their
        // source layers: the interleaved loop keeps every XCD busy
        FELT_SPAN_LOOP(i, rng) {
            if (tag[i] != layer) continue;  // another layer's entry
            out[i] = distance_rec<D>(G, P, list[i], side);
        }
    } else {
        FELT_XCD_SPAN_LOOP(i, rng) out[i] = distance_rec<D>(G, P, list[i], side);
    }
}

// update_distance pass 2 (Surface.hpp:1076-1121): phi = d; tk = new layer.
template <bool Tagged>
__device__ __forceinline__ void dist2_body(const Geo& G, const Pool& P, const uint32_t* list,
                                           const int8_t* tag, const int* rng, int layer,
                                           const float* d, int8_t* tk, int* nan_ctr) {
    if (Tagged) {
        FELT_SPAN_LOOP(i, rng) {
            if (tag[i] != layer) continue;
            const uint32_t v = list[i];
            const float x = d[i];
            P.phi[v] = x;
            tk[i] = clamp_k(layer_of(x, G.L, nan_ctr));
        }
    } else {
        FELT_XCD_SPAN_LOOP(i, rng) {
            const float x = d[i];
            P.phi[list[i]] = x;
            tk[i] = clamp_k(layer_of(x, G.L, nan_ctr));
        }
    }
}

// First round of a global step: the span is the layer's segment of the band list.
template <int D>
__global__ void k_dist1_band(Geo G, Pool P, const uint32_t* list, const int* rng, int layer,
                             float* out, const int* live) {
    if (!*live) return;
    dist1_body<D, false>(G, P, list, nullptr, rng, layer, out);
}

__global__ void k_dist2_band(Geo G, Pool P, const uint32_t* list, const int* rng, int layer,
                             const float* d, int8_t* tk, int* nan_ctr, const int* live) {
    if (!*live) return;
    dist2_body<false>(G, P, list, nullptr, rng, layer, d, tk, nan_ctr);
}

// Later rounds and bounded updates: a lookup holding every layer, filtered by tag.
template <int D>
__global__ void k_dist1_lookup(Geo G, Pool P, const uint32_t* list, const int8_t* tag,
                               const int* rng, int layer, float* out, const int* live) {
    if (!*live) return;
    dist1_body<D, true>(G, P, list, tag, rng, layer, out);
}

// The first distance pass of a lookup round with the round's preparation in the same launch: tk =
// tag over the lookup's span, and the first nmask words of smask cleared (its status-change
// partitions start empty).  The pass reads neither tk nor smask, so the two need no order.
template <int D>
__global__ void k_dist1_lookup_prep(Geo G, Pool P, const uint32_t* list, const int8_t* tag,
                                    const int* rng, int layer, float* out, int8_t* tk,
                                    unsigned* smask, int nmask, const int* live) {
    if (!*live) return;
    const int lo = rng[0], n = max(rng[1] - lo, nmask);
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        if (lo + i < rng[1]) tk[lo + i] = tag[lo + i];
        if (i < nmask) smask[i] = 0u;
    }
    dist1_body<D, true>(G, P, list, tag, rng, layer, out);
}

__global__ void k_dist2_lookup(Geo G, Pool P, const uint32_t* list, const int8_t* tag,
                               const int* rng, int layer, const float* d, int8_t* tk,
                               int* nan_ctr, const int* live) {
    if (!*live) return;
    dist2_body<true>(G, P, list, tag, rng, layer, d, tk, nan_ctr);
}

__device__ __forceinline__ bool expanding(const Geo& G, int from, int to) {
    const int side = from < 0 ? -1 : 1;
    return (from == G.L || from == -G.L) && to == G.L * side - side;
}

// Change compaction, pass 1: the list is cut into tiles of BT * ITEMS entries, each thread taking
// ITEMS consecutive entries (one 16-byte load of the old and of the new layer bytes); per tile the
// number of status changes, of in-band targets and of expansions.  Raises the round flag for a
// change out of a non-zero layer (Surface.hpp:1116).  Blocks stride over the tiles the span
// needs, so a short span costs few blocks whatever the grid.
constexpr int kTile = BT * ITEMS;

// The layer bytes (old, new) of a thread's ITEMS entries starting at i0; entries past n read as
// unchanged.  The lists are allocation bases (256-byte aligned), so whole groups load as vectors.
__device__ __forceinline__ void chg_load(const int8_t* tag, const int8_t* tk, int i0, int n,
                                         uint32_t a[ITEMS / 4], uint32_t b[ITEMS / 4]) {
    static_assert(ITEMS == 16, "one 16-byte load per array");
    if (i0 + ITEMS <= n) {
        const uint4 x = *reinterpret_cast<const uint4*>(tag + i0);
        const uint4 y = *reinterpret_cast<const uint4*>(tk + i0);
        a[0] = x.x, a[1] = x.y, a[2] = x.z, a[3] = x.w;
        b[0] = y.x, b[1] = y.y, b[2] = y.z, b[3] = y.w;
    } else {
#pragma unroll
        for (int w = 0; w < ITEMS / 4; w++) a[w] = b[w] = 0u;
        for (int j = 0; j < ITEMS && i0 + j < n; j++) {
            const int sh = 8 * (j & 3);
            a[j >> 2] |= (uint32_t)(uint8_t)tag[i0 + j] << sh;
            b[j >> 2] |= (uint32_t)(uint8_t)tk[i0 + j] << sh;
        }
    }
}

__device__ __forceinline__ int byte_of(const uint32_t w[], int j) {
    return (int)(int8_t)(w[j >> 2] >> (8 * (j & 3)));
}

// Per-thread counts of the status changes, in-band targets and expansions among its entries.
__device__ __forceinline__ void chg_counts(const Geo& G, const uint32_t a[], const uint32_t b[],
                                           int cnt[3], bool* fl) {
    cnt[0] = cnt[1] = cnt[2] = 0;
#pragma unroll
    for (int w = 0; w < ITEMS / 4; w++) {
        if (a[w] == b[w]) continue;
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const int x = byte_of(a, 4 * w + q), y = byte_of(b, 4 * w + q);
            if (x == y) continue;
            cnt[0]++;
            if (y >= -G.L && y <= G.L) cnt[1]++;
            if (x != 0) *fl = true;
            if (expanding(G, x, y)) cnt[2]++;
        }
    }
}

// Pass 2 (single block): exclusive scan of the per-tile counts; the totals become the round's
// status-change count and the buffer lookup's span.
__device__ __forceinline__ void chg_scan_block(int* counts, const int* rng, Ctr* ctr, int buf) {
    const int nb = (rng[1] + kTile - 1) / kTile;
    int carry[3] = {0, 0, 0};
    for (int b0 = 0; b0 < nb; b0 += BT) {
        const int b = b0 + threadIdx.x;
        int v[3], e[3], t[3];
#pragma unroll
        for (int c = 0; c < 3; c++) v[c] = b < nb ? counts[3 * b + c] : 0;
        block_excl_scan3(v, e, t);
        if (b < nb)
#pragma unroll
            for (int c = 0; c < 3; c++) counts[3 * b + c] = carry[c] + e[c];
#pragma unroll
        for (int c = 0; c < 3; c++) carry[c] += t[c];
    }
    if (threadIdx.x == 0) {
        ctr->sc_n = carry[0];
        ctr->lu_span[buf][0] = 0;
        ctr->lu_span[buf][1] = carry[1];
        ctr->exp_n = carry[2];
    }
}

__global__ void __launch_bounds__(BT) k_chg_count(Geo G, const int8_t* tag, const int8_t* tk,
                                                  const int* rng, int* counts, Ctr* ctr,
                                                  const int* live) {
    if (!*live) return;
    const int n = rng[1], ntiles = (n + kTile - 1) / kTile;
    for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const int i0 = t * kTile + threadIdx.x * ITEMS;
        uint32_t a[ITEMS / 4], b[ITEMS / 4];
        chg_load(tag, tk, i0, n, a, b);
        int cnt[3], ex[3], tot[3];
        bool fl = false;
        chg_counts(G, a, b, cnt, &fl);
        block_excl_scan3(cnt, ex, tot);
        if (__ballot(fl) != 0ull && fl) ctr->flag = 1;
        if (threadIdx.x == 0) {
            counts[3 * t] = tot[0];
            counts[3 * t + 1] = tot[1];
            counts[3 * t + 2] = tot[2];
        }
    }
}

// Pass 2 as a kernel of its own (a block that finished last could scan instead, but the
// device-scope fence each block would need before taking its ticket writes back the L2 of its
// XCD, which costs far more than a launch).
__global__ void __launch_bounds__(BT) k_chg_scan(int* counts, const int* rng, Ctr* ctr, int buf,
                                                 const int* live) {
    if (!*live) return;
    chg_scan_block(counts, rng, ctr, buf);
}

// flush_status_change (Surface.hpp:1183-1244) with Untrack::retrack / untrack
// (PartitionedMixin.hpp:489-558) of one status change per lane (p = -1: none) -- voxel v of
// partition p from layer index dec to layer `to`: its layer byte (and, for an untrack, its value
// Distance(to)), then per distinct partition of the wavefront one atomic per layer count that
// changes, the from-layers of its changes into smask (its status_change(k) lists), and for
// untracks the key of its new background (Distance(to) of the untrack that comes last in the
// reference's flush order: the highest from-layer) and a flag for the emptiness check.  Every
// lane of the wavefront calls it.
__device__ __forceinline__ void flush_change(const Geo& G, const Pool& P, uint32_t v, int p, int dec,
                                             int to, unsigned* smask, int* ukey, int* tflag) {
    int inc = -1, key = -1;
    if (p >= 0) {
        if (to >= -G.L && to <= G.L) {
            P.lay[v] = (int8_t)to;
            inc = to + G.L;
        } else {
            P.lay[v] = NONE;
            P.phi[v] = (float)to;
            key = (dec << 16) | ((to + 32768) & 0xffff);
        }
    }
    const int lane = __lane_id();
    unsigned long long todo = __ballot(p >= 0);
    while (todo) {
        const int leader = __ffsll((long long)todo) - 1;
        const int p0 = __shfl(p, leader);
        const bool mine = p >= 0 && p == p0;
        // only the layers the partition's records touch (usually two or three)
        unsigned lm = mine ? ((1u << dec) | (inc >= 0 ? 1u << inc : 0u)) : 0u;
        for (int o = 32; o > 0; o >>= 1) lm |= __shfl_xor(lm, o);
        unsigned from_bits = 0;
        while (lm) {
            const int k = __ffs(lm) - 1;
            lm &= lm - 1u;
            const unsigned long long md = __ballot(mine && dec == k);
            const int d = __popcll(__ballot(mine && inc == k)) - __popcll(md);
            if (md) from_bits |= 1u << k;
            if (lane == leader && d != 0) atomicAdd(&P.pcnt[p0 * G.NL + k], d);
        }
        if (lane == leader) atomicOr(&smask[p0], from_bits);
        if (__ballot(mine && key >= 0)) {
            int kmax = mine ? key : -1;
            for (int o = 32; o > 0; o >>= 1) kmax = max(kmax, __shfl_xor(kmax, o));
            if (lane == leader) {
                atomicMax(&ukey[p0], kmax);
                tflag[p0] = 1;
            }
        }
        todo &= ~__ballot(mine);
    }
}

// Pass 3: the status changes (status_change, Surface.hpp:1138-1178) flushed, and the buffer
// entries of the in-band targets and the expansion records, in list order.  Each lane stages the
// wavefront-local indices of its changed entries in LDS, in list order; the wavefront then reads
// them back 64 at a time (the entries' values come again from the cache), so the flush's layer
// byte stores and the buffer's entries keep list order and every store instruction of the
// buffer writes one contiguous run.
__global__ void __launch_bounds__(BT) k_chg_scatter(Geo G, Pool P, const uint32_t* lv,
                                                    const int8_t* tag, const int8_t* tk,
                                                    const int* rng, const int* offsets, SC sc,
                                                    Lu buf, unsigned* smask, int* ukey,
                                                    int* tflag, const int* live) {
    if (!*live) return;
    __shared__ uint16_t stage[BT / 64][64 * ITEMS];
    const int lane = __lane_id(), wv = threadIdx.x >> 6;
    const int n = rng[1], ntiles = (n + kTile - 1) / kTile;
    for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const int i0 = t * kTile + threadIdx.x * ITEMS;
        uint32_t a[ITEMS / 4], b[ITEMS / 4];
        chg_load(tag, tk, i0, n, a, b);
        int cnt[3], ex[3], tot[3];
        bool fl = false;
        chg_counts(G, a, b, cnt, &fl);
        block_excl_scan3(cnt, ex, tot);
        // the wavefront's first record of each kind, and its number of status changes
        const int w0 = __shfl(ex[0], 0);
        const int nw = __shfl(ex[0] + cnt[0], 63) - w0;
        if (nw == 0) continue;
        int obf = __shfl(ex[1], 0) + offsets[3 * t + 1], oex = __shfl(ex[2], 0) + offsets[3 * t + 2];
        int at = ex[0] - w0;
#pragma unroll
        for (int w = 0; w < ITEMS / 4; w++) {
            if (a[w] == b[w]) continue;
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const int j = 4 * w + q;
                if (byte_of(a, j) != byte_of(b, j)) stage[wv][at++] = (uint16_t)(lane * ITEMS + j);
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        const int e0 = t * kTile + wv * 64 * ITEMS;  // the wavefront's first entry
        for (int k0 = 0; k0 < nw; k0 += 64) {
            const int k = k0 + lane;
            const bool ok = k < nw;
            const int e = ok ? e0 + stage[wv][k] : e0;
            const uint32_t v = lv[e];
            const int x = tag[e], y = tk[e];
            const bool inb = ok && y >= -G.L && y <= G.L;
            const bool xp = ok && expanding(G, x, y);
            flush_change(G, P, v, ok ? P.slot_part[fdiv(v, G.fpv)] : -1, x + G.L, y, smask, ukey,
                         tflag);
            const unsigned long long m1 = __ballot(inb), m2 = __ballot(xp);
            if (inb) {
                buf.v[obf + lanes_below(m1)] = v;
                buf.tag[obf + lanes_below(m1)] = (int8_t)y;
            }
            if (xp) {
                sc.xv[oex + lanes_below(m2)] = v;
                sc.xside[oex + lanes_below(m2)] = x < 0 ? -1 : 1;
            }
            obf += __popcll(m1);
            oex += __popcll(m2);
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
    }
}

// Loops over a device-sized list with a grid stride, keeping every wavefront's lanes together
// (the wave-wide helpers need all lanes).
#define FELT_WAVE_LOOP(i, valid, n)                                                            \
    for (int i##_b = blockIdx.x * blockDim.x, i##_n = (n); i##_b < i##_n;                      \
         i##_b += gridDim.x * blockDim.x)                                                      \
        for (int i = i##_b + (int)threadIdx.x, valid = i < i##_n, i##_once = 1; i##_once;     \
             i##_once = 0)

// Partitions flagged by the flush: a partition left with no voxel is deactivated with the
// background of its key and its brick returned to the pool.  Also starts the expansion's
// activation list empty.
__global__ void k_deactivate(Geo G, Pool P, int* ukey, int* tflag, int* free_stack, int* dlist,
                             Ctr* ctr, const int* live) {
    if (!*live) return;
    if (blockIdx.x == 0 && threadIdx.x == 0) ctr->need_n = 0;
    for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < G.np; p += gridDim.x * blockDim.x) {
        if (!tflag[p]) continue;
        tflag[p] = 0;
        const int key = ukey[p];
        ukey[p] = -1;
        int sum = 0;
        for (int k = 0; k < G.NL; k++) sum += P.pcnt[p * G.NL + k];
        const int s = P.page[p];
        if (sum == 0 && s >= 0) {
            P.page[p] = -1;
            P.slot_part[s] = -1;
            P.pbg[p] = (float)((key & 0xffff) - 32768);
            const int t = atomicAdd(&ctr->free_top, 1);
            free_stack[t] = s;
            dlist[atomicAdd(&ctr->dead_n, 1)] = p;
        }
    }
}


// expand_narrow_band (Surface.hpp:1249-1355), phase 1: queue inactive partitions that an
// expanding voxel's out-of-band neighbours fall in.  The neighbours' partitions, then their
// pages, then the backgrounds of the inactive ones are loaded level by level.
template <int D>
__global__ void k_expand_mark(Geo G, Pool P, SC sc, const Ctr* ctrc, int* nflag, int* nlist,
                              Ctr* ctr, const int* live) {
    if (!*live) return;
    constexpr int NN = 2 * D;
    FELT_WAVE_LOOP(i, valid, ctrc->exp_n) {
        const bool cand = valid;
        int pq[NN];
        bool q_new[NN];
#pragma unroll
        for (int j = 0; j < NN; j++) {
            pq[j] = -1;
            q_new[j] = false;
        }
        if (cand) {
            VoxCtx x;
            vox_ctx<D>(G, P, sc.xv[i], x);
#pragma unroll
            for (int a = 0; a < D; a++)
#pragma unroll
                for (int h = 0; h < 2; h++) {
                    const int sd = h ? 1 : -1, j = 2 * a + h;
                    const int qa = x.g[a] + sd, la = x.loc[a] + sd;
                    // only neighbours in another partition can be in an inactive one
                    if (qa >= 0 && qa < x.sz[a] && (la < 0 || la >= G.part[a]))
                        pq[j] = x.own + sd * x.cs[a];
                }
            int pg[NN];
#pragma unroll
            for (int j = 0; j < NN; j++) pg[j] = pq[j] >= 0 ? P.page[pq[j]] : 0;
            float bg[NN];
#pragma unroll
            for (int j = 0; j < NN; j++) bg[j] = (pq[j] >= 0 && pg[j] < 0) ? P.pbg[pq[j]] : 0.0f;
#pragma unroll
            for (int j = 0; j < NN; j++)
                if (pq[j] >= 0 && pg[j] < 0 &&
                    !in_band(layer_of(bg[j], G.L, &ctr->nan_events), G.L))
                    q_new[j] = atomicExch(&nflag[pq[j]], 1) == 0;
        }
#pragma unroll
        for (int j = 0; j < NN; j++) {
            const int at = wave_append(&ctr->need_n, q_new[j]);
            if (q_new[j]) nlist[at] = pq[j];
        }
    }
}

// Partition activation from the free-slot stack: the brick is filled with the partition's
// background (GridMixin.hpp:75-81) and no voxel is tracked.
__global__ void k_alloc(Geo G, Pool P, const int* nlist, const Ctr* ctrc, int* nflag,
                        int* free_stack, Ctr* ctr, const int* live) {
    if (!*live) return;
    __shared__ int slot_sh;
    const int n = ctrc->need_n;
    for (int j = blockIdx.x; j < n; j += gridDim.x) {
        const int p = nlist[j];
        if (threadIdx.x == 0) {
            const int t = atomicSub(&ctr->free_top, 1) - 1;
            int s = -1;
            if (t >= 0) s = free_stack[t];
            else {
                atomicAdd(&ctr->free_top, 1);
                ctr->err = 1;
            }
            slot_sh = s;
            if (s >= 0) {
                P.page[p] = s;
                P.slot_part[s] = p;
                const int m = P.part_mem[p];
                const Mem& M = P.mem[m];
                int pc[3] = {0, 0, 0};
                part_coords<3>(M, p - M.pbase, pc);
                bool edge = false;
                for (int a = 0; a < G.D; a++)
                    if (pc[a] == 0 || (pc[a] + 1) * G.part[a] >= M.size[a]) edge = true;
                P.slot_org[s] = make_int4(pc[0] * G.part[0], pc[1] * G.part[1],
                                          pc[2] * G.part[2], m | (edge ? kEdgeBrick : 0));
            }
            nflag[p] = 0;
        }
        __syncthreads();
        const int s = slot_sh;
        if (s >= 0) {
            const float bg = P.pbg[p];
            float* ph = P.phi + (size_t)s * G.pv;
            int8_t* ly = P.lay + (size_t)s * G.pv;
            for (int i = threadIdx.x; i < G.pv; i += blockDim.x) {
                ph[i] = bg;
                ly[i] = NONE;
            }
            if (threadIdx.x < G.NL) P.pcnt[p * G.NL + threadIdx.x] = 0;
        }
        __syncthreads();
    }
}

// Brings the face-neighbour records (felt_dev.h) of the listed partitions and of their active
// face neighbours in line with the page table: the first n_a partitions of la (activated) and
// the first n_b of lb (deactivated).  Any partition may be listed, in either list, any number
// of times: every record field written is a function of the page table, which no kernel changes
// while this one runs, so racing writers of a field write the same value.
template <int D>
__global__ void k_nbr_fix(Geo G, Pool P, const int* la, const int* n_a, const int* lb,
                          const int* n_b, const int* live) {
    if (!*live) return;
    const int na = *n_a, n = na + *n_b;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const int p = i < na ? la[i] : lb[i - na];
        const int m = P.part_mem[p];
        const Mem& M = P.mem[m];
        int pc[3] = {0, 0, 0};
        part_coords<D>(M, p - M.pbase, pc);
        const int s = P.page[p];
        const int self = s >= 0 ? s : kNbrInactive - p;
        int rec[6] = {kNbrOut, kNbrOut, kNbrOut, kNbrOut, kNbrOut, kNbrOut};
        int* words = reinterpret_cast<int*>(P.nbr);
#pragma unroll
        for (int a = 0; a < D; a++)
#pragma unroll
            for (int h = 0; h < 2; h++) {
                const int sd = h ? 1 : -1, j = 2 * a + h;
                const int c = pc[a] + sd;
                if (c < 0 || c >= M.csize[a]) continue;
                const int q = p + sd * M.cstr[a];
                const int t = P.page[q];
                // a partition of the children grid that starts past the grid's end (the +1
                // rule, PartitionedMixin.hpp:131-141) holds no grid position: kNbrOut
                if (c * G.part[a] < M.size[a]) rec[j] = t >= 0 ? t : kNbrInactive - q;
                // the neighbour's field that faces p
                if (t >= 0) words[8 * (size_t)t + (j ^ 1)] = self;
            }
        if (s >= 0) {
            int ext[3] = {1, 1, 1};
#pragma unroll
            for (int a = 0; a < D; a++)
                ext[a] = max(0, min(G.part[a], M.size[a] - pc[a] * G.part[a]));
            P.nbr[2 * (size_t)s] = make_int4(rec[0], rec[1], rec[2], rec[3]);
            P.nbr[2 * (size_t)s + 1] = make_int4(rec[4], rec[5], ext[0] | (ext[1] << 16), ext[2]);
        }
    }
}

// Counts the active bricks whose face-neighbour record differs from what the page table implies
// (a consistency check for tests: k_nbr_fix keeps every active brick's record in step).
template <int D>
__global__ void k_nbr_check(Geo G, Pool P, int cap, int* bad) {
    for (int s = blockIdx.x * blockDim.x + threadIdx.x; s < cap; s += gridDim.x * blockDim.x) {
        const int p = P.slot_part[s];
        if (p < 0) continue;
        const Mem& M = P.mem[P.part_mem[p]];
        int pc[3] = {0, 0, 0};
        part_coords<D>(M, p - M.pbase, pc);
        int want[8] = {kNbrOut, kNbrOut, kNbrOut, kNbrOut, kNbrOut, kNbrOut, 0, 1};
        for (int a = 0; a < D; a++)
            for (int h = 0; h < 2; h++) {
                const int sd = h ? 1 : -1, c = pc[a] + sd;
                if (c < 0 || c >= M.csize[a] || c * G.part[a] >= M.size[a]) continue;
                const int q = p + sd * M.cstr[a];
                const int t = P.page[q];
                want[2 * a + h] = t >= 0 ? t : kNbrInactive - q;
            }
        int ext[3] = {1, 1, 1};
        for (int a = 0; a < D; a++) ext[a] = max(0, min(G.part[a], M.size[a] - pc[a] * G.part[a]));
        want[6] = ext[0] | (ext[1] << 16);
        want[7] = ext[2];
        const int4 ra = P.nbr[2 * (size_t)s], rb = P.nbr[2 * (size_t)s + 1];
        const int have[8] = {ra.x, ra.y, ra.z, ra.w, rb.x, rb.y, rb.z, rb.w};
        bool ok = true;
        for (int j = 0; j < 8; j++)
            if ((j < 2 * D || j >= 6) && have[j] != want[j]) ok = false;
        if (!ok) atomicAdd(bad, 1);
    }
}

// expand_narrow_band phase 2 (Surface.hpp:1249-1355): every neighbour whose value is outside the
// band gets its distance, and joins layer +-L if it is in no list (the value does not depend on
// which expanding voxel claims it: the claimed voxel's nearest neighbour towards the zero curve
// is in the band already, so claims racing for one voxel write the same value).  The neighbours'
// bricks come from the expanding voxel's brick record, then their values and the claims are
// issued level by level; each claimed voxel's distance gathers its own face neighbours through
// its brick's record.
template <int D>
__global__ void k_expand_claim(Geo G, Pool P, SC sc, const Ctr* ctrc, Ctr* ctr, const int* live) {
    if (!*live) return;
    // the round's deactivations are in the neighbour records by now (k_nbr_fix ran)
    if (blockIdx.x == 0 && threadIdx.x == 0) ctr->dead_n = 0;
    constexpr int NN = 2 * D;
    FELT_WAVE_LOOP(i, valid, ctrc->exp_n) {
        const bool cand = valid;
        const int side = cand ? sc.xside[i] : 1;
        const int8_t target = (int8_t)(G.L * side);
        int cp[NN];
#pragma unroll
        for (int j = 0; j < NN; j++) cp[j] = -1;
        if (cand) {
            // the expanding voxel's face neighbours through its brick's record: in this brick, or
            // in an active neighbour brick (an inactive one failed activation -- pool exhausted,
            // flagged -- and is skipped, as one outside the grid)
            const uint32_t vx = sc.xv[i];
            const uint32_t slot = fdiv(vx, G.fpv);
            int loc[3] = {0, 0, 0};
            leaf_coords<D>(G, vx - slot * (uint32_t)G.pv, loc);
            const int4 ra = P.nbr[2 * (size_t)slot], rb = P.nbr[2 * (size_t)slot + 1];
            const int rec[6] = {ra.x, ra.y, ra.z, ra.w, rb.x, rb.y};
            const int ext[3] = {(int)((unsigned)rb.z & 0xffffu), (int)((unsigned)rb.z >> 16), rb.w};
            const int own = P.slot_part[slot];
            bool in[NN];
            uint32_t vq[NN];
            int sq[NN];
#pragma unroll
            for (int a = 0; a < D; a++)
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
                    in[j] = cross ? t >= 0 : la < ext[a];
                    sq[j] = cross ? max(t, 0) : (int)slot;
                    vq[j] = cross ? (uint32_t)max(t, 0) * (uint32_t)G.pv + (uint32_t)leaf_of<D>(G, lc)
                                  : vx + (uint32_t)step_offset<D>(G, loc, d);
                }
            float dq[NN];
#pragma unroll
            for (int j = 0; j < NN; j++)
                dq[j] = in[j] ? __hip_atomic_load(&P.phi[vq[j]], __ATOMIC_RELAXED,
                                                  __HIP_MEMORY_SCOPE_AGENT)
                              : 0.0f;
            // As the reference, a neighbour is claimed when its value is outside the band, even
            // if it is tracked already (a value of exactly +-(L+0.5) rounds outside): its value
            // is recomputed, and it joins the list only if it was in none (first list wins).
            bool cl[NN], fresh[NN];
#pragma unroll
            for (int j = 0; j < NN; j++) {
                cl[j] = in[j] && !in_band(layer_of(dq[j], G.L, &ctr->nan_events), G.L);
                fresh[j] = cl[j] && cas_byte(&P.lay[vq[j]], NONE, target);
            }
#pragma unroll
            for (int j = 0; j < NN; j++) {
                if (!cl[j]) continue;
                const float nd = distance_rec_self<D>(G, P, vq[j], dq[j], (float)side);
                __hip_atomic_store(&P.phi[vq[j]], nd, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                if (fresh[j]) cp[j] = sq[j] == (int)slot ? own : P.slot_part[sq[j]];
            }
        }
#pragma unroll
        for (int j = 0; j < NN; j++)
            agg_pcnt(P.pcnt, G.NL, cp[j], -1, cp[j] >= 0 ? target + G.L : -1);
    }
}

// Tracked voxels in total (sum of pcnt), for sizing the next step's lists.
__global__ void k_band_total2(Geo G, Pool P, Ctr* c) {
    int s = 0;
    for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < G.np; p += gridDim.x * blockDim.x)
        for (int k = 0; k < G.NL; k++) s += P.pcnt[p * G.NL + k];
    for (int o = 32; o > 0; o >>= 1) s += __shfl_down(s, o);
    if ((threadIdx.x & 63) == 0 && s) atomicAdd(&c->band_total, s);
}

__global__ void k_zero_int(int* p) { *p = 0; }

}  // namespace step
}  // namespace felt
