// felt_ckpt.hip — Surface::save / Surface::load in the reference's on-disk format: the IsoGrid as
// a cereal binary archive (Surface.hpp:180-196, PartitionedMixin.hpp:692-761).
//
// Archive layout (cereal writes arithmetic values as raw little-endian bytes and container sizes
// as uint64; Felt writes an Eigen vector as int32 rows, int32 cols and a std::vector of its
// elements):
//   f32 background;  vec child_size                         Access, Children (PartitionedMixin.hpp:582, 93)
//   children grid:   child bg_child; u64 n; child[n]; children lookup; 3 vec sizes
//   3 vec sizes (size, offset, offset+size);  f32 dx; vec pos_min; vec pos_max   Size, Spatial
//   child   := f32 bg; u64 n; f32[n] values; child lookup; 3 vec sizes          (Tracked.hpp:53-60)
//   child lookup := u64 null; u64 n; u64[n] list index per voxel; lists; 3 vec  (Lookup.hpp:170-177)
//   children lookup := tuple null; u64 n; tuple[n] list index per partition; lists; 3 vec
//   vec := i32 D, i32 1, u64 D, i32[D];  tuple := i32 N, i32 1, u64 N, u64[N]
//   lists := i32 N, i32 1, u64 N, then per list u64 length and u64 entries (N = 2L+1)
// A child is active exactly when it has values.  The device keeps no list order, so save writes
// every list in ascending position order (any order is a valid archive: load rebuilds the lists
// from it); load takes layer membership from the child lists.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "felt_polys.h"
#include "felt_surface.h"

using namespace felt;

namespace {

constexpr int TPB = 256;
constexpr uint64_t kNull = ~0ull;  // Felt::null_idx (Common.hpp:102)

#define HIPX(expr)                                                                              \
    do {                                                                                       \
        hipError_t e_ = (expr);                                                                \
        if (e_ != hipSuccess)                                                                  \
            return felt_fail(FELT_E_DEVICE, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

// Member m back to an empty grid: bricks returned to the free stack, every partition inactive
// with background `bg`, no layer counts, no delta or status partitions.
__global__ void k_member_clear(Geo G, Pool P, int pbase, int np, float bg, int* free_stack,
                               int* dlist, Ctr* ctr, unsigned* dmask, unsigned* smask) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < np; i += gridDim.x * blockDim.x) {
        const int p = pbase + i;
        const int s = P.page[p];
        if (s >= 0) {
            P.page[p] = -1;
            P.slot_part[s] = -1;
            const int t = atomicAdd(&ctr->free_top, 1);
            free_stack[t] = s;
            dlist[atomicAdd(&ctr->dead_n, 1)] = p;
        }
        P.pbg[p] = bg;
        for (int k = 0; k < G.NL; k++) P.pcnt[(size_t)p * G.NL + k] = 0;
        dmask[p] = 0u;
        smask[p] = 0u;
    }
}

// Bricks of the listed partitions to / from contiguous staging arrays, one workgroup per brick.
__global__ void k_bricks_out(Geo G, Pool P, const int* parts, int n, float* phi, int8_t* lay) {
    for (int j = blockIdx.x; j < n; j += gridDim.x) {
        const int s = P.page[parts[j]];
        if (s < 0) continue;
        const size_t src = (size_t)s * G.pv, dst = (size_t)j * G.pv;
        for (int l = threadIdx.x; l < G.pv; l += blockDim.x) {
            phi[dst + l] = P.phi[src + l];
            lay[dst + l] = P.lay[src + l];
        }
    }
}

__global__ void k_bricks_in(Geo G, Pool P, const int* parts, int n, const float* phi,
                            const int8_t* lay) {
    for (int j = blockIdx.x; j < n; j += gridDim.x) {
        const int s = P.page[parts[j]];
        if (s < 0) continue;
        const size_t dst = (size_t)s * G.pv, src = (size_t)j * G.pv;
        for (int l = threadIdx.x; l < G.pv; l += blockDim.x) {
            P.phi[dst + l] = phi[src + l];
            P.lay[dst + l] = lay[src + l];
        }
    }
}

// Device staging that frees itself.
struct Stage {
    void* p = nullptr;
    ~Stage() {
        if (p) (void)hipFree(p);
    }
};

struct Out {
    uint8_t* b;
    size_t at = 0;
    template <class T>
    void put(T v) {
        std::memcpy(b + at, &v, sizeof(T));
        at += sizeof(T);
    }
    void vec(int D, const int* v) {
        put<int32_t>(D);
        put<int32_t>(1);
        put<uint64_t>((uint64_t)D);
        for (int a = 0; a < D; a++) put<int32_t>(v[a]);
    }
    void sizes(int D, const int* sz, const int* off) {
        int end[3];
        for (int a = 0; a < D; a++) end[a] = off[a] + sz[a];
        vec(D, sz);
        vec(D, off);
        vec(D, end);
    }
    void tuple_head(int N) {
        put<int32_t>(N);
        put<int32_t>(1);
        put<uint64_t>((uint64_t)N);
    }
};

struct In {
    const uint8_t* b;
    size_t n, at = 0;
    bool ok = true;
    template <class T>
    T get() {
        T v{};
        if (!ok || n - at < sizeof(T)) {
            ok = false;
            return v;
        }
        std::memcpy(&v, b + at, sizeof(T));
        at += sizeof(T);
        return v;
    }
    // Start of an array of cnt elements of `bytes` each; skips it.
    size_t skip(uint64_t cnt, size_t bytes) {
        const size_t here = at;
        if (!ok || cnt > (n - at) / bytes) {
            ok = false;
            return here;
        }
        at += (size_t)cnt * bytes;
        return here;
    }
    bool vec(int D, int* v) {
        if (get<int32_t>() != D || get<int32_t>() != 1 || get<uint64_t>() != (uint64_t)D) ok = false;
        for (int a = 0; a < D && ok; a++) v[a] = get<int32_t>();
        return ok;
    }
    bool sizes_are(int D, const int* sz, const int* off) {
        int a[3], b2[3], c[3];
        if (!vec(D, a) || !vec(D, b2) || !vec(D, c)) return false;
        for (int i = 0; i < D; i++)
            if (a[i] != sz[i] || b2[i] != off[i] || c[i] != off[i] + sz[i]) return ok = false;
        return true;
    }
    bool skip_sizes(int D) {
        int v[3];
        return vec(D, v) && vec(D, v) && vec(D, v);
    }
    bool tuple_head(int N) {
        if (get<int32_t>() != N || get<int32_t>() != 1 || get<uint64_t>() != (uint64_t)N) ok = false;
        return ok;
    }
};

size_t vec_bytes(int D) { return 16 + 4 * (size_t)D; }
size_t lists_bytes(int N, uint64_t entries) { return 16 + 8 * (size_t)N + 8 * (size_t)entries; }

// perm[r] = storage leaf of the reference's row-major leaf PosIdx r.
std::vector<int> leaf_perm(const Geo& G) {
    std::vector<int> perm(G.pv);
    for (int r = 0; r < G.pv; r++) {
        int lc[3] = {0, 0, 0}, t = r;
        for (int a = G.D - 1; a >= 0; a--) {
            lc[a] = t % G.part[a];
            t /= G.part[a];
        }
        perm[r] = G.D == 2 ? leaf_of<2>(G, lc) : leaf_of<3>(G, lc);
    }
    return perm;
}

// World offset of member-local partition lp.
void child_offset(const Geo& G, const Mem& M, int lp, int* off) {
    int r = lp;
    for (int a = G.D - 1; a >= 0; a--) {
        const int c = r % M.csize[a];
        r /= M.csize[a];
        off[a] = c * G.part[a] + M.offset[a];
    }
}

}  // namespace

// (D, L, size, part) of an archive: the child size is its first vector, the background child's
// list tuple holds N = 2L+1 lists, and the grid size is the isogrid's Size, written just before
// the trailing Spatial fields.
int felt_archive_geometry(const uint8_t* src, size_t n, int* D, int* L, int* size, int* part) {
    In in{src, n};
    in.get<float>();
    const int32_t d = in.get<int32_t>();
    if (!in.ok || (d != 2 && d != 3)) return felt_fail(FELT_E_ARG, "not a surface archive");
    in.at -= 4;
    if (!in.vec(d, part)) return felt_fail(FELT_E_ARG, "not a surface archive");
    in.get<float>();                              // background child: bg, values, lookup
    in.skip(in.get<uint64_t>(), 4);
    in.get<uint64_t>();
    in.skip(in.get<uint64_t>(), 8);
    const int32_t N = in.get<int32_t>();
    const size_t V = vec_bytes(d);
    if (!in.ok || N < 3 || N % 2 == 0 || N > 17 || n < 9 * V + 4)
        return felt_fail(FELT_E_ARG, "not a surface archive");
    In tail{src, n, n - (4 + 2 * V) - 3 * V};
    if (!tail.vec(d, size)) return felt_fail(FELT_E_ARG, "not a surface archive");
    *D = d;
    *L = (N - 1) / 2;
    return FELT_OK;
}

int felt_member_clear_impl(felt_surface* s, int m) {
    const Mem& M = s->mems[m];
    HIPX(hipMemsetAsync(&s->ctr->dead_n, 0, sizeof(int), s->st));
    k_member_clear<<<std::min(1024, (M.np + TPB - 1) / TPB), TPB, 0, s->st>>>(
        s->G, s->P, M.pbase, M.np, (float)(s->G.L + 1), s->free_stack, s->dlist, s->ctr, s->dmask,
        s->smask);
    HIPX(hipGetLastError());
    int rc;
    if ((rc = felt_nbr_fix(s))) return rc;
    HIPX(hipStreamSynchronize(s->st));
    s->band_known = false;
    s->hint = -1;
    return FELT_OK;
}

int felt_save_impl(felt_surface* s, int m, uint8_t* dst, size_t cap, size_t* nout) {
    const Geo& G = s->G;
    const Mem& M = s->mems[m];
    const int D = G.D, N = G.NL, np = M.np, pv = G.pv;
    std::vector<int> page(np), cnt((size_t)np * N);
    std::vector<float> bg(np);
    HIPX(hipMemcpyAsync(page.data(), s->P.page + M.pbase, np * sizeof(int), hipMemcpyDeviceToHost, s->st));
    HIPX(hipMemcpyAsync(bg.data(), s->P.pbg + M.pbase, np * sizeof(float), hipMemcpyDeviceToHost, s->st));
    HIPX(hipMemcpyAsync(cnt.data(), s->P.pcnt + (size_t)M.pbase * N, cnt.size() * sizeof(int),
                        hipMemcpyDeviceToHost, s->st));
    HIPX(hipStreamSynchronize(s->st));
    const size_t V = vec_bytes(D);
    std::vector<int> act;
    std::vector<uint64_t> plen(N, 0);
    size_t total = 4 + V;
    const size_t empty_child = 4 + 8 + 8 + 8 + lists_bytes(N, 0) + 6 * V;
    total += empty_child + 8;
    for (int p = 0; p < np; p++) {
        uint64_t e = 0;
        for (int k = 0; k < N; k++) {
            const int c = cnt[(size_t)p * N + k];
            e += (uint64_t)c;
            if (c > 0) plen[k]++;
        }
        if (page[p] >= 0) {
            act.push_back(p);
            total += 4 + 8 + 4 * (size_t)pv + 8 + 8 + 8 * (size_t)pv + lists_bytes(N, e) + 6 * V;
        } else {
            if (e) return felt_fail(FELT_E_STATE, "inactive partition with tracked voxels");
            total += empty_child;
        }
    }
    uint64_t pent = 0;
    for (int k = 0; k < N; k++) pent += plen[k];
    total += (16 + 8 * (size_t)N) + 8 + (size_t)np * (16 + 8 * (size_t)N) + lists_bytes(N, pent) +
             6 * V + 3 * V + 4 + 2 * V;
    *nout = total;
    if (!dst) return FELT_OK;
    if (cap < total) return felt_fail(FELT_E_CAPACITY, "archive buffer too small");

    // bricks of the active partitions, ascending
    const int na = (int)act.size();
    std::vector<float> phi((size_t)na * pv);
    std::vector<int8_t> lay((size_t)na * pv);
    if (na) {
        Stage st;
        const size_t bytes = (size_t)na * pv * 5 + (size_t)na * 4 + 16;
        HIPX(hipMalloc(&st.p, bytes));
        float* d_phi = (float*)st.p;
        int8_t* d_lay = (int8_t*)(d_phi + (size_t)na * pv);
        int* d_parts = (int*)(((uintptr_t)(d_lay + (size_t)na * pv) + 15) & ~(uintptr_t)15);
        std::vector<int> parts(na);
        for (int j = 0; j < na; j++) parts[j] = M.pbase + act[j];
        HIPX(hipMemcpyAsync(d_parts, parts.data(), na * sizeof(int), hipMemcpyHostToDevice, s->st));
        k_bricks_out<<<std::min(na, 4096), TPB, 0, s->st>>>(G, s->P, d_parts, na, d_phi, d_lay);
        HIPX(hipGetLastError());
        HIPX(hipMemcpyAsync(phi.data(), d_phi, phi.size() * sizeof(float), hipMemcpyDeviceToHost, s->st));
        HIPX(hipMemcpyAsync(lay.data(), d_lay, lay.size(), hipMemcpyDeviceToHost, s->st));
        HIPX(hipStreamSynchronize(s->st));
    }

    Out o{dst};
    const int zero[3] = {0, 0, 0};
    o.put<float>((float)(G.L + 1));
    o.vec(D, G.part);
    // background child of the children grid (its sizes are never set by the reference)
    o.put<float>((float)(G.L + 1));
    o.put<uint64_t>(0);
    o.put<uint64_t>(kNull);
    o.put<uint64_t>(0);
    o.tuple_head(N);
    for (int k = 0; k < N; k++) o.put<uint64_t>(0);
    for (int i = 0; i < 6; i++) o.vec(D, zero);
    o.put<uint64_t>((uint64_t)np);
    std::vector<uint64_t> look(pv);
    std::vector<float> row(pv);
    std::vector<std::vector<uint32_t>> lists(N);
    const std::vector<int> perm = leaf_perm(G);
    int j = 0;
    for (int p = 0; p < np; p++) {
        int off[3] = {0, 0, 0};
        child_offset(G, M, p, off);
        o.put<float>(bg[p]);
        const bool active = page[p] >= 0;
        if (active) {
            const float* ph = &phi[(size_t)j * pv];
            const int8_t* ly = &lay[(size_t)j * pv];
            for (auto& l : lists) l.clear();
            for (int r = 0; r < pv; r++) {  // the reference's leaf order
                const int l = perm[r];
                row[r] = ph[l];
                look[r] = kNull;
                if (ly[l] >= -G.L && ly[l] <= G.L) {
                    auto& lst = lists[ly[l] + G.L];
                    look[r] = lst.size();
                    lst.push_back((uint32_t)r);
                }
            }
            for (int k = 0; k < N; k++)
                if ((int)lists[k].size() != cnt[(size_t)p * N + k])
                    return felt_fail(FELT_E_STATE, "layer counts and layer lists disagree");
            o.put<uint64_t>((uint64_t)pv);
            std::memcpy(o.b + o.at, row.data(), (size_t)pv * 4);
            o.at += (size_t)pv * 4;
            o.put<uint64_t>(kNull);
            o.put<uint64_t>((uint64_t)pv);
            std::memcpy(o.b + o.at, look.data(), (size_t)pv * 8);
            o.at += (size_t)pv * 8;
            o.tuple_head(N);
            for (int k = 0; k < N; k++) {
                o.put<uint64_t>(lists[k].size());
                for (uint32_t l : lists[k]) o.put<uint64_t>(l);
            }
            j++;
        } else {
            o.put<uint64_t>(0);
            o.put<uint64_t>(kNull);
            o.put<uint64_t>(0);
            o.tuple_head(N);
            for (int k = 0; k < N; k++) o.put<uint64_t>(0);
        }
        o.sizes(D, G.part, off);
        o.sizes(D, G.part, off);
    }
    // children lookup: background tuple, per-partition list indices, parent lists, sizes
    o.tuple_head(N);
    for (int k = 0; k < N; k++) o.put<uint64_t>(kNull);
    o.put<uint64_t>((uint64_t)np);
    std::vector<uint64_t> next(N, 0);
    for (int p = 0; p < np; p++) {
        o.tuple_head(N);
        for (int k = 0; k < N; k++)
            o.put<uint64_t>(cnt[(size_t)p * N + k] > 0 ? next[k]++ : kNull);
    }
    o.tuple_head(N);
    for (int k = 0; k < N; k++) {
        o.put<uint64_t>(plen[k]);
        for (int p = 0; p < np; p++)
            if (cnt[(size_t)p * N + k] > 0) o.put<uint64_t>((uint64_t)p);
    }
    int coff[3] = {0, 0, 0};
    for (int a = 0; a < D; a++) coff[a] = M.offset[a] / G.part[a];
    o.sizes(D, M.csize, coff);
    o.sizes(D, M.csize, coff);
    o.sizes(D, M.size, M.offset);
    o.put<float>(1.0f);
    int hi[3] = {0, 0, 0};
    for (int a = 0; a < D; a++) hi[a] = M.offset[a] + M.size[a];
    o.vec(D, M.offset);
    o.vec(D, hi);
    if (o.at != total) return felt_fail(FELT_E_STATE, "archive size mismatch");
    return FELT_OK;
}

int felt_load_impl(felt_surface* s, int m, const uint8_t* src, size_t n) {
    const Geo& G = s->G;
    const Mem& M = s->mems[m];
    const int D = G.D, N = G.NL, np = M.np, pv = G.pv;
    auto bad = [](const char* what) {
        return felt_fail(FELT_E_ARG, std::string("archive: ") + what);
    };
    In in{src, n};
    if (in.get<float>() != (float)(G.L + 1)) return bad("grid background is not L+1");
    int part[3] = {0, 0, 0};
    if (!in.vec(D, part)) return bad("malformed child size");
    for (int a = 0; a < D; a++)
        if (part[a] != G.part[a]) return bad("partition size differs from the surface's");
    // background child
    in.get<float>();
    in.skip(in.get<uint64_t>(), 4);
    in.get<uint64_t>();
    in.skip(in.get<uint64_t>(), 8);
    if (!in.tuple_head(N)) return bad("list count differs from the surface's");
    for (int k = 0; k < N; k++) in.skip(in.get<uint64_t>(), 8);
    for (int i = 0; i < 2; i++) in.skip_sizes(D);
    if (!in.ok || in.get<uint64_t>() != (uint64_t)np) return bad("partition count differs");

    std::vector<float> bg(np);
    std::vector<int> act, cnt((size_t)np * N, 0);
    std::vector<size_t> data_at;
    std::vector<int8_t> lay;
    std::vector<uint8_t> seen(pv);
    const std::vector<int> perm = leaf_perm(G);
    for (int p = 0; p < np; p++) {
        int off[3] = {0, 0, 0};
        child_offset(G, M, p, off);
        bg[p] = in.get<float>();
        const uint64_t nd = in.get<uint64_t>();
        const size_t at = in.skip(nd, 4);
        in.get<uint64_t>();
        const uint64_t nl = in.get<uint64_t>();
        in.skip(nl, 8);
        if (!in.ok) return bad("truncated child");
        const bool active = nd != 0;
        if (active && (nd != (uint64_t)pv || nl != (uint64_t)pv)) return bad("child of the wrong size");
        if (!active && nl != 0) return bad("inactive child with a lookup");
        if (!in.tuple_head(N)) return bad("malformed child lists");
        size_t base = 0;
        if (active) {
            act.push_back(p);
            data_at.push_back(at);
            base = lay.size();
            lay.resize(base + pv, NONE);
            std::fill(seen.begin(), seen.end(), 0);
        }
        for (int k = 0; k < N; k++) {
            const uint64_t len = in.get<uint64_t>();
            const size_t la = in.skip(len, 8);
            if (!in.ok) return bad("truncated child list");
            if (len && !active) return bad("inactive child with tracked voxels");
            for (uint64_t i = 0; i < len; i++) {
                uint64_t l;
                std::memcpy(&l, src + la + 8 * i, 8);
                if (l >= (uint64_t)pv || seen[l]) return bad("bad or repeated leaf in a child list");
                seen[l] = 1;
                int q[3] = {0, 0, 0};
                int r = (int)l;
                for (int a = D - 1; a >= 0; a--) {
                    q[a] = r % G.part[a] + off[a] - M.offset[a];
                    r /= G.part[a];
                }
                for (int a = 0; a < D; a++)
                    if (q[a] < 0 || q[a] >= M.size[a]) return bad("tracked voxel outside the grid");
                lay[base + perm[l]] = (int8_t)(k - G.L);
            }
            cnt[(size_t)p * N + k] = (int)len;
        }
        if (!in.sizes_are(D, G.part, off) || !in.sizes_are(D, G.part, off))
            return bad("child geometry differs from the surface's");
    }
    // children lookup: list indices are rebuilt from the child lists; parent lists must agree
    if (!in.tuple_head(N)) return bad("malformed children lookup");
    in.skip(N, 8);
    if (in.get<uint64_t>() != (uint64_t)np) return bad("children lookup of the wrong size");
    for (int p = 0; p < np; p++) {
        if (!in.tuple_head(N)) return bad("malformed children lookup");
        in.skip(N, 8);
    }
    if (!in.tuple_head(N)) return bad("malformed parent lists");
    for (int k = 0; k < N; k++) {
        const uint64_t len = in.get<uint64_t>();
        const size_t la = in.skip(len, 8);
        if (!in.ok) return bad("truncated parent list");
        uint64_t want = 0;
        for (int p = 0; p < np; p++) want += cnt[(size_t)p * N + k] > 0;
        if (len != want) return bad("parent lists disagree with the child lists");
        for (uint64_t i = 0; i < len; i++) {
            uint64_t p;
            std::memcpy(&p, src + la + 8 * i, 8);
            if (p >= (uint64_t)np || cnt[(size_t)p * N + k] == 0)
                return bad("parent lists disagree with the child lists");
        }
    }
    int coff[3] = {0, 0, 0};
    for (int a = 0; a < D; a++) coff[a] = M.offset[a] / G.part[a];
    if (!in.sizes_are(D, M.csize, coff) || !in.sizes_are(D, M.csize, coff))
        return bad("children grid differs from the surface's");
    if (!in.sizes_are(D, M.size, M.offset)) return bad("grid size or offset differs from the surface's");
    int hi[3] = {0, 0, 0}, v[3];
    for (int a = 0; a < D; a++) hi[a] = M.offset[a] + M.size[a];
    if (in.get<float>() != 1.0f) return bad("grid spacing is not 1");
    if (!in.vec(D, v) || !std::equal(v, v + D, M.offset)) return bad("bad position bounds");
    if (!in.vec(D, v) || !std::equal(v, v + D, hi)) return bad("bad position bounds");
    if (in.at != n) return bad("trailing bytes");

    // device: empty the member, set backgrounds, activate, fill the bricks, set the counts
    int rc;
    if ((rc = felt_member_clear_impl(s, m))) return rc;
    HIPX(hipMemcpyAsync(s->P.pbg + M.pbase, bg.data(), np * sizeof(float), hipMemcpyHostToDevice, s->st));
    const int na = (int)act.size();
    if (na) {
        Stage st;
        const size_t bytes = (size_t)na * pv * 5 + (size_t)na * 4 + 16;
        HIPX(hipMalloc(&st.p, bytes));
        float* d_phi = (float*)st.p;
        int8_t* d_lay = (int8_t*)(d_phi + (size_t)na * pv);
        int* d_parts = (int*)(((uintptr_t)(d_lay + (size_t)na * pv) + 15) & ~(uintptr_t)15);
        std::vector<int> parts(na);
        std::vector<float> vals((size_t)na * pv);
        for (int j = 0; j < na; j++) {
            parts[j] = M.pbase + act[j];
            const uint8_t* d = src + data_at[j];
            float* out = &vals[(size_t)j * pv];
            for (int r = 0; r < pv; r++) std::memcpy(&out[perm[r]], d + 4 * (size_t)r, 4);
        }
        HIPX(hipMemcpyAsync(d_phi, vals.data(), vals.size() * 4, hipMemcpyHostToDevice, s->st));
        HIPX(hipMemcpyAsync(d_lay, lay.data(), lay.size(), hipMemcpyHostToDevice, s->st));
        HIPX(hipMemcpyAsync(d_parts, parts.data(), na * sizeof(int), hipMemcpyHostToDevice, s->st));
        HIPX(hipMemcpyAsync(s->nlist, parts.data(), na * sizeof(int), hipMemcpyHostToDevice, s->st));
        HIPX(hipMemcpyAsync(&s->ctr->need_n, &na, sizeof(int), hipMemcpyHostToDevice, s->st));
        HIPX(hipMemsetAsync(&s->ctr->err, 0, sizeof(int), s->st));
        if ((rc = felt_alloc_listed(s, std::min(1024, na)))) return rc;
        k_bricks_in<<<std::min(na, 4096), TPB, 0, s->st>>>(G, s->P, d_parts, na, d_phi, d_lay);
        HIPX(hipGetLastError());
        HIPX(hipStreamSynchronize(s->st));
        Ctr hc;
        HIPX(hipMemcpy(&hc, s->ctr, sizeof(Ctr), hipMemcpyDeviceToHost));
        if (hc.err == 1) {
            (void)felt_member_clear_impl(s, m);
            return felt_fail(FELT_E_NOMEM, "brick pool exhausted");
        }
    }
    HIPX(hipMemcpyAsync(s->P.pcnt + (size_t)M.pbase * N, cnt.data(), cnt.size() * sizeof(int),
                        hipMemcpyHostToDevice, s->st));
    HIPX(hipStreamSynchronize(s->st));
    s->band_known = false;
    s->hint = -1;
    if (s->polys) return felt_polys_invalidate_impl(s->polys);
    return FELT_OK;
}
