/*
 * felt_oracle.hpp — TEST INFRASTRUCTURE, NOT PRODUCT CODE.
 *
 * A CPU restatement of Felt's sparse-field level-set surface and its polygonisation, written
 * from the reference's semantics so that the HIP implementation can be checked against it.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it, and only as
 * the checker.  The reference itself (feltech/Felt, header-only C++14) cannot be compiled here:
 * it needs Eigen3 and Boost, which the image lacks.
 *
 * What it restates, with the reference lines it follows:
 *   - grid geometry and index conventions: Util.hpp:23-83, PartitionedMixin.hpp:59-141, 221-245
 *   - lazily activated partition children, per-partition tracking lists, parent (children-grid)
 *     lists with swap-with-last untrack and first-list-wins track: LookupMixin.hpp:225-576,
 *     TrackedMixin.hpp:19-63, 250-420, PartitionedMixin.hpp:145-615, GridMixin.hpp:565-622
 *   - Surface: seed, update(fn), update(lo,hi,fn), update_start/delta/update_end, round loop,
 *     calc_affected, zero-layer apply, layered distance transform, status change, flush,
 *     expand: Surface.hpp:163-459, 739-1440, 1677-1742
 *   - numeric stencils: curv, grad (int and float positions), interp: NumericMixin.hpp:72-231,
 *     390-500
 *   - Polys / Poly::Single marching squares and cubes: Polys.hpp:29-201, Poly.hpp:17-368,
 *     PolyMixin.hpp:28-672
 *
 * Every list is a RawList: a vector whose pop_back leaves its storage untouched, as
 * std::vector<size_t> does, because the reference's MultiList::reset iterates a parent list
 * while removing from it (PartitionedMixin.hpp:336-360) and reads the stale tail.
 *
 * Arithmetic is fp32 (Real = float) or fp64 (Real = double, the FP64 error oracle), built with
 * -ffp-contract=off.  Pinned third-party behaviour (Eigen and Boost are not vendored):
 *   - boost::math::iround (Surface.hpp:1696): exact round-half-away-from-zero (std::round);
 *   - Eigen Vector3::sum() and squaredNorm(): v0 + (v1 + v2) (Eigen's recursive-halving
 *     unroller for a 3-vector);
 *   - Eigen normalize(): n = squaredNorm(); if n > 0 then v /= sqrt(n).
 * layer_id(NaN) has no reference value (Boost raises rounding_error); here it is L+1 and is
 * counted in nan_events.
 */
#pragma once

#include <algorithm>
#include <array>
#include <cfloat>
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <limits>
#include <mutex>
#include <stdexcept>
#include <utility>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace felt_oracle {

using Idx = std::size_t;
constexpr Idx NUL = std::numeric_limits<Idx>::max();
constexpr Idx OMP_MIN_CHUNK = 32;  // FELT_SURFACE_OMP_MIN_CHUNK_SIZE (Surface.hpp:19-27)

template <int D> using Vi = std::array<int, D>;
template <int D, class R> using Vr = std::array<R, D>;

/* Vector whose pop_back leaves storage untouched, like std::vector of a trivial type. */
struct RawList {
    std::vector<Idx> mem;
    Idx n = 0;
    Idx size() const { return n; }
    Idx operator[](Idx i) const { return mem[i]; }
    Idx& operator[](Idx i) { return mem[i]; }
    void push_back(Idx v) {
        if (n < mem.size()) mem[n] = v; else mem.push_back(v);
        ++n;
    }
    void pop_back() { --n; }
    void clear() { n = 0; }
    void release() { mem.clear(); mem.shrink_to_fit(); n = 0; }
    const Idx* begin() const { return mem.data(); }
    const Idx* end() const { return mem.data() + n; }
};

/* Row-major index, last axis fastest (Util.hpp:23-40). */
template <int D>
Idx lin(const Vi<D>& p, const Vi<D>& size) {
    long long idx = 0;
    for (int i = 0; i < D; i++) idx = idx * size[i] + p[i];
    return Idx(idx);
}
/* Inverse of lin (Util.hpp:56-83). */
template <int D>
Vi<D> unlin(Idx idx, const Vi<D>& size) {
    Vi<D> p{};
    for (int a = D - 1; a > 0; a--) {
        p[a] = int(idx % Idx(size[a]));
        idx /= Idx(size[a]);
    }
    p[0] = int(idx % Idx(size[0]));
    return p;
}

template <int D>
struct Geom {
    Vi<D> size{}, part{}, offset{}, csize{}, coff{};
    Idx np = 0, pv = 0;
    Geom() = default;
    Geom(const Vi<D>& size_, const Vi<D>& part_) : size(size_), part(part_) {
        bool uneven = false;
        for (int a = 0; a < D; a++) {
            offset[a] = -1 * size[a] / 2;                 // Surface.hpp:1716-1719
            csize[a] = size[a] / part[a];
            if (csize[a] * part[a] != size[a]) uneven = true;
            coff[a] = offset[a] / part[a];                // PartitionedMixin.hpp:66
        }
        if (uneven)                                       // PartitionedMixin.hpp:131-141
            for (int a = 0; a < D; a++) csize[a] += 1;
        np = 1; pv = 1;
        for (int a = 0; a < D; a++) { np *= Idx(csize[a]); pv *= Idx(part[a]); }
    }
    /* Half-open bounds compared as float (Util.hpp:94-106). */
    template <class P>
    bool inside(const P& p) const {
        for (int a = 0; a < D; a++) {
            if (float(p[a]) >= float(offset[a] + size[a])) return false;
            if (float(p[a]) < float(offset[a])) return false;
        }
        return true;
    }
    Idx child_of(const Vi<D>& pos) const {               // PartitionedMixin.hpp:221-245
        Vi<D> c{};
        for (int a = 0; a < D; a++) c[a] = (pos[a] - offset[a]) / part[a];
        return lin<D>(c, csize);
    }
    Vi<D> child_offset(Idx p) const {                    // PartitionedMixin.hpp:71-85
        Vi<D> c = unlin<D>(p, csize), o{};
        for (int a = 0; a < D; a++) o[a] = c[a] * part[a] + offset[a];
        return o;
    }
    Idx leaf_of(Idx p, const Vi<D>& pos) const {
        Vi<D> o = child_offset(p), l{};
        for (int a = 0; a < D; a++) l[a] = pos[a] - o[a];
        return lin<D>(l, part);
    }
    Vi<D> pos_of(Idx p, Idx l) const {
        Vi<D> o = child_offset(p), q = unlin<D>(l, part);
        for (int a = 0; a < D; a++) q[a] += o[a];
        return q;
    }
    /* Children-grid position (with children offset) of partition p. */
    Vi<D> child_pos(Idx p) const {
        Vi<D> c = unlin<D>(p, csize);
        for (int a = 0; a < D; a++) c[a] += coff[a];
        return c;
    }
    Idx child_index(const Vi<D>& cpos) const {
        Vi<D> c{};
        for (int a = 0; a < D; a++) c[a] = cpos[a] - coff[a];
        return lin<D>(c, csize);
    }
};

/* One lazily activated child grid: values (tracked grids only), a lookup of list indices and
 * NL tracking lists (Tracked.hpp:16-79, Lookup.hpp:136-194). */
template <class T>
struct Kid {
    bool active = false;
    T bg{};
    std::vector<T> data;
    std::vector<Idx> look;
    std::vector<RawList> lists;

    T get(Idx l) const { return active ? data[l] : bg; }
    Idx lookup(Idx l) const { return look.empty() ? NUL : look[l]; }
    /* MultiList::SingleIdx::track (LookupMixin.hpp:298-331). */
    bool ltrack(Idx l, int k) {
        if (look[l] != NUL) return false;
        look[l] = lists[k].size();
        lists[k].push_back(l);
        return true;
    }
    /* MultiList::SingleIdx::untrack (LookupMixin.hpp:333-372). */
    void luntrack(Idx l, int k) {
        Idx at = look[l];
        if (at == NUL) return;
        RawList& lst = lists[k];
        Idx last = lst.size() - 1;
        if (at < last) {
            Idx moved = lst[last];
            lst[at] = moved;
            look[moved] = at;
        }
        look[l] = NUL;
        lst.pop_back();
    }
    void lreset() {
        for (RawList& lst : lists) {
            for (Idx l : lst) look[l] = NUL;
            lst.clear();
        }
    }
};

/* Partitioned grid: children with per-partition lists, plus the children grid's own per-list
 * partition lists (MultiIdx, LookupMixin.hpp:390-576).  HasValue selects Tracked (value + lookup)
 * or Lookup (lookup only) children. */
template <int D, class T, bool HasValue>
struct PartGrid {
    const Geom<D>* g = nullptr;
    int NL = 0;
    T grid_bg{};
    std::vector<Kid<T>> kids;
    std::vector<Idx> pidx;          // np * NL parent list indices
    std::vector<RawList> plists;    // NL parent lists
    std::mutex mtx;

    void init(const Geom<D>* geom, int nl, T bg) {
        g = geom; NL = nl; grid_bg = bg;
        kids.assign(g->np, Kid<T>{});
        for (Kid<T>& k : kids) { k.bg = bg; k.lists.assign(NL, RawList{}); }
        pidx.assign(g->np * Idx(NL), NUL);
        plists.assign(NL, RawList{});
    }
    bool ptracked(Idx p, int k) const { return pidx[p * NL + k] != NUL; }
    bool ptracked(Idx p) const {
        for (int k = 0; k < NL; k++) if (ptracked(p, k)) return true;
        return false;
    }
    bool ptrack(Idx p, int k) {
        Idx& at = pidx[p * NL + k];
        if (at != NUL) return false;
        at = plists[k].size();
        plists[k].push_back(p);
        return true;
    }
    void puntrack(Idx p, int k) {
        Idx& at = pidx[p * NL + k];
        if (at == NUL) return;
        RawList& lst = plists[k];
        Idx last = lst.size() - 1;
        if (at < last) {
            Idx moved = lst[last];
            lst[at] = moved;
            pidx[moved * NL + k] = at;
        }
        at = NUL;
        lst.pop_back();
    }
    /* Grid::Activate / Tracked::Activate (GridMixin.hpp:75-81, TrackedMixin.hpp:19-63). */
    void activate(Idx p) {
        Kid<T>& kid = kids[p];
        if (HasValue) kid.data.assign(g->pv, kid.bg);
        kid.look.assign(g->pv, NUL);
        kid.active = true;
    }
    void deactivate(Idx p) {
        Kid<T>& kid = kids[p];
        kid.data.clear(); kid.data.shrink_to_fit();
        kid.look.clear(); kid.look.shrink_to_fit();
        for (RawList& l : kid.lists) l.release();
        kid.active = false;
    }
    /* Child reset: values of tracked leaves back to background, lookup cleared. */
    void reset_kid(Idx p) {
        Kid<T>& kid = kids[p];
        if (HasValue)
            for (RawList& lst : kid.lists)
                for (Idx l : lst) kid.data[l] = kid.bg;
        kid.lreset();
    }
    /* Leafs::track_child (PartitionedMixin.hpp:201-213). */
    void track_child(Idx p, int k) {
        if (ptracked(p, k)) return;
        std::lock_guard<std::mutex> lock(mtx);
        if (ptracked(p, k)) return;
        if (!kids[p].active) activate(p);
        ptrack(p, k);
    }
    /* Tracked::track (PartitionedMixin.hpp:441-470). */
    void track(T val, Idx p, Idx l, int k) {
        track_child(p, k);
        Kid<T>& kid = kids[p];
        kid.data[l] = val;
        kid.ltrack(l, k);
    }
    void track(T val, const Vi<D>& pos, int k) {
        Idx p = g->child_of(pos);
        track(val, p, g->leaf_of(p, pos), k);
    }
    /* Lookup::track (PartitionedMixin.hpp:388-413). */
    void ltrack(Idx p, Idx l, int k) {
        track_child(p, k);
        kids[p].ltrack(l, k);
    }
    void ltrack(const Vi<D>& pos, int k) {
        Idx p = g->child_of(pos);
        ltrack(p, g->leaf_of(p, pos), k);
    }
    /* Untrack::untrack (PartitionedMixin.hpp:489-516). */
    void untrack(T bg, Idx p, Idx l, int k) {
        Kid<T>& kid = kids[p];
        kid.luntrack(l, k);
        kid.data[l] = bg;
        if (kid.lists[k].size() == 0) {
            {
                std::lock_guard<std::mutex> lock(mtx);
                puntrack(p, k);
            }
            if (!ptracked(p)) {
                kid.bg = bg;
                deactivate(p);
            }
        }
    }
    /* Untrack::retrack (PartitionedMixin.hpp:520-558). */
    void retrack(Idx p, Idx l, int from, int to) {
        Kid<T>& kid = kids[p];
        kid.luntrack(l, from);
        kid.ltrack(l, to);
        if (!ptracked(p, to) || kid.lists[from].size() == 0) {
            std::lock_guard<std::mutex> lock(mtx);
            ptrack(p, to);
            if (kid.lists[from].size() == 0) puntrack(p, from);
        }
    }
    /* Reset::MultiList::track_children (PartitionedMixin.hpp:319-334). */
    template <class Mask>
    void track_children(const Mask& mask) {
        for (int k = 0; k < mask.NL; k++)
            for (Idx p : mask.plists[k]) {
                if (ptracked(p, k)) continue;
                if (!kids[p].active) activate(p);
                ptrack(p, k);
            }
    }
    /* Reset::MultiList::reset (PartitionedMixin.hpp:336-360): iterates each parent list while
     * removing from it, visiting the stale tail as std::vector iteration would. */
    template <class Mask>
    void reset(const Mask& mask) {
        for (int k = 0; k < NL; k++) {
            Idx n0 = plists[k].size();
            for (Idx i = 0; i < n0; i++) {
                Idx p = plists[k].mem[i];
                puntrack(p, k);
                if (!mask.ptracked(p)) deactivate(p);
                else reset_kid(p);
            }
        }
    }
    T get(const Vi<D>& pos) const {                      // PartitionedMixin.hpp:594-601
        if (!g->inside(pos)) return grid_bg;
        Idx p = g->child_of(pos);
        return kids[p].get(g->leaf_of(p, pos));
    }
    void set(const Vi<D>& pos, T v) {
        Idx p = g->child_of(pos);
        kids[p].data[g->leaf_of(p, pos)] = v;
    }
    template <class Fn>
    void leafs(int k, Fn&& fn) const {                   // PartitionedMixin.hpp:176-191
        for (Idx p : plists[k])
            for (Idx l : kids[p].lists[k]) fn(g->pos_of(p, l));
    }
};

/* ---- cereal binary archive of an IsoGrid (Surface.hpp:180-196, PartitionedMixin.hpp:692-761) ----
 * cereal's BinaryOutputArchive writes arithmetic values as their raw bytes and container sizes as
 * uint64 (cereal/macros.hpp:61-70, types/vector.hpp:40-85); Felt writes an Eigen vector as int32
 * rows, int32 cols and a std::vector of its elements (PartitionedMixin.hpp:719-735). */
struct ArcOut {
    std::vector<uint8_t> b;
    template <class T>
    void put(T v) {
        const uint8_t* p = reinterpret_cast<const uint8_t*>(&v);
        b.insert(b.end(), p, p + sizeof(T));
    }
    template <class T>
    void put_n(const T* v, size_t n) {
        const uint8_t* p = reinterpret_cast<const uint8_t*>(v);
        b.insert(b.end(), p, p + n * sizeof(T));
    }
    template <int D>
    void vecdi(const Vi<D>& v) {                          // Eigen VecDi
        put<int32_t>(D); put<int32_t>(1); put<uint64_t>(D);
        for (int a = 0; a < D; a++) put<int32_t>(v[a]);
    }
    void lists(const std::vector<RawList>& ls, int nl) {  // Tuple<PosIdxList, N>
        put<int32_t>(nl); put<int32_t>(1); put<uint64_t>(uint64_t(nl));
        for (int k = 0; k < nl; k++) {
            const Idx n = k < int(ls.size()) ? ls[k].size() : 0;
            put<uint64_t>(n);
            for (Idx i = 0; i < n; i++) put<uint64_t>(ls[k][i]);
        }
    }
};

struct ArcIn {
    const uint8_t* p;
    size_t n, at = 0;
    bool ok = true;
    template <class T>
    T get() {
        T v{};
        if (at + sizeof(T) > n) { ok = false; return v; }
        std::memcpy(&v, p + at, sizeof(T));
        at += sizeof(T);
        return v;
    }
    template <class T>
    void get_n(std::vector<T>& out, uint64_t cnt) {
        if (!ok || cnt > (n - at) / sizeof(T)) { ok = false; return; }
        out.resize(size_t(cnt));
        std::memcpy(out.data(), p + at, size_t(cnt) * sizeof(T));
        at += size_t(cnt) * sizeof(T);
    }
    template <int D>
    Vi<D> vecdi() {
        Vi<D> v{};
        if (get<int32_t>() != D || get<int32_t>() != 1 || get<uint64_t>() != uint64_t(D)) ok = false;
        for (int a = 0; a < D && ok; a++) v[a] = get<int32_t>();
        return v;
    }
    std::vector<std::vector<uint64_t>> lists(int nl) {
        std::vector<std::vector<uint64_t>> ls;
        if (get<int32_t>() != nl || get<int32_t>() != 1 || get<uint64_t>() != uint64_t(nl)) {
            ok = false;
            return ls;
        }
        ls.resize(size_t(nl));
        for (int k = 0; k < nl && ok; k++) get_n(ls[size_t(k)], get<uint64_t>());
        return ls;
    }
};

/* Dimension and list count of an archive: the child size is the first Eigen vector, and the
 * background child's list tuple gives N = 2L+1. */
inline bool archive_shape(const uint8_t* src, size_t n, int* D, int* NL) {
    ArcIn in{src, n};
    in.get<float>();
    const int32_t rows = in.get<int32_t>();
    if (!in.ok || (rows != 2 && rows != 3)) return false;
    in.at += 4 + 8 + 4 * size_t(rows);                  // cols, count, elements
    in.get<float>();                                    // background child: bg
    std::vector<float> d;
    in.get_n(d, in.get<uint64_t>());                    // its data
    in.get<uint64_t>();                                 // its lookup's background
    std::vector<uint64_t> l;
    in.get_n(l, in.get<uint64_t>());                    // its lookup data
    const int32_t nl = in.get<int32_t>();
    if (!in.ok || nl < 3 || nl % 2 == 0) return false;
    *D = rows;
    *NL = nl;
    return true;
}

/* Child size (the archive's first Eigen vector) and grid size (the isogrid's Size, written just
 * before Spatial's dx and two vectors at the end). */
template <int D>
bool archive_geometry(const uint8_t* src, size_t n, Vi<D>* size, Vi<D>* part) {
    const size_t V = 16 + 4 * size_t(D);
    if (n < 4 + V + 4 + 2 * V + 3 * V) return false;
    ArcIn a{src, n, 4};
    *part = a.vecdi<D>();
    ArcIn b{src, n, n - (4 + 2 * V) - 3 * V};
    *size = b.vecdi<D>();
    return a.ok && b.ok;
}

inline long long g_round_trace = 0, g_round_cap = 0, g_round_dump = 0;

/* SplitMix64 (Steele, Lea & Flood 2014). */
inline uint64_t splitmix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ULL;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ULL;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBULL;
    return x ^ (x >> 31);
}

template <int D, class R = float>
class Surface {
public:
    using Vf = Vr<D, R>;
    int L, NL;
    Geom<D> g;
    PartGrid<D, R, true> iso;          // IsoGrid, background L+1
    PartGrid<D, R, true> delta_;       // DeltaIsoGrid, background 0
    PartGrid<D, int, true> status;     // StatusChangeGrid, background L+1
    PartGrid<D, char, false> aff;      // AffectedLookupGrid
    PartGrid<D, char, false> affbuf;
    uint64_t nan_events = 0;
    uint64_t last_updates = 0, last_rounds = 0;
    bool omp = true;

    Surface(int L_, const Vi<D>& size, const Vi<D>& part) : L(L_), NL(2 * L_ + 1), g(size, part) {
        iso.init(&g, NL, R(L + 1));
        delta_.init(&g, NL, R(0));
        status.init(&g, NL, L + 1);
        aff.init(&g, NL, 0);
        affbuf.init(&g, NL, 0);
    }

    int layer_idx(int id) const { return id + L; }
    bool inside_band(int id) const { return std::abs(id) <= L; }
    /* iround(v + FLT_EPSILON) in the working precision (Surface.hpp:1694-1697). */
    int layer_id(R v) {
        R t = v + R(FLT_EPSILON);
        if (std::isnan(t)) {
#pragma omp atomic
            nan_events++;
            return L + 1;
        }
        return int(std::round(t));
    }
    int layer_id(const Vi<D>& pos) { return layer_id(iso.get(pos)); }
    /* The same without counting NaN (read-only queries such as ray). */
    int layer_id_peek(R v) const {
        R t = v + R(FLT_EPSILON);
        if (std::isnan(t)) return L + 1;
        return int(std::round(t));
    }

    template <class Fn>
    static void neighs(Vi<D> pos, Fn&& fn) {            // NumericMixin.hpp:489-500
        for (int a = 0; a < D; a++) {
            pos[a] -= 1; fn(pos);
            pos[a] += 2; fn(pos);
            pos[a] -= 1;
        }
    }

    /* ---- seed (Surface.hpp:207-241) ---- */
    void seed(const Vi<D>& c) {
        Vi<D> lo{}, win{};
        for (int a = 0; a < D; a++) { lo[a] = c[a] - L; win[a] = 2 * L + 1; }
        Idx n = 1;
        for (int a = 0; a < D; a++) n *= Idx(win[a]);
        for (Idx i = 0; i <= n; i++) {                  // inclusive bound as in the reference
            Vi<D> pos = unlin<D>(i, win);
            int dist = 0;
            for (int a = 0; a < D; a++) { pos[a] += lo[a]; dist += std::abs(pos[a] - c[a]); }
            int k = layer_id(R(dist));
            if (inside_band(k) && g.inside(pos)) iso.track(R(dist), pos, layer_idx(k));
        }
    }

    /* Dense load (the rule felt_surface_load_dense documents). */
    void load_dense(const float* phi) {
        std::vector<char> has(g.np, 0);
        Idx total = 1;
        for (int a = 0; a < D; a++) total *= Idx(g.size[a]);
        for (Idx i = 0; i < total; i++) {
            Vi<D> pos = unlin<D>(i, g.size);
            for (int a = 0; a < D; a++) pos[a] += g.offset[a];
            if (inside_band(layer_id(R(phi[i])))) has[g.child_of(pos)] = 1;
        }
        std::vector<char> seen(g.np, 0);
        for (Idx i = 0; i < total; i++) {
            Vi<D> pos = unlin<D>(i, g.size);
            for (int a = 0; a < D; a++) pos[a] += g.offset[a];
            Idx p = g.child_of(pos);
            if (!seen[p]) {
                seen[p] = 1;
                if (has[p]) iso.activate(p);
                else iso.kids[p].bg = phi[i] < 0 ? R(-(L + 1)) : R(L + 1);
            }
            if (has[p]) iso.kids[p].data[g.leaf_of(p, pos)] = R(phi[i]);
        }
        for (Idx i = 0; i < total; i++) {
            Vi<D> pos = unlin<D>(i, g.size);
            for (int a = 0; a < D; a++) pos[a] += g.offset[a];
            int k = layer_id(R(phi[i]));
            if (inside_band(k)) iso.track(R(phi[i]), pos, layer_idx(k));
        }
    }

    /* ---- numeric stencils (NumericMixin.hpp:72-500) ----
     * Each is templated on the position type as the reference's are: an integer position reads
     * get(), a real one reads interp() (NumericMixin.hpp:390-393).  Vector sums are Eigen's
     * v0 + (v1 + v2); derivatives divide by the grid spacing dx (1 for a Surface's isogrid). */
    R at(const Vi<D>& p) const { return iso.get(p); }
    R at(const Vf& p) const { return interp(p); }
    /* Mean curvature by difference of normals (NumericMixin.hpp:72-156). */
    template <class P>
    R curv(const P& pos) const {
        const R c = at(pos);
        P dir = pos;
        Vf nf{}, nb{};
        for (int a = 0; a < D; a++) {
            dir[a] += 1;
            const R va = at(dir) - c;
            R sq = 0;
            for (int b = 0; b < D; b++) {
                if (b == a) continue;
                P dn = dir;
                dn[b] -= 1;
                const R lo = at(dn);
                dn[b] += 2;
                const R hi = at(dn);
                const R v = (hi - lo) / R(2);
                sq += v * v;
            }
            nf[a] = va / std::sqrt(va * va + sq);
            dir[a] -= 1;
        }
        for (int a = 0; a < D; a++) {
            dir[a] -= 1;
            const R va = c - at(dir);
            R sq = 0;
            for (int b = 0; b < D; b++) {
                if (b == a) continue;
                P dn = dir;
                dn[b] -= 1;
                const R lo = at(dn);
                dn[b] += 2;
                const R hi = at(dn);
                const R v = (hi - lo) / R(2);
                sq += v * v;
            }
            nb[a] = va / std::sqrt(va * va + sq);
            dir[a] += 1;
        }
        Vf dn{};
        for (int a = 0; a < D; a++) dn[a] = nf[a] - nb[a];
        return sum(dn) / R(2);
    }
    static R sum(const Vf& v) {
        if (D == 3) return v[0] + (v[1] + v[2]);
        return v[0] + v[1];
    }
    /* N-linear interpolation (NumericMixin.hpp:403-481). */
    R interp(const Vf& pos) const {
        Vi<D> fl{};
        Vf t{};
        for (int a = 0; a < D; a++) { R f = std::floor(pos[a]); fl[a] = int(f); t[a] = pos[a] - f; }
        R vals[1 << D];
        for (int i = 0; i < (1 << D); i++) {
            Vi<D> c = fl;
            for (int a = 0; a < D; a++) c[a] += (i >> a) & 1;
            vals[i] = iso.get(c);
        }
        int n = 1 << D;
        int axis = 0;
        while (n > 1) {
            const R w = t[axis];
            for (int i = 0; i < n / 2; i++) vals[i] = w * vals[2 * i + 1] + (R(1) - w) * vals[2 * i];
            n /= 2;
            axis++;
        }
        return vals[0];
    }
    /* Safe gradient (NumericMixin.hpp:188-231): central where both neighbours are inside the grid,
     * one-sided where one is, 0 where none is; a real position's probes step -1, +2, -1 in working
     * precision as the reference's do. */
    template <class P>
    Vf grad(const P& pos, R dx = R(1)) const {
        Vf gv{};
        P pt = pos;
        const R centre = at(pos);
        for (int a = 0; a < D; a++) {
            R back = centre, fwd = centre;
            int order = 0;
            pt[a] -= 1;
            if (g.inside(pt)) { back = at(pt); order++; }
            pt[a] += 2;
            if (g.inside(pt)) { fwd = at(pt); order++; }
            pt[a] -= 1;
            gv[a] = order != 0 ? (fwd - back) / R(order) : R(0);
        }
        for (int a = 0; a < D; a++) gv[a] = gv[a] / dx;
        return gv;
    }
    /* Upwind gradient (NumericMixin.hpp:242-273): max(c - back, 0) + min(fwd - c, 0) per axis. */
    template <class P>
    Vf gradE(const P& pos, R dx = R(1)) const {
        Vf gv{};
        P pt = pos;
        const R centre = at(pos);
        for (int a = 0; a < D; a++) {
            pt[a] -= 1;
            R back = at(pt);
            pt[a] += 2;
            R fwd = at(pt);
            pt[a] -= 1;
            back = std::max(centre - back, R(0));
            fwd = std::min(fwd - centre, R(0));
            gv[a] = fwd + back;
        }
        for (int a = 0; a < D; a++) gv[a] = gv[a] / dx;
        return gv;
    }
    /* Forward, backward and central differences (NumericMixin.hpp:279-358). */
    template <class P>
    Vf gradF(const P& pos, R dx = R(1)) const {
        Vf gv{};
        P pt = pos;
        const R centre = at(pos);
        for (int a = 0; a < D; a++) {
            pt[a] += 1;
            gv[a] = at(pt) - centre;
            pt[a] -= 1;
        }
        for (int a = 0; a < D; a++) gv[a] = gv[a] / dx;
        return gv;
    }
    template <class P>
    Vf gradB(const P& pos, R dx = R(1)) const {
        Vf gv{};
        P pt = pos;
        const R centre = at(pos);
        for (int a = 0; a < D; a++) {
            pt[a] -= 1;
            gv[a] = centre - at(pt);
            pt[a] += 1;
        }
        for (int a = 0; a < D; a++) gv[a] = gv[a] / dx;
        return gv;
    }
    template <class P>
    Vf gradC(const P& pos, R dx = R(1)) const {
        Vf gv{};
        P pt = pos;
        for (int a = 0; a < D; a++) {
            pt[a] -= 1;
            const R back = at(pt);
            pt[a] += 2;
            const R fwd = at(pt);
            pt[a] -= 1;
            gv[a] = (fwd - back) / R(2);
        }
        for (int a = 0; a < D; a++) gv[a] = gv[a] / dx;
        return gv;
    }
    /* Divergence of the gradient: sum(gradF - gradB) / dx^2 (NumericMixin.hpp:164-175). */
    template <class P>
    R divergence(const P& pos, R dx = R(1)) const {
        const Vf f = gradF(pos, dx), b = gradB(pos, dx);
        Vf d{};
        for (int a = 0; a < D; a++) d[a] = f[a] - b[a];
        return sum(d) / (dx * dx);
    }

    /* ---- raycasting (Surface.hpp:468-616, 1450-1598) ---- */
    static constexpr R TINY = R(0.00001f);                // Surface.hpp:67
    static Vf ray_miss() {                                // Surface.hpp:1805-1807
        Vf m{};
        m.fill(std::numeric_limits<R>::max());
        return m;
    }
    static R dotv(const Vf& a, const Vf& b) {             // Eigen dot: v0 + (v1 + v2)
        Vf p{};
        for (int i = 0; i < D; i++) p[i] = a[i] * b[i];
        return sum(p);
    }
    static void normalize(Vf& n) {
        Vf p{};
        for (int i = 0; i < D; i++) p[i] = n[i] * n[i];
        const R z = sum(p);
        if (z > R(0)) {
            const R r = std::sqrt(z);
            for (int i = 0; i < D; i++) n[i] = n[i] / r;
        }
    }
    bool child_inside(Idx p, const Vf& pos) const {       // Util.hpp:94-106 on the child
        const Vi<D> o = g.child_offset(p);
        for (int a = 0; a < D; a++) {
            if (float(pos[a]) >= float(o[a] + g.part[a])) return false;
            if (float(pos[a]) < float(o[a])) return false;
        }
        return true;
    }
    /* Ray within one child: half-voxel march until the zero layer, then converge. */
    Vf ray_child(const Vf& origin, const Vf& dir, Idx p) const {
        Vf ps = origin;
        R t = R(0);
        while (child_inside(p, ps)) {
            if (layer_id_peek(interp(ps)) == 0) {
                Vf n = grad(ps);
                normalize(n);
                if (dotv(n, dir) < R(0)) {
                    for (int step = 0; step < 100; step++) {
                        const R dist = interp(ps);
                        for (int i = 0; i < D; i++) ps[i] = ps[i] - n[i] * dist;
                        if (!g.inside(ps)) return ray_miss();
                        if (std::abs(dist) <= TINY || dotv(n, dir) >= R(0)) break;
                        n = grad(ps);
                        normalize(n);
                    }
                    return ps;
                }
            }
            t += R(0.5f);
            for (int i = 0; i < D; i++) ps[i] = origin[i] + dir[i] * t;
        }
        return ray_miss();
    }
    R round_to_next(int dim, R dir, R pos) const {
        R v = (pos - R(g.offset[dim])) / R(g.part[dim]);
        v = dir == R(-1) ? std::floor(v) : std::ceil(v);
        v *= R(g.part[dim]);
        v += R(g.offset[dim]);
        return v;
    }
    struct ChildHit {
        Vf pos;
        Idx p;
    };
    /* Intersect the ray with the plane normal . x + d = 0 (normal -dir_dim along dim), step
     * TINY past it, and keep the child there if it tracks layer 0 or +-1. */
    bool ray_check_track_child(std::vector<ChildHit>& hits, const Vf& o, const Vf& dir, int dim,
                               R normal_dim, R d) const {
        const R t = -(d + normal_dim * o[dim]) / (normal_dim * dir[dim]);
        Vf pi{};
        for (int i = 0; i < D; i++) pi[i] = (o[i] + dir[i] * t) + dir[i] * TINY;
        for (int i = 0; i < D; i++)
            if ((dir[i] > R(0) && pi[i] > R(g.size[i])) || (dir[i] < R(0) && pi[i] < R(g.offset[i])))
                return false;
        if (!g.inside(pi)) return true;
        Vi<D> pf{};
        for (int i = 0; i < D; i++) pf[i] = int(std::floor(pi[i]));
        const Idx p = g.child_of(pf);
        const Kid<R>& k = iso.kids[p];
        if (k.lists[layer_idx(0)].size() || k.lists[layer_idx(1)].size() ||
            k.lists[layer_idx(-1)].size())
            hits.push_back(ChildHit{pi, p});
        return true;
    }
    Vf ray(const Vf& o, const Vf& dir) const {
        if (g.inside(o)) {
            Vi<D> oi{};
            for (int a = 0; a < D; a++) oi[a] = int(o[a]);    // Eigen cast: truncation
            const Vf h = ray_child(o, dir, g.child_of(oi));
            if (h != ray_miss()) return h;
        }
        std::vector<ChildHit> hits;
        for (int dim = 0; dim < D; dim++) {
            const R dir_dim = R((R(0) < dir[dim]) - (dir[dim] < R(0)));
            if (dir_dim == R(0)) continue;
            R ppd = round_to_next(dim, dir_dim, o[dim]);
            Vf plane{};
            plane[dim] = ppd;
            if (!g.inside(plane)) {
                R gd;
                if (dir_dim == R(-1)) {
                    gd = R(g.offset[dim] + g.size[dim]);
                    if (ppd < gd) continue;
                } else {
                    gd = R(g.offset[dim]);
                    if (ppd > gd) continue;
                }
                plane[dim] = gd;
            }
            const R normal_dim = -dir_dim;
            ray_check_track_child(hits, o, dir, dim, normal_dim, plane[dim] * dir_dim);
            ppd = round_to_next(dim, dir_dim, plane[dim]);
            if (ppd != plane[dim]) {
                plane[dim] = ppd;
                if (!ray_check_track_child(hits, o, dir, dim, normal_dim, plane[dim] * dir_dim))
                    continue;
            }
            const R cs = R(g.part[dim]);
            while (true) {
                plane[dim] += dir_dim * cs;
                if (!ray_check_track_child(hits, o, dir, dim, normal_dim, plane[dim] * dir_dim))
                    break;
            }
        }
        // front to back (ties keep insertion order), then drop repeats of a child
        auto dist2 = [&](const ChildHit& h) {
            Vf dv{};
            for (int i = 0; i < D; i++) dv[i] = h.pos[i] - o[i];
            return dotv(dv, dv);
        };
        std::stable_sort(hits.begin(), hits.end(), [&](const ChildHit& a, const ChildHit& b) {
            return dist2(a) < dist2(b);
        });
        hits.erase(std::unique(hits.begin(), hits.end(),
                               [](const ChildHit& a, const ChildHit& b) { return a.p == b.p; }),
                   hits.end());
        for (const ChildHit& h : hits) {
            const Vf hit = ray_child(h.pos, dir, h.p);
            if (hit != ray_miss()) return hit;
        }
        return ray_miss();
    }

    /* ---- save / load (Surface.hpp:180-196) ----
     * The isogrid is written as the reference's IsoGrid::serialize does: Access background,
     * Children (child size, then the children grid: its background child, every child, its lookup
     * and its size), Size, Spatial (Partitioned.hpp:151-158).  A child is Activate, Data,
     * LookupInterface, Size (Tracked.hpp:53-60); a lookup is Activate, Data, lists, Size
     * (Lookup.hpp:170-177, 229-236).  The background child's sizes are never set by the reference
     * (uninitialised Eigen vectors); they are written as zeros. */
    void save(std::vector<uint8_t>& out) const {
        ArcOut o;
        const Vi<D> zero{};
        const std::vector<RawList> none;
        auto sizes = [&](const Vi<D>& sz, const Vi<D>& off) {
            Vi<D> end{};
            for (int a = 0; a < D; a++) end[a] = off[a] + sz[a];
            o.vecdi<D>(sz); o.vecdi<D>(off); o.vecdi<D>(end);
        };
        o.put<float>(float(iso.grid_bg));
        o.vecdi<D>(g.part);
        // children grid: background child
        o.put<float>(float(L + 1));
        o.put<uint64_t>(0);
        o.put<uint64_t>(NUL); o.put<uint64_t>(0); o.lists(none, NL);
        o.vecdi<D>(zero); o.vecdi<D>(zero); o.vecdi<D>(zero);
        o.vecdi<D>(zero); o.vecdi<D>(zero); o.vecdi<D>(zero);
        o.put<uint64_t>(g.np);
        for (Idx p = 0; p < g.np; p++) {
            const Kid<R>& k = iso.kids[p];
            const Vi<D> off = g.child_offset(p);
            o.put<float>(float(k.bg));
            o.put<uint64_t>(k.data.size());
            for (R v : k.data) o.put<float>(float(v));
            o.put<uint64_t>(NUL);
            o.put<uint64_t>(k.look.size());
            o.put_n(k.look.data(), k.look.size());
            o.lists(k.lists, NL);
            sizes(g.part, off);
            sizes(g.part, off);
        }
        // children lookup: background tuple, per-partition tuples, parent lists, size
        o.put<int32_t>(NL); o.put<int32_t>(1); o.put<uint64_t>(uint64_t(NL));
        for (int k = 0; k < NL; k++) o.put<uint64_t>(NUL);
        o.put<uint64_t>(g.np);
        for (Idx p = 0; p < g.np; p++) {
            o.put<int32_t>(NL); o.put<int32_t>(1); o.put<uint64_t>(uint64_t(NL));
            for (int k = 0; k < NL; k++) o.put<uint64_t>(iso.pidx[p * NL + k]);
        }
        o.lists(iso.plists, NL);
        sizes(g.csize, g.coff);
        sizes(g.csize, g.coff);
        sizes(g.size, g.offset);
        o.put<float>(1.0f);                                // Spatial: dx, pos_min, pos_max
        Vi<D> hi{};
        for (int a = 0; a < D; a++) hi[a] = g.offset[a] + g.size[a];
        o.vecdi<D>(g.offset); o.vecdi<D>(hi);
        out.swap(o.b);
    }
    /* Fills a fresh surface of the archive's geometry; false if the archive does not match it. */
    bool load(const uint8_t* src, size_t n) {
        ArcIn in{src, n};
        auto sizes = [&](const Vi<D>& sz, const Vi<D>& off, bool check) {
            const Vi<D> a = in.vecdi<D>(), b = in.vecdi<D>(), c = in.vecdi<D>();
            if (!check) return;
            for (int i = 0; i < D; i++)
                if (a[i] != sz[i] || b[i] != off[i] || c[i] != off[i] + sz[i]) in.ok = false;
        };
        iso.grid_bg = R(in.get<float>());
        const Vi<D> part = in.vecdi<D>();
        if (part != g.part) return false;
        std::vector<float> fv;
        std::vector<uint64_t> uv;
        in.get<float>();                                   // background child
        in.get_n(fv, in.get<uint64_t>());
        in.get<uint64_t>();
        in.get_n(uv, in.get<uint64_t>());
        in.lists(NL);
        for (int i = 0; i < 6; i++) in.vecdi<D>();
        if (!in.ok || in.get<uint64_t>() != g.np) return false;
        for (Idx p = 0; p < g.np && in.ok; p++) {
            Kid<R>& k = iso.kids[p];
            k.bg = R(in.get<float>());
            in.get_n(fv, in.get<uint64_t>());
            in.get<uint64_t>();
            in.get_n(uv, in.get<uint64_t>());
            auto ls = in.lists(NL);
            if (!in.ok) return false;
            k.active = !fv.empty();
            if ((k.active && (fv.size() != g.pv || uv.size() != g.pv)) ||
                (!k.active && !uv.empty()))
                return false;
            k.data.assign(fv.begin(), fv.end());
            k.look.assign(uv.begin(), uv.end());
            k.lists.assign(NL, RawList{});
            for (int j = 0; j < NL; j++) {
                if (!k.active && !ls[size_t(j)].empty()) return false;
                for (uint64_t l : ls[size_t(j)]) {
                    if (l >= g.pv) return false;
                    k.lists[j].push_back(Idx(l));
                }
            }
            const Vi<D> off = g.child_offset(p);
            sizes(g.part, off, true);
            sizes(g.part, off, true);
        }
        if (!in.ok) return false;
        if (in.get<int32_t>() != NL || in.get<int32_t>() != 1 || in.get<uint64_t>() != uint64_t(NL))
            return false;
        for (int k = 0; k < NL; k++) in.get<uint64_t>();
        if (in.get<uint64_t>() != g.np) return false;
        for (Idx p = 0; p < g.np && in.ok; p++) {
            if (in.get<int32_t>() != NL || in.get<int32_t>() != 1 ||
                in.get<uint64_t>() != uint64_t(NL))
                return false;
            for (int k = 0; k < NL; k++) iso.pidx[p * NL + k] = Idx(in.get<uint64_t>());
        }
        auto pl = in.lists(NL);
        if (!in.ok) return false;
        iso.plists.assign(NL, RawList{});
        for (int k = 0; k < NL; k++)
            for (uint64_t p : pl[size_t(k)]) {
                if (p >= g.np) return false;
                iso.plists[k].push_back(Idx(p));
            }
        sizes(g.csize, g.coff, true);
        sizes(g.csize, g.coff, true);
        sizes(g.size, g.offset, true);
        in.get<float>();
        in.vecdi<D>(); in.vecdi<D>();
        return in.ok && in.at == n;
    }

    /* ---- update API ---- */
    /* update(fn) (Surface.hpp:254-296). fn(pos, surface) -> delta. */
    template <class Fn>
    void update(Fn&& fn) {
        update_start();
        delta_.track_children(iso);
        const RawList& parts = iso.plists[layer_idx(0)];
        const Idx np = parts.size();
        uint64_t nband = band_size();
#pragma omp parallel for if (omp && np >= OMP_MIN_CHUNK)
        for (Idx i = 0; i < np; i++) {
            const Idx p = parts[i];
            Kid<R>& kid = iso.kids[p];
            Kid<R>& dk = delta_.kids[p];
            for (Idx l : kid.lists[layer_idx(0)]) {
                const R d = fn(g.pos_of(p, l), *this);
                dk.data[l] = d;
                dk.ltrack(l, layer_idx(0));
            }
        }
        update_end_global();
        last_updates = nband;
    }

    /* update(lo, hi, fn) (Surface.hpp:310-386). */
    template <class Fn>
    void update(const Vi<D>& lo, const Vi<D>& hi, Fn&& fn) {
        Vi<D> glo = g.offset, ghi{}, clo{}, chi{}, cbox{}, leaf_hi{};
        for (int a = 0; a < D; a++) ghi[a] = g.offset[a] + g.size[a];
        Vi<D> ghm1{};
        for (int a = 0; a < D; a++) ghm1[a] = ghi[a] - 1;
        const Vi<D> cgl = pos_child(glo), cgh = pos_child(ghm1), cl = pos_child(lo), ch = pos_child(hi);
        Idx nbox = 1;
        for (int a = 0; a < D; a++) {
            clo[a] = std::max(cgl[a], cl[a]);
            chi[a] = std::min(cgh[a], ch[a]);
            cbox[a] = chi[a] - clo[a] + 1;
            leaf_hi[a] = std::min(ghi[a], hi[a] + 1);
        }
        bool empty = false;
        for (int a = 0; a < D; a++) if (cbox[a] <= 0) empty = true;
        if (!empty) for (int a = 0; a < D; a++) nbox *= Idx(cbox[a]);
        else nbox = 0;
        update_start();
#pragma omp parallel for if (omp && nbox >= OMP_MIN_CHUNK)
        for (Idx ci = 0; ci < nbox; ci++) {
            Vi<D> cp = unlin<D>(ci, cbox);
            for (int a = 0; a < D; a++) cp[a] += clo[a];
            const Idx p = g.child_index(cp);
            Kid<R>& kid = iso.kids[p];
            for (Idx l : kid.lists[layer_idx(0)]) {
                const Vi<D> pos = g.pos_of(p, l);
                if (!inside_box(pos, lo, leaf_hi)) continue;
                const R d = fn(pos, *this);
                delta_.track(d, p, l, layer_idx(0));
            }
        }
        update_end();
    }
    static bool inside_box(const Vi<D>& p, const Vi<D>& lo, const Vi<D>& hi) {
        for (int a = 0; a < D; a++) {
            if (float(p[a]) >= float(hi[a])) return false;
            if (float(p[a]) < float(lo[a])) return false;
        }
        return true;
    }
    Vi<D> pos_child(const Vi<D>& pos) const {
        Vi<D> c{};
        for (int a = 0; a < D; a++) {
            c[a] = (pos[a] - g.offset[a]) / g.part[a] + g.coff[a];
        }
        return c;
    }

    /* update_start (Surface.hpp:393-399). */
    void update_start() {
        delta_.reset(iso);
        aff.reset(iso);
        affbuf.reset(iso);
        status.reset(iso);
    }
    /* delta(pos, val) (Surface.hpp:441-459). */
    void delta(const Vi<D>& pos, R val) { delta_.track(val, pos, layer_idx(0)); }

    /* update_end (Surface.hpp:404-430). */
    void update_end() {
        calc_affected();
        uint64_t n = 0;
        for (int k = 0; k < NL; k++)
            for (Idx p : aff.plists[k]) n += aff.kids[p].lists[k].size();
        iso.track_children(aff);
        delta_.track_children(aff);
        update_zero_layer(&affbuf);
        PartGrid<D, char, false>* look = &aff;
        PartGrid<D, char, false>* buf = &affbuf;
        bool changed = update_distance_lookup(look, buf);
        flush_status_change();
        expand_narrow_band();
        uint64_t rounds = 1;
        while (changed) {
            look->reset(iso);
            std::swap(look, buf);
            status.reset(aff);
            changed = update_distance_lookup(look, buf);
            flush_status_change();
            expand_narrow_band();
            rounds++;
            round_check(rounds);
        }
        last_updates = n;
        last_rounds = rounds;
    }

    /* Round diagnostics of the test harness (felt_oracle_set_round_diag): every g_round_trace-th
     * round of a step's round loop is printed to stderr, rounds g_round_dump.. +2 print their
     * status changes, and a loop of more than g_round_cap rounds stops with std::runtime_error
     * (the reference's loop has no bound).  Zero turns each off. */
    void round_check(uint64_t rounds) {
        const long long every = g_round_trace, cap = g_round_cap, dump = g_round_dump;
        if (dump > 0 && rounds >= (uint64_t)dump && rounds < (uint64_t)dump + 3) {
            int shown = 0;
            for (int from = -L; from <= L && shown < 40; from++) {
                const int kf = layer_idx(from);
                for (Idx p : status.plists[kf]) {
                    for (Idx l : status.kids[p].lists[kf]) {
                        if (shown++ >= 40) break;
                        const Vi<D> pos = g.pos_of(p, l);
                        std::fprintf(stderr, "[oracle] round %llu change (",
                                     (unsigned long long)rounds);
                        for (int a = 0; a < D; a++)
                            std::fprintf(stderr, a ? ",%d" : "%d", pos[a]);
                        std::fprintf(stderr, ") %d -> %d, phi %.9g\n", from,
                                     status.kids[p].data[l], double(iso.get(pos)));
                    }
                }
            }
            std::fflush(stderr);
        }
        if (every > 0 && rounds % (uint64_t)every == 0) {
            std::fprintf(stderr, "[oracle] round %llu\n", (unsigned long long)rounds);
            std::fflush(stderr);
        }
        if (cap > 0 && rounds > (uint64_t)cap) throw std::runtime_error("round cap reached");
    }

    /* update_end_global (Surface.hpp:739-762). */
    void update_end_global() {
        update_zero_layer(&affbuf);
        bool changed = update_distance_iso(&affbuf);
        flush_status_change();
        expand_narrow_band();
        uint64_t rounds = 1;
        if (changed) {
            PartGrid<D, char, false>* look = &affbuf;
            PartGrid<D, char, false>* buf = &aff;
            while (changed) {
                status.reset(iso);
                changed = update_distance_lookup(look, buf);
                flush_status_change();
                expand_narrow_band();
                look->reset(iso);
                std::swap(look, buf);
                rounds++;
                round_check(rounds);
            }
        }
        last_rounds = rounds;
    }

    /* calc_affected (Surface.hpp:780-913): L rounds of breadth-first growth over face
     * neighbours inside the band, seeded with the delta list. */
    void calc_affected() {
        const int z = layer_idx(0);
        for (Idx p : delta_.plists[z])
            for (Idx l : delta_.kids[p].lists[z]) aff.ltrack(p, l, z);
        std::vector<std::vector<Idx>> first(NL), last(NL);
        for (int ud = 1; ud <= L; ud++) {
            for (int k = 0; k < NL; k++) {
                const Idx nc = aff.plists[k].size();
                last[k].resize(nc);
                first[k].resize(nc, 0);
                for (Idx ic = 0; ic < nc; ic++) {
                    const Idx p = aff.plists[k][ic];
                    last[k][ic] = aff.kids[p].lists[k].size();
                }
            }
            for (int k = 0; k < NL; k++) {
                for (Idx ic = 0; ic < first[k].size(); ic++) {
                    const Idx p = aff.plists[k][ic];
                    for (Idx in = first[k][ic]; in < last[k][ic]; in++) {
                        const Idx l = aff.kids[p].lists[k][in];
                        const Vi<D> pc = g.pos_of(p, l);
                        neighs(pc, [&](const Vi<D>& q) {
                            const int kq = layer_id(q);
                            if (inside_band(kq)) aff.ltrack(q, layer_idx(kq));
                        });
                    }
                }
            }
            for (int k = 0; k < NL; k++)
                for (Idx ic = 0; ic < first[k].size(); ic++) first[k][ic] = last[k][ic];
        }
    }

    /* update_zero_layer (Surface.hpp:920-979). */
    void update_zero_layer(PartGrid<D, char, false>* buf) {
        const int z = layer_idx(0);
        const RawList& parts = delta_.plists[z];
        const Idx np = parts.size();
#pragma omp parallel for if (omp && np >= OMP_MIN_CHUNK)
        for (Idx i = 0; i < np; i++) {
            const Idx p = parts[i];
            Kid<R>& dk = delta_.kids[p];
            Kid<R>& ik = iso.kids[p];
            for (Idx l : dk.lists[z]) {
                const R prev = ik.data[l];
                const R d = dk.data[l];
                const R nv = prev + d;
                const int knew = layer_id(nv);
                ik.data[l] = nv;
                status_change(p, l, 0, knew, buf);
            }
        }
    }

    /* update_distance over the isogrid's own lists (global first round). */
    bool update_distance_iso(PartGrid<D, char, false>* buf) {
        bool changed = false;
        for (int k = -1; k >= -L; k--) changed |= update_distance_layer(k, -1, iso.plists, iso.kids, buf);
        for (int k = 1; k <= L; k++) changed |= update_distance_layer(k, 1, iso.plists, iso.kids, buf);
        return changed;
    }
    bool update_distance_lookup(PartGrid<D, char, false>* look, PartGrid<D, char, false>* buf) {
        bool changed = false;
        for (int k = -1; k >= -L; k--) changed |= update_distance_layer(k, -1, look->plists, look->kids, buf);
        for (int k = 1; k <= L; k++) changed |= update_distance_layer(k, 1, look->plists, look->kids, buf);
        return changed;
    }
    /* update_distance(layer, side) (Surface.hpp:1014-1125): Jacobi within a layer. */
    template <class KidT>
    bool update_distance_layer(int layer, int side, const std::vector<RawList>& plists,
                               const std::vector<KidT>& kids, PartGrid<D, char, false>* buf) {
        const int k = layer_idx(layer);
        const RawList& parts = plists[k];
        const Idx np = parts.size();
#pragma omp parallel for if (omp && np >= OMP_MIN_CHUNK)
        for (Idx i = 0; i < np; i++) {
            const Idx p = parts[i];
            Kid<R>& dk = delta_.kids[p];
            for (Idx l : kids[p].lists[k]) {
                const R d = distance(p, l, side);
                dk.data[l] = d;
                dk.ltrack(l, k);
            }
        }
        bool changed = false;
#pragma omp parallel for if (omp && np >= OMP_MIN_CHUNK) reduction(|| : changed)
        for (Idx i = 0; i < np; i++) {
            const Idx p = parts[i];
            Kid<R>& ik = iso.kids[p];
            Kid<R>& dk = delta_.kids[p];
            for (Idx l : kids[p].lists[k]) {
                const R d = dk.data[l];
                const int knew = layer_id(d);
                ik.data[l] = d;
                changed = status_change(p, l, layer, knew, buf) || changed;
            }
        }
        return changed;
    }

    /* status_change (Surface.hpp:1138-1178). */
    bool status_change(Idx p, Idx l, int from, int to, PartGrid<D, char, false>* buf) {
        if (from == to) return false;
        Kid<int>& sk = status.kids[p];
        const int cur = sk.get(l);
        if (cur != L + 1) sk.data[l] = to;
        else status.track(to, p, l, layer_idx(from));
        if (inside_band(to)) buf->ltrack(p, l, layer_idx(to));
        return true;
    }

    /* flush_status_change (Surface.hpp:1183-1244). */
    void flush_status_change() {
        for (int from = -L; from <= L; from++) {
            const int kf = layer_idx(from);
            const RawList& parts = status.plists[kf];
            const Idx np = parts.size();
#pragma omp parallel for if (omp && np >= OMP_MIN_CHUNK)
            for (Idx i = 0; i < np; i++) {
                const Idx p = parts[i];
                const Kid<int>& sk = status.kids[p];
                for (Idx l : sk.lists[kf]) {
                    const int to = sk.data[l];
                    if (inside_band(to)) iso.retrack(p, l, kf, layer_idx(to));
                    else iso.untrack(R(to), p, l, kf);
                }
            }
        }
    }

    /* expand_narrow_band (Surface.hpp:1249-1355), serial as in the reference. */
    void expand_narrow_band() {
        for (int layer : {-L, L}) {
            const int k = layer_idx(layer);
            const int side = layer > 0 ? 1 : -1;
            const RawList& parts = status.plists[k];
            for (Idx ic = 0; ic < parts.size(); ic++) {
                const Idx p = parts[ic];
                const Kid<int>& sk = status.kids[p];
                for (Idx l : sk.lists[k]) {
                    if (sk.data[l] != L * side - side) continue;
                    const Vi<D> pos = g.pos_of(p, l);
                    neighs(pos, [&](const Vi<D>& q) {
                        if (!g.inside(q)) return;
                        R dq = iso.get(q);
                        if (inside_band(layer_id(dq))) return;
                        dq = distance(q, dq, R(side));
                        iso.track(dq, q, k);
                    });
                }
            }
        }
    }

    /* distance (Surface.hpp:1364-1440). */
    R distance(Idx p, Idx l, int side) const {
        return distance(g.pos_of(p, l), iso.kids[p].data[l], R(side));
    }
    R distance(const Vi<D>& pos, R dist, R dir) const {
        dist *= dir;
        neighs(pos, [&](const Vi<D>& q) {
            if (!g.inside(q)) return;
            const R dn = iso.get(q);
            if (dn * dir < dist) dist = dn * dir;
        });
        const R dn = dist * dir;
        return dn + dir;
    }

    /* ---- queries ---- */
    uint64_t band_size() const {
        uint64_t n = 0;
        for (int k = 0; k < NL; k++)
            for (Idx p : iso.plists[k]) n += iso.kids[p].lists[k].size();
        return n;
    }
    uint64_t layer_size(int layer) const {
        const int k = layer_idx(layer);
        uint64_t n = 0;
        for (Idx p : iso.plists[k]) n += iso.kids[p].lists[k].size();
        return n;
    }
    /* Surface::Stats (Surface.hpp:1732-1742). */
    template <class G>
    Idx num_active_partitions(const G& grid) const {
        Idx n = 0;
        for (int k = 0; k < NL; k++) n = std::max(n, grid.plists[k].size());
        return n;
    }
};

/* ---- polygonisation ---- */
static const char* const kTri3[256] = {
    "", "083", "019", "183981", "12a", "08312a", "92a029", "2832a8a98",
    "3b2", "0b28b0", "19023b", "1b219b98b", "3a1ba3", "0a108a8ba", "3903b9ba9", "98aa8b",
    "478", "430734", "019847", "419471731", "12a847", "34730412a", "92a902847", "2a9297273794",
    "8473b2", "b47b24204", "90184723b", "47b94b9b2921", "3a13ba784", "1ba14b1047b4", "47890b9bab03", "47b4b99ba",
    "954", "954083", "054150", "854835315", "12a954", "30812a495", "52a542402", "2a5325354348",
    "95423b", "0b208b495", "05401523b", "21525828b485", "a3ba13954", "4950818a18ba", "54050b5bab03", "54858aa8b",
    "978579", "930953573", "078017157", "153357", "978957a12", "a12950530573", "802825857a52", "2a5253357",
    "7957893b2", "95797292027b", "23b018178157", "b21b17715", "958857a13a3b", "5705097b010aba0", "ba0b03a50807570", "ba57b5",
    "a65", "0835a6", "9015a6", "1831985a6", "165261", "165126308", "965906026", "598582526328",
    "23ba65", "b08b20a65", "01923b5a6", "5a61929b298b", "63b653513", "08b0b50515b6", "3b6036065059", "65969bb98",
    "5a6478", "43047365a", "1905a6847", "a65197173794", "612651478", "125526304347", "847905065026", "739794329596269",
    "3b2784a65", "5a647242027b", "01947823b5a6", "9219b294b7b45a6", "8473b53515b6", "51b5b610b7b404b", "059065036b63847", "65969b4797b9",
    "a4964a", "4a649a083", "a01a60640", "83181686461a", "149124264", "308129249264", "024426", "832824426",
    "a49a64b23", "08228b49a4a6", "3b201606461a", "64161a48121b8b1", "964936913b63", "8b1810b61914641", "3b6360064", "648b68",
    "7a678a89a", "0730a709a67a", "a671a7178180", "a67a71173", "126168189867", "269291679093739", "780706602", "732672",
    "23ba68a89867", "20727b09767a9a7", "1801781a767a23b", "b21b17a61671", "896867916b63136", "091b67", "7807063b0b60", "7b6",
    "76b", "308b76", "019b76", "819831b76", "a126b7", "12a3086b7", "2902a96b7", "6b72a3a83a98",
    "723627", "708760620", "276237019", "162186198876", "a76a17137", "a7617a187108", "03707a0a96a7", "76a7a88a9",
    "684b86", "36b306046", "86b846901", "946963931b36", "6846b82a1", "12a30b06b046", "4b846b0292a9", "a93a32943b36463",
    "823842462", "042462", "190234246438", "194142246", "8138618466a1", "a10a06604", "4634386a3039a93", "a946a4",
    "49576b", "083495b76", "50154076b", "b76834354315", "954a1276b", "6b712a083495", "76b54a42a402", "348354325a52b76",
    "723762549", "954086062687", "362376150540", "628687218485158", "954a16176137", "16a176107870954", "40a4a503a6a737a", "76a7a854a48a",
    "6956b9b89", "36b063056095", "0b805b01556b", "6b3635531", "12a95b9b8b56", "0b306b09656912a", "b85b56805a52025", "6b36352a3a53",
    "589528562382", "956960062", "158180568382628", "156216", "13616a386569896", "a10a06950560", "03856a", "a56",
    "b5a75b", "b5ab75830", "5b75ab190", "a75ab7981831", "b12b71751", "08312717572b", "9759279022b7", "75272b592328982",
    "25a235375", "820852875a25", "9015a35373a2", "982921872a25752", "135375", "087071175", "903935537", "987597",
    "5845a8ab8", "5045b05abb30", "01984a8aba45", "ab4a45b34941314", "2512852b8458", "04b0b345b2b151b", "0250592b5458b85", "9452b3",
    "25a352345384", "5a2524420", "3a235a385458019", "5a2524192942", "845853351", "045105", "845853905035", "945",
    "4b749b9ab", "0834979b79ab", "1ab1b414074b", "3143481a474bab4", "4b79b492b912", "9749b791b2b1083", "b74b42240", "b74b42834324",
    "29a279237749", "9a7974a27870207", "37a3a274a1a040a", "1a2874", "491417713", "491417081871", "403743", "487",
    "9a8ab8", "30939bb9a", "01a0a88ab", "31ab3a", "12b1b99b8", "30939b1292b9", "02b80b", "32b",
    "23828aa89", "9a2092", "23828a0181a8", "1a2", "138918", "091", "038", "",
};
static const char* const kSeg2[16] = {
    "", "30", "01", "31", "12", "3012", "02", "32", "23", "20", "2103", "21", "31", "10", "03", ""
};

inline int hexd(char c) { return c <= '9' ? c - '0' : c - 'a' + 10; }

template <int D> struct McGeom;
template <> struct McGeom<3> {
    static constexpr int nedges = 12;
    static constexpr int corners[8][3] = {{0,0,0},{1,0,0},{1,0,-1},{0,0,-1},{0,1,0},{1,1,0},{1,1,-1},{0,1,-1}};
    static constexpr int edges[12][4] = {
        {0,0,0,0},{1,0,-1,2},{0,0,-1,0},{0,0,-1,2},{0,1,0,0},{1,1,-1,2},
        {0,1,-1,0},{0,1,-1,2},{0,0,0,1},{1,0,0,1},{1,0,-1,1},{0,0,-1,1}};
    static constexpr int spx_offset[3] = {0, 0, -1};
    static const char* order(unsigned m) { return kTri3[m]; }
};
template <> struct McGeom<2> {
    static constexpr int nedges = 4;
    static constexpr int corners[4][2] = {{0,0},{1,0},{1,1},{0,1}};
    static constexpr int edges[4][3] = {{0,0,0},{1,0,1},{0,1,0},{0,0,1}};
    static constexpr int spx_offset[2] = {0, 0};
    static const char* order(unsigned m) { return kSeg2[m]; }
};

template <int D>
struct Vertex {
    std::array<float, D> pos{};
    std::array<float, D> norm{};
};
template <int D>
struct Simplex {
    std::array<uint32_t, D> idx{};
};

/* Poly::Single (Poly.hpp:17-368): per-partition polygoniser with an edge-vertex lookup over
 * the partition grown by one voxel on every side. */
template <int D>
struct PolyChild {
    bool active = false;
    Vi<D> off{}, size{};
    std::vector<std::array<Idx, D>> look;
    RawList tracked;              // lookup positions that hold a vertex index
    std::vector<Idx> tracked_at;  // SingleListSingleIdx lookup of those positions
    std::vector<Vertex<D>> vtx;
    std::vector<Simplex<D>> spx;
    Idx npos() const { Idx n = 1; for (int a = 0; a < D; a++) n *= Idx(size[a]); return n; }
    void activate() {
        std::array<Idx, D> nul; nul.fill(NUL);
        look.assign(npos(), nul);
        tracked_at.assign(npos(), NUL);
        active = true;
    }
    void deactivate() {
        look.clear(); look.shrink_to_fit();
        tracked_at.clear(); tracked_at.shrink_to_fit();
        tracked.release();
        vtx.clear(); vtx.shrink_to_fit();
        spx.clear(); spx.shrink_to_fit();
        active = false;
    }
    void reset() {
        std::array<Idx, D> nul; nul.fill(NUL);
        for (Idx i : tracked) { look[i] = nul; tracked_at[i] = NUL; }
        tracked.clear();
        vtx.clear();
        spx.clear();
    }
    Idx at(const Vi<D>& pos) const {
        Vi<D> r{};
        for (int a = 0; a < D; a++) r[a] = pos[a] - off[a];
        return lin<D>(r, size);
    }
};

template <int D, class R = float>
struct Polys {
    Surface<D, R>* s;
    std::vector<PolyChild<D>> kids;
    struct Changes {
        std::vector<Idx> look;
        RawList list;
        void track(Idx p) { if (look[p] != NUL) return; look[p] = list.size(); list.push_back(p); }
        void untrack(Idx p) {
            Idx at = look[p];
            if (at == NUL) return;
            Idx last = list.size() - 1;
            if (at < last) { Idx m = list[last]; list[at] = m; look[m] = at; }
            look[p] = NUL;
            list.pop_back();
        }
        void reset() { for (Idx p : list) look[p] = NUL; list.clear(); }
    };
    Changes a, b;
    Changes* pending = &a;
    Changes* done = &b;

    explicit Polys(Surface<D, R>& surface) : s(&surface) {
        kids.resize(s->g.np);
        for (Idx p = 0; p < s->g.np; p++) {
            Vi<D> o = s->g.child_offset(p);
            for (int a = 0; a < D; a++) { kids[p].off[a] = o[a] - 1; kids[p].size[a] = s->g.part[a] + 2; }
        }
        a.look.assign(s->g.np, NUL);
        b.look.assign(s->g.np, NUL);
    }

    /* notify (Polys.hpp:92-133): outermost list indices 0 and 2L. */
    void notify() {
        const int nl = s->NL;
        for (int k = 0; k <= nl; k += nl - 1) {
            for (Idx p : s->delta_.plists[k]) {
                bool act = kids[p].active;
                for (int k2 = 0; !act && k2 <= nl; k2 += nl - 1) act = s->iso.ptracked(p, k2);
                if (act) pending->track(p);
                else pending->untrack(p);
            }
        }
        for (int k = 0; k <= nl; k += nl - 1)
            for (Idx p : s->status.plists[k])
                if (kids[p].active) pending->track(p);
    }
    /* march (Polys.hpp:138-165). */
    void march() {
        const Idx n = pending->list.size();
#pragma omp parallel for if (s->omp)
        for (Idx i = 0; i < n; i++) {
            const Idx p = pending->list[i];
            PolyChild<D>& kid = kids[p];
            if (s->iso.kids[p].active) {
                if (kid.active) kid.reset(); else kid.activate();
                march_child(p);
            } else if (kid.active) {
                kid.deactivate();
            }
        }
        std::swap(pending, done);
        pending->reset();
    }
    /* invalidate (Polys.hpp:171-191).  The polys children grid's own tracking list is never
     * populated by the reference, so only the isogrid's outermost-layer partitions are marked. */
    void invalidate() {
        pending->reset();
        const int nl = s->NL;
        for (int k = 0; k <= nl; k += nl - 1)
            for (Idx p : s->iso.plists[k]) pending->track(p);
    }

    void march_child(Idx p) {
        const Kid<R>& ik = s->iso.kids[p];
        for (int k = 0; k < s->NL; k++)
            for (Idx l : ik.lists[k]) spx(p, s->g.pos_of(p, l));
    }
    unsigned mask(const Vi<D>& pc) const {
        unsigned m = 0;
        for (int i = 0; i < (1 << D); i++) {
            Vi<D> c = pc;
            for (int a = 0; a < D; a++) c[a] += McGeom<D>::corners[i][a];
            const R v = s->iso.get(c);
            m |= unsigned(v > 0) << i;
        }
        return m;
    }
    void spx(Idx p, const Vi<D>& pos) {
        PolyChild<D>& kid = kids[p];
        Vi<D> pc = pos;
        for (int a = 0; a < D; a++) pc[a] -= McGeom<D>::spx_offset[a];
        const unsigned m = mask(pc);
        const char* ord = McGeom<D>::order(m);
        if (!*ord) return;
        unsigned emask = 0;
        for (const char* c = ord; *c; c++) emask |= 1u << hexd(*c);
        uint32_t vi[12] = {0};
        for (int e = 0; e < McGeom<D>::nedges; e++) {
            if (!((emask >> e) & 1)) continue;
            Vi<D> pa = pc;
            for (int a = 0; a < D; a++) pa[a] += McGeom<D>::edges[e][a];
            vi[e] = uint32_t(idx(kid, pa, McGeom<D>::edges[e][D]));
        }
        for (const char* c = ord; *c; c += D) {
            Simplex<D> sx;
            for (int j = 0; j < D; j++) sx.idx[j] = vi[hexd(c[j])];
            kid.spx.push_back(sx);
        }
    }
    /* Single::idx (Poly.hpp:296-346). */
    Idx idx(PolyChild<D>& kid, const Vi<D>& pa, int axis) {
        const Idx at = kid.at(pa);
        const Idx got = kid.look[at][axis];
        if (got != NUL) return got;
        Vi<D> pb = pa;
        pb[axis] += 1;
        const R va = s->iso.get(pa), vb = s->iso.get(pb);
        const R eps = R(FLT_EPSILON);
        Vertex<D> v;
        if (std::abs(va) <= eps) v = make_vertex(pa);
        else if (std::abs(vb) <= eps) v = make_vertex(pb);
        else {
            R mu;
            if (std::abs(va - vb) <= eps) mu = R(0.5f);
            else mu = va / (va - vb);
            Vr<D, R> c{};
            for (int a = 0; a < D; a++) {
                const R fa = R(pa[a]), fb = R(pb[a]);
                c[a] = fa + (fb - fa) * mu;
            }
            v = make_vertex(c);
        }
        const Idx id = kid.vtx.size();
        kid.vtx.push_back(v);
        kid.look[at][axis] = id;
        if (kid.tracked_at[at] == NUL) { kid.tracked_at[at] = kid.tracked.size(); kid.tracked.push_back(at); }
        return id;
    }
    template <class P>
    Vertex<D> make_vertex(const P& pos) const {
        Vertex<D> v;
        for (int a = 0; a < D; a++) v.pos[a] = float(pos[a]);
        if (D == 3) {
            Vr<D, R> n = s->grad(pos);
            R sq = (D == 3) ? n[0] * n[0] + (n[1] * n[1] + n[D - 1] * n[D - 1]) : R(0);
            if (sq > R(0)) {
                const R r = std::sqrt(sq);
                for (int a = 0; a < D; a++) n[a] = n[a] / r;
            }
            for (int a = 0; a < D; a++) v.norm[a] = float(n[a]);
        }
        return v;
    }
};

}  // namespace felt_oracle
