// Felt::Surface<D,L> and Felt::Polys<S> over the felt C ABI (include/felt_amd.h).
//
// Drop-in for the reference's header-only classes on the update / polygonisation path:
//   Surface(size, size_partition)           Surface.hpp:163-173
//   seed(pos)                               Surface.hpp:207-241
//   update(fn), update(lo, hi, fn)          Surface.hpp:254-386
//   update_start(), delta(pos, v), update_end()   Surface.hpp:393-459
//   isogrid().snapshot(), isogrid().get(pos)      PartitionedMixin.hpp:594-660
//   delta(k), status_change(k), affected(), layer_idx(id), stats(), is_intersected(p)
//                                            Surface.hpp:623-696
//   ray(origin, dir), ray_miss                 Surface.hpp:468-616, 1805-1807
//   save(ostream), load(istream)               Surface.hpp:180-196 (the same cereal archive)
//   Polys(surface), notify(), march(), invalidate(), changes(), children().get(p)
//                                            Polys.hpp:56-201, Poly.hpp:176-189
// Functors: a felt_speed descriptor runs on the device; so does any callable wrapped as
// device_fn(fn), compiled into the caller's own kernel (device.hpp); a host callable with one of
// the reference's call shapes -- fn(pos, isogrid), fn(pos) or fn() returning float -- is
// evaluated on the host for every zero-layer voxel before the step, as the reference does on its
// CPU.
// Errors raise std::domain_error, as the reference does when FELT_EXCEPTIONS is defined.
#pragma once

#include <algorithm>
#include <array>
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <istream>
#include <iterator>
#include <limits>
#include <ostream>
#include <memory>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

#include "../felt_amd.h"
#include "types.hpp"

namespace Felt {

namespace detail {
inline void check(int rc) {
    if (rc != FELT_OK) throw std::domain_error(felt_last_error());
}
template <int D>
VecDi<D> constant(int v) {
    VecDi<D> a{};
    a.fill(v);
    return a;
}
template <int D>
VecDf<D> constant_f(float v) {
    VecDf<D> a{};
    a.fill(v);
    return a;
}
// Calls q(buffer, capacity, &n) for positions of D ints, once more with a larger buffer when
// the first call reports the count it needed.
template <int D, class Q>
std::vector<int32_t> query_positions(Q&& q) {
    std::size_t n = 0;
    std::vector<int32_t> out(64 * D);
    int rc = q(out.data(), out.size() / D, &n);
    if (rc == FELT_E_CAPACITY) {
        out.resize((n + 1) * D);
        rc = q(out.data(), out.size() / D, &n);
    }
    check(rc);
    out.resize(n * D);
    return out;
}
inline PosIdxList partition_list(felt_surface* h, int kind, TupleIdx k) {
    std::size_t n = 0;
    std::vector<uint64_t> buf(64);
    int rc = felt_surface_partition_list(h, kind, k, buf.data(), buf.size(), &n);
    if (rc == FELT_E_CAPACITY) {
        buf.resize(n + 1);
        rc = felt_surface_partition_list(h, kind, k, buf.data(), buf.size(), &n);
    }
    check(rc);
    return PosIdxList(buf.begin(), buf.begin() + std::ptrdiff_t(n));
}
// Row-major index of pos in a grid of `size` at `offset`, the last axis fastest (Util.hpp:23-40),
// and its inverse (Util.hpp:56-80).
template <int D>
PosIdx index_of(const VecDi<D>& pos, const VecDi<D>& size, const VecDi<D>& offset) {
    int64_t idx = 0;
    for (int a = 0; a < D; a++) idx = idx * size[a] + (pos[a] - offset[a]);
    return PosIdx(idx);
}
template <int D>
VecDi<D> position_of(PosIdx idx, const VecDi<D>& size, const VecDi<D>& offset) {
    VecDi<D> pos{};
    for (int a = D - 1; a >= 0; a--) {
        pos[a] = int32_t(idx % PosIdx(size[a])) + offset[a];
        idx /= PosIdx(size[a]);
    }
    return pos;
}
}  // namespace detail

// A functor evaluated on the GPU.  update(device_fn(fn)) and update(lo, hi, device_fn(fn))
// compile fn into a kernel of the caller's own HIP translation unit, which must include
// <felt_amd/device.hpp> and be built with hipcc -ffp-contract=off.  fn takes one of the
// reference's call shapes (Surface.hpp:1754-1802) with the grid as Felt::device::Grid<D>:
// fn(pos, grid), fn(pos) or fn(), returning float.
template <class F>
struct DeviceFn {
    F fn;
};
template <class F>
DeviceFn<std::decay_t<F>> device_fn(F&& f) {
    return DeviceFn<std::decay_t<F>>{std::forward<F>(f)};
}
namespace detail {
template <class T>
struct is_device_fn : std::false_type {};
template <class F>
struct is_device_fn<DeviceFn<F>> : std::true_type {};
// Launches the speed kernel of a device functor for a step's zero layer (device.hpp).
template <int D, class F>
struct DeviceSpeed;
}  // namespace detail

// Host view of the isogrid for host functors and read-back, with the reference's grid API for
// functors (Partitioned.hpp:160-189; NumericMixin.hpp:73-500): get at integer and real
// positions, interp, the gradients, curvature, divergence, dx, inside, neighs.  While a host
// functor runs, reads come from one dense copy taken before the step (the functor sees the grid
// as it was before the update, as in the reference); otherwise each get reads the device.
template <int D>
class IsoGridView {
public:
    IsoGridView(felt_surface* h, const VecDi<D>& size, const VecDi<D>& child_size, int L)
        : m_h(h), m_size(size), m_child_size(child_size), m_L(L) {
        for (int a = 0; a < D; a++) m_offset[a] = -1 * size[a] / 2;
        std::size_t st = 1;
        for (int a = D - 1; a >= 0; a--) {
            m_stride[a] = st;
            st *= std::size_t(size[a]);
        }
        int32_t cs[3] = {1, 1, 1};
        uint64_t np = 0;
        detail::check(felt_surface_geometry(h, cs, &np));
        for (int a = 0; a < D; a++) {
            m_children_size[a] = cs[a];
            m_children_offset[a] = m_offset[a] / child_size[a];  // truncating, as C++ int division
        }
    }
    const VecDi<D>& size() const { return m_size; }
    const VecDi<D>& offset() const { return m_offset; }
    Distance dx() const { return 1.0f; }
    Distance get(const VecDi<D>& pos) const {
        if (m_pinned && inside(pos)) {
            std::size_t i = 0;
            for (int a = 0; a < D; a++) i += std::size_t(pos[a] - m_offset[a]) * m_stride[a];
            return m_cache[i];
        }
        float v = 0.f;
        detail::check(felt_surface_get(m_h, 1, pos.data(), &v));
        return v;
    }
    // Real position: N-linear interpolation (NumericMixin.hpp:390-481).
    Distance get(const VecDf<D>& pos) const { return interp(pos); }
    Distance interp(const VecDf<D>& pos) const {
        VecDi<D> fl{};
        float t[D];
        for (int a = 0; a < D; a++) {
            const float f = std::floor(pos[a]);
            fl[a] = int(f);
            t[a] = pos[a] - f;
        }
        float val[1 << D];
        for (int i = 0; i < (1 << D); i++) {
            VecDi<D> q = fl;
            for (int a = 0; a < D; a++) q[a] += (i >> a) & 1;
            val[i] = get(q);
        }
        int n = 1 << D;
        for (int a = 0; a < D; a++) {
            for (int i = 0; i < n / 2; i++) val[i] = t[a] * val[2 * i + 1] + (1.0f - t[a]) * val[2 * i];
            n /= 2;
        }
        return val[0];
    }
    bool inside(const VecDi<D>& pos) const {
        for (int a = 0; a < D; a++)
            if (pos[a] < m_offset[a] || pos[a] >= m_offset[a] + m_size[a]) return false;
        return true;
    }
    bool inside(const VecDf<D>& pos) const {  // Util.hpp:94-106
        for (int a = 0; a < D; a++)
            if (!(pos[a] < float(m_offset[a] + m_size[a]) && pos[a] >= float(m_offset[a])))
                return false;
        return true;
    }
    // Face neighbours, -1 then +1 along each axis in turn (NumericMixin.hpp:489-500).
    template <class Fn>
    static void neighs(VecDi<D> pos, Fn fn) {
        for (int a = 0; a < D; a++) {
            pos[a] -= 1;
            fn(pos);
            pos[a] += 2;
            fn(pos);
            pos[a] -= 1;
        }
    }
    // Safe gradient: central difference where both neighbours are inside, else one-sided
    // (NumericMixin.hpp:188-231).
    template <class V>
    VecDf<D> grad(const V& pos) const {
        VecDf<D> g{};
        V pt = pos;
        const float centre = get(pos);
        for (int a = 0; a < D; a++) {
            float back = centre, fwd = centre;
            int order = 0;
            pt[a] -= 1;
            if (inside(pt)) {
                back = get(pt);
                order++;
            }
            pt[a] += 2;
            if (inside(pt)) {
                fwd = get(pt);
                order++;
            }
            pt[a] -= 1;
            g[a] = order != 0 ? (fwd - back) / float(order) : 0.0f;
            g[a] = g[a] / dx();
        }
        return g;
    }
    // Upwind (entropy satisfying) gradient (NumericMixin.hpp:243-273).
    template <class V>
    VecDf<D> gradE(const V& pos) const {
        VecDf<D> g{};
        V pt = pos;
        const float centre = get(pos);
        for (int a = 0; a < D; a++) {
            pt[a] -= 1;
            float back = get(pt);
            pt[a] += 2;
            float fwd = get(pt);
            pt[a] -= 1;
            back = std::max(centre - back, 0.0f);
            fwd = std::min(fwd - centre, 0.0f);
            g[a] = (fwd + back) / dx();
        }
        return g;
    }
    // Forward, backward and central differences (NumericMixin.hpp:279-358).
    template <class V>
    VecDf<D> gradF(const V& pos) const {
        VecDf<D> g{};
        V pt = pos;
        const float centre = get(pos);
        for (int a = 0; a < D; a++) {
            pt[a] += 1;
            g[a] = (get(pt) - centre) / dx();
            pt[a] -= 1;
        }
        return g;
    }
    template <class V>
    VecDf<D> gradB(const V& pos) const {
        VecDf<D> g{};
        V pt = pos;
        const float centre = get(pos);
        for (int a = 0; a < D; a++) {
            pt[a] -= 1;
            g[a] = (centre - get(pt)) / dx();
            pt[a] += 1;
        }
        return g;
    }
    template <class V>
    VecDf<D> gradC(const V& pos) const {
        VecDf<D> g{};
        V pt = pos;
        for (int a = 0; a < D; a++) {
            pt[a] -= 1;
            const float back = get(pt);
            pt[a] += 2;
            const float fwd = get(pt);
            pt[a] -= 1;
            g[a] = ((fwd - back) / 2) / dx();
        }
        return g;
    }
    // Mean curvature by the difference of normals (NumericMixin.hpp:73-158); vector sums in
    // Eigen's order, v0 + (v1 + v2).
    template <class V>
    Distance curv(const V& pos) const {
        const float c = get(pos);
        float nf[D], nb[D];
        V dir = pos;
        for (int a = 0; a < D; a++) {
            dir[a] += 1;
            const float va = get(dir) - c;
            float sq = 0;
            for (int b = 0; b < D; b++) {
                if (b == a) continue;
                V dn = dir;
                dn[b] -= 1;
                const float lo = get(dn);
                dn[b] += 2;
                const float hi = get(dn);
                const float t = (hi - lo) / 2;
                sq += t * t;
            }
            nf[a] = va / std::sqrt(va * va + sq);
            dir[a] -= 1;
        }
        for (int a = 0; a < D; a++) {
            dir[a] -= 1;
            const float va = c - get(dir);
            float sq = 0;
            for (int b = 0; b < D; b++) {
                if (b == a) continue;
                V dn = dir;
                dn[b] -= 1;
                const float lo = get(dn);
                dn[b] += 2;
                const float hi = get(dn);
                const float t = (hi - lo) / 2;
                sq += t * t;
            }
            nb[a] = va / std::sqrt(va * va + sq);
            dir[a] += 1;
        }
        float d[D];
        for (int a = 0; a < D; a++) d[a] = nf[a] - nb[a];
        return sum(d) / 2;
    }
    // Divergence of the gradient, sum(gradF - gradB) / dx^2 (NumericMixin.hpp:165-176).
    template <class V>
    Distance divergence(const V& pos) const {
        const VecDf<D> f = gradF(pos), b = gradB(pos);
        float d[D];
        for (int a = 0; a < D; a++) d[a] = f[a] - b[a];
        return sum(d) / (dx() * dx());
    }
    // Dense copy of the isogrid, row-major with the last axis fastest (Grid::Snapshot).
    std::vector<Distance> snapshot() const {
        std::size_t n = 1;
        for (int a = 0; a < D; a++) n *= std::size_t(m_size[a]);
        std::vector<Distance> out(n);
        detail::check(felt_surface_snapshot(m_h, out.data(), out.size()));
        return out;
    }

    // ---- partitions (Partitioned.hpp:160-166; PartitionedMixin.hpp:57-124, 177-244) ----------
    // Each child (partition) has child_size(); child c sits at (c - children().offset()) *
    // child_size() + offset() in world space.
    class ChildView {
    public:
        ChildView(const IsoGridView* g, PosIdx pos_idx) : m_g(g), m_pos_idx(pos_idx) {
            const VecDi<D> c = detail::position_of<D>(pos_idx, g->m_children_size,
                                                      g->m_children_offset);
            for (int a = 0; a < D; a++)
                m_offset[a] = (c[a] - g->m_children_offset[a]) * g->m_child_size[a] +
                              g->m_offset[a];
        }
        const VecDi<D>& size() const { return m_g->m_child_size; }
        const VecDi<D>& offset() const { return m_offset; }
        bool inside(const VecDi<D>& pos) const {
            for (int a = 0; a < D; a++)
                if (pos[a] < m_offset[a] || pos[a] >= m_offset[a] + size()[a]) return false;
            return true;
        }
        PosIdx index(const VecDi<D>& pos) const { return detail::index_of<D>(pos, size(), m_offset); }
        VecDi<D> index(PosIdx idx) const { return detail::position_of<D>(idx, size(), m_offset); }
        // Whether the partition holds storage (Tracked.hpp:16-79); an inactive one reads as its
        // background.
        bool is_active() const { return m_g->children().lookup().is_active(m_pos_idx); }
        Distance get(const VecDi<D>& pos) const { return m_g->get(pos); }
        // Leaf PosIdx (in this child) of the voxels in tracking list k = layer + L.
        PosIdxList list(TupleIdx k) const {
            const std::vector<int32_t> pos = m_g->layer_positions(int(k) - m_g->m_L);
            PosIdxList out;
            for (std::size_t i = 0; i * D < pos.size(); i++) {
                VecDi<D> p{};
                for (int a = 0; a < D; a++) p[a] = pos[i * D + a];
                if (inside(p)) out.push_back(index(p));
            }
            return out;
        }

    private:
        const IsoGridView* m_g;
        PosIdx m_pos_idx;
        VecDi<D> m_offset{};
    };
    class ChildrenView {
    public:
        explicit ChildrenView(const IsoGridView* g) : m_g(g) {}
        const VecDi<D>& size() const { return m_g->m_children_size; }
        const VecDi<D>& offset() const { return m_g->m_children_offset; }
        bool inside(const VecDi<D>& pos) const {
            for (int a = 0; a < D; a++)
                if (pos[a] < offset()[a] || pos[a] >= offset()[a] + size()[a]) return false;
            return true;
        }
        PosIdx index(const VecDi<D>& pos) const { return detail::index_of<D>(pos, size(), offset()); }
        VecDi<D> index(PosIdx idx) const { return detail::position_of<D>(idx, size(), offset()); }
        ChildView get(PosIdx pos_idx) const { return ChildView(m_g, pos_idx); }
        ChildView get(const VecDi<D>& pos) const { return ChildView(m_g, index(pos)); }
        // The children grid's tracking lists: partitions with voxels in list k.
        struct Lookup {
            const IsoGridView* g;
            PosIdxList list(TupleIdx k) const {
                return detail::partition_list(g->m_h, FELT_LIST_ISOGRID, k);
            }
            bool is_tracked(PosIdx pos_idx, TupleIdx k) const {
                uint64_t n = 0;
                detail::check(felt_surface_partition_layer_size(g->m_h, pos_idx,
                                                                int(k) - g->m_L, &n));
                return n != 0;
            }
            bool is_active(PosIdx pos_idx) const {
                std::vector<uint8_t> act(g->num_children());
                detail::check(felt_surface_partition_active(g->m_h, act.data(), act.size()));
                return pos_idx < act.size() && act[pos_idx] != 0;
            }
        };
        Lookup lookup() const { return Lookup{m_g}; }

    private:
        const IsoGridView* m_g;
    };
    ChildrenView children() const { return ChildrenView(this); }
    const VecDi<D>& child_size() const { return m_child_size; }
    // Partition holding a leaf position (PartitionedMixin.hpp:221-244).
    VecDi<D> pos_child(const VecDi<D>& pos_leaf) const {
        VecDi<D> c{};
        for (int a = 0; a < D; a++)
            c[a] = (pos_leaf[a] - m_offset[a]) / m_child_size[a] + m_children_offset[a];
        return c;
    }
    PosIdx pos_idx_child(const VecDi<D>& pos_leaf) const {
        return detail::index_of<D>(pos_child(pos_leaf), m_children_size, m_children_offset);
    }
    // fn(pos) for every voxel of tracking list k = layer + L, partition by partition in the
    // order of children().lookup().list(k) (PartitionedMixin.hpp:177-190); the order within a
    // partition is unspecified, as the reference's tracking order is.
    template <class Fn>
    void leafs(TupleIdx k, Fn&& fn) const {
        const PosIdxList parts = children().lookup().list(k);
        const std::vector<int32_t> pos = layer_positions(int(k) - m_L);
        std::vector<std::pair<PosIdx, std::size_t>> order;
        order.reserve(pos.size() / D);
        for (std::size_t i = 0; i * D < pos.size(); i++) {
            VecDi<D> p{};
            for (int a = 0; a < D; a++) p[a] = pos[i * D + a];
            order.emplace_back(pos_idx_child(p), i);
        }
        std::sort(order.begin(), order.end());
        for (PosIdx pc : parts) {
            auto it = std::lower_bound(order.begin(), order.end(), std::make_pair(pc, std::size_t(0)));
            for (; it != order.end() && it->first == pc; ++it) {
                VecDi<D> p{};
                for (int a = 0; a < D; a++) p[a] = pos[it->second * D + a];
                fn(p);
            }
        }
    }
    std::size_t num_children() const {
        std::size_t n = 1;
        for (int a = 0; a < D; a++) n *= std::size_t(m_children_size[a]);
        return n;
    }
    // Positions tracked in layer `layer` (-L..L), D ints each, order unspecified.
    std::vector<int32_t> layer_positions(int layer) const {
        return detail::query_positions<D>([&](int32_t* out, std::size_t cap, std::size_t* n) {
            return felt_surface_layer_positions(m_h, layer, out, cap, n);
        });
    }

    // Reads from one dense copy until unpin(): for host functors during an update.
    void pin() {
        m_cache = snapshot();
        m_pinned = true;
    }
    void unpin() {
        m_pinned = false;
        std::vector<Distance>().swap(m_cache);
    }

private:
    static float sum(const float* v) { return D == 3 ? v[0] + (v[1] + v[2]) : v[0] + v[1]; }
    felt_surface* m_h;
    VecDi<D> m_size, m_offset{}, m_child_size{}, m_children_size{}, m_children_offset{};
    int m_L;
    std::array<std::size_t, D> m_stride{};
    std::vector<Distance> m_cache;
    bool m_pinned = false;
};

template <int D, LayerId L>
class Surface {
public:
    using IsoGrid = IsoGridView<D>;
    struct Stats {
        ListIdx active_isogrid_partitions;
        ListIdx active_delta_partitions;
    };
    // What ray() returns for a miss: every component the largest float.
    static inline const VecDf<D> ray_miss =
        detail::constant_f<D>(std::numeric_limits<Distance>::max());

    explicit Surface(const VecDi<D>& size, const VecDi<D>& size_partition = detail::constant<D>(8))
        : m_size(size), m_part(size_partition) {
        felt_surface* h = nullptr;
        detail::check(felt_surface_create(D, L, size.data(), size_partition.data(), &h));
        m_h.reset(h);
        m_iso.reset(new IsoGrid(h, size, size_partition, L));
    }

    // Write the isogrid to a stream in the reference's on-disk format (a cereal binary archive).
    void save(std::ostream& out) const {
        std::size_t n = 0;
        detail::check(felt_surface_save(m_h.get(), nullptr, 0, &n));
        std::vector<uint8_t> buf(n);
        detail::check(felt_surface_save(m_h.get(), buf.data(), buf.size(), &n));
        out.write(reinterpret_cast<const char*>(buf.data()), std::streamsize(n));
        out.flush();
    }
    // A surface from such a stream (read to its end).
    static Surface load(std::istream& in) {
        std::vector<uint8_t> buf{std::istreambuf_iterator<char>(in), std::istreambuf_iterator<char>()};
        felt_surface* h = nullptr;
        detail::check(felt_surface_load(buf.data(), buf.size(), &h));
        std::unique_ptr<felt_surface, Del> owned(h);
        int d = 0, l = 0;
        int32_t size[3] = {0, 0, 0}, part[3] = {0, 0, 0};
        detail::check(felt_surface_shape(h, &d, &l, size, part));
        if (d != D || l != L) throw std::domain_error("archive holds a surface of another D or L");
        VecDi<D> sz{}, pt{};
        for (int a = 0; a < D; a++) {
            sz[a] = size[a];
            pt[a] = part[a];
        }
        return Surface(std::move(owned), sz, pt);
    }
    Surface(Surface&&) = default;

    void seed(const VecDi<D>& pos) { detail::check(felt_surface_seed(m_h.get(), pos.data())); }

    void update(const felt_speed& fn) { detail::check(felt_surface_update(m_h.get(), &fn)); }
    void update(const VecDi<D>& lo, const VecDi<D>& hi, const felt_speed& fn) {
        detail::check(felt_surface_update_box(m_h.get(), lo.data(), hi.data(), &fn));
    }
    template <class F>
    void update(DeviceFn<F> f) {
        felt_zero_job job{};
        detail::check(felt_surface_update_begin(m_h.get(), &job));
        run_device_job(job, f.fn);
    }
    template <class F>
    void update(const VecDi<D>& lo, const VecDi<D>& hi, DeviceFn<F> f) {
        felt_zero_job job{};
        detail::check(felt_surface_update_box_begin(m_h.get(), lo.data(), hi.data(), &job));
        run_device_job(job, f.fn);
    }
    template <class Fn, class = std::enable_if_t<!std::is_same<std::decay_t<Fn>, felt_speed>::value &&
                                                 !detail::is_device_fn<std::decay_t<Fn>>::value>>
    void update(Fn&& fn) {
        std::vector<int32_t> pos = layer_positions(0);
        std::vector<float> val = evaluate(fn, pos);
        detail::check(felt_surface_update_deltas(m_h.get(), val.size(), pos.data(), val.data()));
    }
    template <class Fn, class = std::enable_if_t<!std::is_same<std::decay_t<Fn>, felt_speed>::value &&
                                                 !detail::is_device_fn<std::decay_t<Fn>>::value>>
    void update(const VecDi<D>& lo, const VecDi<D>& hi, Fn&& fn) {
        std::vector<int32_t> pos = query([&](int32_t* out, std::size_t cap, std::size_t* n) {
            return felt_surface_zero_layer_in_box(m_h.get(), lo.data(), hi.data(), out, cap, n);
        });
        std::vector<float> val = evaluate(fn, pos);
        update_start();
        if (!val.empty())
            detail::check(felt_surface_delta(m_h.get(), val.size(), pos.data(), val.data()));
        update_end();
    }

    void update_start() { detail::check(felt_surface_update_start(m_h.get())); }
    void delta(const VecDi<D>& pos, Distance val) {
        detail::check(felt_surface_delta(m_h.get(), 1, pos.data(), &val));
    }
    void update_end() { detail::check(felt_surface_update_end(m_h.get())); }

    const IsoGrid& isogrid() const { return *m_iso; }
    PosIdxList delta(TupleIdx layer_idx_) const { return parts(FELT_LIST_DELTA, layer_idx_); }
    PosIdxList status_change(TupleIdx layer_idx_) const {
        return parts(FELT_LIST_STATUS_CHANGE, layer_idx_);
    }
    static constexpr TupleIdx layer_idx(LayerId id) { return TupleIdx(id + L); }
    // affected() (Surface.hpp:693): the affected set of the last bounded update, per layer index
    // as the partitions of its lookup list (children().lookup().list(k)) and as positions.
    struct Affected {
        const Surface* s;
        PosIdxList list(TupleIdx k) const { return s->parts(FELT_LIST_AFFECTED, k); }
        std::vector<int32_t> leafs(TupleIdx k) const {
            return query([&](int32_t* out, std::size_t cap, std::size_t* n) {
                return felt_surface_affected_positions(s->m_h.get(), k, out, cap, n);
            });
        }
    };
    Affected affected() const { return Affected{this}; }
    // Cast a ray to the zero layer; ray_miss when nothing is hit.  dir is a unit vector.
    VecDf<D> ray(const VecDf<D>& origin, const VecDf<D>& dir) const {
        VecDf<D> hit{};
        detail::check(felt_surface_ray(m_h.get(), 1, origin.data(), dir.data(), hit.data()));
        return hit;
    }
    // Many rays in one launch.
    std::vector<VecDf<D>> ray(const std::vector<VecDf<D>>& origins,
                              const std::vector<VecDf<D>>& dirs) const {
        if (origins.size() != dirs.size())
            throw std::domain_error("origins and directions differ in count");
        std::vector<VecDf<D>> hits(origins.size());
        if (hits.empty()) return hits;
        detail::check(felt_surface_ray(m_h.get(), origins.size(), origins.data()->data(),
                                       dirs.data()->data(), hits.data()->data()));
        return hits;
    }
    Stats stats() const {
        felt_stats st{};
        detail::check(felt_surface_stats(m_h.get(), &st));
        return Stats{ListIdx(st.active_isogrid_partitions), ListIdx(st.active_delta_partitions)};
    }
    bool is_intersected(PosIdx pos_idx_child) const {
        uint64_t n = 0;
        detail::check(felt_surface_partition_layer_size(m_h.get(), pos_idx_child, 0, &n));
        return n != 0;
    }
    ListIdx layer_size(LayerId id) const {
        uint64_t n = 0;
        detail::check(felt_surface_layer_size(m_h.get(), id, &n));
        return ListIdx(n);
    }
    std::vector<int32_t> layer_positions(LayerId id) const {
        return query([&](int32_t* out, std::size_t cap, std::size_t* n) {
            return felt_surface_layer_positions(m_h.get(), id, out, cap, n);
        });
    }
    felt_surface* handle() const { return m_h.get(); }

private:
    struct Del {
        void operator()(felt_surface* h) const { felt_surface_destroy(h); }
    };
    // The caller's speed kernel between begin and commit; a failed launch abandons the step.
    template <class F>
    void run_device_job(const felt_zero_job& job, F& fn) {
        const std::string err = detail::DeviceSpeed<D, F>::launch(job, fn);
        if (!err.empty()) {
            felt_surface_update_cancel(m_h.get());
            throw std::domain_error("device functor launch failed: " + err);
        }
        detail::check(felt_surface_update_commit(m_h.get()));
    }
    template <class Q>
    static std::vector<int32_t> query(Q&& q) {
        return detail::query_positions<D>(std::forward<Q>(q));
    }
    PosIdxList parts(int kind, TupleIdx k) const {
        return detail::partition_list(m_h.get(), kind, k);
    }
    static VecDi<D> at(const std::vector<int32_t>& pos, std::size_t i) {
        VecDi<D> v{};
        for (int a = 0; a < D; a++) v[a] = pos[i * D + a];
        return v;
    }
    // The functor at every position of pos; a functor that reads the grid reads one copy of it.
    template <class Fn>
    std::vector<float> evaluate(Fn& fn, const std::vector<int32_t>& pos) {
        std::vector<float> val(pos.size() / D);
        constexpr bool reads_grid = takes_grid<Fn>(0);
        if (reads_grid && !val.empty()) m_iso->pin();
        try {
            for (std::size_t i = 0; i < val.size(); i++) val[i] = call(fn, at(pos, i));
        } catch (...) {
            m_iso->unpin();
            throw;
        }
        m_iso->unpin();
        return val;
    }
    template <class Fn>
    static constexpr auto takes_grid(int)
        -> decltype(std::declval<Fn&>()(std::declval<const VecDi<D>&>(),
                                        std::declval<const IsoGrid&>()),
                    bool()) {
        return true;
    }
    template <class Fn>
    static constexpr bool takes_grid(...) {
        return false;
    }
    // The reference's three functor shapes (Surface.hpp:1754-1802).
    template <class Fn>
    auto call(Fn& fn, const VecDi<D>& p) -> decltype(fn(p, std::declval<const IsoGrid&>())) {
        return fn(p, *m_iso);
    }
    template <class Fn>
    auto call(Fn& fn, const VecDi<D>& p) -> decltype(fn(p)) {
        return fn(p);
    }
    template <class Fn>
    auto call(Fn& fn, const VecDi<D>&) -> decltype(fn()) {
        return fn();
    }

    Surface(std::unique_ptr<felt_surface, Del> h, const VecDi<D>& size, const VecDi<D>& part)
        : m_size(size), m_part(part), m_h(std::move(h)) {
        m_iso.reset(new IsoGrid(m_h.get(), size, part, L));
    }

    VecDi<D> m_size, m_part;
    std::unique_ptr<felt_surface, Del> m_h;
    std::unique_ptr<IsoGrid> m_iso;
};

// Polygonisation of one partition (Poly::Single's read interface, Poly.hpp:176-189).
template <int D>
struct PolyChild {
    struct Vertex {
        VecDf<D> pos;
        VecDf<D> norm;  // zero in 2D
    };
    struct Simplex {
        std::array<uint32_t, D> idxs;
    };
    bool active = false;
    std::vector<Vertex> m_vtx;
    std::vector<Simplex> m_spx;
    bool is_active() const { return active; }
    const std::vector<Vertex>& vtxs() const { return m_vtx; }
    const std::vector<Simplex>& spxs() const { return m_spx; }
};

template <class TSurface>
class Polys;

template <int D, LayerId L>
class Polys<Surface<D, L>> {
public:
    using Child = PolyChild<D>;
    explicit Polys(const Surface<D, L>& surface) {
        felt_polys* p = nullptr;
        detail::check(felt_polys_create(surface.handle(), &p));
        m_p.reset(p);
    }
    void notify() { detail::check(felt_polys_notify(m_p.get())); }
    void march() { detail::check(felt_polys_march(m_p.get())); }
    void invalidate() { detail::check(felt_polys_invalidate(m_p.get())); }
    PosIdxList changes() const {
        std::size_t n = 0;
        std::vector<uint64_t> buf(1024);
        int rc = felt_polys_changes(m_p.get(), buf.data(), buf.size(), &n);
        if (rc == FELT_E_CAPACITY) {
            buf.resize(n + 1);
            rc = felt_polys_changes(m_p.get(), buf.data(), buf.size(), &n);
        }
        detail::check(rc);
        return PosIdxList(buf.begin(), buf.begin() + std::ptrdiff_t(n));
    }
    // children().get(pos_idx) as a copied-out child.
    Child child(PosIdx pos_idx) const {
        int act = 0;
        uint64_t nv = 0, ns = 0;
        detail::check(felt_polys_child_info(m_p.get(), pos_idx, &act, &nv, &ns));
        Child c;
        c.active = act != 0;
        const std::size_t vw = D == 3 ? 6 : 2;
        std::vector<float> v(std::size_t(nv) * vw + 1);
        std::vector<uint32_t> t(std::size_t(ns) * D + 1);
        detail::check(felt_polys_child_mesh(m_p.get(), pos_idx, v.data(), v.size(), t.data(),
                                            t.size()));
        c.m_vtx.resize(std::size_t(nv));
        for (std::size_t i = 0; i < nv; i++)
            for (int a = 0; a < D; a++) {
                c.m_vtx[i].pos[a] = v[i * vw + a];
                c.m_vtx[i].norm[a] = D == 3 ? v[i * vw + 3 + a] : 0.f;
            }
        c.m_spx.resize(std::size_t(ns));
        for (std::size_t i = 0; i < ns; i++)
            for (int a = 0; a < D; a++) c.m_spx[i].idxs[a] = t[i * D + a];
        return c;
    }
    // children() (Polys.hpp:56): get(pos_idx) returns the child as child(pos_idx) does.
    struct Children {
        const Polys* p;
        Child get(PosIdx pos_idx) const { return p->child(pos_idx); }
    };
    Children children() const { return Children{this}; }

private:
    struct Del {
        void operator()(felt_polys* p) const { felt_polys_destroy(p); }
    };
    std::unique_ptr<felt_polys, Del> m_p;
};

// Device speed descriptors (the functor set the device path evaluates).
inline felt_speed speed_constant(float a) { return felt_speed{FELT_SPEED_CONSTANT, a, 0.f, 0.f, 0, 0, 0}; }
inline felt_speed speed_curvature(float b) { return felt_speed{FELT_SPEED_CURVATURE, 0.f, b, 0.f, 0, 0, 0}; }
inline felt_speed speed_const_curv(float a, float b) {
    return felt_speed{FELT_SPEED_CONST_CURV, a, b, 0.f, 0, 0, 0};
}
inline felt_speed speed_noise_curv(float a, float b, float c, uint64_t seed, int cell) {
    return felt_speed{FELT_SPEED_NOISE_CURV, a, b, c, seed, cell, 0};
}

}  // namespace Felt
