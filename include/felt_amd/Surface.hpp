// Felt::Surface<D,L> and Felt::Polys<S> over the felt C ABI (include/felt_amd.h).
//
// Drop-in for the reference's header-only classes on the update / polygonisation path:
//   Surface(size, size_partition)           Surface.hpp:163-173
//   seed(pos)                               Surface.hpp:207-241
//   update(fn), update(lo, hi, fn)          Surface.hpp:254-386
//   update_start(), delta(pos, v), update_end()   Surface.hpp:393-459
//   isogrid().snapshot(), isogrid().get(pos)      PartitionedMixin.hpp:594-660
//   delta(k), status_change(k), layer_idx(id), stats(), is_intersected(p)  Surface.hpp:623-696
//   ray(origin, dir), ray_miss                 Surface.hpp:468-616, 1805-1807
//   save(ostream), load(istream)               Surface.hpp:180-196 (the same cereal archive)
//   Polys(surface), notify(), march(), invalidate(), changes(), children().get(p)
//                                            Polys.hpp:56-201, Poly.hpp:176-189
// Functors: a felt_speed descriptor runs on the device; a host callable with one of the
// reference's call shapes -- fn(pos, isogrid), fn(pos) or fn() returning float -- is evaluated
// on the host for every zero-layer voxel before the step, as the reference does on its CPU.
// Errors raise std::domain_error, as the reference does when FELT_EXCEPTIONS is defined.
#pragma once

#include <array>
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

namespace Felt {

using Distance = float;
using LayerId = int;
using PosIdx = std::size_t;
using TupleIdx = int;
using ListIdx = std::size_t;
using PosIdxList = std::vector<PosIdx>;
template <int D>
using VecDi = std::array<int, D>;
template <int D>
using VecDf = std::array<float, D>;

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
}  // namespace detail

// Host view of the isogrid for host functors and read-back.
template <int D>
class IsoGridView {
public:
    explicit IsoGridView(felt_surface* h, const VecDi<D>& size) : m_h(h), m_size(size) {
        for (int a = 0; a < D; a++) m_offset[a] = -1 * size[a] / 2;
    }
    const VecDi<D>& size() const { return m_size; }
    const VecDi<D>& offset() const { return m_offset; }
    Distance get(const VecDi<D>& pos) const {
        float v = 0.f;
        detail::check(felt_surface_get(m_h, 1, pos.data(), &v));
        return v;
    }
    bool inside(const VecDi<D>& pos) const {
        for (int a = 0; a < D; a++)
            if (pos[a] < m_offset[a] || pos[a] >= m_offset[a] + m_size[a]) return false;
        return true;
    }
    // Dense copy of the isogrid, row-major with the last axis fastest (Grid::Snapshot).
    std::vector<Distance> snapshot() const {
        std::size_t n = 1;
        for (int a = 0; a < D; a++) n *= std::size_t(m_size[a]);
        std::vector<Distance> out(n);
        detail::check(felt_surface_snapshot(m_h, out.data(), out.size()));
        return out;
    }

private:
    felt_surface* m_h;
    VecDi<D> m_size, m_offset{};
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
        m_iso.reset(new IsoGrid(h, size));
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
    template <class Fn, class = std::enable_if_t<!std::is_same<std::decay_t<Fn>, felt_speed>::value>>
    void update(Fn&& fn) {
        std::vector<int32_t> pos = layer_positions(0);
        std::vector<float> val(pos.size() / D);
        for (std::size_t i = 0; i < val.size(); i++) val[i] = call(fn, at(pos, i));
        detail::check(felt_surface_update_deltas(m_h.get(), val.size(), pos.data(), val.data()));
    }
    template <class Fn, class = std::enable_if_t<!std::is_same<std::decay_t<Fn>, felt_speed>::value>>
    void update(const VecDi<D>& lo, const VecDi<D>& hi, Fn&& fn) {
        std::vector<int32_t> pos = query([&](int32_t* out, std::size_t cap, std::size_t* n) {
            return felt_surface_zero_layer_in_box(m_h.get(), lo.data(), hi.data(), out, cap, n);
        });
        std::vector<float> val(pos.size() / D);
        for (std::size_t i = 0; i < val.size(); i++) val[i] = call(fn, at(pos, i));
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
    template <class Q>
    static std::vector<int32_t> query(Q&& q) {
        std::size_t n = 0;
        std::vector<int32_t> out(64 * D);
        int rc = q(out.data(), out.size() / D, &n);
        if (rc == FELT_E_CAPACITY) {
            out.resize((n + 1) * D);
            rc = q(out.data(), out.size() / D, &n);
        }
        detail::check(rc);
        out.resize(n * D);
        return out;
    }
    PosIdxList parts(int kind, TupleIdx k) const {
        std::size_t n = 0;
        std::vector<uint64_t> buf(64);
        int rc = felt_surface_partition_list(m_h.get(), kind, k, buf.data(), buf.size(), &n);
        if (rc == FELT_E_CAPACITY) {
            buf.resize(n + 1);
            rc = felt_surface_partition_list(m_h.get(), kind, k, buf.data(), buf.size(), &n);
        }
        detail::check(rc);
        return PosIdxList(buf.begin(), buf.begin() + std::ptrdiff_t(n));
    }
    static VecDi<D> at(const std::vector<int32_t>& pos, std::size_t i) {
        VecDi<D> v{};
        for (int a = 0; a < D; a++) v[a] = pos[i * D + a];
        return v;
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
        m_iso.reset(new IsoGrid(m_h.get(), size));
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
