/*
 * felt_oracle_capi.cpp — TEST INFRASTRUCTURE, NOT PRODUCT CODE.
 *
 * C ABI over the CPU restatement in felt_oracle.hpp.  It mirrors include/felt_amd.h with the
 * prefix felt_oracle_ so that tests can drive the oracle and the HIP library with the same
 * calls, plus a few oracle-only calls (precision choice, thread count, list-order reads).
 */
#include "felt_oracle.hpp"

#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>

#include "../include/felt_amd.h"

using namespace felt_oracle;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
    g_err = msg;
    return code;
}

/* Value noise shared with the device speed functor: lattice values U[-1,1) from SplitMix64,
 * N-linear interpolation reducing axis 0 first, all fp32 without fused multiply-add. */
template <int D>
float value_noise(const Vi<D>& pos, uint64_t seed, int cell) {
    int base[3] = {0, 0, 0};
    float t[3] = {0.f, 0.f, 0.f};
    for (int a = 0; a < D; a++) {
        int q = pos[a] >= 0 ? pos[a] / cell : -((-pos[a] + cell - 1) / cell);
        base[a] = q;
        t[a] = float(pos[a] - q * cell) / float(cell);
    }
    float v[8];
    for (int i = 0; i < (1 << D); i++) {
        uint64_t key = 0;
        for (int a = 0; a < D; a++) {
            static const uint64_t mul[3] = {0x9E3779B97F4A7C15ULL, 0xC2B2AE3D27D4EB4FULL,
                                            0x165667B19E3779F9ULL};
            key += uint64_t(int64_t(base[a] + ((i >> a) & 1))) * mul[a];
        }
        uint64_t h = splitmix64(seed ^ key);
        float u = float(h >> 40) * (1.0f / 16777216.0f);
        v[i] = u * 2.0f - 1.0f;
    }
    int n = 1 << D;
    int axis = 0;
    while (n > 1) {
        const float w = t[axis];
        for (int i = 0; i < n / 2; i++) v[i] = w * v[2 * i + 1] + (1.0f - w) * v[2 * i];
        n /= 2;
        axis++;
    }
    return v[0];
}

template <int D, class R>
R eval_speed(const felt_speed& sp, const Vi<D>& pos, const Surface<D, R>& s) {
    switch (sp.kind) {
    case FELT_SPEED_CONSTANT:
        return R(sp.a);
    case FELT_SPEED_CURVATURE:
        return R(sp.b) * s.curv(pos);
    case FELT_SPEED_CONST_CURV:
        return R(sp.a) + R(sp.b) * s.curv(pos);
    case FELT_SPEED_NOISE_CURV: {
        const R n = R(value_noise<D>(pos, sp.seed, sp.cell));
        R v = R(sp.a) * (R(1) + R(sp.c) * n) + R(sp.b) * s.curv(pos);
        if (v < R(-1)) v = R(-1);
        else if (v > R(1)) v = R(1);
        return v;
    }
    }
    return R(0);
}

struct Handle {
    virtual ~Handle() = default;
    int dim = 3, L = 1;
    virtual int seed(const int32_t* pos) = 0;
    virtual int load_dense(const float* phi, size_t n) = 0;
    virtual int update(const felt_speed& sp) = 0;
    virtual int update_box(const int32_t* lo, const int32_t* hi, const felt_speed& sp) = 0;
    virtual int update_deltas(size_t n, const int32_t* pos, const float* d) = 0;
    virtual int update_start() = 0;
    virtual int delta(size_t n, const int32_t* pos, const float* d) = 0;
    virtual int update_end() = 0;
    virtual int snapshot(float* dst, size_t n) = 0;
    virtual int get(size_t n, const int32_t* pos, float* out) = 0;
    virtual uint64_t layer_size(int layer) = 0;
    virtual int layer_positions(int layer, int32_t* pos, size_t cap, size_t* n) = 0;
    virtual int zero_in_box(const int32_t* lo, const int32_t* hi, int32_t* pos, size_t cap, size_t* n) = 0;
    virtual int partition_list(int kind, int k, uint64_t* out, size_t cap, size_t* n) = 0;
    virtual int affected_positions(int k, int32_t* pos, size_t cap, size_t* n) = 0;
    virtual int partition_active(uint8_t* out, size_t n) = 0;
    virtual int partition_layer_size(uint64_t p, int layer, uint64_t* n) = 0;
    virtual int stats(felt_stats* st) = 0;
    virtual int geometry(int32_t* cs, uint64_t* np) = 0;
    virtual void set_omp(bool on) = 0;
    virtual int ray(size_t n, const float* origin, const float* dir, float* out) = 0;
    virtual int save(std::vector<uint8_t>& out) = 0;
    virtual int stencil(int op, float dx, size_t n, const int32_t* ipos, const float* fpos,
                        float* out) = 0;
    virtual int speed_eval(const felt_speed& sp, size_t n, const int32_t* pos, float* out) = 0;
    virtual bool load(const uint8_t* src, size_t n) = 0;
    /* polys */
    virtual int polys_create() = 0;
    virtual int polys_notify() = 0;
    virtual int polys_march() = 0;
    virtual int polys_invalidate() = 0;
    virtual int polys_mark(uint64_t p) = 0;
    virtual int polys_changes(uint64_t* out, size_t cap, size_t* n) = 0;
    virtual int polys_child_info(uint64_t p, int* active, uint64_t* nv, uint64_t* ns) = 0;
    virtual int polys_child_mesh(uint64_t p, float* vtx, size_t vcap, uint32_t* spx, size_t scap) = 0;
    virtual int polys_totals(uint64_t* nv, uint64_t* ns) = 0;
};

template <int D, class R>
struct HandleT final : Handle {
    Surface<D, R> s;
    std::unique_ptr<Polys<D, R>> polys;

    HandleT(int L_, const Vi<D>& size, const Vi<D>& part) : s(L_, size, part) {
        this->dim = D;
        this->L = L_;
    }
    Vi<D> vec(const int32_t* p) const {
        Vi<D> v{};
        for (int a = 0; a < D; a++) v[a] = p[a];
        return v;
    }
    void put(int32_t* out, const Vi<D>& v) const {
        for (int a = 0; a < D; a++) out[a] = v[a];
    }
    int check_pos(const Vi<D>& v) const {
        if (!s.g.inside(v)) return fail(FELT_E_ARG, "position outside the grid");
        return FELT_OK;
    }
    using Vf = Vr<D, R>;
    /* Grid operator op at n positions (integer ipos or real fpos), felt_surface_stencil. */
    template <class P>
    void stencil_at(int op, R dx, const P& p, float* out) const {
        Vf v{};
        switch (op) {
        case FELT_STENCIL_GET: *out = float(s.at(p)); return;
        case FELT_STENCIL_CURV: *out = float(s.curv(p)); return;
        case FELT_STENCIL_DIVERGENCE: *out = float(s.divergence(p, dx)); return;
        case FELT_STENCIL_GRAD: v = s.grad(p, dx); break;
        case FELT_STENCIL_GRAD_E: v = s.gradE(p, dx); break;
        case FELT_STENCIL_GRAD_F: v = s.gradF(p, dx); break;
        case FELT_STENCIL_GRAD_B: v = s.gradB(p, dx); break;
        default: v = s.gradC(p, dx); break;
        }
        for (int a = 0; a < D; a++) out[a] = float(v[a]);
    }
    int stencil(int op, float dx, size_t n, const int32_t* ipos, const float* fpos,
                float* out) override {
        if (op < FELT_STENCIL_GET || op > FELT_STENCIL_GRAD_C) return fail(FELT_E_ARG, "unknown stencil operator");
        const size_t w = op >= FELT_STENCIL_GRAD ? size_t(D) : 1;
        for (size_t i = 0; i < n; i++) {
            if (fpos) {
                Vf p{};
                for (int a = 0; a < D; a++) p[a] = R(fpos[i * D + a]);
                stencil_at(op, R(dx), p, out + i * w);
            } else {
                stencil_at(op, R(dx), vec(ipos + i * D), out + i * w);
            }
        }
        return FELT_OK;
    }
    int speed_eval(const felt_speed& sp, size_t n, const int32_t* pos, float* out) override {
        for (size_t i = 0; i < n; i++) {
            const Vi<D> p = vec(pos + i * D);
            if (int rc = check_pos(p)) return rc;
            out[i] = float(eval_speed<D, R>(sp, p, s));
        }
        return FELT_OK;
    }
    int seed(const int32_t* pos) override {
        s.seed(vec(pos));
        return FELT_OK;
    }
    int load_dense(const float* phi, size_t n) override {
        size_t total = 1;
        for (int a = 0; a < D; a++) total *= size_t(s.g.size[a]);
        if (n != total) return fail(FELT_E_ARG, "load_dense: size mismatch");
        s.load_dense(phi);
        return FELT_OK;
    }
    int update(const felt_speed& sp) override {
        s.update([&](const Vi<D>& pos, const Surface<D, R>& ss) { return eval_speed<D, R>(sp, pos, ss); });
        return FELT_OK;
    }
    int update_box(const int32_t* lo, const int32_t* hi, const felt_speed& sp) override {
        s.update(vec(lo), vec(hi),
                 [&](const Vi<D>& pos, const Surface<D, R>& ss) { return eval_speed<D, R>(sp, pos, ss); });
        return FELT_OK;
    }
    int check_zero_layer(size_t n, const int32_t* pos, const float* d) {
        for (size_t i = 0; i < n; i++) {
            Vi<D> v = vec(pos + i * D);
            if (check_pos(v)) return FELT_E_ARG;
            const Idx p = s.g.child_of(v);
            const Kid<R>& kid = s.iso.kids[p];
            if (!kid.active || kid.lookup(s.g.leaf_of(p, v)) == NUL ||
                s.layer_id(kid.get(s.g.leaf_of(p, v))) != 0)
                return fail(FELT_E_DELTA, "delta at a position outside the zero layer");
            const int kd = s.layer_id(R(d[i]));
            if (kd < -1 || kd > 1) return fail(FELT_E_DELTA, "delta would jump more than one layer");
        }
        return FELT_OK;
    }
    int update_deltas(size_t n, const int32_t* pos, const float* d) override {
        if (int rc = check_zero_layer(n, pos, d)) return rc;
        /* map positions to values; fn returns 0 elsewhere */
        std::vector<std::pair<Idx, float>> m;
        m.reserve(n);
        for (size_t i = 0; i < n; i++) {
            Vi<D> v = vec(pos + i * D);
            m.emplace_back(lin<D>(offsetless(v), s.g.size), d[i]);
        }
        std::sort(m.begin(), m.end());
        s.update([&](const Vi<D>& q, const Surface<D, R>&) {
            const Idx key = lin<D>(offsetless(q), s.g.size);
            auto it = std::lower_bound(m.begin(), m.end(), std::make_pair(key, -INFINITY));
            if (it != m.end() && it->first == key) return R(it->second);
            return R(0);
        });
        return FELT_OK;
    }
    Vi<D> offsetless(const Vi<D>& v) const {
        Vi<D> r{};
        for (int a = 0; a < D; a++) r[a] = v[a] - s.g.offset[a];
        return r;
    }
    /* ray(origin, dir) per ray (Surface.hpp:468-616); a miss gives ray_miss (FLT_MAX). */
    int ray(size_t n, const float* origin, const float* dir, float* out) override {
        for (size_t i = 0; i < n; i++) {
            typename Surface<D, R>::Vf o{}, d{};
            for (int a = 0; a < D; a++) {
                o[a] = R(origin[i * D + a]);
                d[a] = R(dir[i * D + a]);
            }
            const auto h = s.ray(o, d);
            for (int a = 0; a < D; a++) out[i * D + a] = float(h[a]);
        }
        return FELT_OK;
    }
    int save(std::vector<uint8_t>& out) override {
        s.save(out);
        return FELT_OK;
    }
    bool load(const uint8_t* src, size_t n) override { return s.load(src, n); }
    int update_start() override {
        s.update_start();
        return FELT_OK;
    }
    int delta(size_t n, const int32_t* pos, const float* d) override {
        if (int rc = check_zero_layer(n, pos, d)) return rc;
        for (size_t i = 0; i < n; i++) s.delta(vec(pos + i * D), R(d[i]));
        return FELT_OK;
    }
    int update_end() override {
        s.update_end();
        return FELT_OK;
    }
    int snapshot(float* dst, size_t n) override {
        size_t total = 1;
        for (int a = 0; a < D; a++) total *= size_t(s.g.size[a]);
        if (n < total) return fail(FELT_E_CAPACITY, "snapshot buffer too small");
        for (size_t i = 0; i < total; i++) {
            Vi<D> v = unlin<D>(i, s.g.size);
            for (int a = 0; a < D; a++) v[a] += s.g.offset[a];
            dst[i] = float(s.iso.get(v));
        }
        return FELT_OK;
    }
    int get(size_t n, const int32_t* pos, float* out) override {
        for (size_t i = 0; i < n; i++) out[i] = float(s.iso.get(vec(pos + i * D)));
        return FELT_OK;
    }
    uint64_t layer_size(int layer) override { return s.layer_size(layer); }
    int layer_positions(int layer, int32_t* pos, size_t cap, size_t* n) override {
        if (layer < -L || layer > L) return fail(FELT_E_ARG, "layer out of range");
        const int k = s.layer_idx(layer);
        size_t cnt = s.layer_size(layer);
        *n = cnt;
        if (cap < cnt) return fail(FELT_E_CAPACITY, "capacity");
        size_t i = 0;
        s.iso.leafs(k, [&](const Vi<D>& v) { put(pos + D * i++, v); });
        return FELT_OK;
    }
    int zero_in_box(const int32_t* lo, const int32_t* hi, int32_t* pos, size_t cap, size_t* n) override {
        Vi<D> l = vec(lo), h = vec(hi), hb{};
        for (int a = 0; a < D; a++) hb[a] = std::min(s.g.offset[a] + s.g.size[a], h[a] + 1);
        std::vector<Vi<D>> found;
        s.iso.leafs(s.layer_idx(0), [&](const Vi<D>& v) {
            if (Surface<D, R>::inside_box(v, l, hb)) found.push_back(v);
        });
        *n = found.size();
        if (cap < found.size()) return fail(FELT_E_CAPACITY, "capacity");
        for (size_t i = 0; i < found.size(); i++) put(pos + D * i, found[i]);
        return FELT_OK;
    }
    template <class G>
    int plist(const G& grid, int k, uint64_t* out, size_t cap, size_t* n) {
        if (k < 0 || k >= s.NL) return fail(FELT_E_ARG, "list index out of range");
        const RawList& lst = grid.plists[k];
        *n = lst.size();
        if (cap < lst.size()) return fail(FELT_E_CAPACITY, "capacity");
        for (size_t i = 0; i < lst.size(); i++) out[i] = lst[i];
        return FELT_OK;
    }
    int partition_list(int kind, int k, uint64_t* out, size_t cap, size_t* n) override {
        switch (kind) {
        case FELT_LIST_ISOGRID: return plist(s.iso, k, out, cap, n);
        case FELT_LIST_DELTA: return plist(s.delta_, k, out, cap, n);
        case FELT_LIST_STATUS_CHANGE: return plist(s.status, k, out, cap, n);
        case FELT_LIST_AFFECTED: return plist(s.aff, k, out, cap, n);
        }
        return fail(FELT_E_ARG, "unknown list kind");
    }
    int affected_positions(int k, int32_t* pos, size_t cap, size_t* n) override {
        if (k < 0 || k >= s.NL) return fail(FELT_E_ARG, "list index out of range");
        std::vector<Vi<D>> found;
        s.aff.leafs(k, [&](const Vi<D>& v) { found.push_back(v); });
        *n = found.size();
        if (cap < found.size()) return fail(FELT_E_CAPACITY, "capacity");
        for (size_t i = 0; i < found.size(); i++) put(pos + D * i, found[i]);
        return FELT_OK;
    }
    int partition_active(uint8_t* out, size_t n) override {
        if (n < s.g.np) return fail(FELT_E_CAPACITY, "capacity");
        for (size_t p = 0; p < s.g.np; p++) out[p] = s.iso.kids[p].active ? 1 : 0;
        return FELT_OK;
    }
    int partition_layer_size(uint64_t p, int layer, uint64_t* n) override {
        if (p >= s.g.np || layer < -L || layer > L) return fail(FELT_E_ARG, "argument out of range");
        *n = s.iso.kids[p].lists[s.layer_idx(layer)].size();
        return FELT_OK;
    }
    int stats(felt_stats* st) override {
        st->active_isogrid_partitions = s.num_active_partitions(s.iso);
        st->active_delta_partitions = s.num_active_partitions(s.delta_);
        st->band_voxels = s.band_size();
        st->last_voxel_updates = s.last_updates;
        st->last_rounds = s.last_rounds;
        st->nan_events = s.nan_events;
        return FELT_OK;
    }
    int geometry(int32_t* cs, uint64_t* np) override {
        for (int a = 0; a < 3; a++) cs[a] = a < D ? s.g.csize[a] : 1;
        *np = s.g.np;
        return FELT_OK;
    }
    void set_omp(bool on) override { s.omp = on; }

    int polys_create() override {
        polys.reset(new Polys<D, R>(s));
        return FELT_OK;
    }
    int need_polys() {
        if (!polys) return fail(FELT_E_STATE, "no polys bound to this surface");
        return FELT_OK;
    }
    int polys_notify() override {
        if (int rc = need_polys()) return rc;
        polys->notify();
        return FELT_OK;
    }
    int polys_march() override {
        if (int rc = need_polys()) return rc;
        polys->march();
        return FELT_OK;
    }
    int polys_invalidate() override {
        if (int rc = need_polys()) return rc;
        polys->invalidate();
        return FELT_OK;
    }
    int polys_mark(uint64_t p) override {
        if (int rc = need_polys()) return rc;
        if (p >= s.g.np) return fail(FELT_E_ARG, "partition out of range");
        polys->pending->track(p);
        return FELT_OK;
    }
    int polys_changes(uint64_t* out, size_t cap, size_t* n) override {
        if (int rc = need_polys()) return rc;
        const RawList& lst = polys->done->list;
        *n = lst.size();
        if (cap < lst.size()) return fail(FELT_E_CAPACITY, "capacity");
        for (size_t i = 0; i < lst.size(); i++) out[i] = lst[i];
        return FELT_OK;
    }
    int polys_child_info(uint64_t p, int* active, uint64_t* nv, uint64_t* ns) override {
        if (int rc = need_polys()) return rc;
        if (p >= s.g.np) return fail(FELT_E_ARG, "partition out of range");
        const PolyChild<D>& k = polys->kids[p];
        *active = k.active ? 1 : 0;
        *nv = k.vtx.size();
        *ns = k.spx.size();
        return FELT_OK;
    }
    int polys_child_mesh(uint64_t p, float* vtx, size_t vcap, uint32_t* spx, size_t scap) override {
        if (int rc = need_polys()) return rc;
        if (p >= s.g.np) return fail(FELT_E_ARG, "partition out of range");
        const PolyChild<D>& k = polys->kids[p];
        const size_t vw = D == 3 ? 6 : 2;
        if (vcap < k.vtx.size() * vw || scap < k.spx.size() * D)
            return fail(FELT_E_CAPACITY, "capacity");
        for (size_t i = 0; i < k.vtx.size(); i++) {
            for (int a = 0; a < D; a++) vtx[i * vw + a] = k.vtx[i].pos[a];
            if (D == 3)
                for (int a = 0; a < D; a++) vtx[i * vw + 3 + a] = k.vtx[i].norm[a];
        }
        for (size_t i = 0; i < k.spx.size(); i++)
            for (int a = 0; a < D; a++) spx[i * D + a] = k.spx[i].idx[a];
        return FELT_OK;
    }
    int polys_totals(uint64_t* nv, uint64_t* ns) override {
        if (int rc = need_polys()) return rc;
        uint64_t v = 0, x = 0;
        for (const PolyChild<D>& k : polys->kids) { v += k.vtx.size(); x += k.spx.size(); }
        *nv = v;
        *ns = x;
        return FELT_OK;
    }
};

}  // namespace

struct felt_oracle_surface {
    std::unique_ptr<Handle> h;
};

#define H(x) ((x)->h)
#define OAPI extern "C" __attribute__((visibility("default")))

OAPI const char* felt_oracle_last_error(void) { return g_err.c_str(); }

/* Round-loop diagnostics for the tests: trace every n-th round, stop loops longer than cap
 * rounds (FELT_E_STATE), print the status changes of rounds dump..dump+2; 0 turns each off. */
OAPI int felt_oracle_set_round_diag(long long trace, long long cap, long long dump) {
    g_round_trace = trace;
    g_round_cap = cap;
    g_round_dump = dump;
    return FELT_OK;
}

OAPI int felt_oracle_set_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
    return FELT_OK;
}

OAPI int felt_oracle_surface_create(int D, int L, const int32_t* size, const int32_t* part,
                                    int fp64, felt_oracle_surface** out) {
    if ((D != 2 && D != 3) || L < 1 || L > 8 || !size || !part || !out)
        return fail(FELT_E_ARG, "bad surface arguments");
    for (int a = 0; a < D; a++)
        if (size[a] < 1 || part[a] < 1) return fail(FELT_E_ARG, "bad size");
    auto* s = new felt_oracle_surface;
    if (D == 2) {
        Vi<2> sz{size[0], size[1]}, pt{part[0], part[1]};
        if (fp64) s->h.reset(new HandleT<2, double>(L, sz, pt));
        else s->h.reset(new HandleT<2, float>(L, sz, pt));
    } else {
        Vi<3> sz{size[0], size[1], size[2]}, pt{part[0], part[1], part[2]};
        if (fp64) s->h.reset(new HandleT<3, double>(L, sz, pt));
        else s->h.reset(new HandleT<3, float>(L, sz, pt));
    }
    *out = s;
    return FELT_OK;
}
OAPI int felt_oracle_surface_destroy(felt_oracle_surface* s) { delete s; return FELT_OK; }
template <int D>
static bool sane_geometry(const Vi<D>& sz, const Vi<D>& pt) {
    double np = 1, pv = 1;
    for (int a = 0; a < D; a++) {
        if (sz[a] < 1 || pt[a] < 1 || sz[a] > (1 << 20)) return false;
        np *= double((sz[a] + pt[a] - 1) / pt[a] + 1);
        pv *= double(pt[a]);
    }
    return np <= double(1 << 26) && pv <= double(1 << 24);
}
/* Surface::save: the isogrid's cereal archive; with dst NULL only *n is set. */
OAPI int felt_oracle_surface_save(felt_oracle_surface* s, uint8_t* dst, size_t cap, size_t* n) {
    std::vector<uint8_t> b;
    int rc = H(s)->save(b);
    if (rc) return rc;
    *n = b.size();
    if (!dst) return FELT_OK;
    if (cap < b.size()) return fail(FELT_E_CAPACITY, "archive buffer too small");
    std::memcpy(dst, b.data(), b.size());
    return FELT_OK;
}
/* Surface::load: a new (fp32) surface from a cereal archive of an isogrid. */
OAPI int felt_oracle_surface_load(const uint8_t* src, size_t n, felt_oracle_surface** out) {
    int D = 0, NL = 0;
    if (!src || !out || !archive_shape(src, n, &D, &NL)) return fail(FELT_E_ARG, "not an isogrid archive");
    const int L = (NL - 1) / 2;
    std::unique_ptr<felt_oracle_surface> s(new felt_oracle_surface);
    if (D == 2) {
        Vi<2> sz{}, pt{};
        if (!archive_geometry<2>(src, n, &sz, &pt) || !sane_geometry<2>(sz, pt))
            return fail(FELT_E_ARG, "bad archive geometry");
        s->h.reset(new HandleT<2, float>(L, sz, pt));
    } else {
        Vi<3> sz{}, pt{};
        if (!archive_geometry<3>(src, n, &sz, &pt) || !sane_geometry<3>(sz, pt))
            return fail(FELT_E_ARG, "bad archive geometry");
        s->h.reset(new HandleT<3, float>(L, sz, pt));
    }
    if (!s->h->load(src, n)) return fail(FELT_E_ARG, "archive does not describe a valid isogrid");
    *out = s.release();
    return FELT_OK;
}
OAPI int felt_oracle_surface_set_omp(felt_oracle_surface* s, int on) { H(s)->set_omp(on != 0); return FELT_OK; }
OAPI int felt_oracle_surface_geometry(felt_oracle_surface* s, int32_t* cs, uint64_t* np) { return H(s)->geometry(cs, np); }
OAPI int felt_oracle_surface_seed(felt_oracle_surface* s, const int32_t* pos) { return H(s)->seed(pos); }
OAPI int felt_oracle_surface_load_dense(felt_oracle_surface* s, const float* phi, size_t n) { return H(s)->load_dense(phi, n); }
// A round loop stopped by FELT_ORACLE_MAX_ROUNDS comes back as FELT_E_STATE.
template <class F>
static int guarded(F&& f) {
    try {
        return f();
    } catch (const std::runtime_error& e) {
        return fail(FELT_E_STATE, e.what());
    }
}
OAPI int felt_oracle_surface_update(felt_oracle_surface* s, const felt_speed* fn) {
    return guarded([&] { return H(s)->update(*fn); });
}
OAPI int felt_oracle_surface_update_box(felt_oracle_surface* s, const int32_t* lo, const int32_t* hi, const felt_speed* fn) {
    return guarded([&] { return H(s)->update_box(lo, hi, *fn); });
}
OAPI int felt_oracle_surface_update_deltas(felt_oracle_surface* s, size_t n, const int32_t* pos, const float* d) {
    return H(s)->update_deltas(n, pos, d);
}
OAPI int felt_oracle_surface_update_start(felt_oracle_surface* s) { return H(s)->update_start(); }
OAPI int felt_oracle_surface_delta(felt_oracle_surface* s, size_t n, const int32_t* pos, const float* d) {
    return H(s)->delta(n, pos, d);
}
OAPI int felt_oracle_surface_update_end(felt_oracle_surface* s) { return H(s)->update_end(); }
OAPI int felt_oracle_surface_snapshot(felt_oracle_surface* s, float* dst, size_t n) { return H(s)->snapshot(dst, n); }
OAPI int felt_oracle_surface_get(felt_oracle_surface* s, size_t n, const int32_t* pos, float* out) { return H(s)->get(n, pos, out); }
OAPI int felt_oracle_surface_layer_size(felt_oracle_surface* s, int layer, uint64_t* n) {
    if (layer < -H(s)->L || layer > H(s)->L) return fail(FELT_E_ARG, "layer out of range");
    *n = H(s)->layer_size(layer);
    return FELT_OK;
}
OAPI int felt_oracle_surface_layer_positions(felt_oracle_surface* s, int layer, int32_t* pos, size_t cap, size_t* n) {
    return H(s)->layer_positions(layer, pos, cap, n);
}
OAPI int felt_oracle_surface_zero_layer_in_box(felt_oracle_surface* s, const int32_t* lo, const int32_t* hi, int32_t* pos, size_t cap, size_t* n) {
    return H(s)->zero_in_box(lo, hi, pos, cap, n);
}
OAPI int felt_oracle_surface_partition_list(felt_oracle_surface* s, int kind, int k, uint64_t* out, size_t cap, size_t* n) {
    return H(s)->partition_list(kind, k, out, cap, n);
}
OAPI int felt_oracle_surface_affected_positions(felt_oracle_surface* s, int k, int32_t* pos, size_t cap, size_t* n) {
    return H(s)->affected_positions(k, pos, cap, n);
}
OAPI int felt_oracle_surface_partition_active(felt_oracle_surface* s, uint8_t* out, size_t n) { return H(s)->partition_active(out, n); }
OAPI int felt_oracle_surface_partition_layer_size(felt_oracle_surface* s, uint64_t p, int layer, uint64_t* n) {
    return H(s)->partition_layer_size(p, layer, n);
}
OAPI int felt_oracle_surface_stats(felt_oracle_surface* s, felt_stats* st) { return H(s)->stats(st); }
OAPI int felt_oracle_surface_ray(felt_oracle_surface* s, size_t n, const float* origin, const float* dir, float* out) {
    return H(s)->ray(n, origin, dir, out);
}
OAPI int felt_oracle_surface_sync(felt_oracle_surface*) { return FELT_OK; }

OAPI int felt_oracle_polys_create(felt_oracle_surface* s) { return H(s)->polys_create(); }
OAPI int felt_oracle_polys_notify(felt_oracle_surface* s) { return H(s)->polys_notify(); }
OAPI int felt_oracle_polys_march(felt_oracle_surface* s) { return H(s)->polys_march(); }
OAPI int felt_oracle_polys_invalidate(felt_oracle_surface* s) { return H(s)->polys_invalidate(); }
OAPI int felt_oracle_polys_mark(felt_oracle_surface* s, uint64_t p) { return H(s)->polys_mark(p); }
OAPI int felt_oracle_polys_changes(felt_oracle_surface* s, uint64_t* out, size_t cap, size_t* n) { return H(s)->polys_changes(out, cap, n); }
OAPI int felt_oracle_polys_child_info(felt_oracle_surface* s, uint64_t p, int* active, uint64_t* nv, uint64_t* ns) {
    return H(s)->polys_child_info(p, active, nv, ns);
}
OAPI int felt_oracle_polys_child_mesh(felt_oracle_surface* s, uint64_t p, float* vtx, size_t vcap, uint32_t* spx, size_t scap) {
    return H(s)->polys_child_mesh(p, vtx, vcap, spx, scap);
}
OAPI int felt_oracle_polys_totals(felt_oracle_surface* s, uint64_t* nv, uint64_t* ns) { return H(s)->polys_totals(nv, ns); }

/* Value noise at a position, for tests of the noise definition. */
OAPI float felt_oracle_value_noise(int D, const int32_t* pos, uint64_t seed, int cell) {
    if (D == 2) return value_noise<2>(Vi<2>{pos[0], pos[1]}, seed, cell);
    return value_noise<3>(Vi<3>{pos[0], pos[1], pos[2]}, seed, cell);
}
/* Grid operators of the current isogrid (felt_surface_stencil / _stencil_real). */
OAPI int felt_oracle_surface_stencil(felt_oracle_surface* s, int op, float dx, size_t n,
                                     const int32_t* pos, float* out) {
    return H(s)->stencil(op, dx, n, pos, nullptr, out);
}
OAPI int felt_oracle_surface_stencil_real(felt_oracle_surface* s, int op, float dx, size_t n,
                                          const float* pos, float* out) {
    return H(s)->stencil(op, dx, n, nullptr, pos, out);
}
/* The speed functor at n positions (felt_surface_speed_eval). */
OAPI int felt_oracle_surface_speed_eval(felt_oracle_surface* s, const felt_speed* fn, size_t n,
                                        const int32_t* pos, float* out) {
    return H(s)->speed_eval(*fn, n, pos, out);
}
/* Row-major index of pos in a grid of the given size and offset, and its inverse
 * (Util.hpp:23-83). */
OAPI int felt_oracle_index(int D, const int32_t* size, const int32_t* offset, const int32_t* pos,
                           uint64_t* idx, int32_t* back) {
    if (D == 2) {
        const Vi<2> sz{size[0], size[1]}, p{pos[0] - offset[0], pos[1] - offset[1]};
        *idx = lin<2>(p, sz);
        const Vi<2> q = unlin<2>(Idx(*idx), sz);
        for (int a = 0; a < 2; a++) back[a] = q[a] + offset[a];
    } else if (D == 3) {
        const Vi<3> sz{size[0], size[1], size[2]};
        const Vi<3> p{pos[0] - offset[0], pos[1] - offset[1], pos[2] - offset[2]};
        *idx = lin<3>(p, sz);
        const Vi<3> q = unlin<3>(Idx(*idx), sz);
        for (int a = 0; a < 3; a++) back[a] = q[a] + offset[a];
    } else {
        return fail(FELT_E_ARG, "D must be 2 or 3");
    }
    return FELT_OK;
}
