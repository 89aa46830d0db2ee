// felt_capi.hip — the extern "C" boundary declared in include/felt_amd.h.
//
// Converts the reference's world coordinates to grid coordinates, validates arguments the way
// the reference's debug checks do (Surface.hpp:278-287, 441-459), and forwards to the device
// implementation.  No call falls back to host computation.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "../../include/felt_amd.h"
#include "felt_polys.h"
#include "felt_surface.h"

using namespace felt;

#define API extern "C" __attribute__((visibility("default")))
#define CHECK_S(s) \
    if (!(s)) return felt_fail(FELT_E_ARG, "null surface handle")
#define CHECK_P(p) \
    if (!(p) || !(p)->s) return felt_fail(FELT_E_ARG, "null polys handle")

namespace {

int layer_host(float v) {
    const float t = v + FLT_EPSILON;
    if (std::isnan(t)) return 1 << 20;
    return (int)std::round(t);
}

// World position -> grid coordinates (3 ints, unused axes 0); false outside the grid.
bool to_grid(const Geo& G, const int32_t* w, int* g) {
    bool in = true;
    for (int a = 0; a < 3; a++) {
        if (a < G.D) {
            g[a] = w[a] - G.offset[a];
            if (g[a] < 0 || g[a] >= G.size[a]) in = false;
        } else {
            g[a] = 0;
        }
    }
    return in;
}

int check_deltas(const Geo& G, size_t n, const int32_t* pos, const float* d,
                 std::vector<int>& gpos) {
    gpos.resize(n * 3);
    for (size_t i = 0; i < n; i++) {
        if (!to_grid(G, pos + i * G.D, gpos.data() + i * 3))
            return felt_fail(FELT_E_ARG, "delta position outside the grid");
        const int k = layer_host(d[i]);
        if (k < -1 || k > 1) return felt_fail(FELT_E_DELTA, "delta would jump more than one layer");
    }
    return FELT_OK;
}

}  // namespace

API const char* felt_last_error(void) { return g_felt_err.c_str(); }

API int felt_device_count(int* n) {
    if (!n) return felt_fail(FELT_E_ARG, "null");
    hipError_t e = hipGetDeviceCount(n);
    if (e != hipSuccess) return felt_fail(FELT_E_DEVICE, hipGetErrorString(e));
    return FELT_OK;
}

API int felt_surface_create(int D, int L, const int32_t* size, const int32_t* part,
                            felt_surface** out) {
    if ((D != 2 && D != 3) || L < 1 || L > 8 || !size || !part || !out)
        return felt_fail(FELT_E_ARG, "bad surface arguments (D in {2,3}, 1 <= L <= 8)");
    Geo G{};
    G.D = D;
    G.L = L;
    G.NL = 2 * L + 1;
    bool uneven = false;
    int64_t np = 1, pv = 1;
    for (int a = 0; a < 3; a++) {
        if (a < D) {
            if (size[a] < 1 || part[a] < 1) return felt_fail(FELT_E_ARG, "sizes must be positive");
            G.size[a] = size[a];
            G.part[a] = part[a];
            G.offset[a] = -1 * size[a] / 2;                       // Surface.hpp:1716-1719
            G.csize[a] = size[a] / part[a];
            if (G.csize[a] * part[a] != size[a]) uneven = true;
        } else {
            G.size[a] = G.part[a] = G.csize[a] = 1;
            G.offset[a] = 0;
        }
    }
    if (uneven)                                                    // PartitionedMixin.hpp:131-141
        for (int a = 0; a < D; a++) G.csize[a] += 1;
    for (int a = 0; a < 3; a++) {
        np *= G.csize[a];
        pv *= G.part[a];
    }
    if (np > (1 << 30) || pv > (1 << 24) || np * pv > (int64_t)0xffffffffll)
        return felt_fail(FELT_E_ARG, "grid too large for 32-bit voxel indices");
    G.np = (int)np;
    G.pv = (int)pv;
    G.lstr[2] = 1;
    G.lstr[1] = G.part[2];
    G.lstr[0] = G.part[1] * G.part[2];
    G.cstr[2] = 1;
    G.cstr[1] = G.csize[2];
    G.cstr[0] = G.csize[1] * G.csize[2];
    for (int a = 0; a < 3; a++) {
        G.fpart[a] = make_fastdiv((uint32_t)G.part[a]);
        G.fcsize[a] = make_fastdiv((uint32_t)G.csize[a]);
    }
    G.fpv = make_fastdiv((uint32_t)G.pv);
    felt_surface* s = new (std::nothrow) felt_surface();
    if (!s) return felt_fail(FELT_E_NOMEM, "host allocation failed");
    s->G = G;
    s->cap = G.np;
    int rc = felt_surface_init(s);
    if (rc) {
        felt_surface_release(s);
        delete s;
        return rc;
    }
    *out = s;
    return FELT_OK;
}

API int felt_surface_destroy(felt_surface* s) {
    if (!s) return FELT_OK;
    felt_surface_release(s);
    delete s;
    return FELT_OK;
}

API int felt_surface_geometry(felt_surface* s, int32_t* cs, uint64_t* np) {
    CHECK_S(s);
    for (int a = 0; a < 3; a++) cs[a] = s->G.csize[a];
    *np = (uint64_t)s->G.np;
    return FELT_OK;
}

// seed (Surface.hpp:207-241): L1 distances on the (2L+1)^D window, tracked by layer.
API int felt_surface_seed(felt_surface* s, const int32_t* pos) {
    CHECK_S(s);
    const Geo& G = s->G;
    const int w = 2 * G.L + 1;
    int64_t n = 1;
    for (int a = 0; a < G.D; a++) n *= w;
    std::vector<int> gp;
    std::vector<float> val;
    std::vector<int> lay;
    for (int64_t i = 0; i <= n; i++) {  // inclusive bound as in the reference
        int off[3] = {0, 0, 0};
        int64_t r = i;
        for (int a = G.D - 1; a > 0; a--) {
            off[a] = (int)(r % w);
            r /= w;
        }
        off[0] = (int)(r % w);
        int wp[3] = {0, 0, 0}, dist = 0;
        for (int a = 0; a < G.D; a++) {
            wp[a] = pos[a] - G.L + off[a];
            dist += std::abs(wp[a] - pos[a]);
        }
        const int k = layer_host((float)dist);
        if (k < -G.L || k > G.L) continue;
        int g[3];
        if (!to_grid(G, wp, g)) continue;
        gp.insert(gp.end(), g, g + 3);
        val.push_back((float)dist);
        lay.push_back(k);
    }
    return felt_track_points(s, gp, val, lay);
}

API int felt_surface_load_dense(felt_surface* s, const float* phi, size_t n) {
    CHECK_S(s);
    size_t total = 1;
    for (int a = 0; a < s->G.D; a++) total *= (size_t)s->G.size[a];
    if (n != total || !phi) return felt_fail(FELT_E_ARG, "load_dense: size mismatch");
    return felt_load_dense_impl(s, phi, n);
}

API int felt_surface_update(felt_surface* s, const felt_speed* fn) {
    CHECK_S(s);
    if (!fn || fn->kind < 0 || fn->kind > FELT_SPEED_NOISE_CURV)
        return felt_fail(FELT_E_ARG, "bad speed descriptor");
    return felt_update_global(s, fn);
}

API int felt_surface_update_box(felt_surface* s, const int32_t* lo, const int32_t* hi,
                                const felt_speed* fn) {
    CHECK_S(s);
    if (!fn || !lo || !hi || fn->kind < 0 || fn->kind > FELT_SPEED_NOISE_CURV)
        return felt_fail(FELT_E_ARG, "bad arguments");
    if (fn->kind == FELT_SPEED_NOISE_CURV && fn->cell <= 0)
        return felt_fail(FELT_E_ARG, "noise cell must be positive");
    int l3[3] = {0, 0, 0}, h3[3] = {0, 0, 0};
    for (int a = 0; a < s->G.D; a++) {
        l3[a] = lo[a];
        h3[a] = hi[a];
    }
    return felt_update_box_impl(s, l3, h3, fn);
}

API int felt_surface_update_deltas(felt_surface* s, size_t n, const int32_t* pos,
                                   const float* delta) {
    CHECK_S(s);
    std::vector<int> g;
    if (int rc = check_deltas(s->G, n, pos, delta, g)) return rc;
    return felt_update_global_deltas(s, n, g.data(), delta);
}

API int felt_surface_update_start(felt_surface* s) {
    CHECK_S(s);
    return felt_update_start_impl(s);
}

API int felt_surface_delta(felt_surface* s, size_t n, const int32_t* pos, const float* delta) {
    CHECK_S(s);
    std::vector<int> g;
    if (int rc = check_deltas(s->G, n, pos, delta, g)) return rc;
    return felt_delta_impl(s, n, g.data(), delta);
}

API int felt_surface_update_end(felt_surface* s) {
    CHECK_S(s);
    return felt_update_end_public(s);
}

API int felt_surface_snapshot(felt_surface* s, float* dst, size_t n) {
    CHECK_S(s);
    size_t total = 1;
    for (int a = 0; a < s->G.D; a++) total *= (size_t)s->G.size[a];
    if (n < total) return felt_fail(FELT_E_CAPACITY, "snapshot buffer too small");
    return felt_snapshot_impl(s, dst, total);
}

API int felt_surface_get(felt_surface* s, size_t n, const int32_t* pos, float* out) {
    CHECK_S(s);
    std::vector<int> g(n * 3);
    std::vector<int> outside;
    for (size_t i = 0; i < n; i++)
        if (!to_grid(s->G, pos + i * s->G.D, g.data() + i * 3)) {
            // out-of-grid reads return the grid background; keep them off the device
            for (int a = 0; a < 3; a++) g[i * 3 + a] = -1;
        }
    return felt_get_impl(s, n, g.data(), out);
}

API int felt_surface_layer_size(felt_surface* s, int layer, uint64_t* n) {
    CHECK_S(s);
    const Geo& G = s->G;
    if (layer < -G.L || layer > G.L) return felt_fail(FELT_E_ARG, "layer out of range");
    std::vector<int> cnt((size_t)G.np * G.NL);
    if (int rc = felt_download(s, cnt.data(), s->P.pcnt, cnt.size() * sizeof(int))) return rc;
    uint64_t t = 0;
    for (int p = 0; p < G.np; p++) t += (uint64_t)cnt[(size_t)p * G.NL + layer + G.L];
    *n = t;
    return FELT_OK;
}

static int emit_positions(const Geo& G, const std::vector<int>& p3, int32_t* pos, size_t cap,
                          size_t* n) {
    const size_t m = p3.size() / 3;
    *n = m;
    if (cap < m) return felt_fail(FELT_E_CAPACITY, "position buffer too small");
    for (size_t i = 0; i < m; i++)
        for (int a = 0; a < G.D; a++) pos[i * G.D + a] = p3[i * 3 + a];
    return FELT_OK;
}

API int felt_surface_layer_positions(felt_surface* s, int layer, int32_t* pos, size_t cap,
                                     size_t* n) {
    CHECK_S(s);
    if (layer < -s->G.L || layer > s->G.L) return felt_fail(FELT_E_ARG, "layer out of range");
    std::vector<int> p3;
    if (int rc = felt_collect_layer_impl(s, layer, p3)) return rc;
    return emit_positions(s->G, p3, pos, cap, n);
}

API int felt_surface_zero_layer_in_box(felt_surface* s, const int32_t* lo, const int32_t* hi,
                                       int32_t* pos, size_t cap, size_t* n) {
    CHECK_S(s);
    const Geo& G = s->G;
    std::vector<int> p3;
    if (int rc = felt_collect_layer_impl(s, 0, p3)) return rc;
    std::vector<int> kept;
    for (size_t i = 0; i < p3.size() / 3; i++) {
        bool in = true;
        for (int a = 0; a < G.D; a++) {
            const int hb = std::min(G.offset[a] + G.size[a], hi[a] + 1);
            if (p3[i * 3 + a] < lo[a] || p3[i * 3 + a] >= hb) in = false;
        }
        if (in) kept.insert(kept.end(), p3.begin() + i * 3, p3.begin() + i * 3 + 3);
    }
    return emit_positions(G, kept, pos, cap, n);
}

API int felt_surface_partition_list(felt_surface* s, int kind, int k, uint64_t* out, size_t cap,
                                    size_t* n) {
    CHECK_S(s);
    const Geo& G = s->G;
    if (k < 0 || k >= G.NL) return felt_fail(FELT_E_ARG, "list index out of range");
    std::vector<uint64_t> parts;
    if (kind == FELT_LIST_ISOGRID) {
        std::vector<int> cnt((size_t)G.np * G.NL);
        if (int rc = felt_download(s, cnt.data(), s->P.pcnt, cnt.size() * sizeof(int))) return rc;
        for (int p = 0; p < G.np; p++)
            if (cnt[(size_t)p * G.NL + k] > 0) parts.push_back(p);
    } else if (kind == FELT_LIST_DELTA || kind == FELT_LIST_STATUS_CHANGE) {
        std::vector<unsigned> m(G.np);
        if (int rc = felt_download(s, m.data(), kind == FELT_LIST_DELTA ? s->dmask : s->smask,
                                   m.size() * sizeof(unsigned)))
            return rc;
        for (int p = 0; p < G.np; p++)
            if ((m[p] >> k) & 1u) parts.push_back(p);
    } else if (kind == FELT_LIST_AFFECTED) {
        std::vector<int> p3;
        if (int rc = felt_lu_positions_impl(s, 0, k - G.L, p3)) return rc;
        std::vector<char> seen(G.np, 0);
        for (size_t i = 0; i < p3.size() / 3; i++) {
            int g[3];
            to_grid(G, p3.data() + i * 3, g);
            int p = 0;
            for (int a = 0; a < G.D; a++) p += (g[a] / G.part[a]) * G.cstr[a];
            if (!seen[p]) {
                seen[p] = 1;
                parts.push_back(p);
            }
        }
        std::sort(parts.begin(), parts.end());
    } else {
        return felt_fail(FELT_E_ARG, "unknown list kind");
    }
    *n = parts.size();
    if (cap < parts.size()) return felt_fail(FELT_E_CAPACITY, "list buffer too small");
    std::copy(parts.begin(), parts.end(), out);
    return FELT_OK;
}

API int felt_surface_affected_positions(felt_surface* s, int k, int32_t* pos, size_t cap,
                                        size_t* n) {
    CHECK_S(s);
    if (k < 0 || k >= s->G.NL) return felt_fail(FELT_E_ARG, "list index out of range");
    std::vector<int> p3;
    if (int rc = felt_lu_positions_impl(s, 0, k - s->G.L, p3)) return rc;
    return emit_positions(s->G, p3, pos, cap, n);
}

API int felt_surface_partition_active(felt_surface* s, uint8_t* out, size_t n) {
    CHECK_S(s);
    const Geo& G = s->G;
    if (n < (size_t)G.np) return felt_fail(FELT_E_CAPACITY, "buffer too small");
    std::vector<int> page(G.np);
    if (int rc = felt_download(s, page.data(), s->P.page, page.size() * sizeof(int))) return rc;
    for (int p = 0; p < G.np; p++) out[p] = page[p] >= 0 ? 1 : 0;
    return FELT_OK;
}

API int felt_surface_partition_layer_size(felt_surface* s, uint64_t p, int layer, uint64_t* n) {
    CHECK_S(s);
    const Geo& G = s->G;
    if (p >= (uint64_t)G.np || layer < -G.L || layer > G.L)
        return felt_fail(FELT_E_ARG, "argument out of range");
    int c = 0;
    if (int rc = felt_download(s, &c, s->P.pcnt + p * G.NL + layer + G.L, sizeof(int))) return rc;
    *n = (uint64_t)c;
    return FELT_OK;
}

// Surface::Stats (Surface.hpp:623-629, 1732-1742): the longest parent list of each grid.
API int felt_surface_stats(felt_surface* s, felt_stats* st) {
    CHECK_S(s);
    const Geo& G = s->G;
    std::vector<int> cnt((size_t)G.np * G.NL);
    std::vector<unsigned> dm(G.np);
    if (int rc = felt_download(s, cnt.data(), s->P.pcnt, cnt.size() * sizeof(int))) return rc;
    if (int rc = felt_download(s, dm.data(), s->dmask, dm.size() * sizeof(unsigned))) return rc;
    if (int rc = felt_download(s, &s->hc, s->ctr, sizeof(Ctr))) return rc;
    uint64_t best_i = 0, best_d = 0, band = 0;
    for (int k = 0; k < G.NL; k++) {
        uint64_t ni = 0, nd = 0;
        for (int p = 0; p < G.np; p++) {
            const int c = cnt[(size_t)p * G.NL + k];
            if (c > 0) ni++;
            band += (uint64_t)c;
            if ((dm[p] >> k) & 1u) nd++;
        }
        best_i = std::max(best_i, ni);
        best_d = std::max(best_d, nd);
    }
    st->active_isogrid_partitions = best_i;
    st->active_delta_partitions = best_d;
    st->band_voxels = band;
    st->last_voxel_updates = s->last_updates;
    st->last_rounds = s->last_rounds;
    st->nan_events = (uint64_t)s->hc.nan_events;
    return FELT_OK;
}

API int felt_surface_set_graph(felt_surface* s, int enable) {
    CHECK_S(s);
    s->use_graph = enable != 0;
    return FELT_OK;
}

API int felt_surface_step_info(felt_surface* s, uint64_t* voxel_updates, uint64_t* rounds) {
    CHECK_S(s);
    *voxel_updates = s->last_updates;
    *rounds = s->last_rounds;
    return FELT_OK;
}

API int felt_surface_profile(felt_surface* s, int enable) {
    CHECK_S(s);
    if (int rc = prof_collect(s)) return rc;
    s->prof.on = enable != 0;
    if (s->prof.on) {
        // 3 ms of the wall clock: longer than the host takes to enqueue one step
        int dev = 0, khz = 0;
        hipError_t e = hipGetDevice(&dev);
        if (e == hipSuccess) e = hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev);
        if (e != hipSuccess) return felt_fail(FELT_E_DEVICE, hipGetErrorString(e));
        s->prof.hold_ticks = 3LL * khz;
    }
    for (int k = 0; k < PK_COUNT; k++) {
        s->prof.launches[k] = 0;
        s->prof.ms[k] = 0.0;
        s->prof.items[k] = 0;
    }
    return FELT_OK;
}

API int felt_surface_profile_read(felt_surface* s, int kernel, uint64_t* launches, double* ms,
                                  uint64_t* items) {
    CHECK_S(s);
    if (kernel < 0 || kernel >= PK_COUNT) return felt_fail(FELT_E_ARG, "unknown kernel id");
    if (int rc = prof_collect(s)) return rc;
    *launches = (uint64_t)s->prof.launches[kernel];
    *ms = s->prof.ms[kernel];
    *items = (uint64_t)s->prof.items[kernel];
    return FELT_OK;
}

API int felt_surface_time_pass(felt_surface* s, int pass, int reps, double* us_per_launch,
                               double* items_per_launch, int* launches) {
    CHECK_S(s);
    if (!us_per_launch || !items_per_launch || !launches)
        return felt_fail(FELT_E_ARG, "null output");
    if (int rc = prof_collect(s)) return rc;
    return felt_time_pass_impl(s, pass, reps, us_per_launch, items_per_launch, launches);
}

API int felt_surface_sync(felt_surface* s) {
    CHECK_S(s);
    return felt_sync_impl(s);
}

API int felt_polys_create(felt_surface* s, felt_polys** out) {
    CHECK_S(s);
    if (!out) return felt_fail(FELT_E_ARG, "null output");
    felt_polys* q = new (std::nothrow) felt_polys();
    if (!q) return felt_fail(FELT_E_NOMEM, "host allocation failed");
    q->s = s;
    if (int rc = felt_polys_init(q)) {
        felt_polys_release(q);
        delete q;
        return rc;
    }
    *out = q;
    return FELT_OK;
}

API int felt_polys_destroy(felt_polys* q) {
    if (!q) return FELT_OK;
    felt_polys_release(q);
    delete q;
    return FELT_OK;
}

API int felt_polys_notify(felt_polys* q) {
    CHECK_P(q);
    return felt_polys_notify_impl(q);
}

API int felt_polys_march(felt_polys* q) {
    CHECK_P(q);
    return felt_polys_march_impl(q);
}

API int felt_polys_invalidate(felt_polys* q) {
    CHECK_P(q);
    return felt_polys_invalidate_impl(q);
}

API int felt_polys_mark(felt_polys* q, uint64_t p) {
    CHECK_P(q);
    if (p >= (uint64_t)q->s->G.np) return felt_fail(FELT_E_ARG, "partition out of range");
    q->pending->track((int64_t)p);
    return FELT_OK;
}

API int felt_polys_changes(felt_polys* q, uint64_t* out, size_t cap, size_t* n) {
    CHECK_P(q);
    const std::vector<int64_t>& l = q->done->list;
    *n = l.size();
    if (cap < l.size()) return felt_fail(FELT_E_CAPACITY, "buffer too small");
    for (size_t i = 0; i < l.size(); i++) out[i] = (uint64_t)l[i];
    return FELT_OK;
}

API int felt_polys_child_info(felt_polys* q, uint64_t p, int* active, uint64_t* nv,
                              uint64_t* ns) {
    CHECK_P(q);
    if (p >= (uint64_t)q->s->G.np) return felt_fail(FELT_E_ARG, "partition out of range");
    const PolyKid& k = q->kids[p];
    *active = k.active ? 1 : 0;
    *nv = (uint64_t)k.nv;
    *ns = (uint64_t)k.nt;
    return FELT_OK;
}

API int felt_polys_child_mesh(felt_polys* q, uint64_t p, float* vtx, size_t vcap, uint32_t* spx,
                              size_t scap) {
    CHECK_P(q);
    if (p >= (uint64_t)q->s->G.np) return felt_fail(FELT_E_ARG, "partition out of range");
    return felt_polys_child_mesh_impl(q, p, vtx, vcap, spx, scap);
}

API int felt_polys_totals(felt_polys* q, uint64_t* nv, uint64_t* ns) {
    CHECK_P(q);
    uint64_t v = 0, t = 0;
    for (const PolyKid& k : q->kids) {
        v += (uint64_t)k.nv;
        t += (uint64_t)k.nt;
    }
    *nv = v;
    *ns = t;
    return FELT_OK;
}

API int felt_mc_table(uint64_t* out, size_t n) {
    if (!out || n < 256) return felt_fail(FELT_E_CAPACITY, "need 256 entries");
    felt_mc_table_host(out);
    return FELT_OK;
}
