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
// A surface accepts calls unless a device-functor step is open on it (felt_surface_update_begin
// until commit or cancel); polys also refuse while an asynchronous batch step is in flight, whose
// later rounds would change partitions after the march.
#define CHECK_S(s)                                                                             \
    if (!(s)) return felt_fail(FELT_E_ARG, "null surface handle");                            \
    if ((s)->job) return felt_fail(FELT_E_STATE, "a device-functor step is open: commit or cancel it")
#define CHECK_P(p)                                                                             \
    if (!(p) || !(p)->s) return felt_fail(FELT_E_ARG, "null polys handle");                  \
    if ((p)->s->job || (p)->s->step_pending)                                                  \
        return felt_fail(FELT_E_STATE, "the surface has a step in flight")

namespace {

int layer_host(float v) {
    const float t = v + FLT_EPSILON;
    if (std::isnan(t)) return 1 << 20;
    return (int)std::round(t);
}

// World position -> member grid coordinates (3 ints, unused axes 0); false outside the grid.
bool to_grid(const Mem& M, int D, const int32_t* w, int* g) {
    bool in = true;
    for (int a = 0; a < 3; a++) {
        if (a < D) {
            g[a] = w[a] - M.offset[a];
            if (g[a] < 0 || g[a] >= M.size[a]) in = false;
        } else {
            g[a] = 0;
        }
    }
    return in;
}

int check_deltas(const felt_surface* s, size_t n, const int32_t* pos, const float* d,
                 std::vector<int>& gpos) {
    gpos.resize(n * 3);
    for (size_t i = 0; i < n; i++) {
        if (!to_grid(s->mems[0], s->G.D, pos + i * s->G.D, gpos.data() + i * 3))
            return felt_fail(FELT_E_ARG, "delta position outside the grid");
        const int k = layer_host(d[i]);
        if (k < -1 || k > 1) return felt_fail(FELT_E_DELTA, "delta would jump more than one layer");
    }
    return FELT_OK;
}

// A pool of n members of sizes[i*D..] sharing D, L and the partition size.  Each member's
// geometry follows Surface(size, part) (Surface.hpp:163-173): offset -size/2, children grid
// size/part, one more partition per axis when any axis does not divide
// (PartitionedMixin.hpp:131-141).  The brick pool holds every partition unless max_bricks
// limits it; voxel ids are 32-bit, so it never exceeds 2^32 / prod(part) bricks.
int make_pool(int D, int L, const int32_t* part, int n, const int32_t* sizes, uint64_t max_bricks,
              felt_surface** out) {
    if ((D != 2 && D != 3) || L < 1 || L > 8 || !sizes || !part || !out || n < 1)
        return felt_fail(FELT_E_ARG, "bad surface arguments (D in {2,3}, 1 <= L <= 8)");
    Geo G{};
    G.D = D;
    G.L = L;
    G.NL = 2 * L + 1;
    G.nmem = n;
    int64_t pv = 1;
    for (int a = 0; a < 3; a++) {
        if (a < D && part[a] < 1) return felt_fail(FELT_E_ARG, "sizes must be positive");
        G.part[a] = a < D ? part[a] : 1;
        pv *= G.part[a];
    }
    if (pv > (1 << 24)) return felt_fail(FELT_E_ARG, "partition too large");
    // a brick's face-neighbour record holds its in-grid extent along axes 0 and 1 in 16 bits
    if (G.part[0] > 65535 || G.part[1] > 65535)
        return felt_fail(FELT_E_ARG, "partition edges along the first two axes are at most 65535");
    G.pv = (int)pv;
    G.lstr[2] = 1;
    G.lstr[1] = G.part[2];
    G.lstr[0] = G.part[1] * G.part[2];
    // brick tiles: 4x4x2 (3D) or 4x8 (2D), each edge cut down to a power of two dividing part
    {
        const int target[2][3] = {{4, 8, 1}, {4, 4, 2}};
        int tv = 1;
        for (int a = 0; a < 3; a++) {
            int t = a < D ? target[D - 2][a] : 1;
            while (t > 1 && G.part[a] % t != 0) t >>= 1;
            G.tl[a] = t;
            G.tmask[a] = t - 1;
            G.tsh[a] = 0;
            while ((1 << G.tsh[a]) < t) G.tsh[a]++;
            tv *= t;
        }
        G.tvsh = 0;
        while ((1 << G.tvsh) < tv) G.tvsh++;
        for (int a = 0; a < 3; a++) {
            int w = 1, ts = tv;
            for (int b = a + 1; b < 3; b++) {
                w *= G.tl[b];
                ts *= G.part[b] / G.tl[b];
            }
            G.wstr[a] = w;
            G.tstr[a] = ts;
            G.fnt[a] = make_fastdiv((uint32_t)(G.part[a] / G.tl[a]));
        }
    }
    for (int a = 0; a < 3; a++) G.fpart[a] = make_fastdiv((uint32_t)G.part[a]);
    G.fpv = make_fastdiv((uint32_t)G.pv);
    std::vector<Mem> mems((size_t)n);
    int64_t np_total = 0;
    for (int i = 0; i < n; i++) {
        Mem& M = mems[i];
        std::memset(&M, 0, sizeof(Mem));
        bool uneven = false;
        for (int a = 0; a < 3; a++) {
            if (a < D) {
                const int sz = sizes[(size_t)i * D + a];
                if (sz < 1) return felt_fail(FELT_E_ARG, "sizes must be positive");
                M.size[a] = sz;
                M.offset[a] = -1 * sz / 2;  // Surface.hpp:1716-1719
                M.csize[a] = sz / G.part[a];
                if (M.csize[a] * G.part[a] != sz) uneven = true;
            } else {
                M.size[a] = M.csize[a] = 1;
                M.offset[a] = 0;
            }
        }
        if (uneven)  // PartitionedMixin.hpp:131-141
            for (int a = 0; a < D; a++) M.csize[a] += 1;
        int64_t np = 1;
        for (int a = 0; a < 3; a++) np *= M.csize[a];
        M.cstr[2] = 1;
        M.cstr[1] = M.csize[2];
        M.cstr[0] = M.csize[1] * M.csize[2];
        for (int a = 0; a < 3; a++) M.fcsize[a] = make_fastdiv((uint32_t)M.csize[a]);
        M.pbase = (int)np_total;
        M.np = (int)std::min<int64_t>(np, 1 << 30);
        np_total += np;
        if (np_total > (1 << 30)) return felt_fail(FELT_E_ARG, "too many partitions");
    }
    if (np_total * G.NL >= ((int64_t)1 << 31)) return felt_fail(FELT_E_ARG, "too many partitions");
    G.np = (int)np_total;
    int64_t cap = np_total;
    const int64_t id_limit = (int64_t)0xffffffffll / pv;
    if (max_bricks > 0) cap = std::min<int64_t>(cap, (int64_t)max_bricks);
    if (n == 1 && cap > id_limit)
        return felt_fail(FELT_E_ARG, "grid too large for 32-bit voxel indices");
    cap = std::min(cap, id_limit);
    felt_surface* s = new (std::nothrow) felt_surface();
    if (!s) return felt_fail(FELT_E_NOMEM, "host allocation failed");
    s->G = G;
    s->mems = std::move(mems);
    s->cap = (int)cap;
    int rc = felt_surface_init(s);
    if (rc) {
        felt_surface_release(s);
        delete s;
        return rc;
    }
    *out = s;
    return FELT_OK;
}

size_t grid_total(const felt_surface* s, int m) {
    size_t t = 1;
    for (int a = 0; a < s->G.D; a++) t *= (size_t)s->mems[m].size[a];
    return t;
}

// seed (Surface.hpp:207-241): L1 distances on the (2L+1)^D window, tracked by layer.
int seed_m(felt_surface* s, int m, const int32_t* pos) {
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
        if (!to_grid(s->mems[m], G.D, wp, g)) continue;
        gp.insert(gp.end(), g, g + 3);
        val.push_back((float)dist);
        lay.push_back(k);
    }
    return felt_track_points(s, m, gp, val, lay);
}

int check_speed(const felt_speed* fn) {
    if (!fn || fn->kind < 0 || fn->kind > FELT_SPEED_NOISE_CURV)
        return felt_fail(FELT_E_ARG, "bad speed descriptor");
    if (fn->kind == FELT_SPEED_NOISE_CURV && fn->cell <= 0)
        return felt_fail(FELT_E_ARG, "noise cell must be positive");
    return FELT_OK;
}

int emit_positions(int D, const std::vector<int>& p3, int32_t* pos, size_t cap, size_t* n) {
    const size_t k = p3.size() / 3;
    *n = k;
    if (cap < k) return felt_fail(FELT_E_CAPACITY, "position buffer too small");
    for (size_t i = 0; i < k; i++)
        for (int a = 0; a < D; a++) pos[i * D + a] = p3[i * 3 + a];
    return FELT_OK;
}

// Member m's slice of the per-partition layer counts.
int member_counts(felt_surface* s, int m, std::vector<int>& cnt) {
    const Mem& M = s->mems[m];
    cnt.resize((size_t)M.np * s->G.NL);
    return felt_download(s, cnt.data(), s->P.pcnt + (size_t)M.pbase * s->G.NL,
                         cnt.size() * sizeof(int));
}

int layer_size_m(felt_surface* s, int m, int layer, uint64_t* n) {
    const Geo& G = s->G;
    if (layer < -G.L || layer > G.L) return felt_fail(FELT_E_ARG, "layer out of range");
    std::vector<int> cnt;
    if (int rc = member_counts(s, m, cnt)) return rc;
    uint64_t t = 0;
    for (int p = 0; p < s->mems[m].np; p++) t += (uint64_t)cnt[(size_t)p * G.NL + layer + G.L];
    *n = t;
    return FELT_OK;
}

int layer_positions_m(felt_surface* s, int m, int layer, int32_t* pos, size_t cap, size_t* n) {
    if (layer < -s->G.L || layer > s->G.L) return felt_fail(FELT_E_ARG, "layer out of range");
    std::vector<int> p3;
    if (int rc = felt_collect_layer_impl(s, m, layer, p3)) return rc;
    return emit_positions(s->G.D, p3, pos, cap, n);
}

// Surface::Stats (Surface.hpp:623-629, 1732-1742): the longest parent list of each grid.
int stats_m(felt_surface* s, int m, felt_stats* st) {
    const Geo& G = s->G;
    const Mem& M = s->mems[m];
    std::vector<int> cnt;
    std::vector<unsigned> dm(M.np);
    if (int rc = member_counts(s, m, cnt)) return rc;
    if (int rc = felt_download(s, dm.data(), s->dmask + M.pbase, dm.size() * sizeof(unsigned)))
        return rc;
    if (int rc = felt_download(s, &s->hc, s->ctr, sizeof(Ctr))) return rc;
    uint64_t best_i = 0, best_d = 0, band = 0;
    for (int k = 0; k < G.NL; k++) {
        uint64_t ni = 0, nd = 0;
        for (int p = 0; p < M.np; p++) {
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
    if (!size) return felt_fail(FELT_E_ARG, "null size");
    return make_pool(D, L, part, 1, size, 0, out);
}

API int felt_surface_destroy(felt_surface* s) {
    if (!s) return FELT_OK;
    felt_surface_release(s);
    delete s;
    return FELT_OK;
}

API int felt_surface_geometry(felt_surface* s, int32_t* cs, uint64_t* np) {
    CHECK_S(s);
    for (int a = 0; a < 3; a++) cs[a] = s->mems[0].csize[a];
    *np = (uint64_t)s->mems[0].np;
    return FELT_OK;
}

API int felt_surface_seed(felt_surface* s, const int32_t* pos) {
    CHECK_S(s);
    if (!pos) return felt_fail(FELT_E_ARG, "null position");
    return seed_m(s, 0, pos);
}

API int felt_surface_load_dense(felt_surface* s, const float* phi, size_t n) {
    CHECK_S(s);
    if (n != grid_total(s, 0) || !phi) return felt_fail(FELT_E_ARG, "load_dense: size mismatch");
    return felt_load_dense_impl(s, 0, phi, n);
}

static int load_spheres_m(felt_surface* s, int m, int32_t n, const double* centres,
                          const double* radii) {
    if (n < 1 || !centres || !radii) return felt_fail(FELT_E_ARG, "need at least one sphere");
    const int D = s->G.D;
    std::vector<double> c3((size_t)n * 3, 0.0);
    for (int k = 0; k < n; k++)
        for (int a = 0; a < D; a++) c3[(size_t)k * 3 + a] = centres[(size_t)k * D + a];
    return felt_load_spheres_impl(s, m, n, c3.data(), radii);
}

API int felt_surface_load_spheres(felt_surface* s, int32_t n, const double* centres,
                                  const double* radii) {
    CHECK_S(s);
    return load_spheres_m(s, 0, n, centres, radii);
}

API int felt_surface_update(felt_surface* s, const felt_speed* fn) {
    CHECK_S(s);
    if (int rc = check_speed(fn)) return rc;
    return felt_update_global(s, fn);
}

API int felt_surface_update_box(felt_surface* s, const int32_t* lo, const int32_t* hi,
                                const felt_speed* fn) {
    CHECK_S(s);
    if (!lo || !hi) return felt_fail(FELT_E_ARG, "bad arguments");
    if (int rc = check_speed(fn)) return rc;
    int l3[3] = {0, 0, 0}, h3[3] = {0, 0, 0};
    for (int a = 0; a < s->G.D; a++) {
        l3[a] = lo[a];
        h3[a] = hi[a];
    }
    const int id = 0;
    return felt_update_boxes_impl(s, 1, &id, l3, h3, fn);
}

API int felt_surface_update_deltas(felt_surface* s, size_t n, const int32_t* pos,
                                   const float* delta) {
    CHECK_S(s);
    std::vector<int> g;
    if (int rc = check_deltas(s, n, pos, delta, g)) return rc;
    return felt_update_global_deltas(s, n, g.data(), delta);
}

API int felt_surface_update_start(felt_surface* s) {
    CHECK_S(s);
    return felt_update_start_impl(s);
}

API int felt_surface_delta(felt_surface* s, size_t n, const int32_t* pos, const float* delta) {
    CHECK_S(s);
    std::vector<int> g;
    if (int rc = check_deltas(s, n, pos, delta, g)) return rc;
    return felt_delta_impl(s, n, g.data(), delta);
}

API int felt_surface_update_end(felt_surface* s) {
    CHECK_S(s);
    return felt_update_end_public(s);
}

API int felt_surface_snapshot(felt_surface* s, float* dst, size_t n) {
    CHECK_S(s);
    const size_t total = grid_total(s, 0);
    if (n < total) return felt_fail(FELT_E_CAPACITY, "snapshot buffer too small");
    return felt_snapshot_impl(s, 0, dst, total);
}

API int felt_surface_get(felt_surface* s, size_t n, const int32_t* pos, float* out) {
    CHECK_S(s);
    std::vector<int> g(n * 3);
    for (size_t i = 0; i < n; i++)
        if (!to_grid(s->mems[0], s->G.D, pos + i * s->G.D, g.data() + i * 3)) {
            // out-of-grid reads return the grid background; keep them off the device
            for (int a = 0; a < 3; a++) g[i * 3 + a] = -1;
        }
    return felt_get_impl(s, 0, n, g.data(), out);
}

// Positions the stencil kernels accept: integers within 2^30 and finite reals within 2^24 of the
// origin, so that no neighbour or interpolation corner overflows.
API int felt_surface_stencil(felt_surface* s, int op, float dx, size_t n, const int32_t* pos,
                             float* out) {
    CHECK_S(s);
    if (n && (!pos || !out)) return felt_fail(FELT_E_ARG, "null position or output array");
    for (size_t i = 0; i < n * (size_t)s->G.D; i++)
        if (pos[i] > (1 << 30) || pos[i] < -(1 << 30))
            return felt_fail(FELT_E_ARG, "stencil position beyond 2^30");
    return felt_stencil_impl(s, 0, op, dx, n, pos, nullptr, out);
}

API int felt_surface_stencil_real(felt_surface* s, int op, float dx, size_t n, const float* pos,
                                  float* out) {
    CHECK_S(s);
    if (n && (!pos || !out)) return felt_fail(FELT_E_ARG, "null position or output array");
    for (size_t i = 0; i < n * (size_t)s->G.D; i++)
        if (!std::isfinite(pos[i]) || std::fabs(pos[i]) > 16777216.0f)
            return felt_fail(FELT_E_ARG, "stencil position not finite or beyond 2^24");
    return felt_stencil_impl(s, 0, op, dx, n, nullptr, pos, out);
}

API int felt_surface_speed_eval(felt_surface* s, const felt_speed* fn, size_t n,
                                const int32_t* pos, float* out) {
    CHECK_S(s);
    if (int rc = check_speed(fn)) return rc;
    if (n && (!pos || !out)) return felt_fail(FELT_E_ARG, "null position or output array");
    std::vector<int> g(n * 3);
    for (size_t i = 0; i < n; i++)
        if (!to_grid(s->mems[0], s->G.D, pos + i * s->G.D, g.data() + i * 3))
            return felt_fail(FELT_E_ARG, "speed_eval position outside the grid");
    return felt_speed_eval_impl(s, 0, fn, n, g.data(), out);
}

API int felt_surface_device_grid(felt_surface* s, felt_device_grid* out) {
    CHECK_S(s);
    if (!out) return felt_fail(FELT_E_ARG, "null output");
    felt_fill_device_grid(s, 0, out);
    return FELT_OK;
}

API int felt_surface_update_begin(felt_surface* s, felt_zero_job* job) {
    CHECK_S(s);
    if (!job) return felt_fail(FELT_E_ARG, "null job");
    if (s->in_update) return felt_fail(FELT_E_STATE, "update(fn) inside update_start / update_end");
    return felt_update_begin_impl(s, job);
}

API int felt_surface_update_box_begin(felt_surface* s, const int32_t* lo, const int32_t* hi,
                                      felt_zero_job* job) {
    CHECK_S(s);
    if (!lo || !hi || !job) return felt_fail(FELT_E_ARG, "bad arguments");
    if (s->in_update) return felt_fail(FELT_E_STATE, "update(fn) inside update_start / update_end");
    int l3[3] = {0, 0, 0}, h3[3] = {0, 0, 0};
    for (int a = 0; a < s->G.D; a++) {
        l3[a] = lo[a];
        h3[a] = hi[a];
    }
    return felt_update_box_begin_impl(s, l3, h3, job);
}

API int felt_surface_update_commit(felt_surface* s) {
    if (!s) return felt_fail(FELT_E_ARG, "null surface handle");
    return felt_update_commit_impl(s);
}

API int felt_surface_update_cancel(felt_surface* s) {
    if (!s) return felt_fail(FELT_E_ARG, "null surface handle");
    return felt_update_cancel_impl(s);
}

API int felt_surface_ray(felt_surface* s, size_t n, const float* origins, const float* dirs,
                         float* out) {
    CHECK_S(s);
    if (n && (!origins || !dirs || !out)) return felt_fail(FELT_E_ARG, "null ray array");
    return felt_ray_impl(s, nullptr, 0, n, origins, dirs, out);
}

API int felt_surface_ray_noalloc(felt_surface* s, size_t n, const float* origins,
                                 const float* dirs, float* out) {
    CHECK_S(s);
    if (n && (!origins || !dirs || !out)) return felt_fail(FELT_E_ARG, "null ray array");
    return felt_ray_noalloc_impl(s, n, origins, dirs, out);
}

API int felt_surface_ray_device(felt_surface* s, size_t n, const float* origins,
                                const float* dirs, float* out) {
    CHECK_S(s);
    if (n && (!origins || !dirs || !out)) return felt_fail(FELT_E_ARG, "null ray array");
    if (n > (size_t)0x7fffffff / 4) return felt_fail(FELT_E_ARG, "too many rays");
    int rc = felt_ray_device(s, nullptr, 0, n, origins, dirs, out);
    if (rc) return rc;
    if (hipStreamSynchronize(s->st) != hipSuccess) return felt_fail(FELT_E_DEVICE, "ray kernel failed");
    return FELT_OK;
}

API int felt_surface_time_ray(felt_surface* s, size_t n, const float* origins, const float* dirs,
                              int reps, double* us) {
    CHECK_S(s);
    if (!origins || !dirs || !us) return felt_fail(FELT_E_ARG, "null argument");
    return felt_time_ray_impl(s, n, origins, dirs, reps, us);
}

API int felt_surface_save(felt_surface* s, uint8_t* dst, size_t cap, size_t* n) {
    CHECK_S(s);
    if (!n) return felt_fail(FELT_E_ARG, "null size output");
    if (s->in_update) return felt_fail(FELT_E_STATE, "save inside update_start / update_end");
    return felt_save_impl(s, 0, dst, cap, n);
}

API int felt_surface_shape(felt_surface* s, int* D, int* L, int32_t* size, int32_t* part) {
    CHECK_S(s);
    if (!D || !L || !size || !part) return felt_fail(FELT_E_ARG, "null output");
    *D = s->G.D;
    *L = s->G.L;
    for (int a = 0; a < s->G.D; a++) {
        size[a] = s->mems[0].size[a];
        part[a] = s->G.part[a];
    }
    return FELT_OK;
}

API int felt_surface_load(const uint8_t* src, size_t n, felt_surface** out) {
    if (!src || !out) return felt_fail(FELT_E_ARG, "null argument");
    int D = 0, L = 0, size[3] = {1, 1, 1}, part[3] = {1, 1, 1};
    int rc = felt_archive_geometry(src, n, &D, &L, size, part);
    if (rc) return rc;
    felt_surface* s = nullptr;
    if ((rc = make_pool(D, L, part, 1, size, 0, &s))) return rc;
    if ((rc = felt_load_impl(s, 0, src, n))) {
        felt_surface_release(s);
        delete s;
        return rc;
    }
    *out = s;
    return FELT_OK;
}

API int felt_surface_layer_size(felt_surface* s, int layer, uint64_t* n) {
    CHECK_S(s);
    return layer_size_m(s, 0, layer, n);
}

API int felt_surface_layer_positions(felt_surface* s, int layer, int32_t* pos, size_t cap,
                                     size_t* n) {
    CHECK_S(s);
    return layer_positions_m(s, 0, layer, pos, cap, n);
}

API int felt_surface_zero_layer_in_box(felt_surface* s, const int32_t* lo, const int32_t* hi,
                                       int32_t* pos, size_t cap, size_t* n) {
    CHECK_S(s);
    const Geo& G = s->G;
    const Mem& M = s->mems[0];
    std::vector<int> p3;
    if (int rc = felt_collect_layer_impl(s, 0, 0, p3)) return rc;
    std::vector<int> kept;
    for (size_t i = 0; i < p3.size() / 3; i++) {
        bool in = true;
        for (int a = 0; a < G.D; a++) {
            const int hb = std::min(M.offset[a] + M.size[a], hi[a] + 1);
            if (p3[i * 3 + a] < lo[a] || p3[i * 3 + a] >= hb) in = false;
        }
        if (in) kept.insert(kept.end(), p3.begin() + i * 3, p3.begin() + i * 3 + 3);
    }
    return emit_positions(G.D, kept, pos, cap, n);
}

API int felt_surface_partition_list(felt_surface* s, int kind, int k, uint64_t* out, size_t cap,
                                    size_t* n) {
    CHECK_S(s);
    const Geo& G = s->G;
    const Mem& M = s->mems[0];
    if (k < 0 || k >= G.NL) return felt_fail(FELT_E_ARG, "list index out of range");
    std::vector<uint64_t> parts;
    if (kind == FELT_LIST_ISOGRID) {
        std::vector<int> cnt;
        if (int rc = member_counts(s, 0, cnt)) return rc;
        for (int p = 0; p < M.np; p++)
            if (cnt[(size_t)p * G.NL + k] > 0) parts.push_back(p);
    } else if (kind == FELT_LIST_DELTA || kind == FELT_LIST_STATUS_CHANGE) {
        std::vector<unsigned> mk(M.np);
        if (int rc = felt_download(s, mk.data(), kind == FELT_LIST_DELTA ? s->dmask : s->smask,
                                   mk.size() * sizeof(unsigned)))
            return rc;
        for (int p = 0; p < M.np; p++)
            if ((mk[p] >> k) & 1u) parts.push_back(p);
    } else if (kind == FELT_LIST_AFFECTED) {
        std::vector<int> p3;
        if (int rc = felt_lu_positions_impl(s, 0, 0, k - G.L, p3)) return rc;
        std::vector<char> seen(M.np, 0);
        for (size_t i = 0; i < p3.size() / 3; i++) {
            int g[3];
            to_grid(M, G.D, p3.data() + i * 3, g);
            int p = 0;
            for (int a = 0; a < G.D; a++) p += (g[a] / G.part[a]) * M.cstr[a];
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
    if (int rc = felt_lu_positions_impl(s, 0, 0, k - s->G.L, p3)) return rc;
    return emit_positions(s->G.D, p3, pos, cap, n);
}

API int felt_surface_partition_active(felt_surface* s, uint8_t* out, size_t n) {
    CHECK_S(s);
    const Mem& M = s->mems[0];
    if (n < (size_t)M.np) return felt_fail(FELT_E_CAPACITY, "buffer too small");
    std::vector<int> page(M.np);
    if (int rc = felt_download(s, page.data(), s->P.page, page.size() * sizeof(int))) return rc;
    for (int p = 0; p < M.np; p++) out[p] = page[p] >= 0 ? 1 : 0;
    return FELT_OK;
}

API int felt_surface_partition_layer_size(felt_surface* s, uint64_t p, int layer, uint64_t* n) {
    CHECK_S(s);
    const Geo& G = s->G;
    if (p >= (uint64_t)s->mems[0].np || layer < -G.L || layer > G.L)
        return felt_fail(FELT_E_ARG, "argument out of range");
    int c = 0;
    if (int rc = felt_download(s, &c, s->P.pcnt + p * G.NL + layer + G.L, sizeof(int))) return rc;
    *n = (uint64_t)c;
    return FELT_OK;
}

API int felt_surface_stats(felt_surface* s, felt_stats* st) {
    CHECK_S(s);
    return stats_m(s, 0, st);
}

API int felt_surface_check_records(felt_surface* s, uint64_t* mismatches) {
    CHECK_S(s);
    if (!mismatches) return felt_fail(FELT_E_ARG, "null argument");
    return felt_nbr_check(s, mismatches);
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

API int felt_surface_autotune(felt_surface* s, int reps, felt_launch_config* out) {
    CHECK_S(s);
    if (!out) return felt_fail(FELT_E_ARG, "null output");
    if (int rc = prof_collect(s)) return rc;
    return felt_autotune_impl(s, reps, out);
}

API int felt_surface_set_launch_config(felt_surface* s, const felt_launch_config* cfg) {
    CHECK_S(s);
    if (!cfg) return felt_fail(FELT_E_ARG, "null config");
    return felt_set_launch_cfg_impl(s, cfg);
}

API int felt_surface_set_round_limit(felt_surface* s, int32_t max_rounds) {
    CHECK_S(s);
    if (max_rounds < 1) return felt_fail(FELT_E_ARG, "max_rounds must be at least 1");
    s->max_rounds = max_rounds;
    return FELT_OK;
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
    if (p >= (uint64_t)q->s->mems[0].np) return felt_fail(FELT_E_ARG, "partition out of range");
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
    if (p >= (uint64_t)q->s->mems[0].np) return felt_fail(FELT_E_ARG, "partition out of range");
    const PolyKid& k = q->kids[p];
    *active = k.active ? 1 : 0;
    *nv = (uint64_t)k.nv;
    *ns = (uint64_t)k.nt;
    return FELT_OK;
}

API int felt_polys_child_mesh(felt_polys* q, uint64_t p, float* vtx, size_t vcap, uint32_t* spx,
                              size_t scap) {
    CHECK_P(q);
    if (p >= (uint64_t)q->s->mems[0].np) return felt_fail(FELT_E_ARG, "partition out of range");
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

// ---- batched surfaces -------------------------------------------------------------------

#define CHECK_B(b) \
    if (!(b) || !(b)->s) return felt_fail(FELT_E_ARG, "null batch handle")
#define CHECK_IDLE(b) \
    if ((b)->s->step_pending) return felt_fail(FELT_E_STATE, "a step is in flight: felt_batch_wait first")
#define CHECK_M(b, i) \
    if ((i) < 0 || (i) >= (b)->s->G.nmem) return felt_fail(FELT_E_ARG, "member out of range")

API int felt_batch_create(int D, int L, const int32_t* part, int32_t n, const int32_t* sizes,
                          uint64_t max_bricks, felt_batch** out) {
    if (!out) return felt_fail(FELT_E_ARG, "null output");
    felt_batch* b = new (std::nothrow) felt_batch();
    if (!b) return felt_fail(FELT_E_NOMEM, "host allocation failed");
    if (int rc = make_pool(D, L, part, n, sizes, max_bricks, &b->s)) {
        delete b;
        return rc;
    }
    b->s->auto_unroll = true;
    *out = b;
    return FELT_OK;
}

API int felt_batch_destroy(felt_batch* b) {
    if (!b) return FELT_OK;
    if (b->s) {
        felt_surface_release(b->s);
        delete b->s;
    }
    delete b;
    return FELT_OK;
}

API int felt_batch_count(felt_batch* b, int32_t* n) {
    CHECK_B(b);
    *n = b->s->G.nmem;
    return FELT_OK;
}

API int felt_batch_memory(felt_batch* b, uint64_t* bytes, uint64_t* used, uint64_t* total) {
    CHECK_B(b);
    CHECK_IDLE(b);
    if (int rc = felt_download(b->s, &b->s->hc, b->s->ctr, sizeof(Ctr))) return rc;
    *bytes = felt_device_bytes(b->s);
    *total = (uint64_t)b->s->cap;
    *used = (uint64_t)(b->s->cap - b->s->hc.free_top);
    return FELT_OK;
}

API int felt_batch_seed(felt_batch* b, int32_t i, const int32_t* pos) {
    CHECK_B(b);
    CHECK_IDLE(b);
    CHECK_M(b, i);
    if (!pos) return felt_fail(FELT_E_ARG, "null position");
    return seed_m(b->s, i, pos);
}

API int felt_batch_load_dense(felt_batch* b, int32_t i, const float* phi, size_t n) {
    CHECK_B(b);
    CHECK_IDLE(b);
    CHECK_M(b, i);
    if (n != grid_total(b->s, i) || !phi) return felt_fail(FELT_E_ARG, "load_dense: size mismatch");
    return felt_load_dense_impl(b->s, i, phi, n);
}

API int felt_batch_load_spheres(felt_batch* b, int32_t i, int32_t n, const double* centres,
                                const double* radii) {
    CHECK_B(b);
    CHECK_IDLE(b);
    CHECK_M(b, i);
    return load_spheres_m(b->s, i, n, centres, radii);
}

API int felt_batch_update(felt_batch* b, const felt_speed* speeds) {
    CHECK_B(b);
    CHECK_IDLE(b);
    if (!speeds) return felt_fail(FELT_E_ARG, "null speeds");
    for (int i = 0; i < b->s->G.nmem; i++)
        if (int rc = check_speed(speeds + i)) return rc;
    return felt_update_global(b->s, speeds);
}

API int felt_batch_update_boxes(felt_batch* b, int32_t n, const int32_t* ids, const int32_t* lo,
                                const int32_t* hi, const felt_speed* speeds) {
    CHECK_B(b);
    CHECK_IDLE(b);
    if (n < 0 || (n > 0 && (!ids || !lo || !hi || !speeds)))
        return felt_fail(FELT_E_ARG, "bad arguments");
    felt_surface* s = b->s;
    const int D = s->G.D;
    std::vector<char> seen(s->G.nmem, 0);
    std::vector<int32_t> l3((size_t)n * 3, 0), h3((size_t)n * 3, 0);
    std::vector<int> id(n);
    for (int j = 0; j < n; j++) {
        if (ids[j] < 0 || ids[j] >= s->G.nmem) return felt_fail(FELT_E_ARG, "member out of range");
        if (seen[ids[j]]) return felt_fail(FELT_E_ARG, "a member appears twice in one step");
        seen[ids[j]] = 1;
        if (int rc = check_speed(speeds + j)) return rc;
        id[j] = ids[j];
        for (int a = 0; a < D; a++) {
            l3[(size_t)j * 3 + a] = lo[(size_t)j * D + a];
            h3[(size_t)j * 3 + a] = hi[(size_t)j * D + a];
        }
    }
    return felt_update_boxes_impl(s, n, id.data(), l3.data(), h3.data(), speeds);
}

API int felt_batch_ray(felt_batch* b, size_t n, const int32_t* members, const float* origins,
                       const float* dirs, float* out) {
    CHECK_B(b);
    CHECK_IDLE(b);
    if (n && (!members || !origins || !dirs || !out)) return felt_fail(FELT_E_ARG, "null ray array");
    return felt_ray_impl(b->s, members, 0, n, origins, dirs, out);
}

API int felt_batch_save(felt_batch* b, int32_t i, uint8_t* dst, size_t cap, size_t* n) {
    CHECK_B(b);
    CHECK_IDLE(b);
    CHECK_M(b, i);
    if (!n) return felt_fail(FELT_E_ARG, "null size output");
    return felt_save_impl(b->s, i, dst, cap, n);
}

API int felt_batch_load(felt_batch* b, int32_t i, const uint8_t* src, size_t n) {
    CHECK_B(b);
    CHECK_IDLE(b);
    CHECK_M(b, i);
    if (!src) return felt_fail(FELT_E_ARG, "null archive");
    int D = 0, L = 0, size[3] = {1, 1, 1}, part[3] = {1, 1, 1};
    int rc = felt_archive_geometry(src, n, &D, &L, size, part);
    if (rc) return rc;
    const felt_surface* s = b->s;
    if (D != s->G.D || L != s->G.L) return felt_fail(FELT_E_ARG, "archive D or L differs from the batch's");
    for (int a = 0; a < D; a++)
        if (size[a] != s->mems[i].size[a] || part[a] != s->G.part[a])
            return felt_fail(FELT_E_ARG, "archive grid differs from the member's");
    return felt_load_impl(b->s, i, src, n);
}

API int felt_batch_clear(felt_batch* b, int32_t i) {
    CHECK_B(b);
    CHECK_IDLE(b);
    CHECK_M(b, i);
    int rc = felt_member_clear_impl(b->s, i);
    if (rc) return rc;
    if (b->s->polys) return felt_polys_invalidate_impl(b->s->polys);
    return FELT_OK;
}

API int felt_batch_snapshot(felt_batch* b, int32_t i, float* dst, size_t n) {
    CHECK_B(b);
    CHECK_IDLE(b);
    CHECK_M(b, i);
    const size_t total = grid_total(b->s, i);
    if (n < total) return felt_fail(FELT_E_CAPACITY, "snapshot buffer too small");
    return felt_snapshot_impl(b->s, i, dst, total);
}

API int felt_batch_layer_size(felt_batch* b, int32_t i, int layer, uint64_t* n) {
    CHECK_B(b);
    CHECK_IDLE(b);
    CHECK_M(b, i);
    return layer_size_m(b->s, i, layer, n);
}

API int felt_batch_layer_positions(felt_batch* b, int32_t i, int layer, int32_t* pos, size_t cap,
                                   size_t* n) {
    CHECK_B(b);
    CHECK_IDLE(b);
    CHECK_M(b, i);
    return layer_positions_m(b->s, i, layer, pos, cap, n);
}

API int felt_batch_check_records(felt_batch* b, uint64_t* mismatches) {
    CHECK_B(b);
    CHECK_IDLE(b);
    if (!mismatches) return felt_fail(FELT_E_ARG, "null argument");
    return felt_nbr_check(b->s, mismatches);
}

API int felt_batch_stats(felt_batch* b, int32_t i, felt_stats* st) {
    CHECK_B(b);
    CHECK_IDLE(b);
    CHECK_M(b, i);
    return stats_m(b->s, i, st);
}

API int felt_batch_update_async(felt_batch* b, const felt_speed* speeds) {
    CHECK_B(b);
    CHECK_IDLE(b);
    if (!speeds) return felt_fail(FELT_E_ARG, "null speeds");
    for (int i = 0; i < b->s->G.nmem; i++)
        if (int rc = check_speed(speeds + i)) return rc;
    return felt_update_global_launch(b->s, speeds);
}

API int felt_batch_wait(felt_batch* b) {
    CHECK_B(b);
    return felt_global_step_finish(b->s);
}

API int felt_group_create(felt_batch* const* batches, int32_t n, felt_group** out) {
    if (!out || !batches || n < 1) return felt_fail(FELT_E_ARG, "need batches and an output");
    *out = nullptr;
    auto* g = new felt_group();
    for (int32_t k = 0; k < n; k++) {
        felt_batch* b = batches[k];
        if (!b || !b->s) {
            delete g;
            return felt_fail(FELT_E_ARG, "null batch handle");
        }
        for (felt_surface* t : g->s)
            if (t == b->s) {
                delete g;
                return felt_fail(FELT_E_ARG, "a batch appears twice in the group");
            }
        if (b->s->device != batches[0]->s->device) {
            delete g;
            return felt_fail(FELT_E_ARG, "the batches of a group must be on one device");
        }
        g->s.push_back(b->s);
    }
    g->ver.assign(g->s.size(), 0);
    g->ready.assign(g->s.size(), nullptr);
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e == hipSuccess) e = hipSetDevice(g->s[0]->device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&g->st, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&g->done, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&g->fork, hipEventDisableTiming);
    for (hipEvent_t& ev : g->ready)
        if (e == hipSuccess) e = hipEventCreateWithFlags(&ev, hipEventDisableTiming);
    (void)hipSetDevice(dev);
    if (e != hipSuccess) {
        felt_group_release(g);
        delete g;
        return felt_fail(FELT_E_DEVICE, hipGetErrorString(e));
    }
    *out = g;
    return FELT_OK;
}

API int felt_group_destroy(felt_group* g) {
    if (!g) return FELT_OK;
    if (g->pending) (void)felt_group_wait_impl(g);
    felt_group_release(g);
    delete g;
    return FELT_OK;
}

API int felt_group_update_async(felt_group* g, const felt_speed* const* speeds) {
    if (!g) return felt_fail(FELT_E_ARG, "null group handle");
    if (!speeds) return felt_fail(FELT_E_ARG, "null speeds");
    if (g->pending) return felt_fail(FELT_E_STATE, "a step is in flight: felt_group_wait first");
    for (size_t k = 0; k < g->s.size(); k++) {
        if (g->s[k]->step_pending || g->s[k]->in_update)
            return felt_fail(FELT_E_STATE, "a batch of the group is busy");
        if (!speeds[k]) return felt_fail(FELT_E_ARG, "null speeds");
        for (int i = 0; i < g->s[k]->G.nmem; i++)
            if (int rc = check_speed(speeds[k] + i)) return rc;
    }
    return felt_group_launch_impl(g, speeds);
}

API int felt_group_wait(felt_group* g) {
    if (!g) return felt_fail(FELT_E_ARG, "null group handle");
    return felt_group_wait_impl(g);
}

API int felt_batch_step_info(felt_batch* b, uint64_t* voxel_updates, uint64_t* rounds) {
    CHECK_B(b);
    *voxel_updates = b->s->last_updates;
    *rounds = b->s->last_rounds;
    return FELT_OK;
}

API int felt_batch_set_graph(felt_batch* b, int enable) {
    CHECK_B(b);
    CHECK_IDLE(b);
    b->s->use_graph = enable != 0;
    return FELT_OK;
}

API int felt_batch_autotune(felt_batch* b, int reps, felt_launch_config* out) {
    CHECK_B(b);
    CHECK_IDLE(b);
    return felt_surface_autotune(b->s, reps, out);
}

API int felt_batch_set_launch_config(felt_batch* b, const felt_launch_config* cfg) {
    CHECK_B(b);
    CHECK_IDLE(b);
    return felt_surface_set_launch_config(b->s, cfg);
}

API int felt_batch_set_round_limit(felt_batch* b, int32_t max_rounds) {
    CHECK_B(b);
    CHECK_IDLE(b);
    return felt_surface_set_round_limit(b->s, max_rounds);
}

API int felt_batch_time_pass(felt_batch* b, int pass, int reps, double* us_per_launch,
                             double* items_per_launch, int* launches) {
    CHECK_B(b);
    CHECK_IDLE(b);
    return felt_surface_time_pass(b->s, pass, reps, us_per_launch, items_per_launch, launches);
}

API int felt_batch_sync(felt_batch* b) {
    CHECK_B(b);
    return felt_sync_impl(b->s);
}

API int felt_batch_polys_create(felt_batch* b, felt_polys** out) {
    CHECK_B(b);
    CHECK_IDLE(b);
    return felt_polys_create(b->s, out);
}

// Pool partition of member i's partition pos_idx.
static int pool_part(felt_polys* q, int32_t i, uint64_t pos_idx, uint64_t* p) {
    if (i < 0 || i >= q->s->G.nmem) return felt_fail(FELT_E_ARG, "member out of range");
    const Mem& M = q->s->mems[i];
    if (pos_idx >= (uint64_t)M.np) return felt_fail(FELT_E_ARG, "partition out of range");
    *p = (uint64_t)M.pbase + pos_idx;
    return FELT_OK;
}

API int felt_batch_polys_changes(felt_polys* q, int32_t i, uint64_t* out, size_t cap, size_t* n) {
    CHECK_P(q);
    if (i < 0 || i >= q->s->G.nmem) return felt_fail(FELT_E_ARG, "member out of range");
    const Mem& M = q->s->mems[i];
    std::vector<uint64_t> mine;
    for (int64_t p : q->done->list)
        if (p >= M.pbase && p < (int64_t)M.pbase + M.np) mine.push_back((uint64_t)(p - M.pbase));
    *n = mine.size();
    if (cap < mine.size()) return felt_fail(FELT_E_CAPACITY, "buffer too small");
    std::copy(mine.begin(), mine.end(), out);
    return FELT_OK;
}

API int felt_batch_polys_child_info(felt_polys* q, int32_t i, uint64_t pos_idx, int* active,
                                    uint64_t* nv, uint64_t* ns) {
    CHECK_P(q);
    uint64_t p;
    if (int rc = pool_part(q, i, pos_idx, &p)) return rc;
    const PolyKid& k = q->kids[p];
    *active = k.active ? 1 : 0;
    *nv = (uint64_t)k.nv;
    *ns = (uint64_t)k.nt;
    return FELT_OK;
}

API int felt_batch_polys_child_mesh(felt_polys* q, int32_t i, uint64_t pos_idx, float* vtx,
                                    size_t vcap, uint32_t* spx, size_t scap) {
    CHECK_P(q);
    uint64_t p;
    if (int rc = pool_part(q, i, pos_idx, &p)) return rc;
    return felt_polys_child_mesh_impl(q, p, vtx, vcap, spx, scap);
}

API int felt_ms_table(uint32_t* out, size_t n) {
    if (!out || n < 16) return felt_fail(FELT_E_CAPACITY, "need 16 entries");
    felt_ms_table_host(out);
    return FELT_OK;
}

API int felt_mc_table(uint64_t* out, size_t n) {
    if (!out || n < 256) return felt_fail(FELT_E_CAPACITY, "need 256 entries");
    felt_mc_table_host(out);
    return FELT_OK;
}
