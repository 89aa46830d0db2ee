// felt_server.hip — a request loop over one batch of resident surfaces (gfx950).
//
// Requests (surface, box, speed, steps) are queued as they arrive.  Each tick admits queued
// requests in arrival order while three limits hold: one request in flight per surface, at
// most `max_active` in flight, and the step lists the in-flight set needs fit the HBM budget
// beside the resident bricks.  Then every in-flight request advances by one bounded update
// (Surface::update(lo, hi, fn), Surface.hpp:310-386) — all of them in one batched step over
// the shared brick pool — followed by repolygonisation of the changed partitions
// (Polys::notify / march, Polys.hpp:92-165).  Requests that have run their steps complete and
// free their slot for the next tick, so work that arrives while others run joins at the next
// step (continuous batching).  No host code touches voxel data.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <deque>
#include <functional>
#include <new>
#include <string>
#include <vector>

#include "../../include/felt_amd.h"
#include "felt_polys.h"
#include "felt_surface.h"

using namespace felt;

#define API extern "C" __attribute__((visibility("default")))

namespace {

double now_s() {
    using clk = std::chrono::steady_clock;
    return std::chrono::duration<double>(clk::now().time_since_epoch()).count();
}

double percentile(std::vector<double> v, double q) {
    if (v.empty()) return 0.0;
    std::sort(v.begin(), v.end());
    const double pos = q / 100.0 * (double)(v.size() - 1);
    const size_t i = (size_t)pos;
    const double f = pos - (double)i;
    return i + 1 < v.size() ? v[i] * (1.0 - f) + v[i + 1] * f : v[i];
}

struct Req {
    felt_request r;
    double t_submit;
    int left;
    double frac;  // the box's share of its grid, clamped to [0, 1]
};

// Tick-time model of the latency target: t = fixed + unit * work, work = sum over in-flight
// requests of band * box share (about the affected set).  Both terms are running averages.
// The target bounds the upper tail of tick time, not its mean: admission keeps the prediction
// times a margin -- the running mean plus 3 standard deviations of actual / predicted time --
// within the target.
struct TickModel {
    double fixed_ms = 2.0;
    double unit_ms = 2e-7;
    double r1 = 1.0, r2 = 1.0;  // running mean of actual / predicted, and of its square
    double predict(double work) const { return fixed_ms + unit_ms * work; }
    double margin() const {
        const double sd = std::sqrt(std::max(0.0, r2 - r1 * r1));
        return std::max(1.0, r1 + 3.0 * sd);
    }
    void observe(double ms, double work) {
        const double r = ms / std::max(1e-3, predict(work));
        r1 = 0.95 * r1 + 0.05 * r;
        r2 = 0.95 * r2 + 0.05 * r * r;
        const double f = std::max(0.0, ms - unit_ms * work);
        fixed_ms = 0.95 * fixed_ms + 0.05 * std::min(f, ms);
        if (work > 0) unit_ms = 0.9 * unit_ms + 0.1 * std::max(0.0, ms - fixed_ms) / work;
    }
};

}  // namespace

struct felt_server {
    felt_batch* batch = nullptr;
    felt_surface* s = nullptr;
    felt_polys* polys = nullptr;
    uint64_t budget = 0;
    int max_active = 0;
    std::deque<Req> queue;
    std::vector<Req> active;
    std::vector<char> busy;            // a request of this surface is in flight
    std::vector<long long> band;       // tracked voxels per surface, refreshed after its steps
    std::vector<uint64_t> done;        // completed request ids not yet polled
    std::vector<uint64_t> failed;      // failed request ids not yet polled
    uint64_t n_failed = 0;
    std::vector<double> tick_ms, req_ms;
    double target_ms = 0.0;            // tick-time target of admission (0: none)
    TickModel model;
    uint64_t ticks = 0, completed = 0, deferred = 0, voxel_updates = 0;
    double busy_s = 0.0;
    double phase_s[4] = {0, 0, 0, 0};  // bounded update, notify, march, band refresh
};

namespace {

int refresh_bands(felt_server* v, const std::vector<int>& ids) {
    std::vector<long long> b;
    if (int rc = felt_member_bands(v->s, ids.data(), (int)ids.size(), b)) return rc;
    for (size_t j = 0; j < ids.size(); j++) v->band[ids[j]] = b[j];
    return FELT_OK;
}

// Step-list bytes after a step of the in-flight set plus a surface of `extra` voxels (the
// lists only grow).
uint64_t list_bytes(const felt_server* v, long long extra) {
    long long sum = extra;
    for (const Req& q : v->active) sum += v->band[q.r.surface];
    return (uint64_t)felt_list_cap_for(v->s, (size_t)sum) * kListBytes;
}

// Bytes of everything but the step lists: bricks, tables, scratch and the polygoniser's mesh
// arenas and march scratch.
uint64_t resident_bytes(const felt_server* v) {
    const felt_surface* s = v->s;
    uint64_t b = felt_device_bytes(s) - (uint64_t)s->list_cap * kListBytes;
    if (v->polys) b += felt_polys_device_bytes(v->polys);
    return b;
}

// A request leaves the server without completing: its surface is free again and its id is
// reported by felt_server_poll_failed.
void fail_request(felt_server* v, const Req& q) {
    v->busy[q.r.surface] = 0;
    v->failed.push_back(q.r.id);
    v->n_failed++;
}

double work_of(const felt_server* v, const Req& q) { return (double)v->band[q.r.surface] * q.frac; }

double active_work(const felt_server* v) {
    double w = 0;
    for (const Req& q : v->active) w += work_of(v, q);
    return w;
}

void admit(felt_server* v) {
    const uint64_t fixed = resident_bytes(v);
    double work = active_work(v);
    for (auto it = v->queue.begin(); it != v->queue.end();) {
        if ((int)v->active.size() >= v->max_active) break;
        const int m = it->r.surface;
        if (v->busy[m]) {
            ++it;
            continue;
        }
        if (fixed + list_bytes(v, v->band[m]) > v->budget) {
            if (v->active.empty()) {
                // it would exceed the budget even alone: it can never run
                fail_request(v, *it);
                it = v->queue.erase(it);
                continue;
            }
            v->deferred++;
            break;  // keep arrival order: later requests wait too
        }
        const double w = work_of(v, *it);
        if (!v->active.empty() && v->target_ms > 0 &&
            v->model.predict(work + w) * v->model.margin() > v->target_ms)
            break;  // the tick would overrun the latency target
        work += w;
        v->busy[m] = 1;
        v->active.push_back(*it);
        it = v->queue.erase(it);
    }
}

}  // namespace

API int felt_server_create(int D, int L, const int32_t* part, int32_t n, const int32_t* sizes,
                           uint64_t budget_bytes, int32_t max_active, int32_t polygonise,
                           felt_server** out) {
    if (!out || max_active < 1 || budget_bytes == 0)
        return felt_fail(FELT_E_ARG, "bad server arguments");
    // bricks take at most three quarters of the budget
    uint64_t pv = 1;
    for (int a = 0; a < D && part; a++) pv *= (uint64_t)std::max(1, part[a]);
    const uint64_t brick_bytes = pv * (sizeof(float) + 2) + 32;
    const uint64_t max_bricks = std::max<uint64_t>(1, budget_bytes / 4 * 3 / brick_bytes);
    felt_server* v = new (std::nothrow) felt_server();
    if (!v) return felt_fail(FELT_E_NOMEM, "host allocation failed");
    if (int rc = felt_batch_create(D, L, part, n, sizes, max_bricks, &v->batch)) {
        delete v;
        return rc;
    }
    v->s = v->batch->s;
    if (felt_device_bytes(v->s) > budget_bytes) {
        felt_batch_destroy(v->batch);
        delete v;
        return felt_fail(FELT_E_NOMEM, "resident surfaces exceed the HBM budget");
    }
    if (polygonise) {
        if (int rc = felt_batch_polys_create(v->batch, &v->polys)) {
            felt_batch_destroy(v->batch);
            delete v;
            return rc;
        }
    }
    v->budget = budget_bytes;
    v->max_active = max_active;
    v->busy.assign(n, 0);
    v->band.assign(n, 0);
    *out = v;
    return FELT_OK;
}

API int felt_server_destroy(felt_server* v) {
    if (!v) return FELT_OK;
    if (v->polys) felt_polys_destroy(v->polys);
    felt_batch_destroy(v->batch);
    delete v;
    return FELT_OK;
}

API int felt_server_batch(felt_server* v, felt_batch** out) {
    if (!v || !out) return felt_fail(FELT_E_ARG, "null argument");
    *out = v->batch;
    return FELT_OK;
}

API int felt_server_polys(felt_server* v, felt_polys** out) {
    if (!v || !out) return felt_fail(FELT_E_ARG, "null argument");
    if (!v->polys) return felt_fail(FELT_E_STATE, "server created without polygonisation");
    *out = v->polys;
    return FELT_OK;
}

// Marks the surfaces' current bands known and, with polygonisation, marches every surface.
API int felt_server_start(felt_server* v) {
    if (!v) return felt_fail(FELT_E_ARG, "null server");
    std::vector<int> all(v->s->G.nmem);
    for (int i = 0; i < v->s->G.nmem; i++) all[i] = i;
    if (int rc = refresh_bands(v, all)) return rc;
    if (v->polys) {
        if (int rc = felt_polys_invalidate_impl(v->polys)) return rc;
        if (int rc = felt_polys_march_impl(v->polys)) return rc;
        if (int rc = felt_polys_reserve(v->polys)) return rc;
    }
    // Step lists for the largest in-flight set (the max_active largest bands), within what the
    // budget leaves beside the resident bytes: a tick then grows no list.
    std::vector<long long> b(v->band);
    std::sort(b.begin(), b.end(), std::greater<long long>());
    long long want = 0;
    for (size_t i = 0; i < b.size() && (int)i < v->max_active; i++) want += b[i];
    const uint64_t fixed = resident_bytes(v);
    const uint64_t room = v->budget > fixed ? (v->budget - fixed) / kListBytes : 0;
    size_t band = (size_t)want;
    while (band > 0 && felt_list_cap_for(v->s, band) > room) band -= band / 8 + 1;
    if (band > 0)
        if (int rc = felt_reserve_lists(v->s, band)) return rc;
    return FELT_OK;
}

API int felt_server_set_target(felt_server* v, double tick_ms) {
    if (!v || tick_ms < 0) return felt_fail(FELT_E_ARG, "bad arguments");
    v->target_ms = tick_ms;
    return FELT_OK;
}

// Every request is checked before any is queued: a call queues all n or none.
API int felt_server_submit(felt_server* v, const felt_request* r, size_t n) {
    if (!v || (n && !r)) return felt_fail(FELT_E_ARG, "null argument");
    for (size_t i = 0; i < n; i++) {
        const felt_request& q = r[i];
        const std::string at = "request " + std::to_string(i) + ": ";
        if (q.surface < 0 || q.surface >= v->s->G.nmem)
            return felt_fail(FELT_E_ARG, at + "surface out of range");
        if (q.steps < 1) return felt_fail(FELT_E_ARG, at + "needs at least one step");
        if (q.speed.kind < 0 || q.speed.kind > FELT_SPEED_NOISE_CURV ||
            (q.speed.kind == FELT_SPEED_NOISE_CURV && q.speed.cell <= 0))
            return felt_fail(FELT_E_ARG, at + "bad speed descriptor");
    }
    const double t = now_s();
    for (size_t i = 0; i < n; i++) {
        const felt_request& q = r[i];
        const Mem& M = v->s->mems[q.surface];
        double frac = 1.0;
        for (int a = 0; a < v->s->G.D; a++) {
            const int lo = std::max(q.lo[a], M.offset[a]);
            const int hi = std::min(q.hi[a], M.offset[a] + M.size[a] - 1);
            frac *= hi >= lo ? (double)(hi - lo + 1) / M.size[a] : 0.0;
        }
        v->queue.push_back(Req{q, t, q.steps, frac});
    }
    return FELT_OK;
}

API int felt_server_tick(felt_server* v, int32_t* stepped) {
    if (!v) return felt_fail(FELT_E_ARG, "null server");
    admit(v);
    if (stepped) *stepped = (int32_t)v->active.size();
    if (v->active.empty()) return FELT_OK;
    const double t0 = now_s();
    const double work = active_work(v);
    const int n = (int)v->active.size();
    const int D = v->s->G.D;
    std::vector<int> ids(n);
    std::vector<int32_t> lo((size_t)n * 3, 0), hi((size_t)n * 3, 0);
    std::vector<felt_speed> sp(n);
    std::vector<long long> bands(n);
    for (int j = 0; j < n; j++) {
        const felt_request& q = v->active[j].r;
        ids[j] = q.surface;
        for (int a = 0; a < D; a++) {
            lo[(size_t)j * 3 + a] = q.lo[a];
            hi[(size_t)j * 3 + a] = q.hi[a];
        }
        sp[j] = q.speed;
        bands[j] = v->band[q.surface];
    }
    double tp = t0;
    auto phase = [&](int k) {
        const double t = now_s();
        v->phase_s[k] += t - tp;
        tp = t;
    };
    // A step the brick pool cannot cover is refused with every surface unchanged: the requests
    // are then stepped one at a time, and only those that cannot run even alone fail.
    std::vector<char> ran(n, 1);
    int rc = felt_update_boxes_impl(v->s, n, ids.data(), lo.data(), hi.data(), sp.data(),
                                    bands.data());
    if (rc == FELT_E_NOMEM) {
        rc = FELT_OK;
        for (int j = 0; j < n && !rc; j++) {
            rc = felt_update_boxes_impl(v->s, 1, &ids[j], &lo[(size_t)j * 3], &hi[(size_t)j * 3],
                                        &sp[j], &bands[j]);
            if (rc == FELT_E_NOMEM) {
                ran[j] = 0;
                rc = FELT_OK;
            }
        }
    }
    phase(0);
    if (!rc && v->polys) rc = felt_polys_notify_impl(v->polys);
    phase(1);
    if (!rc && v->polys) rc = felt_polys_march_impl(v->polys);
    phase(2);
    std::vector<int> stepped_ids;
    for (int j = 0; j < n; j++)
        if (ran[j]) stepped_ids.push_back(ids[j]);
    if (!rc) rc = refresh_bands(v, stepped_ids);
    phase(3);
    if (rc) {
        // the surfaces of the in-flight requests may be partly stepped: none of them continues
        const std::string msg = g_felt_err;
        for (const Req& q : v->active) fail_request(v, q);
        v->active.clear();
        return felt_fail(rc, msg);
    }
    const double t1 = now_s();
    v->tick_ms.push_back((t1 - t0) * 1e3);
    v->model.observe((t1 - t0) * 1e3, work);
    v->busy_s += t1 - t0;
    v->ticks++;
    v->voxel_updates += v->s->last_updates;
    std::vector<Req> keep;
    keep.reserve(v->active.size());
    for (int j = 0; j < n; j++) {
        Req& q = v->active[j];
        if (!ran[j]) {
            fail_request(v, q);
        } else if (--q.left > 0) {
            keep.push_back(q);
        } else {
            v->busy[q.r.surface] = 0;
            v->done.push_back(q.r.id);
            v->req_ms.push_back((t1 - q.t_submit) * 1e3);
            v->completed++;
        }
    }
    v->active.swap(keep);
    return FELT_OK;
}

namespace {
int drain(std::vector<uint64_t>& from, uint64_t* ids, size_t cap, size_t* n) {
    if (!n || (cap && !ids)) return felt_fail(FELT_E_ARG, "null argument");
    const size_t k = std::min(cap, from.size());
    if (k) std::copy(from.begin(), from.begin() + (long)k, ids);
    from.erase(from.begin(), from.begin() + (long)k);
    *n = k;
    return FELT_OK;
}
}  // namespace

API int felt_server_poll(felt_server* v, uint64_t* ids, size_t cap, size_t* n) {
    if (!v) return felt_fail(FELT_E_ARG, "null server");
    return drain(v->done, ids, cap, n);
}

API int felt_server_poll_failed(felt_server* v, uint64_t* ids, size_t cap, size_t* n) {
    if (!v) return felt_fail(FELT_E_ARG, "null server");
    return drain(v->failed, ids, cap, n);
}

// Withdraws a queued or in-flight request (the first with this id); an in-flight one stops
// after the steps it has run.  It is reported neither as done nor as failed.
API int felt_server_cancel(felt_server* v, uint64_t id) {
    if (!v) return felt_fail(FELT_E_ARG, "null server");
    for (auto it = v->queue.begin(); it != v->queue.end(); ++it)
        if (it->r.id == id) {
            v->queue.erase(it);
            return FELT_OK;
        }
    for (auto it = v->active.begin(); it != v->active.end(); ++it)
        if (it->r.id == id) {
            v->busy[it->r.surface] = 0;
            v->active.erase(it);
            return FELT_OK;
        }
    return felt_fail(FELT_E_ARG, "no queued or in-flight request has this id");
}

API int felt_server_read_stats(felt_server* v, felt_server_stats* st) {
    if (!v || !st) return felt_fail(FELT_E_ARG, "null argument");
    std::memset(st, 0, sizeof(*st));
    st->ticks = v->ticks;
    st->requests_done = v->completed;
    st->requests_queued = v->queue.size();
    st->requests_active = v->active.size();
    st->admissions_deferred = v->deferred;
    st->requests_failed = v->n_failed;
    st->voxel_updates = v->voxel_updates;
    st->busy_s = v->busy_s;
    st->step_p50_ms = percentile(v->tick_ms, 50);
    st->step_p99_ms = percentile(v->tick_ms, 99);
    st->request_p50_ms = percentile(v->req_ms, 50);
    st->request_p99_ms = percentile(v->req_ms, 99);
    if (int rc = felt_download(v->s, &v->s->hc, v->s->ctr, sizeof(Ctr))) return rc;
    st->device_bytes = felt_device_bytes(v->s) + (v->polys ? felt_polys_device_bytes(v->polys) : 0);
    st->budget_bytes = v->budget;
    st->bricks_total = (uint64_t)v->s->cap;
    st->bricks_used = (uint64_t)(v->s->cap - v->s->hc.free_top);
    for (int k = 0; k < 4; k++)
        st->phase_ms[k] = v->ticks ? v->phase_s[k] * 1e3 / (double)v->ticks : 0.0;
    return FELT_OK;
}
