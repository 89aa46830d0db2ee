// felt_surface.h — host-side state of one device-resident pool of member surfaces (a single
// surface is a pool of one member).
#pragma once

#include <hip/hip_runtime.h>

#include <array>
#include <cstdint>
#include <map>
#include <string>
#include <utility>
#include <vector>

#include "felt_kernels.h"

// Growable device buffer (contents are not preserved when it grows).
template <class T>
struct DBuf {
    T* ptr = nullptr;
    size_t cap = 0;
    int ensure(size_t n);
    void release();
};

// Grow-only pinned host buffer for transfers on a handle's stream (contents not preserved when
// it grows).  A transfer from or into it is complete once the stream has been synchronised.
template <class T>
struct HBuf {
    T* ptr = nullptr;
    size_t cap = 0;
    int ensure(size_t n);
    void release();
};

// Optional per-kernel timing with HIP events on the surface's stream.
enum ProfKernel { PK_SPEED, PK_APPLY, PK_DIST1, PK_DIST2, PK_FLUSH, PK_EXPAND, PK_COMPACT,
                  PK_AFFECTED, PK_MARCH_COUNT, PK_MARCH_EMIT, PK_DIST1_LATE, PK_DIST2_LATE,
                  PK_COUNT };
struct ProfRec {
    hipEvent_t a, b;
    int kernel;
    long long items;
};
// items >= 0: entries processed; items < 0: band segment -(items+1), resolved after the step.
struct Prof {
    bool on = false;
    long long hold_ticks = 0;  // wall-clock ticks a profiled step waits behind k_hold
    std::vector<ProfRec> pending;
    std::vector<hipEvent_t> spare;
    long long launches[PK_COUNT] = {};
    double ms[PK_COUNT] = {};
    long long items[PK_COUNT] = {};
};

// Launch shape of the per-entry list passes and of band compaction (felt_launch_config).
struct LaunchCfg {
    int tpb = 256;   // threads per block
    int gcap = 4096; // largest grid
    int vpl = 0;     // band compaction voxels per lane; 0: 64 for bricks above 512 voxels, else 8
};

// Device bytes of the step lists per listed voxel: band (index, delta, tag, new layer), two
// lookups (index, tag, distance, new layer), expansion records (voxel, side), zero-layer list
// (index, delta), a bounded step's affected set sorted by layer (index, tag).
constexpr uint64_t kListBytes = 10 + 2 * 10 + 5 + 8 + 5;

struct felt_surface {
    felt::Geo G{};
    felt::Pool P{};
    std::vector<felt::Mem> mems;     // member geometry (host copy of P.mem)
    hipStream_t st = nullptr;
    int device = 0;  // the HIP device the handle was created on
    int cap = 0;     // brick pool slots
    int* free_stack = nullptr;
    felt::Ctr* ctr = nullptr;
    felt::Ctr hc{};            // host copy of the counters
    felt::Ctr* hcp = nullptr;  // pinned staging of that copy
    uint8_t* amark = nullptr;
    int *ukey = nullptr, *tflag = nullptr, *tlist = nullptr, *nflag = nullptr, *nlist = nullptr;
    unsigned* rstamp = nullptr;  // per partition: the last step stamp the reserve count put there
    unsigned *dmask = nullptr, *smask = nullptr;
    int* act = nullptr;
    int* dlist = nullptr;  // partitions deactivated since the neighbour records were fixed
    // band compaction
    DBuf<uint32_t> band;
    DBuf<float> band_d;
    DBuf<int8_t> band_tag, band_tk;
    DBuf<int> chg_cnt;
    DBuf<int> blk_cnt, blk_off, visit;  // band scan: block totals, offsets, visit list
    std::vector<int> band_base;
    // bounded-update zero layer
    DBuf<uint32_t> zlist;
    DBuf<float> zdelta;
    // lookups: 0 = affected, 1 = affected buffer
    DBuf<uint32_t> lu_v[2];
    DBuf<int8_t> lu_tag[2];
    DBuf<float> lu_d[2];
    DBuf<int8_t> lu_tk[2];
    // a bounded step's affected set sorted by layer (first round only)
    DBuf<uint32_t> srt_v;
    DBuf<int8_t> srt_tag;
    // status-change records
    DBuf<uint32_t> sc_xv;
    DBuf<int8_t> sc_xside;
    // scratch
    DBuf<float> vox;
    bool vox_zeroed = false;
    DBuf<int> tmp_i;
    DBuf<float> tmp_f;
    DBuf<uint32_t> tmp_u;
    DBuf<int> step_ids;            // members of the current bounded step (device)
    std::vector<int> h_step_ids;   // and on the host
    int n_step_ids = -1;           // -1: every member
    int* bbase = nullptr;           // band segment starts, NL+1 entries (device)
    felt::Speed* d_speed = nullptr;  // per-member speed descriptors read by the speed kernel
    felt::Speed* h_speed = nullptr;  // pinned staging copy (nmem entries)
    bool speed_valid = false;        // d_speed holds the last global update's functors
    size_t list_cap = 0;             // capacity of every per-step list
    int64_t hint = -1;               // exact entries of the current round's lists, -1: unknown
    std::vector<int> hbase;          // host copy of bbase after a profiled step
    bool band_known = false;         // band_now is the current tracked-voxel count
    size_t band_now = 0;
    int unroll = 3;                  // rounds of a global step launched without host reads
    bool auto_unroll = false;        // raise `unroll` when steps need more rounds (batches)
    bool step_pending = false;       // a global step is queued and not yet finished
    bool use_graph = true;
    LaunchCfg cfg;                   // felt_surface_autotune / set_launch_config
    int max_rounds = 100000;         // rounds of one step before it fails (set_round_limit)
    bool graph_valid = false;
    hipGraphExec_t graph_exec = nullptr;
    hipGraph_t graph = nullptr;       // the captured step behind graph_exec (a group's child)
    unsigned graph_ver = 0;           // counts its captures
    // device-functor steps: the step's kernels before and after the caller's speed kernel, as two
    // graphs (rebuilt whenever graph_valid is)
    bool fn_graph_valid = false;
    hipGraphExec_t gx_pre = nullptr, gx_post = nullptr;
    int job = 0;  // a device-functor step is open: 1 global, 2 bounded

    bool in_update = false;
    std::vector<std::pair<std::array<int, 3>, float>> pending_deltas;
    std::map<std::array<int, 3>, size_t> pending_index;
    uint64_t last_updates = 0, last_rounds = 0;
    struct felt_polys* polys = nullptr;
    Prof prof;
};

// A batch handle: a pool of member surfaces.
struct felt_batch {
    felt_surface* s = nullptr;
};

// Batches stepped by one graph launch (felt_group_*): one graph holding each batch's step
// kernels as an independent chain.
struct felt_group {
    std::vector<felt_surface*> s;
    hipStream_t st = nullptr;
    hipGraphExec_t exec = nullptr;
    std::vector<unsigned> ver;        // each child's graph_ver when `exec` was built
    hipEvent_t done = nullptr;        // end of the group launch, awaited by each batch's stream
    hipEvent_t fork = nullptr;        // start of the group's capture, joined by each batch
    std::vector<hipEvent_t> ready;    // each batch's speed upload
    bool pending = false;
};
int felt_group_launch_impl(felt_group* g, const felt_speed* const* speeds);
int felt_group_wait_impl(felt_group* g);
void felt_group_release(felt_group* g);

// Brackets one launch with events when profiling is on.
void prof_begin(felt_surface* s, int kernel, long long items);
void prof_end(felt_surface* s);
int prof_collect(felt_surface* s);

extern thread_local std::string g_felt_err;
int felt_fail(int code, const std::string& msg);

int felt_surface_init(felt_surface* s);
void felt_surface_release(felt_surface* s);
felt::Lu felt_lu(felt_surface* s, int i);
int felt_update_start_impl(felt_surface* s);
// Global step of every member, member i with speed fns[i].
int felt_update_global(felt_surface* s, const felt_speed* fns);
// The same split in two: queue the step on the handle's stream, then finish it (counter read
// and any rounds beyond the unrolled ones).
int felt_update_global_launch(felt_surface* s, const felt_speed* fns);
int felt_global_step_finish(felt_surface* s);
int felt_autotune_impl(felt_surface* s, int reps, felt_launch_config* out);
int felt_set_launch_cfg_impl(felt_surface* s, const felt_launch_config* c);
int felt_time_pass_impl(felt_surface* s, int pass, int reps, double* us, double* items,
                        int* launches);
int felt_update_global_deltas(felt_surface* s, size_t m, const int32_t* gpos, const float* vals);
// Bounded step of n distinct members ids[j], each over its world box [lo, hi] (3 ints per
// entry) with speed fns[j]; other members are not touched.  `bands` (optional) holds the
// members' current band sizes, which size the step's lists; without it they are counted.
int felt_update_boxes_impl(felt_surface* s, int n, const int* ids, const int32_t* lo,
                           const int32_t* hi, const felt_speed* fns,
                           const long long* bands = nullptr);
int felt_delta_impl(felt_surface* s, size_t m, const int32_t* gpos, const float* vals);
// Device-functor steps (felt_surface_update_begin / _box_begin / _commit / _cancel).
int felt_update_begin_impl(felt_surface* s, felt_zero_job* job);
int felt_update_box_begin_impl(felt_surface* s, const int32_t* lo3, const int32_t* hi3,
                               felt_zero_job* job);
int felt_update_commit_impl(felt_surface* s);
int felt_update_cancel_impl(felt_surface* s);
// The isogrid of member m as device code reads it.
void felt_fill_device_grid(const felt_surface* s, int m, felt_device_grid* out);
// Grid operator op at n world positions of member m (ipos integer, or fpos real).
int felt_stencil_impl(felt_surface* s, int m, int op, float dx, size_t n, const int32_t* ipos,
                      const float* fpos, float* out);
// The speed functor fn at n grid positions gpos (3 ints each) of member m.
int felt_speed_eval_impl(felt_surface* s, int m, const felt_speed* fn, size_t n,
                         const int32_t* gpos, float* out);
int felt_update_end_public(felt_surface* s);
// Per-member operations; positions are member grid coordinates (3 ints per entry) unless
// stated.
int felt_track_points(felt_surface* s, int m, const std::vector<int>& gpos,
                      const std::vector<float>& val, const std::vector<int>& layer);
int felt_load_dense_impl(felt_surface* s, int m, const float* phi, size_t n);
// Loads the field min_k(|x - c_k| - r_k) (world coordinates, fp64) as load_dense would.
int felt_load_spheres_impl(felt_surface* s, int m, int n, const double* centres,
                           const double* radii);
int felt_snapshot_impl(felt_surface* s, int m, float* dst, size_t total);
int felt_get_impl(felt_surface* s, int m, size_t n, const int32_t* gpos, float* out);
// World positions of member m's voxels tracked in `layer`.
int felt_collect_layer_impl(felt_surface* s, int m, int layer, std::vector<int>& out);
// World positions of member m's entries of lookup lu_i with tag `tag`.
int felt_lu_positions_impl(felt_surface* s, int m, int lu_i, int tag, std::vector<int>& out);
int felt_band_total(felt_surface* s, size_t* out);
int felt_download(felt_surface* s, void* dst, const void* src, size_t bytes);
// Tracked voxels of members ids[0..n).
int felt_member_bands(felt_surface* s, const int* ids, int n, std::vector<long long>& out);
// Step lists sized for a step over `band` listed voxels.
int felt_reserve_lists(felt_surface* s, size_t band);
// Capacity the step lists would grow to for a step over `band` listed voxels.
size_t felt_list_cap_for(const felt_surface* s, size_t band);
// Device bytes held by the handle.
uint64_t felt_device_bytes(const felt_surface* s);
int felt_sync_impl(felt_surface* s);
// Surface::ray for n rays of member m0 (or members[i] when given), host arrays in and out.
int felt_ray_impl(felt_surface* s, const int32_t* members, int m0, size_t n, const float* origins,
                  const float* dirs, float* out);
// The same on device arrays, queued on the surface's stream.
int felt_ray_device(felt_surface* s, const int* d_members, int m0, size_t n, const float* d_org,
                    const float* d_dir, float* d_out);
// Checkpoints in the reference's on-disk format (cereal archive of the isogrid).
int felt_archive_geometry(const uint8_t* src, size_t n, int* D, int* L, int* size, int* part);
int felt_save_impl(felt_surface* s, int m, uint8_t* dst, size_t cap, size_t* n);
int felt_load_impl(felt_surface* s, int m, const uint8_t* src, size_t n);
int felt_member_clear_impl(felt_surface* s, int m);
int felt_alloc_listed(felt_surface* s, int grid);
// Face-neighbour records brought in line with the page table for the partitions activated
// (s->nlist) and deactivated (s->dlist) since they were last fixed; queued on the stream.
int felt_nbr_fix(felt_surface* s);
// Active bricks whose face-neighbour record disagrees with the page table (0 when consistent).
int felt_nbr_check(felt_surface* s, uint64_t* bad);
// felt_surface_ray without device allocation: rays staged through a fixed device array.
int felt_ray_noalloc_impl(felt_surface* s, size_t n, const float* origins, const float* dirs,
                          float* out);
// Average device time of one launch over n rays of member 0, reps launches between two events.
int felt_time_ray_impl(felt_surface* s, size_t n, const float* origins, const float* dirs, int reps,
                       double* us);
