/*
 * felt_amd.h — C ABI of the MI355X (gfx950) sparse-field level-set engine.
 *
 * Every entry point replaces one member of Felt's header-only C++ API.  The C++ classes
 * Felt::Surface<D,L> and Felt::Polys<S> in include/felt_amd/Surface.hpp are thin wrappers over
 * these calls; the cited reference lines are the functions each call stands in for.
 *
 * Conventions
 *   - Positions are world coordinates (the reference's VecDi), int32, D values per position,
 *     x first.  A surface of size s has world offset -s/2 (C division), so it spans
 *     [-s/2, s - s/2) per axis (Surface.hpp:1716-1719).
 *   - Partition indices are the reference's children-grid PosIdx: row-major, last axis fastest,
 *     over the children grid (Util.hpp:23-40; PartitionedMixin.hpp:131-141).
 *   - Layer list index k = layer + L (Surface.hpp:672-675).
 *   - Every call returns FELT_OK (0) or a negative FELT_E_* code; felt_last_error() returns a
 *     thread-local message for the last failure.  The reference throws std::domain_error in the
 *     same places when FELT_EXCEPTIONS is defined (Common.hpp:167-172); the C++ wrapper throws it.
 *   - A handle owns one HIP stream; calls on one handle must come from one thread at a time.
 *   - Query calls that fill caller arrays take a capacity and report the full count in *n; when
 *     the capacity is too small they fill nothing and return FELT_E_CAPACITY.
 */
#ifndef FELT_AMD_H
#define FELT_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FELT_OK 0
#define FELT_E_ARG (-1)        /* invalid argument (shape, layer, position outside the grid) */
#define FELT_E_DEVICE (-2)     /* HIP runtime error */
#define FELT_E_NOMEM (-3)      /* device allocation failed or budget exhausted */
#define FELT_E_CAPACITY (-4)   /* output buffer too small; *n holds the required count */
#define FELT_E_STATE (-5)      /* call not valid in the handle's current state */
#define FELT_E_DELTA (-6)      /* |delta| > 1 or a layer jump > 1 (Surface.hpp:278-287, 441-459) */

typedef struct felt_surface felt_surface;
typedef struct felt_polys felt_polys;

/* Speed functor descriptor.  The reference takes any functor returning the zero-layer delta
 * (Surface.hpp:254-296, 1754-1802); the device path evaluates this closed set, all in fp32
 * with no fused multiply-add:
 *   FELT_SPEED_CONSTANT    delta = a
 *   FELT_SPEED_CURVATURE   delta = b * curv(x)
 *   FELT_SPEED_CONST_CURV  delta = a + b * curv(x)
 *   FELT_SPEED_NOISE_CURV  delta = clamp(a * (1 + c * noise(x)) + b * curv(x), -1, 1)
 * curv is Felt's difference-of-normals mean curvature (NumericMixin.hpp:72-156).  noise is
 * value noise on a lattice of `cell` voxels: lattice values are U[-1,1) drawn from
 * SplitMix64(seed ^ hash(lattice point)), trilinearly (bilinearly in 2D) interpolated. */
enum {
    FELT_SPEED_CONSTANT = 0,
    FELT_SPEED_CURVATURE = 1,
    FELT_SPEED_CONST_CURV = 2,
    FELT_SPEED_NOISE_CURV = 3
};
typedef struct felt_speed {
    int32_t kind;
    float a, b, c;
    uint64_t seed;
    int32_t cell;
    int32_t reserved;
} felt_speed;

/* Partition-list kinds for felt_surface_partition_list. */
enum {
    FELT_LIST_ISOGRID = 0,       /* isogrid().children().lookup().list(k) */
    FELT_LIST_DELTA = 1,         /* delta(k)          (Surface.hpp:637-640) */
    FELT_LIST_STATUS_CHANGE = 2, /* status_change(k)  (Surface.hpp:648-651) */
    FELT_LIST_AFFECTED = 3       /* affected().children().lookup().list(k) (Surface.hpp:693-696) */
};

typedef struct felt_stats {
    uint64_t active_isogrid_partitions; /* Surface::Stats (Surface.hpp:142-148, 623-629) */
    uint64_t active_delta_partitions;
    uint64_t band_voxels;               /* sum of the 2L+1 layer sizes now */
    uint64_t last_voxel_updates;        /* voxel updates counted for the last update call */
    uint64_t last_rounds;               /* distance/flush/expand rounds of the last update */
    uint64_t nan_events;                /* layer_id() calls that saw NaN since creation */
} felt_stats;

const char* felt_last_error(void);
/* Number of HIP devices visible, or a negative code. */
int felt_device_count(int* n);

/* ---- Surface lifecycle -------------------------------------------------------------- */
/* Surface(size, size_partition) (Surface.hpp:163-173).  D is 2 or 3, 1 <= L <= 8; a partition
 * holds at most 2^24 voxels and its first two edges are at most 65535. */
int felt_surface_create(int D, int L, const int32_t* size, const int32_t* part, felt_surface** out);
int felt_surface_destroy(felt_surface* s);
/* Geometry: children-grid size (3 values, unused axes 1) and partition count. */
int felt_surface_geometry(felt_surface* s, int32_t* children_size, uint64_t* num_partitions);

/* seed(pos) — singularity seed (Surface.hpp:207-241). */
int felt_surface_seed(felt_surface* s, const int32_t* pos);
/* Load a dense signed-distance field (row-major over the whole grid, world order).  Not a
 * reference member: it builds the state Surface::load (Surface.hpp:180-196) would hold for
 * that field.  Partitions holding a band voxel are active with every voxel set from phi and
 * keep the grid background L+1; band voxels are tracked in row-major order; other partitions
 * stay inactive with background -(L+1) if their first voxel is negative, else L+1. */
int felt_surface_load_dense(felt_surface* s, const float* phi, size_t n);

/* Load the field min_k(|x - c_k| - r_k) over n spheres (world centres, D per sphere),
 * evaluated on the device in fp64, clamped to +-(L+1) and rounded to fp32, exactly as
 * felt_surface_load_dense of that dense field.  Not a reference member (workload setup). */
int felt_surface_load_spheres(felt_surface* s, int32_t n, const double* centres,
                              const double* radii);

/* ---- Stepping ----------------------------------------------------------------------- */
/* update(fn) — global step (Surface.hpp:254-296, 739-762). */
int felt_surface_update(felt_surface* s, const felt_speed* fn);
/* update(pos_lo, pos_hi, fn) — bounded step (Surface.hpp:310-386, 404-430). */
int felt_surface_update_box(felt_surface* s, const int32_t* lo, const int32_t* hi,
                            const felt_speed* fn);
/* update(fn) for a host functor: fn(pos) was evaluated by the caller for the listed zero-layer
 * positions; every other zero-layer voxel gets delta 0 (Surface.hpp:254-296). */
int felt_surface_update_deltas(felt_surface* s, size_t n, const int32_t* pos, const float* delta);
/* Low-level trio update_start / delta / update_end (Surface.hpp:393-459). */
int felt_surface_update_start(felt_surface* s);
int felt_surface_delta(felt_surface* s, size_t n, const int32_t* pos, const float* delta);
int felt_surface_update_end(felt_surface* s);

/* ---- Checkpoint / resume ------------------------------------------------------------ */
/* Surface::save (Surface.hpp:180-183): the isogrid as the reference's cereal binary archive
 * (IsoGrid::write, PartitionedMixin.hpp:692-699), which Surface::load reads back.  With dst NULL
 * only *n is set to the archive's size; otherwise cap must be at least that.  Layer lists are
 * written in ascending position order. */
int felt_surface_save(felt_surface* s, uint8_t* dst, size_t cap, size_t* n);
/* Surface::load (Surface.hpp:192-196): a new surface from such an archive.  D, L, size and
 * partition size come from the archive; its offset must be -size/2 (as every Surface's is), its
 * background L+1, and its parent lists must agree with its child lists, else FELT_E_ARG. */
int felt_surface_load(const uint8_t* src, size_t n, felt_surface** out);
/* D, L, grid size and partition size of a surface (D entries each; e.g. after a load). */
int felt_surface_shape(felt_surface* s, int* D, int* L, int32_t* size, int32_t* part);

/* ---- Read-back ---------------------------------------------------------------------- */
/* isogrid().snapshot() into a caller buffer of prod(size) floats (PartitionedMixin.hpp:642-660). */
int felt_surface_snapshot(felt_surface* s, float* dst, size_t n);
/* get(pos) for n positions, background outside the grid (PartitionedMixin.hpp:594-601). */
int felt_surface_get(felt_surface* s, size_t n, const int32_t* pos, float* out);
/* Total size of layer `layer` (-L..L), summed over partitions. */
int felt_surface_layer_size(felt_surface* s, int layer, uint64_t* n);
/* Positions tracked in layer `layer` (order unspecified). */
int felt_surface_layer_positions(felt_surface* s, int layer, int32_t* pos, size_t cap, size_t* n);
/* Zero-layer positions inside [lo, min(hi+1, grid_hi)) (the voxels update(lo,hi,fn) visits,
 * Surface.hpp:313-353). */
int felt_surface_zero_layer_in_box(felt_surface* s, const int32_t* lo, const int32_t* hi,
                                   int32_t* pos, size_t cap, size_t* n);
/* Partition PosIdx list of a kind (FELT_LIST_*) and list index k = layer + L (order unspecified). */
int felt_surface_partition_list(felt_surface* s, int kind, int k, uint64_t* out, size_t cap,
                                size_t* n);
/* Voxels in the affected lookup of list k (affected().leafs(k, fn)). */
int felt_surface_affected_positions(felt_surface* s, int k, int32_t* pos, size_t cap, size_t* n);
/* 1 per active isogrid partition, 0 otherwise, for all partitions. */
int felt_surface_partition_active(felt_surface* s, uint8_t* out, size_t n);
/* Number of layer-k voxels in partition pos_idx (layer_size helper, test_Surface.cpp:1720-1737). */
int felt_surface_partition_layer_size(felt_surface* s, uint64_t pos_idx, int layer, uint64_t* n);
int felt_surface_stats(felt_surface* s, felt_stats* out);
/* Not a reference member: a consistency check of the device layout, for tests.  Counts the active
 * bricks whose face-neighbour record (the slots of the six neighbouring bricks, or which neighbour
 * partitions are inactive or outside the grid, and the in-grid extent) disagrees with the page
 * table; 0 after every operation. */
int felt_surface_check_records(felt_surface* s, uint64_t* mismatches);
/* ray(origin, dir) for n rays (Surface.hpp:468-616): origins and dirs hold n*D floats in world
 * coordinates, dirs unit vectors; out receives n*D floats, the hit on the zero curve or FLT_MAX
 * in every component for a miss (Surface::ray_miss, Surface.hpp:1805-1807).  Rays that are not
 * finite, start beyond 2^20 or whose direction is not of unit length fail with FELT_E_ARG. */
int felt_surface_ray(felt_surface* s, size_t n, const float* origins, const float* dirs,
                     float* out);
/* felt_surface_ray with no device allocation.  No counterpart in the reference: it is for
 * callers that cannot allocate device memory (allocation-free serving or picking loops, programs
 * that forbid hipMalloc after start-up).  Rays are staged through one fixed device array per GPU
 * (room for 65536 rays, 2.25 MiB, held until the process ends) instead of buffers the library
 * would allocate; n above 65536 fails with FELT_E_ARG.  Every call on a GPU shares that GPU's
 * array, so calls are serialised by a process-wide lock; results equal felt_surface_ray's. */
int felt_surface_ray_noalloc(felt_surface* s, size_t n, const float* origins, const float* dirs,
                             float* out);
/* The same with origins, dirs and out already in device memory (hipMalloc'd, n*D floats each);
 * returns once the results are written. */
int felt_surface_ray_device(felt_surface* s, size_t n, const float* origins, const float* dirs,
                            float* out);
/* Measurement: the average device time in microseconds of one felt_surface_ray launch over the
 * given rays (copied to the device once), from `reps` launches bracketed by two HIP events on the
 * surface's stream. */
int felt_surface_time_ray(felt_surface* s, size_t n, const float* origins, const float* dirs,
                          int reps, double* us);
/* ---- The isogrid's numeric API on the device (NumericMixin.hpp:72-500) ---------------- */
/* Operators a speed functor calls on the grid (Partitioned.hpp:160-189), evaluated on the device
 * at n positions with the code the device functor path compiles (include/felt_amd/device_grid.hpp):
 *   FELT_STENCIL_GET         get(pos)                                      1 value
 *   FELT_STENCIL_CURV        curv(pos), difference of normals (NumericMixin.hpp:72-156)    1 value
 *   FELT_STENCIL_DIVERGENCE  divergence(pos) = sum(gradF - gradB) / dx^2 (:164-175)       1 value
 *   FELT_STENCIL_GRAD        grad(pos), central or one-sided at the grid edge (:188-231) D values
 *   FELT_STENCIL_GRAD_E      gradE(pos), upwind (entropy satisfying) (:242-273)         D values
 *   FELT_STENCIL_GRAD_F, _B, _C   forward, backward, central differences (:279-358)    D values
 * dx is the grid spacing the derivatives divide by (the reference's dx(); 1 for a Surface's
 * isogrid).  felt_surface_stencil takes integer positions (n*D int32); felt_surface_stencil_real
 * takes real positions (n*D floats), every sample of which is then an N-linear interpolation, as
 * the reference's members templated on the position type do (get(VecDf) = interp, :390-481).
 * out receives n values, or n*D for the gradients. */
enum {
    FELT_STENCIL_GET = 0,
    FELT_STENCIL_CURV = 1,
    FELT_STENCIL_DIVERGENCE = 2,
    FELT_STENCIL_GRAD = 3,
    FELT_STENCIL_GRAD_E = 4,
    FELT_STENCIL_GRAD_F = 5,
    FELT_STENCIL_GRAD_B = 6,
    FELT_STENCIL_GRAD_C = 7
};
int felt_surface_stencil(felt_surface* s, int op, float dx, size_t n, const int32_t* pos,
                         float* out);
int felt_surface_stencil_real(felt_surface* s, int op, float dx, size_t n, const float* pos,
                              float* out);
/* fn(pos) of a felt_speed functor at n positions, computed by the device code the update kernel
 * runs for each zero-layer voxel (Surface.hpp:265-294), without taking a step.  Every position
 * must lie in an active partition, else FELT_E_ARG. */
int felt_surface_speed_eval(felt_surface* s, const felt_speed* fn, size_t n, const int32_t* pos,
                            float* out);

/* ---- Device functors (Surface.hpp:254-386, 1754-1802) -------------------------------- */
/* The isogrid as a kernel reads it: plain values and device pointers, valid until the surface's
 * next update or destruction.  A voxel v of the pool is brick slot v / pv, storage leaf v % pv;
 * a brick is cut into tiles of 2^tile_shift[a] voxels per axis, tiles row-major in the brick,
 * voxels row-major in a tile.  include/felt_amd/device_grid.hpp wraps it as Felt's grid API. */
typedef struct felt_device_grid {
    int32_t D, L;
    int32_t size[3];         /* grid size (unused axes 1) */
    int32_t offset[3];       /* world position of grid position 0: -size/2 */
    int32_t part[3];         /* partition size */
    int32_t part_shift[3];   /* log2(part[a]) when a power of two, else -1 */
    int32_t cstride[3];      /* children-grid strides, row-major, last axis fastest */
    int32_t pbase;           /* pool partition of children-grid index 0 */
    int32_t tile_shift[3];   /* log2 of the tile edge */
    int32_t tile_stride[3];  /* leaf stride of one tile step */
    int32_t in_stride[3];    /* leaf stride of one voxel step inside a tile */
    int32_t tiles[3];        /* tiles per axis in a brick */
    int32_t tile_vshift;     /* log2 of the tile volume */
    uint32_t pv;             /* voxels per brick */
    float dx;                /* grid spacing of the derivatives (1) */
    const int32_t* page;     /* pool partition -> brick slot, or -1 when inactive */
    const float* pbg;        /* pool partition -> value of an inactive partition */
    const float* phi;        /* brick pool values */
    const int32_t* slot_org; /* 4 per slot: grid position of the brick's first voxel, flags */
} felt_device_grid;
int felt_surface_device_grid(felt_surface* s, felt_device_grid* out);

/* One update(fn) whose fn runs in the caller's own kernel:
 *   felt_surface_update_begin (global step) or felt_surface_update_box_begin (bounded step,
 *   Surface.hpp:310-386) queues update_start and the listing of the zero-layer voxels fn is
 *   called for on the surface's stream, and describes them in *job: job->voxels[i] for
 *   job->span[0] <= i < job->span[1] (span is device memory), each needing delta[i] = fn(pos of
 *   voxel), at most job->capacity entries;
 *   the caller launches its kernel on job->stream (a hipStream_t);
 *   felt_surface_update_commit queues the rest of the step (apply, distance, flush, expand and
 *   the round loop, Surface.hpp:404-430, 739-762) and finishes it as felt_surface_update does;
 *   felt_surface_update_cancel instead abandons the step, leaving values and layers unchanged.
 * Between begin and commit or cancel every other call on the surface fails with FELT_E_STATE. */
typedef struct felt_zero_job {
    felt_device_grid grid;
    const uint32_t* voxels;
    const int32_t* span;
    float* delta;
    uint64_t capacity;
    void* stream;
} felt_zero_job;
int felt_surface_update_begin(felt_surface* s, felt_zero_job* job);
int felt_surface_update_box_begin(felt_surface* s, const int32_t* lo, const int32_t* hi,
                                  felt_zero_job* job);
int felt_surface_update_commit(felt_surface* s);
int felt_surface_update_cancel(felt_surface* s);

/* Launch a global update(fn) with a device functor as one captured HIP graph (default on):
 * the first three rounds run without host reads, later rounds are launched from the host.
 * Off: the same kernels are launched one by one.  Not a reference member. */
int felt_surface_set_graph(felt_surface* s, int enable);
/* Voxel updates and rounds of the last update call, without device traffic. */
int felt_surface_step_info(felt_surface* s, uint64_t* voxel_updates, uint64_t* rounds);
/* Kernel timing with HIP events on the handle's stream, for measurement.  Not a reference
 * member (the reference has no profiling hooks).  Kernel ids: 0 speed, 1 zero-layer apply,
 * 2 distance pass 1 and 3 distance pass 2 of a step's first round, 4 flush (with the round's
 * lookup and expansion records), 5 expand (three kernels), 6 band count, 7 affected search,
 * 8 march count, 9 march emit, 10 / 11 distance passes of later rounds.  `items` sums the list entries the launches processed (first-round
 * and march kernels).  Events are attached to the dispatches (hipExtLaunchKernelGGL); a
 * profiled global step launches its kernels one by one instead of replaying its graph.
 * Enabling or disabling clears the totals. */
int felt_surface_profile(felt_surface* s, int enable);
int felt_surface_profile_read(felt_surface* s, int kernel, uint64_t* launches, double* ms,
                              uint64_t* items);
/* Times one pass of a global step on the surface's current state.  The band list is rebuilt
 * as a step starts, then `reps` copies of the pass's launches of one step are captured as one
 * graph, which is replayed once to warm up and once between two HIP events on the handle's
 * stream.  Passes:
 * FELT_PASS_SPEED, one launch of the last global update's device functor over the zero layer;
 * FELT_PASS_DIST1, the first-round distance pass 1, one launch per non-zero layer;
 * FELT_PASS_BAND, band compaction (four kernels).  They write only scratch (deltas, distances,
 * the band list), so the surface is unchanged.  Outputs: microseconds per launch, list entries
 * per launch (for FELT_PASS_BAND the pool voxels scanned), launches per replay.  Not a
 * reference member. */
#define FELT_PASS_SPEED 0
#define FELT_PASS_DIST1 1
#define FELT_PASS_BAND 2
int felt_surface_time_pass(felt_surface* s, int pass, int reps, double* us_per_launch,
                           double* items_per_launch, int* launches);
/* Launch shape of a handle's step kernels, chosen by felt_surface_autotune.  The reference's
 * only tuning knob is the OpenMP chunk threshold (Surface.hpp:19-35); this is its counterpart
 * here.  Not a reference member.
 * threads_per_block, grid_cap: block size and largest grid of the per-entry passes over the
 *   layer lists (speed, zero-layer apply, the distance passes, flush, expansion);
 * voxels_per_lane: band compaction, voxels of a brick each lane scans (8, 16, 32 or 64), which
 *   sets the threads per brick: ceil(brick voxels / voxels_per_lane), at most 512;
 * threads_per_brick, us_default, us_tuned: outputs of felt_surface_autotune, the threads per
 *   brick of the chosen compaction and the timed passes (microseconds per step's worth of
 *   them) with the default shape (256, 4096, by brick size) and with the chosen one. */
typedef struct felt_launch_config {
    int32_t threads_per_block;
    int32_t grid_cap;
    int32_t voxels_per_lane;
    int32_t threads_per_brick;
    double us_default;
    double us_tuned;
} felt_launch_config;
/* Times candidate launch shapes on the surface's current state (as felt_surface_time_pass
 * does, `reps` copies per graph; the surface is unchanged) and keeps the fastest: block size x
 * grid cap over the distance passes (and the speed pass when a device functor has run), then
 * voxels per lane over band compaction.  The choice is also kept in a process-wide cache under
 * (device, D, L, partition size, log2 of the band size); with reps == 0 nothing is timed and
 * the cached choice for the current state is applied if there is one (else the handle keeps
 * its shape; `out` then reports it with zero times).  Every shape gives identical results. */
int felt_surface_autotune(felt_surface* s, int reps, felt_launch_config* out);
/* Sets a launch shape directly (threads_per_block a multiple of 64 in [64, 1024], grid_cap >=
 * 1, voxels_per_lane 8, 16, 32 or 64; the outputs are ignored). */
int felt_surface_set_launch_config(felt_surface* s, const felt_launch_config* cfg);
/* Most rounds one update may take (default 100000; at least 1).  The reference's round loop
 * (Surface.hpp:739-762) runs while any voxel changes layer and can cycle forever (an inside
 * blob without a zero layer that touches an out-of-band outside voxel flips sign each round);
 * a step here fails with FELT_E_STATE after this many rounds instead.  Not a reference
 * member. */
int felt_surface_set_round_limit(felt_surface* s, int32_t max_rounds);
/* Block until all queued device work of the handle is done. */
int felt_surface_sync(felt_surface* s);

/* ---- Polygonisation (Polys.hpp:29-201) ---------------------------------------------- */
int felt_polys_create(felt_surface* s, felt_polys** out);
int felt_polys_destroy(felt_polys* p);
int felt_polys_notify(felt_polys* p);
int felt_polys_march(felt_polys* p);
int felt_polys_invalidate(felt_polys* p);
/* Queue one partition for the next march, as driving Impl::Poly::Single::march on that child
 * directly does (Poly.hpp:138-147; test_Poly.cpp:29-60). */
int felt_polys_mark(felt_polys* p, uint64_t pos_idx);
/* changes(): partitions marched by the last march (Polys.hpp:198-201). */
int felt_polys_changes(felt_polys* p, uint64_t* out, size_t cap, size_t* n);
/* children().get(pos_idx): active flag and vertex / simplex counts (Poly.hpp:176-189). */
int felt_polys_child_info(felt_polys* p, uint64_t pos_idx, int* active, uint64_t* nvtx,
                          uint64_t* nspx);
/* Mesh of one child: vtx gets nvtx*(D or 2D) floats (pos, then normal in 3D), spx gets nspx*D
 * vertex indices local to the child (PolyMixin.hpp:34-68, 125-163). */
int felt_polys_child_mesh(felt_polys* p, uint64_t pos_idx, float* vtx, size_t vtx_cap,
                          uint32_t* spx, size_t spx_cap);
/* Totals over all active children. */
int felt_polys_totals(felt_polys* p, uint64_t* nvtx, uint64_t* nspx);

/* ---- Batched surfaces --------------------------------------------------------------- */
/* Not a reference type.  A batch is a pool of n independent surfaces that share D, L and the
 * partition size and each have their own grid size.  Member i behaves exactly as a
 * felt_surface of its size (same layers, values and meshes bit for bit), but one step of many
 * members is one launch sequence over all their bricks (for a global step, one replay of a
 * captured HIP graph).  Member positions are world coordinates of that member's grid
 * (offset -size/2); member partition indices are its own children-grid PosIdx. */
typedef struct felt_batch felt_batch;
/* sizes: n*D grid sizes.  max_bricks: brick pool capacity, 0 for one brick per partition of
 * every member; either way at most 2^32 / prod(part) bricks (32-bit voxel ids). */
int felt_batch_create(int D, int L, const int32_t* part, int32_t n, const int32_t* sizes,
                      uint64_t max_bricks, felt_batch** out);
int felt_batch_destroy(felt_batch* b);
int felt_batch_count(felt_batch* b, int32_t* n);
/* Device bytes held by the batch (bricks, tables, step lists) and brick slots in use / total. */
int felt_batch_memory(felt_batch* b, uint64_t* device_bytes, uint64_t* bricks_used,
                      uint64_t* bricks_total);
int felt_batch_seed(felt_batch* b, int32_t i, const int32_t* pos);
int felt_batch_load_dense(felt_batch* b, int32_t i, const float* phi, size_t n);
/* Loads member i with the field min_k(|x - c_k| - r_k) over n spheres (world centres, D per
 * sphere), evaluated in fp64, clamped to +-(L+1) and rounded to fp32, with the semantics of
 * felt_surface_load_dense for that dense field. */
int felt_batch_load_spheres(felt_batch* b, int32_t i, int32_t n, const double* centres,
                            const double* radii);
/* update(fn) of every member, member i with speeds[i], as one step (Surface.hpp:254-296). */
int felt_batch_update(felt_batch* b, const felt_speed* speeds);
/* update(lo, hi, fn) of the n distinct members ids[j], member ids[j] over its box
 * lo[j*D..], hi[j*D..] with speeds[j], as one step (Surface.hpp:310-386); the other members
 * are not touched. */
int felt_batch_update_boxes(felt_batch* b, int32_t n, const int32_t* ids, const int32_t* lo,
                            const int32_t* hi, const felt_speed* speeds);
/* felt_batch_update split in two: queue the step on the batch's stream and return; then finish
 * it.  Steps of different batches queued before waiting run concurrently on the device.  Until
 * felt_batch_wait, every other call on the batch fails with FELT_E_STATE. */
int felt_batch_update_async(felt_batch* b, const felt_speed* speeds);
int felt_batch_wait(felt_batch* b);
/* A group of batches stepped by one launch: one HIP graph holding each batch's step kernels as
 * an independent chain, replayed once per step on the group's own stream, so a workload split
 * over batches of different partition sizes is one graph launch per step.  The
 * batches must be on one device, distinct, and outlive the group.  Not a reference member. */
typedef struct felt_group felt_group;
int felt_group_create(felt_batch* const* batches, int32_t n, felt_group** out);
int felt_group_destroy(felt_group* g);
/* update(fn) of every member of every batch: speeds[k] holds batch k's per-member speeds (as
 * felt_batch_update_async).  Until felt_group_wait, the batches refuse other calls
 * (FELT_E_STATE), as after felt_batch_update_async. */
int felt_group_update_async(felt_group* g, const felt_speed* const* speeds);
int felt_group_wait(felt_group* g);
int felt_batch_snapshot(felt_batch* b, int32_t i, float* dst, size_t n);
int felt_batch_layer_size(felt_batch* b, int32_t i, int layer, uint64_t* n);
int felt_batch_layer_positions(felt_batch* b, int32_t i, int layer, int32_t* pos, size_t cap,
                               size_t* n);
/* Stats of member i; last_voxel_updates / last_rounds are those of the last batch step. */
int felt_batch_stats(felt_batch* b, int32_t i, felt_stats* out);
/* felt_surface_check_records over every member of the batch. */
int felt_batch_check_records(felt_batch* b, uint64_t* mismatches);
/* felt_surface_save / load for member i of a batch.  A load replaces the member's state (its
 * size, partition size and L must match the archive's); felt_batch_clear returns the member's
 * bricks to the pool and leaves it the empty grid of a newly created surface (for eviction). */
int felt_batch_save(felt_batch* b, int32_t i, uint8_t* dst, size_t cap, size_t* n);
int felt_batch_load(felt_batch* b, int32_t i, const uint8_t* src, size_t n);
int felt_batch_clear(felt_batch* b, int32_t i);
/* felt_surface_ray over the members of a batch: ray i is cast against member members[i]. */
int felt_batch_ray(felt_batch* b, size_t n, const int32_t* members, const float* origins,
                   const float* dirs, float* out);
int felt_batch_step_info(felt_batch* b, uint64_t* voxel_updates, uint64_t* rounds);
int felt_batch_set_graph(felt_batch* b, int enable);
int felt_batch_time_pass(felt_batch* b, int pass, int reps, double* us_per_launch,
                         double* items_per_launch, int* launches);
int felt_batch_autotune(felt_batch* b, int reps, felt_launch_config* out);
int felt_batch_set_launch_config(felt_batch* b, const felt_launch_config* cfg);
int felt_batch_set_round_limit(felt_batch* b, int32_t max_rounds);
int felt_batch_sync(felt_batch* b);
/* Polygonisation of every member: felt_polys_notify / march / invalidate / totals / destroy
 * act on all members at once; the per-member queries below take the member index. */
int felt_batch_polys_create(felt_batch* b, felt_polys** out);
int felt_batch_polys_changes(felt_polys* p, int32_t i, uint64_t* out, size_t cap, size_t* n);
int felt_batch_polys_child_info(felt_polys* p, int32_t i, uint64_t pos_idx, int* active,
                                uint64_t* nvtx, uint64_t* nspx);
int felt_batch_polys_child_mesh(felt_polys* p, int32_t i, uint64_t pos_idx, float* vtx,
                                size_t vtx_cap, uint32_t* spx, size_t spx_cap);

/* ---- Served surfaces ---------------------------------------------------------------- */
/* Not a reference type.  A server keeps n surfaces (one batch) resident and runs requests
 * against them: each request is `steps` bounded updates update(lo, hi, fn) of one surface
 * (Surface.hpp:310-386), each followed by repolygonisation of the changed partitions
 * (Polys.hpp:92-165) when the server polygonises.  felt_server_tick admits queued requests in
 * arrival order — one in flight per surface, at most max_active in flight, and only while the
 * resident bricks plus the step lists of the in-flight set fit budget_bytes — then advances
 * every in-flight request by one step, all of them in one batched step.  Requests submitted
 * while others run join at the next tick.  A surface's result depends only on its own requests
 * and their order, never on what else is in flight. */
typedef struct felt_server felt_server;
typedef struct felt_request {
    int32_t surface;       /* member index */
    int32_t lo[3], hi[3];  /* world box, D values used */
    felt_speed speed;
    int32_t steps;
    int32_t reserved;
    uint64_t id;           /* caller's id, returned by felt_server_poll */
} felt_request;
typedef struct felt_server_stats {
    uint64_t ticks;                /* batched steps run */
    uint64_t requests_done, requests_queued, requests_active;
    uint64_t admissions_deferred;  /* ticks whose admission stopped at the budget */
    uint64_t voxel_updates;        /* sum over ticks of the affected-set sizes */
    double busy_s;                 /* wall time inside ticks' device work */
    double step_p50_ms, step_p99_ms;        /* per-tick latency */
    double request_p50_ms, request_p99_ms;  /* submit to completion */
    uint64_t device_bytes, budget_bytes, bricks_used, bricks_total;
    double phase_ms[4];  /* mean per tick: bounded update, notify, march, band refresh */
    uint64_t requests_failed;      /* requests that left without completing (felt_server_tick) */
} felt_server_stats;
/* sizes: n*D grid sizes.  Bricks get at most three quarters of budget_bytes. */
int felt_server_create(int D, int L, const int32_t* part, int32_t n, const int32_t* sizes,
                       uint64_t budget_bytes, int32_t max_active, int32_t polygonise,
                       felt_server** out);
int felt_server_destroy(felt_server* v);
/* The server's batch (borrowed): load surfaces through it before felt_server_start. */
int felt_server_batch(felt_server* v, felt_batch** out);
/* The server's polygoniser (borrowed), for mesh queries with felt_batch_polys_*. */
int felt_server_polys(felt_server* v, felt_polys** out);
/* Reads every surface's band size and, when polygonising, marches every surface once. */
int felt_server_start(felt_server* v);
/* Latency target for admission: a request joins a tick only while a running model of tick
 * time (a fixed part plus a cost per band voxel inside the in-flight boxes), times a margin for
 * its error (the running mean plus 3 standard deviations of actual / predicted tick time),
 * stays within tick_ms, so that the target bounds the tail of tick time rather than its mean;
 * a request is always admitted into an empty tick.  0 (the default) disables it.
 * felt_server_start also reserves the step lists the in-flight set can need (within the
 * budget), the march's scratch and a spare mesh arena, so that ticks do not allocate. */
int felt_server_set_target(felt_server* v, double tick_ms);
/* Queues n requests; every one is checked first, so a call queues all of them or none. */
int felt_server_submit(felt_server* v, const felt_request* r, size_t n);
/* One batched step; *stepped = requests in the step (0 when nothing is in flight).
 * A request fails -- leaves the server without completing, its id reported by
 * felt_server_poll_failed -- when its step lists alone exceed the budget (at admission), or when
 * its bounded step alone may activate more partitions than the brick pool has free (the step is
 * then refused and its surface left unchanged; the other requests of the tick still run).  Any
 * other error fails every in-flight request, whose surfaces may be partly stepped, and is
 * returned. */
int felt_server_tick(felt_server* v, int32_t* stepped);
/* Ids of requests completed (felt_server_poll) or failed (felt_server_poll_failed) since the
 * last such poll, up to cap; ids may be NULL only when cap is 0. */
int felt_server_poll(felt_server* v, uint64_t* ids, size_t cap, size_t* n);
int felt_server_poll_failed(felt_server* v, uint64_t* ids, size_t cap, size_t* n);
/* Withdraws the queued or in-flight request with this id (the first one submitted); an
 * in-flight request keeps the steps it has run.  FELT_E_ARG when there is none. */
int felt_server_cancel(felt_server* v, uint64_t id);
int felt_server_read_stats(felt_server* v, felt_server_stats* out);

/* ---- Tables ------------------------------------------------------------------------- */
/* Copy of the packed marching-cubes triangle table the kernels use (256 entries). */
int felt_mc_table(uint64_t* out, size_t n);
/* Copy of the packed marching-squares segment table (16 entries: segment count in bits 28-31,
 * edge k of the segment list in bits 4k..4k+3). */
int felt_ms_table(uint32_t* out, size_t n);

#ifdef __cplusplus
}
#endif
#endif /* FELT_AMD_H */
