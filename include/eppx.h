/*
 * eppx.h -- C ABI of libeppx, the MI355X (gfx950) HIP implementation of
 * EtherParallelParticles.jl's SPH time step: cell-list neighbour search,
 * the cubic-spline / Wendland C2 kernels, continuity, momentum (pressure +
 * classic or artificial viscosity), Tait / linear EOS and the integrator.
 *
 * Plain C: opaque handles, POD structs, pointers and sizes; every function
 * returns an int32_t status (EPPX_OK == 0).  Each entry point names the
 * reference function it replaces (paths relative to the reference package).
 *
 * Data crosses the boundary only in the reference's own host layout:
 * column-major (n_capacity, n_columns) Int32 / Float32 property tables plus
 * is_alive and cell_index vectors (src/Class/ParticleSystem/ParticleSystemBase.jl:12-18).
 * The library re-lays them out as cell-sorted structure-of-arrays buffers in
 * HBM and never keeps a caller's pointer after a call returns.
 *
 * Calls on one context are issued in order on the context's HIP stream and
 * return without waiting, except *_download, eppx_sync, eppx_pair_count and
 * eppx_export_pairs, which wait.  Device-side conditions the reference leaves
 * undefined (per-cell or per-particle neighbour capacity exceeded) raise a
 * sticky flag that the next waiting call reports as EPPX_E_CAPACITY.
 * One host thread at a time per context.
 */
#ifndef EPPX_H
#define EPPX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define EPPX_ABI_VERSION 1

/* ---------------------------------------------------------------- status */
enum {
    EPPX_OK = 0,
    EPPX_E_ARG = 1,       /* invalid argument / layout / parameter block      */
    EPPX_E_HIP = 2,       /* HIP runtime error (message in eppx_last_error)    */
    EPPX_E_CAPACITY = 3,  /* cell or neighbour count above max_neighbour_number */
    EPPX_E_NONFINITE = 4, /* non-finite density found by eppx_check_finite     */
    EPPX_E_STATE = 5,     /* call order violated (e.g. interaction before search) */
    EPPX_E_NOMEM = 6      /* device or host allocation failed                  */
};

typedef struct eppx_ctx eppx_ctx; /* device + stream + last error               */
typedef struct eppx_ps eppx_ps;   /* particle system (Class.ParticleSystem)     */
typedef struct eppx_ns eppx_ns;   /* neighbour system (Class.NeighbourSystem)   */

/* Field for field order the reference's struct Domain2D
 * (src/Class/Domain/Domain2D.jl:10-26). */
typedef struct {
    float gap, gap_square;
    int32_t n_x, n_y, n;
    float first_x, last_x, first_y, last_y;
    float span_x, span_y, gap_x, gap_y, gap_x_inv, gap_y_inv;
} eppx_domain2d;

/* Column map of a particle system: the 0-based start column of each named
 * field in the int / float tables, -1 when absent.  It is the reference's
 * NamedIndex offsets minus one (src/Class/ParticleSystem/NamedIndex.jl:50-91)
 * after basic_index_map_dict renaming (BasicIndex.jl:20-44). */
typedef struct {
    int32_t n_int, n_float, max_neighbour_number;
    int32_t Tag, IsMovable, nCount, nIndex;                   /* int columns   */
    int32_t PositionVec, VelocityVec, dVelocityVec, Mass, Volume, Density, dDensity, Pressure, H, Gap;
    int32_t SumWeight, SumWeightedDensity;                    /* float columns */
    int32_t nRVec, nR, nW, nDW, nHInv;                        /* neighbour cols */
    int32_t SumWeightedPressure;                              /* float column   */
} eppx_layout;

/* Search criteria (src/Algorithm/NeighbourSearch.jl:141-164). */
enum { EPPX_CRIT_DEFAULT = 0, EPPX_CRIT_SYMMETRY = 1 };

/* Smoothing kernels (src/SPH/Kernel/CubicSpline.jl, WendlandC2.jl), 2D. */
enum { EPPX_KERNEL_CUBIC = 0, EPPX_KERNEL_WENDLAND_C2 = 1 };
enum { EPPX_EOS_LINEAR = 0, EPPX_EOS_TAIT = 1 };
enum { EPPX_VISC_CLASSIC = 0, EPPX_VISC_ARTIFICIAL = 1 };

/* The reference's actions are Julia closures inlined into its kernels; they
 * cannot cross a C ABI, so the dam-break driver's composite actions
 * (example/collapse_dry/same/main.jl:74-144) and two generic field actions
 * are exposed as IDs with a parameter block. */
enum { /* search actions (fused after each particle's neighbour list) */
    EPPX_SEARCH_NONE = 0,               /* defaultSelfaction!, Action.jl:27-35            */
    EPPX_SEARCH_KERNEL_CONTINUITY = 1   /* sComputeKernelAndiContinuity!, main.jl:74-90    */
};
enum { /* self actions */
    EPPX_SELF_CONTINUITY_EOS = 1,       /* sContinuity!, main.jl:92-105                     */
    EPPX_SELF_GRAVITY_MOVE = 2,         /* sMomentum!, main.jl:120-126                      */
    EPPX_SELF_FILTER = 3,               /* sFilter, main.jl:136-144                         */
    EPPX_SELF_EXTRAPOLATE_PRESSURE = 4, /* sExtrapolatePressure! on wall particles,
                                           Pressure.jl:125-143                             */
    EPPX_SELF_SCALE_FIELD = 10          /* field[I] *= value                                */
};
enum { /* interactions */
    EPPX_INTER_MOMENTUM = 1,            /* iMomentum!, main.jl:107-118                      */
    EPPX_INTER_FILTER = 2,              /* iFilter, main.jl:128-134                         */
    EPPX_INTER_EXTRAPOLATE_PRESSURE = 3,/* iExtrapolatePressure! (2D) for wall i, fluid j,
                                           w = the pair's kernel value, Pressure.jl:81-101 */
    EPPX_INTER_ADD_PER_NEIGHBOUR = 10   /* field[I] += value once per neighbour             */
};

/* Parameter block of every physics action (fields an action does not use
 * are ignored).  Tags are the driver's FLUID_TAG / WALL_TAG. */
typedef struct {
    int32_t fluid_tag, wall_tag, kernel, eos, visc;
    float dt;           /* time step                                   */
    float rho0, p0;     /* reference density, background pressure      */
    float c0;           /* linear EOS sound speed: p = c0^2 (rho - rho0) + p0 */
    float tait_b;       /* Tait: p = tait_b ((rho/rho0)^7 - 1) + p0     */
    float gx, gy;       /* gravity                                     */
    float mu;           /* classic viscosity                           */
    float alpha, beta;  /* artificial viscosity                        */
    float av_c;         /* artificial-viscosity sound speed            */
    float coefficient;  /* pressure correction coefficient (0 in the driver) */
} eppx_physics;

/* Fields addressable by the generic field actions. */
enum {
    EPPX_FIELD_POSITION = 0, EPPX_FIELD_VELOCITY = 1, EPPX_FIELD_DVELOCITY = 2, EPPX_FIELD_MASS = 3,
    EPPX_FIELD_VOLUME = 4, EPPX_FIELD_DENSITY = 5, EPPX_FIELD_DDENSITY = 6, EPPX_FIELD_PRESSURE = 7
};
typedef struct {
    int32_t field, component; /* component 0/1 for vector fields, 0 for scalars */
    float value;
} eppx_field_op;

/* Arithmetic mode of the pair kernels.  EXACT evaluates every pair term
 * with the reference's operation order, correctly rounded divide / sqrt and
 * no FMA; FAST uses FMA and hardware reciprocal / sqrt in the physics terms
 * (the accept / reject decision of the criterion stays exact in both). */
enum { EPPX_MATH_EXACT = 0, EPPX_MATH_FAST = 1 };

/* ------------------------------------------------------------ context */
int32_t eppx_abi_version(void);
/* Replaces KernelAbstractions.device!(Backend, Device) (src/DataIO/Config/Input.jl:17). */
int32_t eppx_ctx_create(int32_t device, eppx_ctx **out);
int32_t eppx_ctx_destroy(eppx_ctx *ctx);
const char *eppx_last_error(const eppx_ctx *ctx);
/* Environment.synchronize / KernelAbstractions.synchronize
 * (src/Environment/Parallel.jl:104-108): waits, then reports sticky flags. */
int32_t eppx_sync(eppx_ctx *ctx);
/* Selects EPPX_MATH_EXACT (default) or EPPX_MATH_FAST for later launches. */
int32_t eppx_set_math_mode(eppx_ctx *ctx, int32_t mode);

/* ------------------------------------------------------------- domain */
/* Domain2D{Int32,Float32}(gap, first_x, first_y, last_x, last_y),
 * src/Class/Domain/Domain2D.jl:28-67 (host arithmetic, no device work). */
int32_t eppx_domain2d_init(double gap, double first_x, double first_y, double last_x, double last_y,
                           eppx_domain2d *out);
/* inside / indexCartesianFromPosition / indexLinearFromPosition,
 * Domain2D.jl:110-137 (host; 1-based indices). */
int32_t eppx_domain2d_locate(const eppx_domain2d *d, float x, float y, int32_t *inside, int32_t *i, int32_t *j,
                             int32_t *linear);

/* ---------------------------------------------------- particle system */
/* ParticleSystemBase(parallel, n_capacity, n_int, n_float),
 * src/Class/ParticleSystem/ParticleSystemBase.jl:52-67. */
int32_t eppx_ps_create(eppx_ctx *ctx, int32_t n_capacity, const eppx_layout *layout, eppx_ps **out);
int32_t eppx_ps_destroy(eppx_ps *ps);
/* Class.asyncto!/syncto!(device, host), ParticleSystemBase.jl:20-50 and
 * ParticleSystem.jl:275-307: host tables -> device. */
int32_t eppx_ps_upload(eppx_ps *ps, int32_t n_particles, const int32_t *is_alive, const int32_t *cell_index,
                       const int32_t *int_props, const float *float_props);
/* Class.asyncto!/syncto!(host, device): device -> host tables.  When ns is
 * given and the layout declares them, nCount / nIndex / nRVec / nR / nW /
 * nDW / nHInv are reconstructed from the last search (nW / nDW / nHInv only
 * where the last search action was EPPX_SEARCH_KERNEL_CONTINUITY). */
int32_t eppx_ps_download(eppx_ps *ps, eppx_ns *ns, int32_t *n_particles, int32_t *is_alive, int32_t *cell_index,
                         int32_t *int_props, float *float_props);
/* No counterpart in the reference.  eppx_ps_download for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving or
 * output loops, harnesses that forbid hipMalloc): the same arguments and
 * results, staged through a fixed 128 MiB device buffer that the library
 * holds on every device it runs on (no hipMalloc / hipFree per call).  Fails
 * with EPPX_E_ARG when the tables need more staging than that:
 * n_capacity * (n_int + n_float + 2) 4-byte words, each table rounded up to
 * 64 words, must not exceed 33,554,432.  Calls on one device share the
 * buffer: calls that overlap on one device race (waits). */
int32_t eppx_ps_download_noalloc(eppx_ps *ps, eppx_ns *ns, int32_t *n_particles, int32_t *is_alive,
                                 int32_t *cell_index, int32_t *int_props, float *float_props);
/* Class.get_alive_n_particles, ParticleSystem.jl:154-158 (waits). */
int32_t eppx_ps_alive_count(eppx_ps *ps, int32_t *out);

/* --------------------------------------------------- neighbour system */
/* Class.NeighbourSystem(parallel, domain, active_pair, NonePeriodicBoundaryPolicy;
 * max_neighbour_number), src/Class/NeighbourSystem/NeighbourSystem.jl:68-86
 * with ActivePair.jl:17-33.  adjacency is the (max_tag, max_tag) column-major
 * Int32 matrix, max_tag <= 32. */
int32_t eppx_ns_create(eppx_ctx *ctx, const eppx_domain2d *domain, const int32_t *adjacency, int32_t max_tag,
                       int32_t max_neighbour_number, eppx_ns **out);
int32_t eppx_ns_destroy(eppx_ns *ns);
/* neighbour_cell_index_count_ / _list_ built by
 * device_initializeNeighbourSystem!, NeighbourSystemBase.jl:28-48:
 * count (n_cells) and list (n_cells, 9) column-major, 1-based (waits). */
int32_t eppx_ns_cell_stencil(eppx_ns *ns, int32_t *count, int32_t *list);

/* --------------------------------------------------------- algorithm */
/* Algorithm.search! / static_search! / dynamic_search!,
 * src/Algorithm/NeighbourSearch.jl:323-413 (clean!, insert :12-47,
 * find :166-227, fused action).  params: eppx_physics for
 * EPPX_SEARCH_KERNEL_CONTINUITY, may be NULL for EPPX_SEARCH_NONE. */
int32_t eppx_search(eppx_ps *ps, eppx_ns *ns, int32_t criterion, int32_t action, const void *params,
                    size_t params_size);
/* Algorithm.selfaction! / static_ / dynamic_, src/Algorithm/Action.jl:39-117.
 * params: eppx_physics, or eppx_field_op for EPPX_SELF_SCALE_FIELD. */
int32_t eppx_selfaction(eppx_ps *ps, int32_t action, const void *params, size_t params_size);
/* Algorithm.interaction! / static_ / dynamic_, src/Algorithm/Action.jl:121-206,
 * over the neighbour lists of the last search on ns.  params: eppx_physics,
 * or eppx_field_op for EPPX_INTER_ADD_PER_NEIGHBOUR. */
int32_t eppx_interaction(eppx_ps *ps, eppx_ns *ns, int32_t action, const void *params, size_t params_size);
/* n_steps iterations of the dam-break loop body (main.jl:166-183):
 * sContinuity!, iMomentum!, sMomentum!, search!(symmetry criterion,
 * sComputeKernelAndiContinuity!) and, when filter_interval > 0, the Shepard
 * filter every filter_interval steps counted from step_offset + 1.  Replays a
 * captured hipGraph; requires one prior eppx_search on (ps, ns). */
int32_t eppx_step(eppx_ps *ps, eppx_ns *ns, const eppx_physics *phys, int32_t n_steps, int32_t filter_interval,
                  int64_t step_offset);

/* Autotuner (no counterpart in the reference, whose kernels take the
 * backend's default work-group size): times the FAST search action and the
 * momentum pass on the current state for each pair-kernel shape (256 or 512
 * threads per block; per-neighbour data gathered from HBM or staged in the
 * block's LDS tile), trials each after one warm-up, keeps the fastest for
 * each kernel, and leaves the state as it was.  choice[4] receives (search
 * block, search staged, momentum block, momentum staged); us[8], when given,
 * the average microseconds (search then momentum; shapes (256, gathered),
 * (512, gathered), (256, staged), (512, staged)).  EXACT math keeps
 * (256, gathered).  Needs a prior eppx_search on ns (waits). */
int32_t eppx_autotune(eppx_ps *ps, eppx_ns *ns, const eppx_physics *phys, int32_t trials, int32_t *choice,
                      double *us);

/* Sets the pair-kernel shapes by hand (choice as eppx_autotune returns it;
 * FAST math; results do not depend on them). */
int32_t eppx_set_pair_shape(eppx_ps *ps, const int32_t *choice);

/* Σ nCount over alive particles after the last search on ns (waits). */
int32_t eppx_pair_count(eppx_ps *ps, eppx_ns *ns, int64_t *out);
/* Running Σ nCount over every search since the last reset (waits). */
int32_t eppx_pair_total(eppx_ps *ps, int64_t *out, int32_t reset);
/* Neighbour lists of the last search in the reference's list order (stencil
 * order, then ascending particle index within a cell), as CSR over particle
 * index 0..n_particles-1: offsets (n_particles + 1) and 0-based js
 * (capacity entries; fails with EPPX_E_CAPACITY when too small) (waits). */
int32_t eppx_export_pairs(eppx_ps *ps, eppx_ns *ns, int64_t *offsets, int32_t *js, int64_t capacity);
/* Number of pair-kernel blocks (search and momentum) that could not stage
 * their neighbourhood in on-chip memory and read it from HBM instead, since
 * the last reset (a performance diagnostic; results do not depend on it). */
int32_t eppx_pair_fallback_blocks(eppx_ps *ps, int64_t *out, int32_t reset);
/* Number of particle walks (search and momentum passes) whose neighbour rows
 * held more than 64 slots each, since the last reset: those walks test the
 * candidates one at a time from HBM instead of through 64-bit accept masks
 * (cells much larger than 2h, or crowded cells; a performance diagnostic,
 * results do not depend on it). */
int32_t eppx_pair_row_fallback(eppx_ps *ps, int64_t *out, int32_t reset);
/* Profiler calibration (no counterpart in the reference): runs `blocks`
 * blocks of 256 threads, each lane issuing iters x 8 independent FP32 FMAs
 * (8 x iters VALU instructions per wave plus loop control), once to warm up
 * and once timed; us receives the timed launch's duration (waits).  The
 * known instruction count scales rocprofv3's VALU counters. */
int32_t eppx_valu_calibration(eppx_ctx *ctx, int32_t blocks, int32_t iters, double *us);
/* Returns EPPX_E_NONFINITE when a live particle's density, pressure,
 * position or velocity is not finite (waits). */
int32_t eppx_check_finite(eppx_ps *ps);
/* Average device duration in microseconds of the named kernel family over
 * the launches timed since the last reset (HIP events on the context's
 * stream; timing is enabled by eppx_set_timing). */
int32_t eppx_set_timing(eppx_ctx *ctx, int32_t enable);
int32_t eppx_kernel_time(eppx_ctx *ctx, const char *kernel, double *avg_us, int64_t *launches, int32_t reset);

/* ---------------------------------------------------------------- batch
 * No counterpart in the reference, whose driver runs one case per process.
 * A batch keeps many independent cases (each with its own Domain2D, H and
 * dt; the rest of eppx_physics shared) resident in one set of device
 * buffers with a segmented cell list, and advances every case that has steps
 * left with one launch per kernel per step (a recorded hipGraph).  Each case
 * stops after exactly its own step budget; cases can be added and removed
 * between steps (continuous batching).  All cases use the layout, adjacency
 * and max_neighbour_number given at creation, symmetrySearchCriterion and the
 * dam-break step (sContinuity!, iMomentum!, sMomentum!, search! with
 * sComputeKernelAndiContinuity!). */
typedef struct eppx_batch eppx_batch;
/* Device bytes per particle slot, for admission by HBM footprint. */
int32_t eppx_batch_bytes_per_particle(int32_t *out);
int32_t eppx_batch_create(eppx_ctx *ctx, const eppx_layout *layout, int32_t particle_capacity, int32_t case_capacity,
                          const int32_t *adjacency, int32_t max_tag, int32_t max_neighbour_number, eppx_batch **out);
int32_t eppx_batch_destroy(eppx_batch *b);
/* Adds one case from its reference tables (n_particles rows, column-major,
 * as eppx_ps_upload) and runs its first search; it then advances for `steps`
 * batch steps.  Fails with EPPX_E_CAPACITY when slots or case slots run out,
 * EPPX_E_ARG when the physics differs from the batch's other than in dt. */
int32_t eppx_batch_add(eppx_batch *b, const eppx_domain2d *domain, const eppx_physics *phys, int32_t n_particles,
                       const int32_t *is_alive, const int32_t *cell_index, const int32_t *int_props,
                       const float *float_props, int32_t steps, int32_t *case_id);
/* n_steps batch steps: every case with steps left advances one step each. */
int32_t eppx_batch_step(eppx_batch *b, int32_t n_steps);
/* Steps left of every case slot (case_capacity entries; -1 when unused) (waits). */
int32_t eppx_batch_steps_left(eppx_batch *b, int32_t *steps_left);
/* A case's reference tables (as eppx_ps_download, without the neighbour
 * columns) (waits). */
int32_t eppx_batch_download(eppx_batch *b, int32_t case_id, int32_t *n_particles, int32_t *is_alive,
                            int32_t *cell_index, int32_t *int_props, float *float_props);
/* Releases a case's slots and case slot. */
int32_t eppx_batch_remove(eppx_batch *b, int32_t case_id);
int32_t eppx_batch_info(eppx_batch *b, int32_t *slots_in_use, int32_t *cases_in_use);
/* eppx_autotune for the batch's current cases (waits). */
int32_t eppx_batch_autotune(eppx_batch *b, int32_t trials, int32_t *choice, double *us);
/* Running Σ nCount over the batch's searches since the last reset (waits). */
int32_t eppx_batch_pair_total(eppx_batch *b, int64_t *out, int32_t reset);

/* ------------------------------------------------------------------- 3D
 * The reference leaves every 3D layer a stub: Domain3D (AbstractDomain.jl:119
 * "TODO: add 3D support"), the 27-cell stencil (NeighbourSystemBase.jl:50-57),
 * insert and find (NeighbourSearch.jl:49, :229), Cuboid lattices
 * (Geometry3D.jl:89-94); its SPH library (Library.jl:21-34, Continuity.jl,
 * Pressure.jl, Viscosity.jl, Motion.jl:61-75, the 3D sGravity! State.jl:49-63)
 * and kernel functions (sigma for N = 3: CubicSpline.jl:15, WendlandC2.jl:15)
 * are written for any dimension.  These entry points extend the 2D ones
 * axis by axis and keep their arguments, results and errors; results are
 * checked against the 3D restatement of the oracle (no reference output
 * exists to pin them).  Vector columns (PositionVec, VelocityVec,
 * dVelocityVec) are 3 wide and nRVec holds 3 floats per neighbour. */
typedef struct eppx_ps3d eppx_ps3d; /* 3D particle system                        */
typedef struct eppx_ns3d eppx_ns3d; /* 3D neighbour system                       */

/* struct Domain2D's fields with a z axis. */
typedef struct {
    float gap, gap_square;
    int32_t n_x, n_y, n_z, n;
    float first_x, last_x, first_y, last_y, first_z, last_z;
    float span_x, span_y, span_z, gap_x, gap_y, gap_z, gap_x_inv, gap_y_inv, gap_z_inv;
} eppx_domain3d;

/* eppx_physics with the z component of gravity (sGravity!, State.jl:49-63). */
typedef struct {
    eppx_physics base;
    float gz;
} eppx_physics3d;

/* Domain2D.jl:28-67 per axis: n_x = floor(span_x / gap), gap_x = span_x / n_x,
 * gap_x_inv = 1 / gap_x (host FP32 arithmetic). */
int32_t eppx_domain3d_init(double gap, double first_x, double first_y, double first_z, double last_x, double last_y,
                           double last_z, eppx_domain3d *out);
/* Domain2D.jl:110-137 per axis; linear = i + n_x (j - 1) + n_x n_y (k - 1). */
int32_t eppx_domain3d_locate(const eppx_domain3d *d, float x, float y, float z, int32_t *inside, int32_t *i,
                             int32_t *j, int32_t *k, int32_t *linear);
/* eppx_ns_create (NeighbourSystem.jl:68-86) over a Domain3D grid. */
int32_t eppx_ns3d_create(eppx_ctx *ctx, const eppx_domain3d *domain, const int32_t *adjacency, int32_t max_tag,
                         int32_t max_neighbour_number, eppx_ns3d **out);
int32_t eppx_ns3d_destroy(eppx_ns3d *ns);
/* The 27-cell stencil, device_initializeNeighbourSystem! (NeighbourSystemBase.jl:28-48)
 * with a z offset innermost: count (n_cells) and list (n_cells, 27) column-major,
 * 1-based, unused entries 0. */
int32_t eppx_ns3d_cell_stencil(eppx_ns3d *ns, int32_t *count, int32_t *list);
/* eppx_ps_create / _destroy / _upload / _download / _alive_count
 * (ParticleSystemBase.jl:52-67, ParticleSystem.jl:154-158, :275-307) for 3D. */
int32_t eppx_ps3d_create(eppx_ctx *ctx, int32_t n_capacity, const eppx_layout *layout, eppx_ps3d **out);
int32_t eppx_ps3d_destroy(eppx_ps3d *ps);
int32_t eppx_ps3d_upload(eppx_ps3d *ps, int32_t n_particles, const int32_t *is_alive, const int32_t *cell_index,
                         const int32_t *int_props, const float *float_props);
int32_t eppx_ps3d_download(eppx_ps3d *ps, eppx_ns3d *ns, int32_t *n_particles, int32_t *is_alive,
                           int32_t *cell_index, int32_t *int_props, float *float_props);
int32_t eppx_ps3d_alive_count(eppx_ps3d *ps, int32_t *out);
/* Algorithm.search! (NeighbourSearch.jl:323-413) in 3D; params: an
 * eppx_physics3d for EPPX_SEARCH_KERNEL_CONTINUITY, may be NULL for
 * EPPX_SEARCH_NONE. */
int32_t eppx_search3d(eppx_ps3d *ps, eppx_ns3d *ns, int32_t criterion, int32_t action, const void *params,
                      size_t params_size);
/* Algorithm.selfaction! (Action.jl:39-117): EPPX_SELF_CONTINUITY_EOS and
 * EPPX_SELF_GRAVITY_MOVE, params an eppx_physics3d. */
int32_t eppx_selfaction3d(eppx_ps3d *ps, int32_t action, const void *params, size_t params_size);
/* Algorithm.interaction! (Action.jl:121-206): EPPX_INTER_MOMENTUM, params an
 * eppx_physics3d, over the pairs of the last search on ns. */
int32_t eppx_interaction3d(eppx_ps3d *ps, eppx_ns3d *ns, int32_t action, const void *params, size_t params_size);
/* n_steps iterations of the loop body main.jl:166-183 (no filter) in 3D. */
int32_t eppx_step3d(eppx_ps3d *ps, eppx_ns3d *ns, const eppx_physics3d *phys, int32_t n_steps);
/* eppx_pair_count, eppx_pair_total, eppx_export_pairs for 3D. */
int32_t eppx_pair_count3d(eppx_ps3d *ps, eppx_ns3d *ns, int64_t *out);
int32_t eppx_pair_total3d(eppx_ps3d *ps, int64_t *out, int32_t reset);
int32_t eppx_export_pairs3d(eppx_ps3d *ps, eppx_ns3d *ns, int64_t *offsets, int32_t *js, int64_t capacity);

#ifdef __cplusplus
}
#endif
#endif
