/*
 * eppx_oracle.h -- CPU restatement of EtherParallelParticles.jl's SPH time step.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the checker the HIP library is compared
 * against.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg may load it; the product path never links or calls it.
 *
 * Every function restates one reference function (file:line cited beside
 * it) on the reference's own data layout: column-major (capacity, columns)
 * int/float property tables, per-cell contained-particle lists of fixed
 * capacity, and materialised per-particle neighbour lists.  Arithmetic is
 * IEEE binary32 in the reference's written operation order (left to right,
 * no FMA contraction, correctly rounded divide and sqrt).
 *
 * Indices in the tables are 1-based exactly as in the reference (cell
 * indices, particle indices stored in nIndex / contained lists); C array
 * offsets are 0-based.
 */
#ifndef EPPX_ORACLE_H
#define EPPX_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Field order follows struct Domain2D, src/Class/Domain/Domain2D.jl:10-26. */
typedef struct {
    float gap, gap_square;
    int32_t n_x, n_y, n;
    float first_x, last_x, first_y, last_y;
    float span_x, span_y, gap_x, gap_y, gap_x_inv, gap_y_inv;
} or_domain2d;

/* 0-based column of each named field in the int / float tables, -1 when the
 * particle system does not declare it (NamedIndex.jl:50-91 gives 1-based
 * offsets; the caller subtracts one). */
typedef struct {
    int32_t n_capacity, n_int, n_float, max_neighbour_number;
    int32_t Tag, IsMovable, nCount, nIndex;
    int32_t PositionVec, VelocityVec, dVelocityVec, Mass, Volume, Density, dDensity, Pressure, H, Gap;
    int32_t SumWeight, SumWeightedDensity;
    int32_t nRVec, nR, nW, nDW, nHInv;
    int32_t SumWeightedPressure;
} or_layout;

enum { OR_KERNEL_CUBIC = 0, OR_KERNEL_WENDLAND_C2 = 1 };
enum { OR_EOS_LINEAR = 0, OR_EOS_TAIT = 1 };
enum { OR_VISC_CLASSIC = 0, OR_VISC_ARTIFICIAL = 1 };
enum { OR_CRIT_DEFAULT = 0, OR_CRIT_SYMMETRY = 1 };

/* Parameters of the dam-break driver's composite actions
 * (example/collapse_dry/same/main.jl:50-70, with the Tait EOS and artificial
 * viscosity the dam-break benchmark specifies). */
typedef struct {
    int32_t fluid_tag, wall_tag, kernel, eos, visc;
    float dt, rho0, p0, c0, tait_b, gx, gy, mu, alpha, beta, av_c, coefficient;
} or_physics;

/* Reference cell-list state (NeighbourSystemBase.jl:12-18, ActivePair.jl). */
typedef struct {
    or_domain2d domain;
    int32_t n_cells, max_contained;
    int32_t *contained_count;      /* (n_cells)                       */
    int32_t *contained_list;       /* (n_cells, max_contained) col-major */
    int32_t *neighbour_cell_count; /* (n_cells)                       */
    int32_t *neighbour_cell_list;  /* (n_cells, 9) col-major            */
    int32_t max_tag;
    int32_t *adjacency;            /* (max_tag, max_tag) col-major      */
} or_neighbour_system;

/* Error bits returned by the search (the reference has undefined behaviour
 * in these cases; the restatement reports them instead). */
enum { OR_ERR_CELL_OVERFLOW = 1, OR_ERR_NEIGHBOUR_OVERFLOW = 2 };

void or_domain2d_init(double gap, double first_x, double first_y, double last_x, double last_y, or_domain2d *d);
void or_index_cartesian_from_position(const or_domain2d *d, float x, float y, int32_t *i, int32_t *j);
int32_t or_index_linear_from_position(const or_domain2d *d, float x, float y);
int32_t or_inside(const or_domain2d *d, float x, float y);

/* Allocates the lists; returns 0 on success. */
int32_t or_ns_create(const or_domain2d *d, int32_t max_contained, const int32_t *adjacency, int32_t max_tag,
                     or_neighbour_system *ns);
void or_ns_destroy(or_neighbour_system *ns);

/* Kernel functions (src/SPH/Kernel/CubicSpline.jl, WendlandC2.jl), 2D. */
float or_kernel_value(int32_t kernel, float r, float hinv);
float or_kernel_gradient(int32_t kernel, float r, float hinv);
float or_kernel_value0(int32_t kernel, float h);
float or_harmonic(float a, float b);
float or_invharmonic(float a, float b);

/* Search (clean + insert + find + fused action).  action: 0 none,
 * 1 kernel+continuity (sComputeKernelAndiContinuity!).  serial_insert != 0
 * inserts in ascending particle order (deterministic slot order); 0 uses
 * OpenMP atomics like the reference's concurrent insert.  n_threads <= 0
 * uses the OpenMP default. */
int32_t or_search(const or_layout *L, int32_t n_particles, int32_t *is_alive, int32_t *cell_index, int32_t *IP,
                  float *FP, or_neighbour_system *ns, int32_t criterion, int32_t action, const or_physics *phys,
                  int32_t serial_insert, int32_t n_threads);

/* Self actions: 1 continuity+volume+EOS (sContinuity!), 2 gravity+move
 * (sMomentum!), 3 filter finish (sFilter), 4 extrapolated wall pressure
 * (sExtrapolatePressure! on wall particles), 10 multiply float column
 * `column` by `value`. */
void or_selfaction(const or_layout *L, int32_t n_particles, const int32_t *is_alive, int32_t *IP, float *FP,
                   int32_t action, const or_physics *phys, int32_t column, float value, int32_t n_threads);

/* Interactions: 1 momentum (iMomentum!), 2 filter sums (iFilter),
 * 3 extrapolation sums (iExtrapolatePressure! for wall I, fluid J),
 * 10 add `value` to float column `column` once per neighbour. */
void or_interaction(const or_layout *L, int32_t n_particles, const int32_t *is_alive, int32_t *IP, float *FP,
                    int32_t action, const or_physics *phys, int32_t column, float value, int32_t n_threads);

/* Summation order of the per-neighbour sums (continuity, momentum): 0 (the
 * default) the reference's list order, 1 reversed.  Global to the library. */
void or_set_pair_order(int32_t reversed);

/* Magnitude bound over I's neighbours of the continuity terms (what = 0,
 * out[I]; after a search with the kernel + continuity action) or the
 * momentum terms (what = 1, out[2 I + k]; before iMomentum!): each term with
 * its inner sums of products taken in absolute value (the scale of the
 * float32 rounding of the sum in any order). */
void or_pair_term_magnitude(const or_layout *L, int32_t n_particles, const int32_t *is_alive, int32_t *IP,
                            float *FP, const or_physics *phys, int32_t what, double *out);

/* n_steps of the dam-break loop body (main.jl:166-183): sContinuity,
 * iMomentum, sMomentum, search (+ filter every filter_interval steps when
 * filter_interval > 0).  Returns the OR of search error bits; writes the
 * total accepted pair count of each step's search into pairs_per_step
 * (may be NULL). */
int32_t or_run_steps(const or_layout *L, int32_t n_particles, int32_t *is_alive, int32_t *cell_index, int32_t *IP,
                     float *FP, or_neighbour_system *ns, const or_physics *phys, int32_t n_steps,
                     int32_t filter_interval, int32_t serial_insert, int32_t n_threads, int64_t *pairs_per_step);

/* ------------------------------------------------------------------ 3D
 * eppx_oracle3d.c: the same algorithm with a z axis.  The reference's 3D
 * layers are stubs (AbstractDomain.jl:119, NeighbourSystemBase.jl:50-57,
 * NeighbourSearch.jl:49, :229), so these functions extend the 2D ones axis
 * by axis and restate the dimension-generic SPH library with N = 3. */

/* struct Domain2D's fields with a z axis. */
typedef struct {
    float gap, gap_square;
    int32_t n_x, n_y, n_z, n;
    float first_x, last_x, first_y, last_y, first_z, last_z;
    float span_x, span_y, span_z, gap_x, gap_y, gap_z, gap_x_inv, gap_y_inv, gap_z_inv;
} or_domain3d;

typedef struct {
    or_domain3d domain;
    int32_t n_cells, max_contained;
    int32_t *contained_count;      /* (n_cells)                          */
    int32_t *contained_list;       /* (n_cells, max_contained) col-major */
    int32_t *neighbour_cell_count; /* (n_cells)                          */
    int32_t *neighbour_cell_list;  /* (n_cells, 27) col-major            */
    int32_t max_tag;
    int32_t *adjacency;
} or_neighbour_system3d;

typedef struct {
    or_physics p;
    float gz; /* sGravity! in 3D, State.jl:49-63 */
} or_physics3d;

void or3_domain_init(double gap, double first_x, double first_y, double first_z, double last_x, double last_y,
                     double last_z, or_domain3d *d);
void or3_index_cartesian_from_position(const or_domain3d *d, float x, float y, float z, int32_t *i, int32_t *j,
                                       int32_t *k);
int32_t or3_index_linear_from_position(const or_domain3d *d, float x, float y, float z);
int32_t or3_inside(const or_domain3d *d, float x, float y, float z);
int32_t or3_ns_create(const or_domain3d *d, int32_t max_contained, const int32_t *adjacency, int32_t max_tag,
                      or_neighbour_system3d *ns);
void or3_ns_destroy(or_neighbour_system3d *ns);
float or3_kernel_value(int32_t kernel, float r, float hinv);
float or3_kernel_gradient(int32_t kernel, float r, float hinv);
float or3_kernel_value0(int32_t kernel, float h);
/* As or_search; PositionVec, VelocityVec and dVelocityVec are 3 wide and
 * nRVec holds 3 floats per neighbour. */
int32_t or3_search(const or_layout *L, int32_t n_particles, int32_t *is_alive, int32_t *cell_index, int32_t *IP,
                   float *FP, or_neighbour_system3d *ns, int32_t criterion, int32_t action, const or_physics3d *phys,
                   int32_t serial_insert, int32_t n_threads);
/* 1 continuity + volume + EOS (sContinuity!), 2 gravity + move (sMomentum!). */
void or3_selfaction(const or_layout *L, int32_t n_particles, const int32_t *is_alive, int32_t *IP, float *FP,
                    int32_t action, const or_physics3d *phys, int32_t n_threads);
/* 1 momentum (iMomentum!). */
void or3_interaction(const or_layout *L, int32_t n_particles, const int32_t *is_alive, int32_t *IP, float *FP,
                     int32_t action, const or_physics3d *phys, int32_t n_threads);
int32_t or3_run_steps(const or_layout *L, int32_t n_particles, int32_t *is_alive, int32_t *cell_index, int32_t *IP,
                      float *FP, or_neighbour_system3d *ns, const or_physics3d *phys, int32_t n_steps,
                      int32_t serial_insert, int32_t n_threads, int64_t *pairs_per_step);
/* or_pair_term_magnitude in 3D: what = 0, out[I]; what = 1, out[3 I + k]. */
void or3_pair_term_magnitude(const or_layout *L, int32_t n_particles, const int32_t *is_alive, int32_t *IP,
                             float *FP, const or_physics3d *phys, int32_t what, double *out);

#ifdef __cplusplus
}
#endif
#endif
