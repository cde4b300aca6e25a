/*
 * eppx_oracle.c -- CPU restatement of EtherParallelParticles.jl's SPH time
 * step.  TEST INFRASTRUCTURE ONLY (see eppx_oracle.h): the checker, never the
 * product.  Built with -O2 -ffp-contract=off -fno-fast-math so every float
 * operation is one IEEE binary32 operation in the order written here, which
 * is the reference's written order (left-to-right products, no FMA).
 *
 * Citations are to the reference package (EtherParallelParticles.jl 0.1.0).
 */
#include "eppx_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* OR_MUTATION plants one deliberate mistake (tests build these variants to
 * show the checks catch them); 0 is the faithful restatement. */
#ifndef OR_MUTATION
#define OR_MUTATION 0
#endif

/* Summation order of the per-neighbour sums (continuity, momentum): 0 the
 * reference's list order, 1 reversed -- a second valid order, for tests that
 * measure how far the order alone moves a float32 sum (the reference's own
 * GPU lists have no fixed order: their slots come from atomic inserts). */
static int32_t g_reversed = 0;
void or_set_pair_order(int32_t reversed) { g_reversed = reversed != 0; }
#define NEIGHBOUR(NI, n) (g_reversed ? (n) - 1 - (NI) : (NI))

#define IPC(L, IP, I, c) (IP)[(size_t)(I) + (size_t)(L)->n_capacity * (size_t)(c)]
#define FPC(L, FP, I, c) (FP)[(size_t)(I) + (size_t)(L)->n_capacity * (size_t)(c)]

/* ---------------------------------------------------------------- Domain */

/* Domain2D{IT,FT}(gap, first_x, first_y, last_x, last_y), Domain2D.jl:28-67 */
void or_domain2d_init(double gap, double first_x, double first_y, double last_x, double last_y, or_domain2d *d) {
    float g = (float)gap;
    d->gap = g;
    d->gap_square = g * g;
    d->first_x = (float)first_x;
    d->first_y = (float)first_y;
    d->last_x = (float)last_x;
    d->last_y = (float)last_y;
    d->span_x = d->last_x - d->first_x;
    d->span_y = d->last_y - d->first_y;
    d->n_x = (int32_t)floorf(d->span_x / g); /* floor(IT, span_x / gap) */
    d->n_y = (int32_t)floorf(d->span_y / g);
    d->n = d->n_x * d->n_y;
    d->gap_x = d->span_x / (float)d->n_x; /* FT / IT promotes IT to FT */
    d->gap_y = d->span_y / (float)d->n_y;
    d->gap_x_inv = 1.0f / d->gap_x;
    d->gap_y_inv = 1.0f / d->gap_y;
}

/* device_floor = unsafe_trunc(IT, floor(x)), AbstractDomain.jl:12-19 */
static inline int32_t device_floor(float x) { return (int32_t)floorf(x); }

/* indexCartesianFromPosition, Domain2D.jl:118-128 (1-based, clamped) */
void or_index_cartesian_from_position(const or_domain2d *d, float x, float y, int32_t *pi, int32_t *pj) {
    int32_t i = device_floor((x - d->first_x) * d->gap_x_inv) + 1;
    if (i > d->n_x) i = d->n_x;
    if (i < 1) i = 1;
    int32_t j = device_floor((y - d->first_y) * d->gap_y_inv) + 1;
    if (j > d->n_y) j = d->n_y;
    if (j < 1) j = 1;
    *pi = i;
    *pj = j;
}

/* indexLinearFromPosition, Domain2D.jl:130-137; indexCartesianToLinear :95-101 */
int32_t or_index_linear_from_position(const or_domain2d *d, float x, float y) {
    int32_t i, j;
    or_index_cartesian_from_position(d, x, y, &i, &j);
    return i + d->n_x * (j - 1);
}

/* inside, Domain2D.jl:110-116 (inclusive bounds; NaN is outside) */
int32_t or_inside(const or_domain2d *d, float x, float y) {
    return d->first_x <= x && x <= d->last_x && d->first_y <= y && y <= d->last_y;
}

/* ------------------------------------------------------- Neighbour system */

/* NeighbourSystemBase(...) with device_initializeNeighbourSystem!,
 * NeighbourSystemBase.jl:28-48 and :76-103; ActivePair.jl:17-33 */
int32_t or_ns_create(const or_domain2d *d, int32_t max_contained, const int32_t *adjacency, int32_t max_tag,
                     or_neighbour_system *ns) {
    memset(ns, 0, sizeof(*ns));
    ns->domain = *d;
    ns->n_cells = d->n;
    ns->max_contained = max_contained;
    size_t nc = (size_t)d->n;
    ns->contained_count = (int32_t *)calloc(nc, sizeof(int32_t));
    ns->contained_list = (int32_t *)calloc(nc * (size_t)max_contained, sizeof(int32_t));
    ns->neighbour_cell_count = (int32_t *)calloc(nc, sizeof(int32_t));
    ns->neighbour_cell_list = (int32_t *)calloc(nc * 9, sizeof(int32_t));
    ns->max_tag = max_tag;
    ns->adjacency = (int32_t *)calloc((size_t)max_tag * (size_t)max_tag, sizeof(int32_t));
    if (!ns->contained_count || !ns->contained_list || !ns->neighbour_cell_count || !ns->neighbour_cell_list ||
        !ns->adjacency) {
        or_ns_destroy(ns);
        return -1;
    }
    memcpy(ns->adjacency, adjacency, sizeof(int32_t) * (size_t)max_tag * (size_t)max_tag);
    for (int32_t I = 1; I <= d->n; ++I) {
        /* indexLinearToCartesian: i = mod1(I, n_x), j = cld(I, n_x) */
        int32_t i = (I - 1) % d->n_x + 1;
        int32_t j = (I + d->n_x - 1) / d->n_x;
        for (int32_t di = -1; di <= 1; ++di) {
            int32_t ii = i + di;
            for (int32_t dj = -1; dj <= 1; ++dj) {
                int32_t jj = j + dj;
                if (ii >= 1 && ii <= d->n_x && jj >= 1 && jj <= d->n_y) {
                    int32_t k = ns->neighbour_cell_count[I - 1]++;
                    ns->neighbour_cell_list[(size_t)(I - 1) + nc * (size_t)k] = ii + d->n_x * (jj - 1);
                }
            }
        }
    }
    return 0;
}

void or_ns_destroy(or_neighbour_system *ns) {
    free(ns->contained_count);
    free(ns->contained_list);
    free(ns->neighbour_cell_count);
    free(ns->neighbour_cell_list);
    free(ns->adjacency);
    memset(ns, 0, sizeof(*ns));
}

/* ---------------------------------------------------------------- Math */

/* Mean.harmonic / invharmonic, src/Math/Mean.jl:28-34 */
float or_harmonic(float a, float b) { return 2.0f * a * b / (a + b); }
float or_invharmonic(float a, float b) { return (a + b) / (2.0f * a * b); }

/* sigma values: CubicSpline.jl:14, WendlandC2.jl:13 (Float64 then FT) */
#define SIGMA_CUBIC_2D ((float)(10.0 / 7.0 / M_PI))
#define SIGMA_WENDLAND_2D ((float)(7.0 / 4.0 / M_PI))

/* _value: CubicSpline.jl:17-31, WendlandC2.jl:16-26 */
float or_kernel_value(int32_t kernel, float r, float hinv) {
    float q = r * hinv;
    if (kernel == OR_KERNEL_CUBIC) {
        if (q < 1.0f) {
            return SIGMA_CUBIC_2D * (hinv * hinv) * (3.0f * q * q * (q - 2.0f) + 4.0f) * 0.25f;
        } else if (q < 2.0f) {
            float to_2 = 2.0f - q;
            return SIGMA_CUBIC_2D * (hinv * hinv) * (to_2 * to_2 * to_2) * 0.25f;
        }
        return 0.0f;
    }
    if (q < 2.0f) {
        float t = 2.0f - q;
        float t2 = t * t;
        return SIGMA_WENDLAND_2D * (t2 * t2) * (hinv * hinv) * (1.0f + 2.0f * q) * 0.0625f;
    }
    return 0.0f;
}

/* _gradient: CubicSpline.jl:41-55, WendlandC2.jl:37-47 */
float or_kernel_gradient(int32_t kernel, float r, float hinv) {
    float q = r * hinv;
    float h3 = hinv * hinv * hinv;
    if (kernel == OR_KERNEL_CUBIC) {
        if (q < 1.0f) {
            return SIGMA_CUBIC_2D * h3 * q * (3.0f * q - 4.0f) * 0.75f;
        } else if (q < 2.0f) {
            float to_2 = 2.0f - q;
            return -SIGMA_CUBIC_2D * h3 * to_2 * to_2 * 0.75f;
        }
        return 0.0f;
    }
    if (q < 2.0f) {
        float t = 2.0f - q;
        if (OR_MUTATION == 5) return -SIGMA_WENDLAND_2D * h3 * 0.625f * q * (t * t);
        return -SIGMA_WENDLAND_2D * h3 * 0.625f * q * (t * t * t);
    }
    return 0.0f;
}

/* value0(h) = _value0(1/h) = sigma * hinv^N, src/SPH/Kernel/Kernel.jl:14-26 */
float or_kernel_value0(int32_t kernel, float h) {
    float hinv = 1.0f / h;
    float s = kernel == OR_KERNEL_CUBIC ? SIGMA_CUBIC_2D : SIGMA_WENDLAND_2D;
    return s * (hinv * hinv);
}

/* EOS: linear c0^2 (rho - rho0) + p0 (main.jl:64); Tait B ((rho/rho0)^7 - 1) + p0
 * with power(x, Val(7)) = (x*x*x)^2 * x (OptimizedFunction.jl:32-34). */
static inline float eos(const or_physics *P, float rho) {
    if (P->eos == OR_EOS_LINEAR) {
        float c02 = OR_MUTATION == 12 ? P->c0 : P->c0 * P->c0;
        return c02 * (rho - P->rho0) + P->p0;
    }
    float x = rho / P->rho0;
    float t = x * x * x;
    return P->tait_b * (t * t * x - 1.0f) + P->p0;
}

/* ---------------------------------------------------------------- Search */

static inline int32_t adjacent(const or_neighbour_system *ns, int32_t ti, int32_t tj) {
    if (ti < 1 || tj < 1 || ti > ns->max_tag || tj > ns->max_tag) return 0; /* out of table: reference UB */
    return ns->adjacency[(size_t)(ti - 1) + (size_t)ns->max_tag * (size_t)(tj - 1)] == 1;
}

/* device_insertParticlesIntoCells!, NeighbourSearch.jl:12-47 (one particle) */
static inline int32_t insert_one(const or_layout *L, int32_t I, int32_t *is_alive, int32_t *cell_index,
                                 const int32_t *IP, const float *FP, or_neighbour_system *ns, int atomic) {
    if (is_alive[I] != 1) return 0;
    int32_t cell = cell_index[I];
    if (!(IPC(L, IP, I, L->IsMovable) == 0 && cell != 0)) {
        float x = FPC(L, FP, I, L->PositionVec);
        float y = FPC(L, FP, I, L->PositionVec + 1);
        if (or_inside(&ns->domain, x, y)) {
            cell = or_index_linear_from_position(&ns->domain, x, y);
            cell_index[I] = cell;
        } else {
            is_alive[I] = 0;
            cell_index[I] = 0;
            return 0;
        }
    }
    int32_t slot;
    if (atomic) {
#pragma omp atomic capture
        slot = ++ns->contained_count[cell - 1];
    } else {
        slot = ++ns->contained_count[cell - 1];
    }
    if (slot > ns->max_contained) return OR_ERR_CELL_OVERFLOW; /* reference writes out of bounds here */
    ns->contained_list[(size_t)(cell - 1) + (size_t)ns->n_cells * (size_t)(slot - 1)] = I + 1;
    return 0;
}

/* sComputeKernelAndiContinuity!, main.jl:74-90, with iValueGradient!
 * (SPH/Library/Kernel.jl:114-128), iClassicContinuity! / iBalancedContinuity!
 * (Continuity.jl:23-47) and vdotx (Library.jl:21-34). */
/* One continuity term of I's neighbour NI (0 for pairs outside the action's
 * classes); stores nHInv / nW / nDW of the pair when store is set. */
static int continuity_term(const or_layout *L, int32_t I, int32_t NI, int32_t *IP, float *FP, const or_physics *P,
                           int store, float *term_out) {
    int32_t J = IPC(L, IP, I, L->nIndex + NI) - 1;
    int32_t ti = IPC(L, IP, I, L->Tag);
    int32_t tj = IPC(L, IP, J, L->Tag);
    int balanced;
    if (ti == P->fluid_tag && tj == P->fluid_tag) balanced = 0;
    else if ((ti == P->fluid_tag && tj == P->wall_tag) || (ti == P->wall_tag && tj == P->fluid_tag)) balanced = 1;
    else return 0;
    float hi = FPC(L, FP, I, L->H), hj = FPC(L, FP, J, L->H);
    float r = FPC(L, FP, I, L->nR + NI);
    float hinv = or_invharmonic(hi, hj);
    float w = or_kernel_value(P->kernel, r, hinv);
    float dw = or_kernel_gradient(P->kernel, r, hinv);
    if (store) {
        if (L->nHInv >= 0) FPC(L, FP, I, L->nHInv + NI) = hinv;
        if (L->nW >= 0) FPC(L, FP, I, L->nW + NI) = w;
        if (L->nDW >= 0) FPC(L, FP, I, L->nDW + NI) = dw;
    }
    float rx = FPC(L, FP, I, L->nRVec + 2 * NI), ry = FPC(L, FP, I, L->nRVec + 2 * NI + 1);
    float vdotx = 0.0f;
    vdotx += rx * (FPC(L, FP, I, L->VelocityVec) - FPC(L, FP, J, L->VelocityVec));
    vdotx += ry * (FPC(L, FP, I, L->VelocityVec + 1) - FPC(L, FP, J, L->VelocityVec + 1));
    float term;
    if (!balanced) term = FPC(L, FP, J, L->Mass) * vdotx * dw / r;
    else term = FPC(L, FP, I, L->Density) * FPC(L, FP, J, L->Volume) * vdotx * dw / r;
    if (OR_MUTATION == 1) term = -term;
    *term_out = term;
    return 1;
}

/* sComputeKernelAndiContinuity!, main.jl:74-90, with iValueGradient!
 * (SPH/Library/Kernel.jl:114-128), iClassicContinuity! / iBalancedContinuity!
 * (Continuity.jl:23-47) and vdotx (Library.jl:21-34). */
static void action_kernel_continuity(const or_layout *L, int32_t I, int32_t *IP, float *FP, const or_physics *P) {
    int32_t n = IPC(L, IP, I, L->nCount);
    for (int32_t k = 0; k < n; ++k) {
        float term;
        if (continuity_term(L, I, NEIGHBOUR(k, n), IP, FP, P, 1, &term)) FPC(L, FP, I, L->dDensity) += term;
    }
}

/* device_findNeighbourParticlesFromCells!, NeighbourSearch.jl:166-227, with
 * defaultSearchCriterion :141-151 and symmetrySearchCriterion :153-164. */
static int32_t find_one(const or_layout *L, int32_t I, const int32_t *is_alive, const int32_t *cell_index,
                        int32_t *IP, float *FP, const or_neighbour_system *ns, int32_t criterion, int32_t action,
                        const or_physics *P) {
    if (is_alive[I] != 1) return 0;
    int32_t err = 0;
    const int32_t M = L->max_neighbour_number;
    const size_t nc = (size_t)ns->n_cells;
    int32_t count = 0;
    int32_t cell = cell_index[I];
    int32_t ti = IPC(L, IP, I, L->Tag);
    float xi = FPC(L, FP, I, L->PositionVec), yi = FPC(L, FP, I, L->PositionVec + 1);
    int32_t ncells = ns->neighbour_cell_count[cell - 1];
    for (int32_t k = 0; k < ncells; ++k) {
        int32_t nb = ns->neighbour_cell_list[(size_t)(cell - 1) + nc * (size_t)k];
        int32_t cnt = ns->contained_count[nb - 1];
        if (cnt > ns->max_contained) cnt = ns->max_contained;
        for (int32_t s = 0; s < cnt; ++s) {
            int32_t J = ns->contained_list[(size_t)(nb - 1) + nc * (size_t)s] - 1;
            if (I == J || !adjacent(ns, ti, IPC(L, IP, J, L->Tag))) continue;
            float dx = xi - FPC(L, FP, J, L->PositionVec);
            float dy = yi - FPC(L, FP, J, L->PositionVec + 1);
            float dr2 = dx * dx + dy * dy;
            int ok;
            if (criterion == OR_CRIT_DEFAULT) {
                ok = dr2 <= ns->domain.gap_square;
            } else {
                float h2 = or_harmonic(FPC(L, FP, I, L->H), FPC(L, FP, J, L->H)) * 2.0f;
                ok = OR_MUTATION == 3 ? dr2 < h2 * h2 : dr2 <= h2 * h2;
            }
            if (!ok) continue;
            if (count >= M) { /* reference writes past the declared columns */
                err |= OR_ERR_NEIGHBOUR_OVERFLOW;
                continue;
            }
            count += 1;
            IPC(L, IP, I, L->nIndex + count - 1) = J + 1;
            if (L->nRVec >= 0) {
                FPC(L, FP, I, L->nRVec + 2 * (count - 1)) = dx;
                FPC(L, FP, I, L->nRVec + 2 * count - 1) = dy;
            }
            if (L->nR >= 0) FPC(L, FP, I, L->nR + count - 1) = sqrtf(dr2);
        }
    }
    IPC(L, IP, I, L->nCount) = count;
    if (action == 1) action_kernel_continuity(L, I, IP, FP, P);
    return err;
}

int32_t or_search(const or_layout *L, int32_t n_particles, int32_t *is_alive, int32_t *cell_index, int32_t *IP,
                  float *FP, or_neighbour_system *ns, int32_t criterion, int32_t action, const or_physics *phys,
                  int32_t serial_insert, int32_t n_threads) {
    int32_t err = 0;
#ifdef _OPENMP
    if (n_threads > 0) omp_set_num_threads(n_threads);
#else
    (void)n_threads;
#endif
    /* Class.clean!(ns): fill!(contained_count, 0), NeighbourSystem.jl:43-54 */
    memset(ns->contained_count, 0, sizeof(int32_t) * (size_t)ns->n_cells);
    if (serial_insert) {
        for (int32_t I = 0; I < n_particles; ++I) err |= insert_one(L, I, is_alive, cell_index, IP, FP, ns, 0);
    } else {
#pragma omp parallel for schedule(static) reduction(| : err)
        for (int32_t I = 0; I < n_particles; ++I) err |= insert_one(L, I, is_alive, cell_index, IP, FP, ns, 1);
    }
#pragma omp parallel for schedule(dynamic, 256) reduction(| : err)
    for (int32_t I = 0; I < n_particles; ++I)
        err |= find_one(L, I, is_alive, cell_index, IP, FP, ns, criterion, action, phys);
    return err;
}

/* ---------------------------------------------------------- Self actions */

static void self_one(const or_layout *L, int32_t I, int32_t *IP, float *FP, int32_t action, const or_physics *P,
                     int32_t column, float value) {
    int32_t tag = IPC(L, IP, I, L->Tag);
    if (action == 1) {
        /* sContinuity! main.jl:92-105 -> Library.sContinuity! Continuity.jl:10-21,
         * sVolume! State.jl:10-19, p = eos(rho) */
        if (tag != P->fluid_tag && tag != P->wall_tag) return;
        FPC(L, FP, I, L->Density) += FPC(L, FP, I, L->dDensity) * P->dt;
        FPC(L, FP, I, L->dDensity) = 0.0f;
        FPC(L, FP, I, L->Volume) = FPC(L, FP, I, L->Mass) / FPC(L, FP, I, L->Density);
        FPC(L, FP, I, L->Pressure) = eos(P, FPC(L, FP, I, L->Density));
    } else if (action == 2) {
        /* sMomentum! main.jl:120-126 -> sGravity! State.jl:35-47,
         * sAccelerateMove! Motion.jl:61-75 */
        if (tag != P->fluid_tag) return;
        FPC(L, FP, I, L->dVelocityVec) += P->gx;
        FPC(L, FP, I, L->dVelocityVec + 1) += P->gy;
        for (int32_t k = 0; k < 2; ++k) {
            float du = FPC(L, FP, I, L->dVelocityVec + k);
            float half = OR_MUTATION == 7 ? 1.0f : 0.5f;
            FPC(L, FP, I, L->PositionVec + k) += (FPC(L, FP, I, L->VelocityVec + k) + du * P->dt * half) * P->dt;
            FPC(L, FP, I, L->VelocityVec + k) += FPC(L, FP, I, L->dVelocityVec + k) * P->dt;
            FPC(L, FP, I, L->dVelocityVec + k) = 0.0f;
        }
    } else if (action == 3) {
        /* sFilter main.jl:136-144 -> sKernelFilter! Filter.jl:24-38, sVolume!, eos */
        if (tag != P->fluid_tag) return;
        float h = FPC(L, FP, I, L->H);
        float w0 = OR_MUTATION == 9 ? 0.0f : or_kernel_value0(P->kernel, h);
        FPC(L, FP, I, L->SumWeight) += w0 * FPC(L, FP, I, L->Volume);
        FPC(L, FP, I, L->SumWeightedDensity) += w0 * FPC(L, FP, I, L->Mass);
        FPC(L, FP, I, L->Density) = FPC(L, FP, I, L->SumWeightedDensity) / FPC(L, FP, I, L->SumWeight);
        FPC(L, FP, I, L->SumWeight) = 0.0f;
        FPC(L, FP, I, L->SumWeightedDensity) = 0.0f;
        FPC(L, FP, I, L->Volume) = FPC(L, FP, I, L->Mass) / FPC(L, FP, I, L->Density);
        FPC(L, FP, I, L->Pressure) = eos(P, FPC(L, FP, I, L->Density));
    } else if (action == 4) {
        /* sExtrapolatePressure! Pressure.jl:125-143 (on wall particles) */
        if (tag != P->wall_tag) return;
        float wv = FPC(L, FP, I, L->SumWeight);
        if (wv > 0.0f) {
            FPC(L, FP, I, L->Pressure) = FPC(L, FP, I, L->SumWeightedPressure) / wv;
            if (OR_MUTATION != 11) {
                FPC(L, FP, I, L->SumWeight) = 0.0f;
                FPC(L, FP, I, L->SumWeightedPressure) = 0.0f;
            }
        } else {
            FPC(L, FP, I, L->Pressure) = P->p0;
        }
    } else if (action == 10) {
        FPC(L, FP, I, column) *= value;
    }
}

void or_selfaction(const or_layout *L, int32_t n_particles, const int32_t *is_alive, int32_t *IP, float *FP,
                   int32_t action, const or_physics *phys, int32_t column, float value, int32_t n_threads) {
#ifdef _OPENMP
    if (n_threads > 0) omp_set_num_threads(n_threads);
#else
    (void)n_threads;
#endif
    /* device_selfaction!, Action.jl:39-51 */
#pragma omp parallel for schedule(static)
    for (int32_t I = 0; I < n_particles; ++I)
        if (is_alive[I] == 1) self_one(L, I, IP, FP, action, phys, column, value);
}

/* ---------------------------------------------------------- Interactions */

/* iMomentum! main.jl:107-118 with iClassicPressure! / iBalancedPressure!
 * (Pressure.jl:24-60) and iClassicViscosity! / sArtificialViscosity!
 * (Viscosity.jl:10-52).  The contributions of neighbour NI of particle I, in
 * the order the reference adds them to du: t[0..1] the pressure term, then
 * (when the viscosity adds one: returns 2, else 1) t[2..3]; 0 for pairs
 * outside the action's classes. */
static int momentum_terms(const or_layout *L, int32_t I, int32_t NI, const float *FP, const int32_t *IP,
                          const or_physics *P, float t[4]) {
    int32_t J = IPC(L, IP, I, L->nIndex + NI) - 1;
    int32_t ti = IPC(L, IP, I, L->Tag), tj = IPC(L, IP, J, L->Tag);
    int classic;
    if (ti == P->fluid_tag && tj == P->fluid_tag) classic = 1;
    else if (ti == P->fluid_tag && tj == P->wall_tag) classic = 0;
    else return 0;
    float r = FPC(L, FP, I, L->nR + NI);
    float dw = FPC(L, FP, I, L->nDW + NI);
    float hinv = FPC(L, FP, I, L->nHInv + NI);
    float rx = FPC(L, FP, I, L->nRVec + 2 * NI), ry = FPC(L, FP, I, L->nRVec + 2 * NI + 1);
    float rho_i = FPC(L, FP, I, L->Density), rho_j = FPC(L, FP, J, L->Density);
    float p_i = FPC(L, FP, I, L->Pressure), p_j = FPC(L, FP, J, L->Pressure);
    float m_j = FPC(L, FP, J, L->Mass);
    /* pressure */
    float p_rho_2;
    if (classic && OR_MUTATION == 2) p_rho_2 = p_i / (rho_j * rho_j) + p_j / (rho_j * rho_j);
    else if (classic) p_rho_2 = p_i / (rho_i * rho_i) + p_j / (rho_j * rho_j);
    else p_rho_2 = (p_i + p_j) / (rho_i * rho_j);
    p_rho_2 += fabsf(p_rho_2) * P->coefficient;
    p_rho_2 *= -m_j * dw / r;
    t[0] = p_rho_2 * rx;
    t[1] = p_rho_2 * ry;
    /* viscosity */
    float dux = FPC(L, FP, I, L->VelocityVec) - FPC(L, FP, J, L->VelocityVec);
    float duy = FPC(L, FP, I, L->VelocityVec + 1) - FPC(L, FP, J, L->VelocityVec + 1);
    if (P->visc == OR_VISC_CLASSIC) {
        float h = 1.0f / hinv;
        float az = r * r + 0.01f * h * h; /* avoidzero, Library.jl:17-19 */
        float m_v = OR_MUTATION == 14 ? FPC(L, FP, I, L->Mass) : m_j;
        float force = (OR_MUTATION == 13 ? 1.0f : 2.0f) * P->mu * m_v * r * dw / (rho_i * rho_j * az);
        t[2] = force * dux;
        t[3] = force * duy;
        return 2;
    }
    float vdotx = 0.0f;
    vdotx += rx * dux;
    vdotx += ry * duy;
    if (OR_MUTATION == 6) vdotx = -vdotx;
    if (vdotx > 0.0f || OR_MUTATION == 4) return 1;
    float rho = (rho_i + rho_j) * 0.5f; /* arithmetic mean, Mean.jl:12-14 (exact scaling) */
    float h = 1.0f / hinv;
    float az = r * r + 0.01f * h * h;
    float phi = vdotx / (hinv * az);
    float force = (-P->alpha * P->av_c + P->beta * phi) * phi / rho;
    force *= -m_j / (rho_i * rho_j * r) * dw;
    t[2] = force * rx;
    t[3] = force * ry;
    return 2;
}

static void momentum_pair(const or_layout *L, int32_t I, int32_t NI, float *FP, const int32_t *IP,
                          const or_physics *P) {
    float t[4];
    int k = momentum_terms(L, I, NI, FP, IP, P, t);
    if (k >= 1) {
        FPC(L, FP, I, L->dVelocityVec) += t[0];
        FPC(L, FP, I, L->dVelocityVec + 1) += t[1];
    }
    if (k == 2) {
        FPC(L, FP, I, L->dVelocityVec) += t[2];
        FPC(L, FP, I, L->dVelocityVec + 1) += t[3];
    }
}

/* iFilter main.jl:128-134 -> iKernelFilter! Filter.jl:10-22 */
static void filter_pair(const or_layout *L, int32_t I, int32_t NI, float *FP, const int32_t *IP,
                        const or_physics *P) {
    int32_t J = IPC(L, IP, I, L->nIndex + NI) - 1;
    if (IPC(L, IP, I, L->Tag) != P->fluid_tag) return;
    if (IPC(L, IP, J, L->Tag) != P->fluid_tag && OR_MUTATION != 8) return;
    float w = FPC(L, FP, I, L->nW + NI);
    FPC(L, FP, I, L->SumWeight) += w * FPC(L, FP, J, L->Volume);
    FPC(L, FP, I, L->SumWeightedDensity) += w * FPC(L, FP, J, L->Mass);
}

/* Julia's max(x, y) of two floats (Base.max for IEEE floats): NaN when
 * either is NaN, and +0 above -0. */
static float jl_max(float x, float y) {
    int take_y = (y > x) || (signbit(y) < signbit(x));
    return take_y ? (isnan(x) ? x : y) : (isnan(y) ? y : x);
}

/* iExtrapolatePressure! (2D) Pressure.jl:81-101 for wall I and fluid J, with
 * w = the pair's kernel value stored by the search (nW), p0 and g from P. */
static void extrapolate_pair(const or_layout *L, int32_t I, int32_t NI, float *FP, const int32_t *IP,
                             const or_physics *P) {
    int32_t J = IPC(L, IP, I, L->nIndex + NI) - 1;
    if (IPC(L, IP, I, L->Tag) != P->wall_tag || IPC(L, IP, J, L->Tag) != P->fluid_tag) return;
    float w = FPC(L, FP, I, L->nW + NI);
    float rx = FPC(L, FP, I, L->nRVec + 2 * NI), ry = FPC(L, FP, I, L->nRVec + 2 * NI + 1);
    float g_dot_x = 0.0f;
    g_dot_x += rx * P->gx;
    g_dot_x += ry * P->gy;
    if (OR_MUTATION == 10) g_dot_x = -g_dot_x;
    float wv = w * FPC(L, FP, J, L->Volume);
    FPC(L, FP, I, L->SumWeight) += wv;
    FPC(L, FP, I, L->SumWeightedPressure) +=
        wv * (jl_max(FPC(L, FP, J, L->Pressure), P->p0) + jl_max(FPC(L, FP, J, L->Density) * g_dot_x, P->p0));
}

void or_interaction(const or_layout *L, int32_t n_particles, const int32_t *is_alive, int32_t *IP, float *FP,
                    int32_t action, const or_physics *phys, int32_t column, float value, int32_t n_threads) {
#ifdef _OPENMP
    if (n_threads > 0) omp_set_num_threads(n_threads);
#else
    (void)n_threads;
#endif
    /* device_interaction!, Action.jl:121-138 */
#pragma omp parallel for schedule(dynamic, 256)
    for (int32_t I = 0; I < n_particles; ++I) {
        if (is_alive[I] != 1) continue;
        int32_t n = IPC(L, IP, I, L->nCount);
        for (int32_t k = 0; k < n; ++k) {
            const int32_t NI = action == 1 ? NEIGHBOUR(k, n) : k;
            if (action == 1) momentum_pair(L, I, NI, FP, IP, phys);
            else if (action == 2) filter_pair(L, I, NI, FP, IP, phys);
            else if (action == 3) extrapolate_pair(L, I, NI, FP, IP, phys);
            else if (action == 10) FPC(L, FP, I, column) += value;
        }
    }
}

/* --------------------------------------------------------------- Driver */

int32_t or_run_steps(const or_layout *L, int32_t n_particles, int32_t *is_alive, int32_t *cell_index, int32_t *IP,
                     float *FP, or_neighbour_system *ns, const or_physics *phys, int32_t n_steps,
                     int32_t filter_interval, int32_t serial_insert, int32_t n_threads, int64_t *pairs_per_step) {
    int32_t err = 0;
    for (int32_t step = 1; step <= n_steps; ++step) {
        or_selfaction(L, n_particles, is_alive, IP, FP, 1, phys, 0, 0.0f, n_threads);
        or_interaction(L, n_particles, is_alive, IP, FP, 1, phys, 0, 0.0f, n_threads);
        or_selfaction(L, n_particles, is_alive, IP, FP, 2, phys, 0, 0.0f, n_threads);
        err |= or_search(L, n_particles, is_alive, cell_index, IP, FP, ns, OR_CRIT_SYMMETRY, 1, phys, serial_insert,
                         n_threads);
        if (filter_interval > 0 && step % filter_interval == 0) {
            or_interaction(L, n_particles, is_alive, IP, FP, 2, phys, 0, 0.0f, n_threads);
            or_selfaction(L, n_particles, is_alive, IP, FP, 3, phys, 0, 0.0f, n_threads);
        }
        if (pairs_per_step) {
            int64_t total = 0;
            for (int32_t I = 0; I < n_particles; ++I)
                if (is_alive[I] == 1) total += IPC(L, IP, I, L->nCount);
            pairs_per_step[step - 1] = total;
        }
    }
    return err;
}

/* Magnitude bounds of the per-neighbour sums: for every term a sum adds, the
 * same expression with each inner sum or difference of products replaced by
 * the sum of the magnitudes (vdotx -> |rx dux| + |ry duy|, p_i/rho_i^2 +
 * p_j/rho_j^2 -> |p_i|/rho_i^2 + |p_j|/rho_j^2), summed over I's neighbours.
 * A float32 evaluation of the same sum in another order, or with another
 * rounding of its terms, differs from this one by a small multiple of eps
 * times this bound.  what = 0: the continuity terms of dDensity (out[I],
 * read after a search with the kernel + continuity action, from the same
 * state); what = 1: the momentum terms of dVelocity (out[2 I + k], read
 * before iMomentum!, from its state). */
/* |f| + r |df/dr| of f = dw / r (dw = _gradient, CubicSpline.jl:41-55,
 * WendlandC2.jl:37-47) in double: dw = C h^3 G(q) with q = r hinv gives
 * r f' = C h^3 (q G'(q) - G(q)) / r, so the bound is |C| h^3 (2 |G| + |q G'|) / r.
 * It covers a relative perturbation of r (a float32 sqrt against a hardware
 * reciprocal square root) near the edge of the support, where G -> 0 faster
 * than its derivative. */
static double grad_over_r_bound(int32_t kernel, double r, double hinv) {
    const double q = r * hinv, h3 = hinv * hinv * hinv;
    double c, g, gp;
    if (kernel == OR_KERNEL_CUBIC) {
        c = 0.75 * (10.0 / 7.0 / M_PI);
        if (q < 1.0) {
            g = q * (3.0 * q - 4.0);
            gp = 6.0 * q - 4.0;
        } else if (q < 2.0) {
            g = -(2.0 - q) * (2.0 - q);
            gp = 2.0 * (2.0 - q);
        } else {
            return 0.0;
        }
    } else {
        if (!(q < 2.0)) return 0.0;
        c = 0.625 * (7.0 / 4.0 / M_PI);
        g = q * (2.0 - q) * (2.0 - q) * (2.0 - q);
        gp = (2.0 - q) * (2.0 - q) * (2.0 - 4.0 * q);
    }
    return c * h3 * (2.0 * fabs(g) + fabs(q * gp)) / r;
}

static double continuity_bound(const or_layout *L, int32_t I, int32_t NI, const int32_t *IP, const float *FP,
                               const or_physics *P) {
    int32_t J = IPC(L, IP, I, L->nIndex + NI) - 1;
    int32_t ti = IPC(L, IP, I, L->Tag), tj = IPC(L, IP, J, L->Tag);
    int balanced;
    if (ti == P->fluid_tag && tj == P->fluid_tag) balanced = 0;
    else if ((ti == P->fluid_tag && tj == P->wall_tag) || (ti == P->wall_tag && tj == P->fluid_tag)) balanced = 1;
    else return 0.0;
    double r = FPC(L, FP, I, L->nR + NI);
    double hinv = or_invharmonic(FPC(L, FP, I, L->H), FPC(L, FP, J, L->H));
    double dw = or_kernel_gradient(P->kernel, (float)r, (float)hinv);
    double rx = FPC(L, FP, I, L->nRVec + 2 * NI), ry = FPC(L, FP, I, L->nRVec + 2 * NI + 1);
    double ax = fabs(rx * ((double)FPC(L, FP, I, L->VelocityVec) - FPC(L, FP, J, L->VelocityVec)));
    double ay = fabs(ry * ((double)FPC(L, FP, I, L->VelocityVec + 1) - FPC(L, FP, J, L->VelocityVec + 1)));
    double m = balanced ? (double)FPC(L, FP, I, L->Density) * FPC(L, FP, J, L->Volume) : FPC(L, FP, J, L->Mass);
    (void)dw;
    return fabs(m) * (ax + ay) * grad_over_r_bound(P->kernel, r, hinv);
}

/* dp / drho of the EOS at rho (main.jl:64 linear; Tait) and the float32 spacing at x */
static double eos_slope(const or_physics *P, double rho) {
    if (P->eos == OR_EOS_LINEAR) return (double)P->c0 * P->c0;
    const double x = rho / P->rho0;
    return 7.0 * P->tait_b * x * x * x * x * x * x / P->rho0;
}
static double ulp32(double x) {
    const float f = (float)fabs(x);
    return (double)nextafterf(f, INFINITY) - (double)f;
}

static void momentum_bound(const or_layout *L, int32_t I, int32_t NI, const int32_t *IP, const float *FP,
                           const or_physics *P, double b[2]) {
    b[0] = b[1] = 0.0;
    int32_t J = IPC(L, IP, I, L->nIndex + NI) - 1;
    int32_t ti = IPC(L, IP, I, L->Tag), tj = IPC(L, IP, J, L->Tag);
    int classic;
    if (ti == P->fluid_tag && tj == P->fluid_tag) classic = 1;
    else if (ti == P->fluid_tag && tj == P->wall_tag) classic = 0;
    else return;
    double r = FPC(L, FP, I, L->nR + NI);
    double hinv = FPC(L, FP, I, L->nHInv + NI);
    double dwr = grad_over_r_bound(P->kernel, r, hinv); /* |dw / r| with its sensitivity to r */
    double rx = FPC(L, FP, I, L->nRVec + 2 * NI), ry = FPC(L, FP, I, L->nRVec + 2 * NI + 1);
    double rho_i = FPC(L, FP, I, L->Density), rho_j = FPC(L, FP, J, L->Density);
    /* |p| plus the pressure change of one float32 spacing of density (dp/drho of
     * the EOS): the bound also covers inputs whose densities differ by an ulp */
    double p_i = fabs((double)FPC(L, FP, I, L->Pressure)) + eos_slope(P, rho_i) * ulp32(rho_i);
    double p_j = fabs((double)FPC(L, FP, J, L->Pressure)) + eos_slope(P, rho_j) * ulp32(rho_j);
    double m_j = FPC(L, FP, J, L->Mass);
    double pr = classic ? p_i / (rho_i * rho_i) + p_j / (rho_j * rho_j) : (p_i + p_j) / (rho_i * rho_j);
    pr *= 1.0 + fabs((double)P->coefficient);
    double radial = pr * fabs(m_j) * dwr;
    b[0] = radial * fabs(rx);
    b[1] = radial * fabs(ry);
    double dux = (double)FPC(L, FP, I, L->VelocityVec) - FPC(L, FP, J, L->VelocityVec);
    double duy = (double)FPC(L, FP, I, L->VelocityVec + 1) - FPC(L, FP, J, L->VelocityVec + 1);
    double h = 1.0 / hinv;
    double az = r * r + 0.01 * h * h;
    if (P->visc == OR_VISC_CLASSIC) {
        double force = fabs(2.0 * P->mu * m_j * r * r / (rho_i * rho_j * az)) * dwr;
        b[0] += force * fabs(dux);
        b[1] += force * fabs(duy);
        return;
    }
    double vdotx = rx * dux + ry * duy, vabs = fabs(rx * dux) + fabs(ry * duy);
    if (vdotx > 1e-6 * vabs) return; /* clearly separating: neither evaluation adds a term */
    double phi = vabs / (hinv * az);
    double force = (fabs((double)P->alpha * P->av_c) + fabs((double)P->beta) * phi) * phi / ((rho_i + rho_j) * 0.5);
    force *= fabs(m_j / (rho_i * rho_j)) * dwr;
    b[0] += force * fabs(rx);
    b[1] += force * fabs(ry);
}

void or_pair_term_magnitude(const or_layout *L, int32_t n_particles, const int32_t *is_alive, int32_t *IP,
                            float *FP, const or_physics *P, int32_t what, double *out) {
#pragma omp parallel for schedule(dynamic, 256)
    for (int32_t I = 0; I < n_particles; ++I) {
        double a[2] = {0.0, 0.0};
        if (is_alive[I] == 1) {
            int32_t n = IPC(L, IP, I, L->nCount);
            for (int32_t NI = 0; NI < n; ++NI) {
                if (what == 0) {
                    a[0] += continuity_bound(L, I, NI, IP, FP, P);
                } else {
                    double b[2];
                    momentum_bound(L, I, NI, IP, FP, P, b);
                    a[0] += b[0];
                    a[1] += b[1];
                }
            }
        }
        if (what == 0) out[I] = a[0];
        else {
            out[2 * (size_t)I] = a[0];
            out[2 * (size_t)I + 1] = a[1];
        }
    }
}
