/*
 * eppx_oracle3d.c -- CPU restatement of the SPH time step in 3D.
 * TEST INFRASTRUCTURE ONLY (see eppx_oracle.h): the checker, never the
 * product.  Built with the flags of eppx_oracle.c (-O2 -ffp-contract=off
 * -fno-fast-math): every float operation is one IEEE binary32 operation in
 * the order written here.
 *
 * The reference leaves its 3D layers as stubs (Domain: AbstractDomain.jl:119;
 * stencil: NeighbourSystemBase.jl:50-57; insert / find: NeighbourSearch.jl:49,
 * :229; Cuboid lattice: Geometry3D.jl:89-94), so this file is the reference's
 * 2D algorithm with a z axis: every 2D step below cites the 2D function it
 * extends, and the SPH library functions (written for any dimension N in
 * the reference) are restated with N = 3.  No reference output pins it; the
 * tests pin it by closed forms, invariants and brute force instead.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "eppx_oracle.h"

#ifndef OR_MUTATION
#define OR_MUTATION 0
#endif

#define IPC(L, IP, I, c) (IP)[(size_t)(I) + (size_t)(L)->n_capacity * (size_t)(c)]
#define FPC(L, FP, I, c) (FP)[(size_t)(I) + (size_t)(L)->n_capacity * (size_t)(c)]

/* ---------------------------------------------------------------- Domain */

/* Domain2D{IT,FT}(gap, ...) (Domain2D.jl:28-67) per axis. */
void or3_domain_init(double gap, double first_x, double first_y, double first_z, double last_x, double last_y,
                     double last_z, or_domain3d *d) {
    float g = (float)gap;
    d->gap = g;
    d->gap_square = g * g;
    d->first_x = (float)first_x;
    d->first_y = (float)first_y;
    d->first_z = (float)first_z;
    d->last_x = (float)last_x;
    d->last_y = (float)last_y;
    d->last_z = (float)last_z;
    d->span_x = d->last_x - d->first_x;
    d->span_y = d->last_y - d->first_y;
    d->span_z = d->last_z - d->first_z;
    d->n_x = (int32_t)floorf(d->span_x / g);
    d->n_y = (int32_t)floorf(d->span_y / g);
    d->n_z = (int32_t)floorf(d->span_z / g);
    d->n = d->n_x * d->n_y * d->n_z;
    d->gap_x = d->span_x / (float)d->n_x;
    d->gap_y = d->span_y / (float)d->n_y;
    d->gap_z = d->span_z / (float)d->n_z;
    d->gap_x_inv = 1.0f / d->gap_x;
    d->gap_y_inv = 1.0f / d->gap_y;
    d->gap_z_inv = 1.0f / d->gap_z;
}

static inline int32_t clamp1(int32_t v, int32_t n) { return v > n ? n : (v < 1 ? 1 : v); }

/* indexCartesianFromPosition (Domain2D.jl:118-128) per axis: 1-based, clamped. */
void or3_index_cartesian_from_position(const or_domain3d *d, float x, float y, float z, int32_t *pi, int32_t *pj,
                                       int32_t *pk) {
    *pi = clamp1((int32_t)floorf((x - d->first_x) * d->gap_x_inv) + 1, d->n_x);
    *pj = clamp1((int32_t)floorf((y - d->first_y) * d->gap_y_inv) + 1, d->n_y);
    *pk = clamp1((int32_t)floorf((z - d->first_z) * d->gap_z_inv) + 1, d->n_z);
}

/* indexLinearFromPosition (Domain2D.jl:130-137) with a z axis: i + n_x (j - 1) + n_x n_y (k - 1). */
int32_t or3_index_linear_from_position(const or_domain3d *d, float x, float y, float z) {
    int32_t i, j, k;
    or3_index_cartesian_from_position(d, x, y, z, &i, &j, &k);
    return i + d->n_x * ((j - 1) + d->n_y * (k - 1));
}

/* inside (Domain2D.jl:110-116) per axis: inclusive bounds, NaN outside. */
int32_t or3_inside(const or_domain3d *d, float x, float y, float z) {
    return d->first_x <= x && x <= d->last_x && d->first_y <= y && y <= d->last_y && d->first_z <= z &&
           z <= d->last_z;
}

/* ------------------------------------------------------- Neighbour system */

/* device_initializeNeighbourSystem! (NeighbourSystemBase.jl:28-48) with a z
 * offset: x offset outer, then y, then z. */
int32_t or3_ns_create(const or_domain3d *d, int32_t max_contained, const int32_t *adjacency, int32_t max_tag,
                      or_neighbour_system3d *ns) {
    memset(ns, 0, sizeof(*ns));
    ns->domain = *d;
    ns->n_cells = d->n;
    ns->max_contained = max_contained;
    size_t nc = (size_t)d->n;
    ns->contained_count = (int32_t *)calloc(nc, sizeof(int32_t));
    ns->contained_list = (int32_t *)calloc(nc * (size_t)max_contained, sizeof(int32_t));
    ns->neighbour_cell_count = (int32_t *)calloc(nc, sizeof(int32_t));
    ns->neighbour_cell_list = (int32_t *)calloc(nc * 27, sizeof(int32_t));
    ns->max_tag = max_tag;
    ns->adjacency = (int32_t *)calloc((size_t)max_tag * (size_t)max_tag, sizeof(int32_t));
    if (!ns->contained_count || !ns->contained_list || !ns->neighbour_cell_count || !ns->neighbour_cell_list ||
        !ns->adjacency) {
        or3_ns_destroy(ns);
        return -1;
    }
    memcpy(ns->adjacency, adjacency, sizeof(int32_t) * (size_t)max_tag * (size_t)max_tag);
    const int32_t nxy = d->n_x * d->n_y;
    for (int32_t I = 1; I <= d->n; ++I) {
        int32_t i = (I - 1) % d->n_x + 1;
        int32_t j = ((I - 1) / d->n_x) % d->n_y + 1;
        int32_t k = (I - 1) / nxy + 1;
        for (int32_t di = -1; di <= 1; ++di)
            for (int32_t dj = -1; dj <= 1; ++dj)
                for (int32_t dk = OR_MUTATION == 18 ? 0 : -1; dk <= 1; ++dk) {
                    int32_t ii = i + di, jj = j + dj, kk = k + dk;
                    if (ii >= 1 && ii <= d->n_x && jj >= 1 && jj <= d->n_y && kk >= 1 && kk <= d->n_z) {
                        int32_t m = ns->neighbour_cell_count[I - 1]++;
                        ns->neighbour_cell_list[(size_t)(I - 1) + nc * (size_t)m] =
                            ii + d->n_x * ((jj - 1) + d->n_y * (kk - 1));
                    }
                }
    }
    return 0;
}

void or3_ns_destroy(or_neighbour_system3d *ns) {
    free(ns->contained_count);
    free(ns->contained_list);
    free(ns->neighbour_cell_count);
    free(ns->neighbour_cell_list);
    free(ns->adjacency);
    memset(ns, 0, sizeof(*ns));
}

/* ---------------------------------------------------------------- Kernels */

/* sigma for N = 3: CubicSpline.jl:15, WendlandC2.jl:15 (Float64, then FT) */
#define SIGMA_CUBIC_3D ((float)(1.0 / M_PI))
#define SIGMA_WENDLAND_3D ((float)(21.0 / 16.0 / M_PI))
#define SIGMA_WENDLAND_2D_AS_3D ((float)(7.0 / 4.0 / M_PI))

/* _value with N = 3 (CubicSpline.jl:17-31, WendlandC2.jl:16-26):
 * power(hinv, Val(3)) = hinv * hinv * hinv (OptimizedFunction.jl:10-31). */
float or3_kernel_value(int32_t kernel, float r, float hinv) {
    float q = r * hinv;
    float h3 = hinv * hinv * hinv;
    if (kernel == OR_KERNEL_CUBIC) {
        if (q < 1.0f) return SIGMA_CUBIC_3D * h3 * (3.0f * q * q * (q - 2.0f) + 4.0f) * 0.25f;
        if (q < 2.0f) {
            float to_2 = 2.0f - q;
            return SIGMA_CUBIC_3D * h3 * (to_2 * to_2 * to_2) * 0.25f;
        }
        return 0.0f;
    }
    if (q < 2.0f) {
        float t = 2.0f - q;
        float t2 = t * t;
        float s = OR_MUTATION == 16 ? SIGMA_WENDLAND_2D_AS_3D : SIGMA_WENDLAND_3D;
        return s * (t2 * t2) * h3 * (1.0f + 2.0f * q) * 0.0625f;
    }
    return 0.0f;
}

/* _gradient with N = 3 (CubicSpline.jl:41-55, WendlandC2.jl:37-47):
 * power(hinv, Val(4)) = (hinv * hinv) * (hinv * hinv). */
float or3_kernel_gradient(int32_t kernel, float r, float hinv) {
    float q = r * hinv;
    float hh = hinv * hinv;
    float h4 = OR_MUTATION == 17 ? hinv * hinv * hinv : hh * hh;
    if (kernel == OR_KERNEL_CUBIC) {
        if (q < 1.0f) return SIGMA_CUBIC_3D * h4 * q * (3.0f * q - 4.0f) * 0.75f;
        if (q < 2.0f) {
            float to_2 = 2.0f - q;
            return -SIGMA_CUBIC_3D * h4 * to_2 * to_2 * 0.75f;
        }
        return 0.0f;
    }
    if (q < 2.0f) {
        float t = 2.0f - q;
        return -SIGMA_WENDLAND_3D * h4 * 0.625f * q * (t * t * t);
    }
    return 0.0f;
}

/* value0(h) = sigma * hinv^3 (Kernel.jl:14-26) */
float or3_kernel_value0(int32_t kernel, float h) {
    float hinv = 1.0f / h;
    float s = kernel == OR_KERNEL_CUBIC ? SIGMA_CUBIC_3D : SIGMA_WENDLAND_3D;
    return s * (hinv * hinv * hinv);
}

static inline float eos3(const or_physics *P, float rho) {
    if (P->eos == OR_EOS_LINEAR) return P->c0 * P->c0 * (rho - P->rho0) + P->p0;
    float x = rho / P->rho0;
    float t = x * x * x;
    return P->tait_b * (t * t * x - 1.0f) + P->p0;
}

/* ----------------------------------------------------------------- Search */

static inline int32_t adjacent3(const or_neighbour_system3d *ns, int32_t ti, int32_t tj) {
    if (ti < 1 || tj < 1 || ti > ns->max_tag || tj > ns->max_tag) return 0;
    return ns->adjacency[(size_t)(ti - 1) + (size_t)ns->max_tag * (size_t)(tj - 1)] == 1;
}

/* device_insertParticlesIntoCells! (NeighbourSearch.jl:12-47) with a z axis. */
static inline int32_t insert3_one(const or_layout *L, int32_t I, int32_t *is_alive, int32_t *cell_index,
                                  const int32_t *IP, const float *FP, or_neighbour_system3d *ns, int atomic) {
    if (is_alive[I] != 1) return 0;
    int32_t cell = cell_index[I];
    if (!(IPC(L, IP, I, L->IsMovable) == 0 && cell != 0)) {
        float x = FPC(L, FP, I, L->PositionVec);
        float y = FPC(L, FP, I, L->PositionVec + 1);
        float z = FPC(L, FP, I, L->PositionVec + 2);
        if (or3_inside(&ns->domain, x, y, z)) {
            cell = or3_index_linear_from_position(&ns->domain, x, y, z);
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
    if (slot > ns->max_contained) return OR_ERR_CELL_OVERFLOW;
    ns->contained_list[(size_t)(cell - 1) + (size_t)ns->n_cells * (size_t)(slot - 1)] = I + 1;
    return 0;
}

/* vdotx (Library.jl:21-34) with N = 3: 0 + rx (u_i - u_j) + ry (v_i - v_j) + rz (w_i - w_j). */
static inline float vdotx3(const or_layout *L, const float *FP, int32_t I, int32_t J, const float r[3]) {
    float v = 0.0f;
    for (int c = 0; c < (OR_MUTATION == 21 ? 2 : 3); ++c)
        v += r[c] * (FPC(L, FP, I, L->VelocityVec + c) - FPC(L, FP, J, L->VelocityVec + c));
    return v;
}

/* sComputeKernelAndiContinuity! (main.jl:74-90) with iValueGradient!
 * (Library/Kernel.jl:114-128) and iClassicContinuity! / iBalancedContinuity!
 * (Continuity.jl:23-47), N = 3. */
static void action3_kernel_continuity(const or_layout *L, int32_t I, int32_t *IP, float *FP, const or_physics *P) {
    int32_t n = IPC(L, IP, I, L->nCount);
    int32_t ti = IPC(L, IP, I, L->Tag);
    for (int32_t NI = 0; NI < n; ++NI) {
        int32_t J = IPC(L, IP, I, L->nIndex + NI) - 1;
        int32_t tj = IPC(L, IP, J, L->Tag);
        int balanced;
        if (ti == P->fluid_tag && tj == P->fluid_tag) balanced = 0;
        else if ((ti == P->fluid_tag && tj == P->wall_tag) || (ti == P->wall_tag && tj == P->fluid_tag)) balanced = 1;
        else continue;
        float r = FPC(L, FP, I, L->nR + NI);
        float hinv = or_invharmonic(FPC(L, FP, I, L->H), FPC(L, FP, J, L->H));
        float w = or3_kernel_value(P->kernel, r, hinv);
        float dw = or3_kernel_gradient(P->kernel, r, hinv);
        if (L->nHInv >= 0) FPC(L, FP, I, L->nHInv + NI) = hinv;
        if (L->nW >= 0) FPC(L, FP, I, L->nW + NI) = w;
        if (L->nDW >= 0) FPC(L, FP, I, L->nDW + NI) = dw;
        float rv[3] = {FPC(L, FP, I, L->nRVec + 3 * NI), FPC(L, FP, I, L->nRVec + 3 * NI + 1),
                       FPC(L, FP, I, L->nRVec + 3 * NI + 2)};
        float vdotx = vdotx3(L, FP, I, J, rv);
        float term;
        if (!balanced) term = FPC(L, FP, J, L->Mass) * vdotx * dw / r;
        else term = FPC(L, FP, I, L->Density) * FPC(L, FP, J, L->Volume) * vdotx * dw / r;
        FPC(L, FP, I, L->dDensity) += term;
    }
}

/* device_findNeighbourParticlesFromCells! (NeighbourSearch.jl:166-227) with a
 * z axis; dr2 = dx*dx + dy*dy + dz*dz left to right. */
static int32_t find3_one(const or_layout *L, int32_t I, const int32_t *is_alive, const int32_t *cell_index,
                         int32_t *IP, float *FP, const or_neighbour_system3d *ns, int32_t criterion, int32_t action,
                         const or_physics *P) {
    if (is_alive[I] != 1) return 0;
    int32_t err = 0;
    const int32_t M = L->max_neighbour_number;
    const size_t nc = (size_t)ns->n_cells;
    int32_t count = 0;
    int32_t cell = cell_index[I];
    int32_t ti = IPC(L, IP, I, L->Tag);
    float xi = FPC(L, FP, I, L->PositionVec), yi = FPC(L, FP, I, L->PositionVec + 1);
    float zi = FPC(L, FP, I, L->PositionVec + 2);
    int32_t ncells = ns->neighbour_cell_count[cell - 1];
    for (int32_t k = 0; k < ncells; ++k) {
        int32_t nb = ns->neighbour_cell_list[(size_t)(cell - 1) + nc * (size_t)k];
        int32_t cnt = ns->contained_count[nb - 1];
        if (cnt > ns->max_contained) cnt = ns->max_contained;
        for (int32_t s = 0; s < cnt; ++s) {
            int32_t J = ns->contained_list[(size_t)(nb - 1) + nc * (size_t)s] - 1;
            if (I == J || !adjacent3(ns, ti, IPC(L, IP, J, L->Tag))) continue;
            float dx = xi - FPC(L, FP, J, L->PositionVec);
            float dy = yi - FPC(L, FP, J, L->PositionVec + 1);
            float dz = zi - FPC(L, FP, J, L->PositionVec + 2);
            float dr2 = OR_MUTATION == 15 ? dx * dx + dy * dy : dx * dx + dy * dy + dz * dz;
            int ok;
            if (criterion == OR_CRIT_DEFAULT) {
                ok = dr2 <= ns->domain.gap_square;
            } else {
                float h2 = or_harmonic(FPC(L, FP, I, L->H), FPC(L, FP, J, L->H)) * 2.0f;
                ok = dr2 <= h2 * h2;
            }
            if (!ok) continue;
            if (count >= M) {
                err |= OR_ERR_NEIGHBOUR_OVERFLOW;
                continue;
            }
            count += 1;
            IPC(L, IP, I, L->nIndex + count - 1) = J + 1;
            if (L->nRVec >= 0) {
                FPC(L, FP, I, L->nRVec + 3 * (count - 1)) = dx;
                FPC(L, FP, I, L->nRVec + 3 * (count - 1) + 1) = dy;
                FPC(L, FP, I, L->nRVec + 3 * (count - 1) + 2) = dz;
            }
            if (L->nR >= 0) FPC(L, FP, I, L->nR + count - 1) = sqrtf(dr2);
        }
    }
    IPC(L, IP, I, L->nCount) = count;
    if (action == 1) action3_kernel_continuity(L, I, IP, FP, P);
    return err;
}

int32_t or3_search(const or_layout *L, int32_t n_particles, int32_t *is_alive, int32_t *cell_index, int32_t *IP,
                   float *FP, or_neighbour_system3d *ns, int32_t criterion, int32_t action, const or_physics3d *phys,
                   int32_t serial_insert, int32_t n_threads) {
    int32_t err = 0;
#ifdef _OPENMP
    if (n_threads > 0) omp_set_num_threads(n_threads);
#else
    (void)n_threads;
#endif
    memset(ns->contained_count, 0, sizeof(int32_t) * (size_t)ns->n_cells);
    if (serial_insert) {
        for (int32_t I = 0; I < n_particles; ++I) err |= insert3_one(L, I, is_alive, cell_index, IP, FP, ns, 0);
    } else {
#pragma omp parallel for schedule(static) reduction(| : err)
        for (int32_t I = 0; I < n_particles; ++I) err |= insert3_one(L, I, is_alive, cell_index, IP, FP, ns, 1);
    }
#pragma omp parallel for schedule(dynamic, 64) reduction(| : err)
    for (int32_t I = 0; I < n_particles; ++I)
        err |= find3_one(L, I, is_alive, cell_index, IP, FP, ns, criterion, action, &phys->p);
    return err;
}

/* ----------------------------------------------------------- Self actions */

void or3_selfaction(const or_layout *L, int32_t n_particles, const int32_t *is_alive, int32_t *IP, float *FP,
                    int32_t action, const or_physics3d *phys, int32_t n_threads) {
#ifdef _OPENMP
    if (n_threads > 0) omp_set_num_threads(n_threads);
#else
    (void)n_threads;
#endif
    const or_physics *P = &phys->p;
#pragma omp parallel for schedule(static)
    for (int32_t I = 0; I < n_particles; ++I) {
        if (is_alive[I] != 1) continue;
        int32_t tag = IPC(L, IP, I, L->Tag);
        if (action == 1) {
            /* sContinuity! main.jl:92-105: Continuity.jl:10-21, sVolume! State.jl:10-19, p = eos(rho) */
            if (tag != P->fluid_tag && tag != P->wall_tag) continue;
            FPC(L, FP, I, L->Density) += FPC(L, FP, I, L->dDensity) * P->dt;
            FPC(L, FP, I, L->dDensity) = 0.0f;
            FPC(L, FP, I, L->Volume) = FPC(L, FP, I, L->Mass) / FPC(L, FP, I, L->Density);
            FPC(L, FP, I, L->Pressure) = eos3(P, FPC(L, FP, I, L->Density));
        } else if (action == 2) {
            /* sMomentum! main.jl:120-126: the 3D sGravity! State.jl:49-63, sAccelerateMove! Motion.jl:61-75 */
            if (tag != P->fluid_tag) continue;
            FPC(L, FP, I, L->dVelocityVec) += P->gx;
            FPC(L, FP, I, L->dVelocityVec + 1) += OR_MUTATION == 19 ? phys->gz : P->gy;
            FPC(L, FP, I, L->dVelocityVec + 2) += OR_MUTATION == 19 ? P->gy : phys->gz;
            for (int32_t k = 0; k < 3; ++k) {
                float du = FPC(L, FP, I, L->dVelocityVec + k);
                FPC(L, FP, I, L->PositionVec + k) += (FPC(L, FP, I, L->VelocityVec + k) + du * P->dt * 0.5f) * P->dt;
                FPC(L, FP, I, L->VelocityVec + k) += FPC(L, FP, I, L->dVelocityVec + k) * P->dt;
                FPC(L, FP, I, L->dVelocityVec + k) = 0.0f;
            }
        }
    }
}

/* ----------------------------------------------------------- Interaction */

/* iMomentum! (main.jl:107-118) with iClassicPressure! / iBalancedPressure!
 * (Pressure.jl:24-60) and iClassicViscosity! / sArtificialViscosity!
 * (Viscosity.jl:10-52), N = 3, in the order the reference adds to du. */
static void momentum3_pair(const or_layout *L, int32_t I, int32_t NI, float *FP, const int32_t *IP,
                           const or_physics *P) {
    int32_t J = IPC(L, IP, I, L->nIndex + NI) - 1;
    int32_t ti = IPC(L, IP, I, L->Tag), tj = IPC(L, IP, J, L->Tag);
    int classic;
    if (ti == P->fluid_tag && tj == P->fluid_tag) classic = 1;
    else if (ti == P->fluid_tag && tj == P->wall_tag) classic = 0;
    else return;
    float r = FPC(L, FP, I, L->nR + NI);
    float dw = FPC(L, FP, I, L->nDW + NI);
    float hinv = FPC(L, FP, I, L->nHInv + NI);
    float rv[3] = {FPC(L, FP, I, L->nRVec + 3 * NI), FPC(L, FP, I, L->nRVec + 3 * NI + 1),
                   FPC(L, FP, I, L->nRVec + 3 * NI + 2)};
    float rho_i = FPC(L, FP, I, L->Density), rho_j = FPC(L, FP, J, L->Density);
    float p_i = FPC(L, FP, I, L->Pressure), p_j = FPC(L, FP, J, L->Pressure);
    float m_j = FPC(L, FP, J, L->Mass);
    float p_rho_2 = classic ? p_i / (rho_i * rho_i) + p_j / (rho_j * rho_j) : (p_i + p_j) / (rho_i * rho_j);
    p_rho_2 += fabsf(p_rho_2) * P->coefficient;
    p_rho_2 *= -m_j * dw / r;
    for (int c = 0; c < (OR_MUTATION == 20 ? 2 : 3); ++c) FPC(L, FP, I, L->dVelocityVec + c) += p_rho_2 * rv[c];
    float du[3];
    for (int c = 0; c < 3; ++c) du[c] = FPC(L, FP, I, L->VelocityVec + c) - FPC(L, FP, J, L->VelocityVec + c);
    float h = 1.0f / hinv;
    float az = r * r + 0.01f * h * h; /* avoidzero, Library.jl:17-19 */
    if (P->visc == OR_VISC_CLASSIC) {
        float force = 2.0f * P->mu * m_j * r * dw / (rho_i * rho_j * az);
        for (int c = 0; c < 3; ++c) FPC(L, FP, I, L->dVelocityVec + c) += force * du[c];
        return;
    }
    float vdotx = 0.0f;
    for (int c = 0; c < 3; ++c) vdotx += rv[c] * du[c];
    if (vdotx > 0.0f) return;
    float rho = (rho_i + rho_j) * 0.5f;
    float phi = vdotx / (hinv * az);
    float force = (-P->alpha * P->av_c + P->beta * phi) * phi / rho;
    force *= -m_j / (rho_i * rho_j * r) * dw;
    for (int c = 0; c < 3; ++c) FPC(L, FP, I, L->dVelocityVec + c) += force * rv[c];
}

void or3_interaction(const or_layout *L, int32_t n_particles, const int32_t *is_alive, int32_t *IP, float *FP,
                     int32_t action, const or_physics3d *phys, int32_t n_threads) {
#ifdef _OPENMP
    if (n_threads > 0) omp_set_num_threads(n_threads);
#else
    (void)n_threads;
#endif
    if (action != 1) return;
#pragma omp parallel for schedule(dynamic, 64)
    for (int32_t I = 0; I < n_particles; ++I) {
        if (is_alive[I] != 1) continue;
        int32_t n = IPC(L, IP, I, L->nCount);
        for (int32_t k = 0; k < n; ++k) momentum3_pair(L, I, k, FP, IP, &phys->p);
    }
}

/* ---------------------------------------------------------------- Driver */

int32_t or3_run_steps(const or_layout *L, int32_t n_particles, int32_t *is_alive, int32_t *cell_index, int32_t *IP,
                      float *FP, or_neighbour_system3d *ns, const or_physics3d *phys, int32_t n_steps,
                      int32_t serial_insert, int32_t n_threads, int64_t *pairs_per_step) {
    int32_t err = 0;
    for (int32_t step = 1; step <= n_steps; ++step) {
        or3_selfaction(L, n_particles, is_alive, IP, FP, 1, phys, n_threads);
        or3_interaction(L, n_particles, is_alive, IP, FP, 1, phys, n_threads);
        or3_selfaction(L, n_particles, is_alive, IP, FP, 2, phys, n_threads);
        err |= or3_search(L, n_particles, is_alive, cell_index, IP, FP, ns, OR_CRIT_SYMMETRY, 1, phys, serial_insert,
                          n_threads);
        if (pairs_per_step) {
            int64_t total = 0;
            for (int32_t I = 0; I < n_particles; ++I)
                if (is_alive[I] == 1) total += IPC(L, IP, I, L->nCount);
            pairs_per_step[step - 1] = total;
        }
    }
    return err;
}

/* ------------------------------------------------------ magnitude bounds */

/* |f| + r |df/dr| of f = dw / r in double, N = 3 (dw = C h^4 G(q)): the bound
 * |C| h^4 (2 |G| + |q G'|) / r, as the 2D bound of eppx_oracle.c. */
static double grad3_over_r_bound(int32_t kernel, double r, double hinv) {
    const double q = r * hinv, h4 = hinv * hinv * hinv * hinv;
    double c, g, gp;
    if (kernel == OR_KERNEL_CUBIC) {
        c = 0.75 * (1.0 / M_PI);
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
        c = 0.625 * (21.0 / 16.0 / M_PI);
        g = q * (2.0 - q) * (2.0 - q) * (2.0 - q);
        gp = (2.0 - q) * (2.0 - q) * (2.0 - 4.0 * q);
    }
    return c * h4 * (2.0 * fabs(g) + fabs(q * gp)) / r;
}

static double eos3_slope(const or_physics *P, double rho) {
    if (P->eos == OR_EOS_LINEAR) return (double)P->c0 * P->c0;
    const double x = rho / P->rho0;
    return 7.0 * P->tait_b * x * x * x * x * x * x / P->rho0;
}

static double ulp32_3(double x) {
    const float f = (float)fabs(x);
    return (double)nextafterf(f, INFINITY) - (double)f;
}

/* The 3D counterpart of or_pair_term_magnitude: what = 0, the continuity
 * terms (out[I]); what = 1, the momentum terms (out[3 I + k]). */
void or3_pair_term_magnitude(const or_layout *L, int32_t n_particles, const int32_t *is_alive, int32_t *IP,
                             float *FP, const or_physics3d *phys, int32_t what, double *out) {
    const or_physics *P = &phys->p;
#pragma omp parallel for schedule(dynamic, 64)
    for (int32_t I = 0; I < n_particles; ++I) {
        double a[3] = {0.0, 0.0, 0.0};
        if (is_alive[I] == 1) {
            int32_t n = IPC(L, IP, I, L->nCount);
            int32_t ti = IPC(L, IP, I, L->Tag);
            for (int32_t NI = 0; NI < n; ++NI) {
                int32_t J = IPC(L, IP, I, L->nIndex + NI) - 1;
                int32_t tj = IPC(L, IP, J, L->Tag);
                double r = FPC(L, FP, I, L->nR + NI);
                double rv[3], du[3];
                for (int c = 0; c < 3; ++c) {
                    rv[c] = FPC(L, FP, I, L->nRVec + 3 * NI + c);
                    du[c] = (double)FPC(L, FP, I, L->VelocityVec + c) - FPC(L, FP, J, L->VelocityVec + c);
                }
                double vabs = fabs(rv[0] * du[0]) + fabs(rv[1] * du[1]) + fabs(rv[2] * du[2]);
                if (what == 0) {
                    int balanced;
                    if (ti == P->fluid_tag && tj == P->fluid_tag) balanced = 0;
                    else if ((ti == P->fluid_tag && tj == P->wall_tag) || (ti == P->wall_tag && tj == P->fluid_tag))
                        balanced = 1;
                    else continue;
                    double hinv = or_invharmonic(FPC(L, FP, I, L->H), FPC(L, FP, J, L->H));
                    double m = balanced ? (double)FPC(L, FP, I, L->Density) * FPC(L, FP, J, L->Volume)
                                        : FPC(L, FP, J, L->Mass);
                    a[0] += fabs(m) * vabs * grad3_over_r_bound(P->kernel, r, hinv);
                    continue;
                }
                int classic;
                if (ti == P->fluid_tag && tj == P->fluid_tag) classic = 1;
                else if (ti == P->fluid_tag && tj == P->wall_tag) classic = 0;
                else continue;
                double hinv = FPC(L, FP, I, L->nHInv + NI);
                double dwr = grad3_over_r_bound(P->kernel, r, hinv);
                double rho_i = FPC(L, FP, I, L->Density), rho_j = FPC(L, FP, J, L->Density);
                double p_i = fabs((double)FPC(L, FP, I, L->Pressure)) + eos3_slope(P, rho_i) * ulp32_3(rho_i);
                double p_j = fabs((double)FPC(L, FP, J, L->Pressure)) + eos3_slope(P, rho_j) * ulp32_3(rho_j);
                double m_j = FPC(L, FP, J, L->Mass);
                double pr = classic ? p_i / (rho_i * rho_i) + p_j / (rho_j * rho_j) : (p_i + p_j) / (rho_i * rho_j);
                pr *= 1.0 + fabs((double)P->coefficient);
                double radial = pr * fabs(m_j) * dwr;
                double h = 1.0 / hinv;
                double az = r * r + 0.01 * h * h;
                for (int c = 0; c < 3; ++c) a[c] += radial * fabs(rv[c]);
                if (P->visc == OR_VISC_CLASSIC) {
                    double force = fabs(2.0 * P->mu * m_j * r * r / (rho_i * rho_j * az)) * dwr;
                    for (int c = 0; c < 3; ++c) a[c] += force * fabs(du[c]);
                    continue;
                }
                double vdotx = rv[0] * du[0] + rv[1] * du[1] + rv[2] * du[2];
                if (vdotx > 1e-6 * vabs) continue;
                double phi = vabs / (hinv * az);
                double force =
                    (fabs((double)P->alpha * P->av_c) + fabs((double)P->beta) * phi) * phi / ((rho_i + rho_j) * 0.5);
                force *= fabs(m_j / (rho_i * rho_j)) * dwr;
                for (int c = 0; c < 3; ++c) a[c] += force * fabs(rv[c]);
            }
        }
        if (what == 0) out[I] = a[0];
        else
            for (int c = 0; c < 3; ++c) out[3 * (size_t)I + c] = a[c];
    }
}
