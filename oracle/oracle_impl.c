/*
 * CPU oracle implementation: TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Compiled twice: with ORACLE_F64 undefined (FP32 reference semantics) and
 * defined (FP64 oracle). Every expression keeps the reference's operand order;
 * the build uses -ffp-contract=off so no FMA is formed, and float/double
 * arithmetic is IEEE (x86-64 SSE2, FLT_EVAL_METHOD == 0).
 *
 * Each function cites the reference file:line it restates; paths are relative
 * to fluids/liquid/particle/ of Xayah-Graphics/physica.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#ifdef ORACLE_F64
typedef double real;
#define SUF(name) name##_f64
#define R_SQRT sqrt
#define R_FLOOR floor
#define R_POW pow
#define R_LOG log
#define R_FMIN fmin
#define R_FMAX fmax
#define R_PI 3.14159265358979323846
#else
typedef float real;
#define SUF(name) name##_f32
#define R_SQRT sqrtf
#define R_FLOOR floorf
#define R_POW powf
#define R_LOG logf
#define R_FMIN fminf
#define R_FMAX fmaxf
#define R_PI 3.14159265358979323846F
#endif
#define K(c) ((real)(c))

/* ---------------------------------------------------------------- vectors */
/* domain/device.cuh:8-82 */
typedef struct {
    real x, y, z;
} v3;

static inline v3 v3_make(real x, real y, real z) {
    v3 r = {x, y, z};
    return r;
}
static inline v3 v3_load(const real* const f[3], uint32_t i) { return v3_make(f[0][i], f[1][i], f[2][i]); }
static inline void v3_store(real* const f[3], uint32_t i, v3 v) {
    f[0][i] = v.x;
    f[1][i] = v.y;
    f[2][i] = v.z;
}
static inline v3 v3_add(v3 a, v3 b) { return v3_make(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline v3 v3_sub(v3 a, v3 b) { return v3_make(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline v3 v3_scale(v3 v, real f) { return v3_make(f * v.x, f * v.y, f * v.z); }
static inline real v3_dot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline v3 v3_cross(v3 a, v3 b) {
    return v3_make(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
static inline real v3_length(v3 a) { return R_SQRT(v3_dot(a, a)); }

/* ------------------------------------------------------------ SPH kernels */
/* density/device.cuh:10-16 */
static real poly6(v3 d, real h) {
    const real sq = v3_dot(d, d);
    const real hr = h * h;
    if (sq >= hr) return K(0.0);
    const real diff = hr - sq;
    return K(315.0) * diff * diff * diff / (K(64.0) * R_PI * hr * hr * hr * h * h * h);
}

/* density/device.cuh:18-24 */
static v3 poly6_gradient(v3 d, real h) {
    const real sq = v3_dot(d, d);
    const real hr = h * h;
    if (sq >= hr) return v3_make(0, 0, 0);
    const real diff = hr - sq;
    return v3_scale(d, K(-945.0) * diff * diff / (K(32.0) * R_PI * hr * hr * hr * h * h * h));
}

/* density/device.cuh:26-34 */
static real cubic(v3 d, real h) {
    const real dist = v3_length(d);
    if (dist >= h) return K(0.0);
    const real q = dist / h;
    const real coef = K(8.0) / (R_PI * h * h * h);
    if (q <= K(0.5)) return coef * (K(6.0) * q * q * q - K(6.0) * q * q + K(1.0));
    const real diff = K(1.0) - q;
    return coef * K(2.0) * diff * diff * diff;
}

/* density/device.cuh:36-43 */
static v3 cubic_gradient(v3 d, real h) {
    const real dist = v3_length(d);
    if (dist == K(0.0) || dist >= h) return v3_make(0, 0, 0);
    const real q = dist / h;
    const real coef = K(8.0) / (R_PI * h * h * h);
    const real deriv = q <= K(0.5) ? K(18.0) * q * q - K(12.0) * q : K(-6.0) * (K(1.0) - q) * (K(1.0) - q);
    return v3_scale(d, coef * deriv / (h * dist));
}

/* density/device.cuh:72-76 */
static real viscosity_laplacian(v3 d, real h) {
    const real dist = v3_length(d);
    if (dist >= h) return K(0.0);
    return K(45.0) * (h - dist) / (R_PI * h * h * h * h * h * h);
}

/* pbf.cu:26-32 (the coefficient -45/(pi h^6) is a per-call constant) */
static real spiky_coefficient(real h) { return K(-45.0) / (R_PI * R_POW(h, K(6.0))); }

static v3 spiky_gradient(v3 d, real h, real coefficient) {
    const real dist = v3_length(d);
    if (dist == K(0.0) || dist >= h) return v3_make(0, 0, 0);
    const real diff = h - dist;
    return v3_scale(d, coefficient * diff * diff / dist);
}

/* pbf.cu:56-71, value only (the forward correction uses nothing else) */
static real artificial_value(v3 d, real h, real strength, real exponent, real radius) {
    const real reference = poly6(v3_make(radius, K(0.0), K(0.0)), h);
    const real kernel = poly6(d, h);
    if (kernel == K(0.0)) return K(0.0);
    const real ratio = kernel / reference;
    const real power = R_POW(ratio, exponent);
    return -strength * power;
}

/* ------------------------------------------------------------------- grid */
#ifndef ORACLE_F64
void oracle_grid_resolution(const oracle_domain* d, uint32_t res[3]) {
    /* neighborhood.cpp:25-30 and neighborhood.cu:31-33 (FP32 in both) */
    for (int a = 0; a < 3; ++a) res[a] = (uint32_t)ceilf((d->box_maximum[a] - d->box_minimum[a]) / d->support_radius);
}
#endif

typedef struct {
    uint32_t res[3];
    uint32_t cells;
    real origin[3];
    real cell_size;
    real time;
    uint32_t n, nb;
    uint64_t* keys;
    uint32_t* idx;
    uint32_t* off;
    uint64_t* bkeys;
    uint32_t* bidx;
    uint32_t* boff;
    const oracle_domain* d;
} grid_t;

/* min(C-1, max(0, __float2int_rd(v))) of neighborhood.cu:34-36 and
 * neighborhood/device.cuh:44-48 */
static int cell_clamp(real v, uint32_t count) {
    const real f = R_FLOOR(v);
    if (!(f > K(0.0))) return 0;
    if (f >= (real)(count - 1u)) return (int)count - 1;
    return (int)f;
}

/* key_kernel, neighborhood.cu:20-39 */
static uint64_t grid_key(const oracle_domain* d, const uint32_t res[3], real time, real x, real y, real z) {
    const real h = (real)d->support_radius;
    const int cx = cell_clamp((x - (real)d->box_minimum[0] - time * (real)d->box_velocity[0]) / h, res[0]);
    const int cy = cell_clamp((y - (real)d->box_minimum[1] - time * (real)d->box_velocity[1]) / h, res[1]);
    const int cz = cell_clamp((z - (real)d->box_minimum[2] - time * (real)d->box_velocity[2]) / h, res[2]);
    return ((uint64_t)cz * res[1] + (uint64_t)cy) * res[0] + (uint64_t)cx;
}

/* lower_bound, neighborhood.cu:9-18 */
static uint32_t lower_bound(const uint64_t* keys, uint32_t count, uint64_t key) {
    uint32_t first = 0u, last = count;
    while (first < last) {
        const uint32_t middle = first + (last - first) / 2u;
        if (keys[middle] < key)
            first = middle + 1u;
        else
            last = middle;
    }
    return first;
}

/* stable sort of (key, index) pairs by key; keys < cells. A sequential
 * counting sort places equal keys in input order, which is exactly the
 * stable-sort semantics of cub::DeviceRadixSort::SortPairs used at
 * neighborhood.cu:57,68. */
static void stable_sort(const uint64_t* keys, uint32_t n, uint32_t cells, uint64_t* sorted_keys, uint32_t* sorted_idx) {
    uint32_t* start = (uint32_t*)calloc((size_t)cells + 1u, sizeof(uint32_t));
    for (uint32_t i = 0; i < n; ++i) start[keys[i] + 1u]++;
    for (uint32_t c = 0; c < cells; ++c) start[c + 1u] += start[c];
    for (uint32_t i = 0; i < n; ++i) {
        const uint32_t p = start[keys[i]]++;
        sorted_keys[p] = keys[i];
        sorted_idx[p] = i;
    }
    free(start);
}

/* build_neighborhood, neighborhood.cu:54-70, plus the host-side origin of
 * neighborhood.cpp:83-88 */
static void grid_build(grid_t* g, const oracle_domain* d, uint64_t step, const real* px, const real* py, const real* pz) {
    memset(g, 0, sizeof(*g));
    g->d = d;
    oracle_grid_resolution(d, g->res);
    g->cells = g->res[0] * g->res[1] * g->res[2];
    g->n = d->particle_count;
    g->nb = d->boundary_count;
    const real time = (real)step * (real)d->time_step;
    g->time = time;
    for (int a = 0; a < 3; ++a) g->origin[a] = (real)d->box_minimum[a] + time * (real)d->box_velocity[a];
    g->cell_size = (real)d->support_radius;
    uint64_t* unsorted = (uint64_t*)malloc(sizeof(uint64_t) * (g->n > g->nb ? g->n : g->nb) + 8u);
    g->keys = (uint64_t*)malloc(sizeof(uint64_t) * g->n + 8u);
    g->idx = (uint32_t*)malloc(sizeof(uint32_t) * g->n + 4u);
    g->off = (uint32_t*)malloc(sizeof(uint32_t) * ((size_t)g->cells + 1u));
    g->bkeys = (uint64_t*)malloc(sizeof(uint64_t) * g->nb + 8u);
    g->bidx = (uint32_t*)malloc(sizeof(uint32_t) * g->nb + 4u);
    g->boff = (uint32_t*)malloc(sizeof(uint32_t) * ((size_t)g->cells + 1u));
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (int64_t)g->n; ++i) unsorted[i] = grid_key(d, g->res, time, px[i], py[i], pz[i]);
    stable_sort(unsorted, g->n, g->cells, g->keys, g->idx);
#pragma omp parallel for schedule(static)
    for (int64_t c = 0; c <= (int64_t)g->cells; ++c) g->off[c] = lower_bound(g->keys, g->n, (uint64_t)c);
    if (g->nb == 0u) {
        memset(g->boff, 0, sizeof(uint32_t) * ((size_t)g->cells + 1u));
    } else {
        for (uint32_t i = 0; i < g->nb; ++i) {
            /* key_kernel with moving_positions = true: x += t v first */
            const real x = (real)d->boundary_position[0][i] + time * (real)d->boundary_velocity[0][i];
            const real y = (real)d->boundary_position[1][i] + time * (real)d->boundary_velocity[1][i];
            const real z = (real)d->boundary_position[2][i] + time * (real)d->boundary_velocity[2][i];
            unsorted[i] = grid_key(d, g->res, time, x, y, z);
        }
        stable_sort(unsorted, g->nb, g->cells, g->bkeys, g->bidx);
        for (uint32_t c = 0; c <= g->cells; ++c) g->boff[c] = lower_bound(g->bkeys, g->nb, (uint64_t)c);
    }
    free(unsorted);
}

static void grid_free(grid_t* g) {
    free(g->keys);
    free(g->idx);
    free(g->off);
    free(g->bkeys);
    free(g->bidx);
    free(g->boff);
}

typedef struct {
    uint32_t first, last, bfirst, blast;
} cell_range;

/* 27-cell walk of every neighbour kernel: particle_cell + cell_range,
 * neighborhood/device.cuh:28-48, visited offset_x outer, offset_z inner
 * (e.g. density.cu:25-27). Returns the valid ranges in visiting order. */
static int stencil(const grid_t* g, v3 topo, cell_range out[27]) {
    const int cx = cell_clamp((topo.x - g->origin[0]) / g->cell_size, g->res[0]);
    const int cy = cell_clamp((topo.y - g->origin[1]) / g->cell_size, g->res[1]);
    const int cz = cell_clamp((topo.z - g->origin[2]) / g->cell_size, g->res[2]);
    int n = 0;
    for (int ox = -1; ox <= 1; ++ox)
        for (int oy = -1; oy <= 1; ++oy)
            for (int oz = -1; oz <= 1; ++oz) {
                const int x = cx + ox, y = cy + oy, z = cz + oz;
                if (x < 0 || x >= (int)g->res[0] || y < 0 || y >= (int)g->res[1] || z < 0 || z >= (int)g->res[2]) continue;
                const uint32_t c = ((uint32_t)z * g->res[1] + (uint32_t)y) * g->res[0] + (uint32_t)x;
                out[n].first = g->off[c];
                out[n].last = g->off[c + 1u];
                out[n].bfirst = g->boff[c];
                out[n].blast = g->boff[c + 1u];
                ++n;
            }
    return n;
}

/* boundary_position / boundary_velocity, neighborhood/device.cuh:16-26 */
static v3 boundary_position(const grid_t* g, uint32_t b) {
    const oracle_domain* d = g->d;
    return v3_make((real)d->boundary_position[0][b] + g->time * (real)d->boundary_velocity[0][b],
                   (real)d->boundary_position[1][b] + g->time * (real)d->boundary_velocity[1][b],
                   (real)d->boundary_position[2][b] + g->time * (real)d->boundary_velocity[2][b]);
}
static v3 boundary_velocity(const grid_t* g, uint32_t b) {
    const oracle_domain* d = g->d;
    return v3_make((real)d->boundary_velocity[0][b], (real)d->boundary_velocity[1][b], (real)d->boundary_velocity[2][b]);
}
static real boundary_volume(const grid_t* g, uint32_t b) { return (real)g->d->boundary_volume[b]; }

/* ---------------------------------------------------------------- helpers */
static real* alloc_real(uint32_t n) { return (real*)malloc(sizeof(real) * (size_t)n + 16u); }
typedef struct {
    real* c[3];
} vfield;
static vfield alloc_vfield(uint32_t n) {
    vfield f;
    for (int a = 0; a < 3; ++a) f.c[a] = alloc_real(n);
    return f;
}
static void free_vfield(vfield* f) {
    for (int a = 0; a < 3; ++a) free(f->c[a]);
}
#define CV(f) ((const real* const*)(f).c)

/* Box of the solvers' collision_box, sph.cpp:21-35 and pbf.cpp:17-31 */
typedef struct {
    real lo[3], hi[3], vel[3];
    int no_slip;
} cbox;
static cbox collision_box(const oracle_domain* d, uint64_t step) {
    cbox b;
    const real time = (real)step * (real)d->time_step;
    const real rp = (real)d->particle_radius;
    for (int a = 0; a < 3; ++a) {
        b.lo[a] = (real)d->box_minimum[a] + time * (real)d->box_velocity[a] + rp;
        b.hi[a] = (real)d->box_maximum[a] + time * (real)d->box_velocity[a] - rp;
        b.vel[a] = (real)d->box_velocity[a];
    }
    b.no_slip = d->no_slip != 0u;
    return b;
}

/* ------------------------------------------------------------- density */
/* density_forward_kernel, density.cu:18-40 */
static void density_kernel(const grid_t* g, int pbf_kernel, const real* const topo[3], const real* const pos[3],
                           const real* masses, const real* rest_densities, real* out) {
    const real h = (real)g->d->support_radius;
#pragma omp parallel for schedule(static)
    for (int64_t ii = 0; ii < (int64_t)g->n; ++ii) {
        const uint32_t i = (uint32_t)ii;
        const v3 p = v3_load(pos, i);
        cell_range r[27];
        const int nr = stencil(g, v3_load(topo, i), r);
        real result = K(0.0);
        for (int k = 0; k < nr; ++k) {
            for (uint32_t s = r[k].first; s < r[k].last; ++s) {
                const uint32_t j = g->idx[s];
                const v3 d = v3_sub(p, v3_load(pos, j));
                result += masses[j] * (pbf_kernel ? poly6(d, h) : cubic(d, h));
            }
            for (uint32_t s = r[k].bfirst; s < r[k].blast; ++s) {
                const uint32_t b = g->bidx[s];
                const v3 d = v3_sub(p, boundary_position(g, b));
                result += rest_densities[i] * boundary_volume(g, b) * (pbf_kernel ? poly6(d, h) : cubic(d, h));
            }
        }
        out[i] = result;
    }
}

/* ------------------------------------------------------------ SPH common */
/* non_pressure_forward_kernel, sph/common.cu:14-47 */
static void non_pressure_kernel(const grid_t* g, const real* const pos[3], const real* const vel[3],
                                const real* const ctl[3], const real* masses, const real* viscosities,
                                const real* densities, real* const out[3]) {
    const oracle_domain* dm = g->d;
    const real h = (real)dm->support_radius;
    const v3 gravity = v3_make((real)dm->gravity[0], (real)dm->gravity[1], (real)dm->gravity[2]);
#pragma omp parallel for schedule(static)
    for (int64_t ii = 0; ii < (int64_t)g->n; ++ii) {
        const uint32_t i = (uint32_t)ii;
        const v3 p = v3_load(pos, i);
        const v3 v = v3_load(vel, i);
        v3 acc = v3_add(gravity, v3_load(ctl, i));
        cell_range r[27];
        const int nr = stencil(g, p, r);
        for (int k = 0; k < nr; ++k) {
            for (uint32_t s = r[k].first; s < r[k].last; ++s) {
                const uint32_t j = g->idx[s];
                if (j == i) continue;
                const v3 d = v3_sub(p, v3_load(pos, j));
                const real visc = K(0.5) * (viscosities[i] + viscosities[j]);
                const real w = visc * masses[j] * viscosity_laplacian(d, h) / densities[j];
                acc = v3_add(acc, v3_scale(v3_sub(v3_load(vel, j), v), w));
            }
            for (uint32_t s = r[k].bfirst; s < r[k].blast; ++s) {
                const uint32_t b = g->bidx[s];
                const v3 d = v3_sub(p, boundary_position(g, b));
                const real w = viscosities[i] * boundary_volume(g, b) * viscosity_laplacian(d, h);
                acc = v3_add(acc, v3_scale(v3_sub(boundary_velocity(g, b), v), w));
            }
        }
        v3_store(out, i, acc);
    }
}

/* pressure_forward_kernel, sph/common.cu:165-194 */
static void pressure_kernel(const grid_t* g, const real* const pos[3], const real* masses, const real* rest_densities,
                            const real* densities, const real* pressures, real* const out[3]) {
    const real h = (real)g->d->support_radius;
#pragma omp parallel for schedule(static)
    for (int64_t ii = 0; ii < (int64_t)g->n; ++ii) {
        const uint32_t i = (uint32_t)ii;
        const v3 p = v3_load(pos, i);
        v3 acc = v3_make(0, 0, 0);
        const real first = pressures[i] / (densities[i] * densities[i]);
        cell_range r[27];
        const int nr = stencil(g, p, r);
        for (int k = 0; k < nr; ++k) {
            for (uint32_t s = r[k].first; s < r[k].last; ++s) {
                const uint32_t j = g->idx[s];
                if (j == i) continue;
                const real second = pressures[j] / (densities[j] * densities[j]);
                acc = v3_add(acc, v3_scale(cubic_gradient(v3_sub(p, v3_load(pos, j)), h), -masses[j] * (first + second)));
            }
            for (uint32_t s = r[k].bfirst; s < r[k].blast; ++s) {
                const uint32_t b = g->bidx[s];
                acc = v3_add(acc, v3_scale(cubic_gradient(v3_sub(p, boundary_position(g, b)), h),
                                           -rest_densities[i] * boundary_volume(g, b) * first));
            }
        }
        v3_store(out, i, acc);
    }
}

/* add_kernel, sph/common.cu:368-374 */
static void add_kernel(uint32_t n, const real* const a[3], const real* const b[3], real* const out[3]) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (int64_t)n; ++i)
        for (int c = 0; c < 3; ++c) out[c][i] = a[c][i] + b[c][i];
}

/* predict_forward_kernel, sph/dfsph.cu:26-44 and pcisph.cu:26-44 / iisph.cu
 * (a = a1 + a2); with a2 == NULL it is integrate_forward_kernel,
 * sph/common.cu:302-320 (a = a1, no addition). */
static void sph_predict_kernel(uint32_t n, real dt, cbox box, const real* const pos[3], const real* const vel[3],
                               const real* const a1[3], const real* const a2[3], real* const out_pos[3],
                               real* const out_vel[3]) {
#pragma omp parallel for schedule(static)
    for (int64_t ii = 0; ii < (int64_t)n; ++ii) {
        const uint32_t i = (uint32_t)ii;
        const v3 a = a2 ? v3_add(v3_load(a1, i), v3_load(a2, i)) : v3_load(a1, i);
        v3 v = v3_add(v3_load(vel, i), v3_scale(a, dt));
        v3 x = v3_add(v3_load(pos, i), v3_scale(v, dt));
        const int cx = x.x < box.lo[0] || x.x > box.hi[0];
        const int cy = x.y < box.lo[1] || x.y > box.hi[1];
        const int cz = x.z < box.lo[2] || x.z > box.hi[2];
        x.x = R_FMIN(box.hi[0], R_FMAX(box.lo[0], x.x));
        x.y = R_FMIN(box.hi[1], R_FMAX(box.lo[1], x.y));
        x.z = R_FMIN(box.hi[2], R_FMAX(box.lo[2], x.z));
        if (box.no_slip && (cx || cy || cz)) {
            v = v3_make(box.vel[0], box.vel[1], box.vel[2]);
        } else {
            if (cx) v.x = box.vel[0];
            if (cy) v.y = box.vel[1];
            if (cz) v.z = box.vel[2];
        }
        if (out_pos) v3_store(out_pos, i, x);
        if (out_vel) v3_store(out_vel, i, v);
    }
}

/* projection_update_forward_kernel, sph/dfsph.cu:104-110 */
static void projection_update_kernel(uint32_t n, real dt, real G, const real* masses, const real* densities,
                                     const real* targets, const real* predicted, const real* relaxation,
                                     real* pressures) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (int64_t)n; ++i) {
        const real volume = masses[i] / densities[i];
        const real delta = K(1.0) / (K(2.0) * dt * dt * volume * volume * G);
        pressures[i] = R_FMAX(K(0.0), pressures[i] + relaxation[i] * delta * (predicted[i] - targets[i]));
    }
}

/* compute_reference_gradient_norm, sph/sph.cpp:37-58 */
real SUF(oracle_reference_gradient_norm)(const oracle_domain* d) {
    const real h = (real)d->support_radius;
    const real diameter = K(2.0) * (real)d->particle_radius;
    const real coef = K(8.0) / (R_PI * h * h * h);
    v3 gs = v3_make(0, 0, 0);
    real sq = K(0.0);
    for (real x = -h; x <= h; x += diameter)
        for (real y = -h; y <= h; y += diameter)
            for (real z = -h; z <= h; z += diameter) {
                const real dist = R_SQRT(x * x + y * y + z * z);
                if (dist == K(0.0) || dist >= h) continue;
                const real q = dist / h;
                const real deriv = q <= K(0.5) ? K(18.0) * q * q - K(12.0) * q : K(-6.0) * (K(1.0) - q) * (K(1.0) - q);
                const real scale = coef * deriv / (h * dist);
                gs.x += scale * x;
                gs.y += scale * y;
                gs.z += scale * z;
                sq += scale * scale * (x * x + y * y + z * z);
            }
    return gs.x * gs.x + gs.y * gs.y + gs.z * gs.z + sq;
}

/* -------------------------------------------------------------- exports */
void SUF(oracle_build_neighborhood)(const oracle_domain* d, uint64_t step, const real* px, const real* py,
                                    const real* pz, uint64_t* sorted_keys, uint32_t* sorted_indices,
                                    uint32_t* cell_offsets, uint64_t* sorted_boundary_keys,
                                    uint32_t* sorted_boundary_indices, uint32_t* boundary_cell_offsets) {
    grid_t g;
    grid_build(&g, d, step, px, py, pz);
    if (sorted_keys) memcpy(sorted_keys, g.keys, sizeof(uint64_t) * g.n);
    if (sorted_indices) memcpy(sorted_indices, g.idx, sizeof(uint32_t) * g.n);
    if (cell_offsets) memcpy(cell_offsets, g.off, sizeof(uint32_t) * ((size_t)g.cells + 1u));
    if (sorted_boundary_keys) memcpy(sorted_boundary_keys, g.bkeys, sizeof(uint64_t) * g.nb);
    if (sorted_boundary_indices) memcpy(sorted_boundary_indices, g.bidx, sizeof(uint32_t) * g.nb);
    if (boundary_cell_offsets) memcpy(boundary_cell_offsets, g.boff, sizeof(uint32_t) * ((size_t)g.cells + 1u));
    grid_free(&g);
}

void SUF(oracle_kernel_values)(int kind, float support_radius, uint32_t count, const real* dx, const real* dy,
                               const real* dz, real* value, real* gx, real* gy, real* gz) {
    const real h = (real)support_radius;
    const real sc = spiky_coefficient(h);
    for (uint32_t i = 0; i < count; ++i) {
        const v3 d = v3_make(dx[i], dy[i], dz[i]);
        v3 gr = v3_make(0, 0, 0);
        real w = K(0.0);
        switch (kind) {
        case ORACLE_POLY6:
            w = poly6(d, h);
            gr = poly6_gradient(d, h);
            break;
        case ORACLE_CUBIC:
            w = cubic(d, h);
            gr = cubic_gradient(d, h);
            break;
        case ORACLE_SPIKY:
            gr = spiky_gradient(d, h, sc);
            break;
        case ORACLE_VISCOSITY_LAPLACIAN:
            w = viscosity_laplacian(d, h);
            break;
        default:
            break;
        }
        if (value) value[i] = w;
        if (gx) gx[i] = gr.x;
        if (gy) gy[i] = gr.y;
        if (gz) gz[i] = gr.z;
    }
}

void SUF(oracle_density)(const oracle_domain* d, uint64_t grid_step, int pbf_kernel, const real* topology[3],
                         const real* positions[3], const real* masses, const real* rest_densities, real* densities) {
    grid_t g;
    grid_build(&g, d, grid_step, topology[0], topology[1], topology[2]);
    density_kernel(&g, pbf_kernel, topology, positions, masses, rest_densities, densities);
    grid_free(&g);
}

static void density_jvp_kernel(const grid_t* g, int pbf, const real* const topo[3], const real* const pos[3],
                               const real* const pos_t[3], const real* masses, const real* masses_t,
                               const real* rest, const real* rest_t, real* out);
static void density_vjp_kernel(const grid_t* g, int pbf, const real* const topo[3], const real* const pos[3],
                               const real* masses, const real* rest, const double* rho_a, double* const pos_a[3],
                               double* masses_a, double* rest_a);

/* density_jvp / density_vjp (density.cu:42-98): the tangent of the density
 * (FP32 or FP64, the primal precision) and its VJP in FP64, accumulated
 * (+=) into pos_a, masses_a and rest_a, in gather form; the grid is built
 * from the topology positions at grid_step. */
void SUF(oracle_density_jvp)(const oracle_domain* d, uint64_t grid_step, int pbf_kernel, const real* topology[3],
                             const real* positions[3], const real* position_tangent[3], const real* masses,
                             const real* mass_tangent, const real* rest_densities, const real* rest_tangent,
                             real* out) {
    grid_t g;
    grid_build(&g, d, grid_step, topology[0], topology[1], topology[2]);
    density_jvp_kernel(&g, pbf_kernel, topology, positions, position_tangent, masses, mass_tangent, rest_densities,
                       rest_tangent, out);
    grid_free(&g);
}

void SUF(oracle_density_vjp)(const oracle_domain* d, uint64_t grid_step, int pbf_kernel, const real* topology[3],
                             const real* positions[3], const real* masses, const real* rest_densities,
                             const double* density_adjoint, double* position_adjoint[3], double* mass_adjoint,
                             double* rest_adjoint) {
    grid_t g;
    grid_build(&g, d, grid_step, topology[0], topology[1], topology[2]);
    density_vjp_kernel(&g, pbf_kernel, topology, positions, masses, rest_densities, density_adjoint, position_adjoint,
                       mass_adjoint, rest_adjoint);
    grid_free(&g);
}

void SUF(oracle_non_pressure)(const oracle_domain* d, uint64_t grid_step, const real* positions[3],
                              const real* velocities[3], const real* controls[3], const real* masses,
                              const real* viscosities, const real* densities, real* out[3]) {
    grid_t g;
    grid_build(&g, d, grid_step, positions[0], positions[1], positions[2]);
    non_pressure_kernel(&g, positions, velocities, controls, masses, viscosities, densities, out);
    grid_free(&g);
}

void SUF(oracle_pressure)(const oracle_domain* d, uint64_t grid_step, const real* positions[3], const real* masses,
                          const real* rest_densities, const real* densities, const real* pressures, real* out[3]) {
    grid_t g;
    grid_build(&g, d, grid_step, positions[0], positions[1], positions[2]);
    pressure_kernel(&g, positions, masses, rest_densities, densities, pressures, out);
    grid_free(&g);
}

/* SPH<DFSPH>::forward_step (sph.ixx:350-354) = build(step, x) -> density(x, x)
 * -> DFSPH::forward (sph.cpp:572-610), fixed iteration counts. */
void SUF(oracle_dfsph_step)(const oracle_domain* d, uint32_t nd, uint32_t ne, int warm_start, uint64_t step,
                            const real* pos[3], const real* vel[3], const real* ctl[3], const real* masses,
                            const real* rest_densities, const real* viscosities, const real* div_relax,
                            const real* den_relax, const real* warm_div, const real* warm_den, real* next_pos[3],
                            real* next_vel[3], real* next_warm_div, real* next_warm_den, real* densities_out,
                            real* non_pressure_out[3]) {
    const uint32_t n = d->particle_count;
    const real dt = (real)d->time_step;
    const real G = SUF(oracle_reference_gradient_norm)(d);
    const cbox box = collision_box(d, step);
    grid_t g;
    grid_build(&g, d, step, pos[0], pos[1], pos[2]);
    real* rho = alloc_real(n);
    real* rho_star = alloc_real(n);
    real* p = alloc_real(n);
    vfield anp = alloc_vfield(n), ap = alloc_vfield(n), adiv = alloc_vfield(n), total = alloc_vfield(n);
    vfield xs = alloc_vfield(n), vs = alloc_vfield(n);
    density_kernel(&g, 0, pos, pos, masses, rest_densities, rho);
    non_pressure_kernel(&g, pos, vel, ctl, masses, viscosities, rho, anp.c);
    /* divergence phase, sph.cpp:574-589 */
    for (uint32_t i = 0; i < n; ++i) p[i] = warm_start ? warm_div[i] : K(0.0);
    pressure_kernel(&g, pos, masses, rest_densities, rho, p, ap.c);
    for (uint32_t it = 1; it <= nd; ++it) {
        sph_predict_kernel(n, dt, box, pos, vel, CV(anp), CV(ap), xs.c, vs.c);
        density_kernel(&g, 0, pos, CV(xs), masses, rest_densities, rho_star);
        projection_update_kernel(n, dt, G, masses, rho, rho, rho_star, div_relax, p);
        pressure_kernel(&g, pos, masses, rest_densities, rho, p, ap.c);
    }
    for (int c = 0; c < 3; ++c) memcpy(adiv.c[c], ap.c[c], sizeof(real) * n);
    memcpy(next_warm_div, p, sizeof(real) * n);
    /* density phase, sph.cpp:591-607 */
    for (uint32_t i = 0; i < n; ++i) p[i] = warm_start ? warm_den[i] : K(0.0);
    pressure_kernel(&g, pos, masses, rest_densities, rho, p, ap.c);
    for (uint32_t it = 1; it <= ne; ++it) {
        add_kernel(n, CV(adiv), CV(ap), total.c);
        sph_predict_kernel(n, dt, box, pos, vel, CV(anp), CV(total), xs.c, vs.c);
        density_kernel(&g, 0, pos, CV(xs), masses, rest_densities, rho_star);
        projection_update_kernel(n, dt, G, masses, rho, rest_densities, rho_star, den_relax, p);
        pressure_kernel(&g, pos, masses, rest_densities, rho, p, ap.c);
    }
    add_kernel(n, CV(adiv), CV(ap), total.c);
    memcpy(next_warm_den, p, sizeof(real) * n);
    sph_predict_kernel(n, dt, box, pos, vel, CV(anp), CV(total), next_pos, next_vel);
    if (densities_out) memcpy(densities_out, rho, sizeof(real) * n);
    if (non_pressure_out)
        for (int c = 0; c < 3; ++c)
            if (non_pressure_out[c]) memcpy(non_pressure_out[c], anp.c[c], sizeof(real) * n);
    free(rho);
    free(rho_star);
    free(p);
    free_vfield(&anp);
    free_vfield(&ap);
    free_vfield(&adiv);
    free_vfield(&total);
    free_vfield(&xs);
    free_vfield(&vs);
    grid_free(&g);
}

/* Tolerance mode (an extension: the reference's DFSPH runs fixed iteration
 * counts, sph.cpp:580, 597). Each phase iterates until the mean compression
 * error of its prediction, (1/N) sum_i max(0, rho*_i - target_i) / rho0_i
 * summed in FP64, is at most `tolerance`, or `max_iterations` is reached;
 * the update of the iteration that meets it is still applied. The error is
 * that of the pressures the projection starts from (the targets: rho in the
 * divergence phase, rho0 in the density phase), and the pressures clamp at
 * zero, so expansion is not counted. */
static double compression_error(uint32_t n, const real* predicted, const real* targets, const real* rest) {
    double sum = 0.0;
    for (uint32_t i = 0; i < n; ++i) sum += fmax(0.0, (double)predicted[i] - (double)targets[i]) / (double)rest[i];
    return n > 0 ? sum / n : 0.0;
}

/* SPH<DFSPH>::forward_step in tolerance mode; iterations_out receives the
 * executed (divergence, density) counts. The result equals oracle_dfsph_step
 * with those counts. */
void SUF(oracle_dfsph_step_tolerance)(const oracle_domain* d, double tolerance, uint32_t max_iterations,
                                      int warm_start, uint64_t step, const real* pos[3], const real* vel[3],
                                      const real* ctl[3], const real* masses, const real* rest_densities,
                                      const real* viscosities, const real* div_relax, const real* den_relax,
                                      const real* warm_div, const real* warm_den, real* next_pos[3],
                                      real* next_vel[3], real* next_warm_div, real* next_warm_den,
                                      uint32_t iterations_out[2]) {
    const uint32_t n = d->particle_count;
    const real dt = (real)d->time_step;
    const real G = SUF(oracle_reference_gradient_norm)(d);
    const cbox box = collision_box(d, step);
    grid_t g;
    grid_build(&g, d, step, pos[0], pos[1], pos[2]);
    real* rho = alloc_real(n);
    real* rho_star = alloc_real(n);
    real* p = alloc_real(n);
    vfield anp = alloc_vfield(n), ap = alloc_vfield(n), adiv = alloc_vfield(n), total = alloc_vfield(n);
    vfield xs = alloc_vfield(n), vs = alloc_vfield(n);
    density_kernel(&g, 0, pos, pos, masses, rest_densities, rho);
    non_pressure_kernel(&g, pos, vel, ctl, masses, viscosities, rho, anp.c);
    iterations_out[0] = iterations_out[1] = 0u;
    for (uint32_t i = 0; i < n; ++i) p[i] = warm_start ? warm_div[i] : K(0.0);
    pressure_kernel(&g, pos, masses, rest_densities, rho, p, ap.c);
    for (uint32_t it = 1; it <= max_iterations; ++it) {
        sph_predict_kernel(n, dt, box, pos, vel, CV(anp), CV(ap), xs.c, vs.c);
        density_kernel(&g, 0, pos, CV(xs), masses, rest_densities, rho_star);
        const double error = compression_error(n, rho_star, rho, rest_densities);
        projection_update_kernel(n, dt, G, masses, rho, rho, rho_star, div_relax, p);
        pressure_kernel(&g, pos, masses, rest_densities, rho, p, ap.c);
        iterations_out[0] = it;
        if (error <= tolerance) break;
    }
    for (int c = 0; c < 3; ++c) memcpy(adiv.c[c], ap.c[c], sizeof(real) * n);
    memcpy(next_warm_div, p, sizeof(real) * n);
    for (uint32_t i = 0; i < n; ++i) p[i] = warm_start ? warm_den[i] : K(0.0);
    pressure_kernel(&g, pos, masses, rest_densities, rho, p, ap.c);
    for (uint32_t it = 1; it <= max_iterations; ++it) {
        add_kernel(n, CV(adiv), CV(ap), total.c);
        sph_predict_kernel(n, dt, box, pos, vel, CV(anp), CV(total), xs.c, vs.c);
        density_kernel(&g, 0, pos, CV(xs), masses, rest_densities, rho_star);
        const double error = compression_error(n, rho_star, rest_densities, rest_densities);
        projection_update_kernel(n, dt, G, masses, rho, rest_densities, rho_star, den_relax, p);
        pressure_kernel(&g, pos, masses, rest_densities, rho, p, ap.c);
        iterations_out[1] = it;
        if (error <= tolerance) break;
    }
    add_kernel(n, CV(adiv), CV(ap), total.c);
    memcpy(next_warm_den, p, sizeof(real) * n);
    sph_predict_kernel(n, dt, box, pos, vel, CV(anp), CV(total), next_pos, next_vel);
    free(rho);
    free(rho_star);
    free(p);
    free_vfield(&anp);
    free_vfield(&ap);
    free_vfield(&adiv);
    free_vfield(&total);
    free_vfield(&xs);
    free_vfield(&vs);
    grid_free(&g);
}

/* -------------------------------------------------------- PCISPH, IISPH */
/* SPH<PCISPH|IISPH>::forward_step (sph.ixx:350-354) = build(step, x) ->
 * density(x, x) -> PCISPH::forward (sph.cpp:296-315) or IISPH::forward
 * (sph.cpp:409-428). The pressure updates pressure_update_forward_kernel
 * (pcisph.cu:104-110, V = m / rho0) and jacobi_update_forward_kernel
 * (iisph.cu:104-110, V = m / rho) share projection_update_kernel with the
 * volume density passed in; both target rho0. */
/* The loop of both: with tolerance < 0 exactly `iterations` iterations;
 * else tolerance mode (an extension, like oracle_dfsph_step_tolerance): stop
 * after the iteration whose prediction has a mean compression error
 * (1/N) sum_i max(0, rho*_i - rho0_i) / rho0_i (FP64) within the tolerance,
 * at most `iterations`; that iteration's update and pressure force still
 * apply. *executed receives the iterations run. */
static void pressure_solver(const oracle_domain* d, int iisph, uint32_t iterations, double tolerance, uint64_t step,
                            const real* pos[3], const real* vel[3], const real* ctl[3], const real* masses,
                            const real* rest_densities, const real* viscosities, const real* relaxation,
                            real* next_pos[3], real* next_vel[3], real* densities_out, real* non_pressure_out[3],
                            real* pressures_out, uint32_t* executed) {
    const uint32_t n = d->particle_count;
    const real dt = (real)d->time_step;
    const real G = SUF(oracle_reference_gradient_norm)(d);
    const cbox box = collision_box(d, step);
    grid_t g;
    grid_build(&g, d, step, pos[0], pos[1], pos[2]);
    real* rho = alloc_real(n);
    real* rho_star = alloc_real(n);
    real* p = alloc_real(n);
    vfield anp = alloc_vfield(n), ap = alloc_vfield(n), xs = alloc_vfield(n), vs = alloc_vfield(n);
    density_kernel(&g, 0, pos, pos, masses, rest_densities, rho);
    non_pressure_kernel(&g, pos, vel, ctl, masses, viscosities, rho, anp.c);
    /* clear_iteration (sph.cpp:298): p = 0, a_p = 0 */
    for (uint32_t i = 0; i < n; ++i) {
        p[i] = K(0.0);
        for (int c = 0; c < 3; ++c) ap.c[c][i] = K(0.0);
    }
    uint32_t run = 0;
    for (uint32_t it = 1; it <= iterations; ++it) {
        sph_predict_kernel(n, dt, box, pos, vel, CV(anp), CV(ap), xs.c, vs.c);
        density_kernel(&g, 0, pos, CV(xs), masses, rest_densities, rho_star);
        const double error = tolerance >= 0.0 ? compression_error(n, rho_star, rest_densities, rest_densities) : 0.0;
        projection_update_kernel(n, dt, G, masses, iisph ? rho : rest_densities, rest_densities, rho_star, relaxation,
                                 p);
        pressure_kernel(&g, pos, masses, rest_densities, rho, p, ap.c);
        run = it;
        if (tolerance >= 0.0 && error <= tolerance) break;
    }
    if (executed) *executed = run;
    sph_predict_kernel(n, dt, box, pos, vel, CV(anp), CV(ap), next_pos, next_vel);
    if (densities_out) memcpy(densities_out, rho, sizeof(real) * n);
    if (pressures_out) memcpy(pressures_out, p, sizeof(real) * n);
    if (non_pressure_out)
        for (int c = 0; c < 3; ++c)
            if (non_pressure_out[c]) memcpy(non_pressure_out[c], anp.c[c], sizeof(real) * n);
    free(rho);
    free(rho_star);
    free(p);
    free_vfield(&anp);
    free_vfield(&ap);
    free_vfield(&xs);
    free_vfield(&vs);
    grid_free(&g);
}

void SUF(oracle_pressure_solver_step)(const oracle_domain* d, int iisph, uint32_t iterations, uint64_t step,
                                      const real* pos[3], const real* vel[3], const real* ctl[3], const real* masses,
                                      const real* rest_densities, const real* viscosities, const real* relaxation,
                                      real* next_pos[3], real* next_vel[3], real* densities_out,
                                      real* non_pressure_out[3], real* pressures_out) {
    pressure_solver(d, iisph, iterations, -1.0, step, pos, vel, ctl, masses, rest_densities, viscosities, relaxation,
                    next_pos, next_vel, densities_out, non_pressure_out, pressures_out, NULL);
}

void SUF(oracle_pressure_solver_step_tolerance)(const oracle_domain* d, int iisph, double tolerance,
                                                uint32_t max_iterations, uint64_t step, const real* pos[3],
                                                const real* vel[3], const real* ctl[3], const real* masses,
                                                const real* rest_densities, const real* viscosities,
                                                const real* relaxation, real* next_pos[3], real* next_vel[3],
                                                real* densities_out, real* pressures_out, uint32_t* iterations_out) {
    pressure_solver(d, iisph, max_iterations, tolerance, step, pos, vel, ctl, masses, rest_densities, viscosities,
                    relaxation, next_pos, next_vel, densities_out, NULL, pressures_out, iterations_out);
}

/* ---------------------------------------------------------------- WCSPH */
/* eos_forward_kernel, wcsph.cu:92-100 (Tait equation of state) */
static void eos_kernel(uint32_t n, const real* densities, const real* rest_densities, const real* speed,
                       const real* exponent, real* pressures) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (int64_t)n; ++i) {
        const real rho0 = rest_densities[i];
        const real c = speed[i];
        const real gamma = exponent[i];
        const real ratio = densities[i] / rho0;
        pressures[i] = rho0 * c * c * (R_POW(ratio, gamma) - K(1.0)) / gamma;
    }
}

/* viscosity_forward_kernel, wcsph.cu:136-174 (artificial viscosity on
 * approaching pairs) */
static void artificial_viscosity_kernel(const grid_t* g, const real* const pos[3], const real* const vel[3],
                                        const real* masses, const real* rest_densities, const real* viscosities,
                                        const real* speed, const real* densities, real* const out[3]) {
    const real h = (real)g->d->support_radius;
#pragma omp parallel for schedule(static)
    for (int64_t ii = 0; ii < (int64_t)g->n; ++ii) {
        const uint32_t i = (uint32_t)ii;
        const v3 p = v3_load(pos, i);
        const v3 v = v3_load(vel, i);
        v3 acc = v3_make(0, 0, 0);
        cell_range r[27];
        const int nr = stencil(g, p, r);
        for (int k = 0; k < nr; ++k) {
            for (uint32_t s = r[k].first; s < r[k].last; ++s) {
                const uint32_t j = g->idx[s];
                if (j == i) continue;
                const v3 d = v3_sub(p, v3_load(pos, j));
                const v3 dv = v3_sub(v, v3_load(vel, j));
                const real compression = v3_dot(dv, d);
                if (compression >= K(0.0)) continue;
                const real alpha = K(0.5) * (viscosities[i] + viscosities[j]);
                const real c = K(0.5) * (speed[i] + speed[j]);
                const real nu = K(2.0) * alpha * h * c / (densities[i] + densities[j]);
                const real pi = -nu * compression / (v3_dot(d, d) + K(0.01) * h * h);
                acc = v3_add(acc, v3_scale(cubic_gradient(d, h), -masses[j] * pi));
            }
            for (uint32_t s = r[k].bfirst; s < r[k].blast; ++s) {
                const uint32_t b = g->bidx[s];
                const v3 d = v3_sub(p, boundary_position(g, b));
                const v3 dv = v3_sub(v, boundary_velocity(g, b));
                const real compression = v3_dot(dv, d);
                if (compression >= K(0.0)) continue;
                const real nu = viscosities[i] * h * speed[i] / densities[i];
                const real pi = -nu * compression / (v3_dot(d, d) + K(0.01) * h * h);
                acc = v3_add(acc, v3_scale(cubic_gradient(d, h), -rest_densities[i] * boundary_volume(g, b) * pi));
            }
        }
        v3_store(out, i, acc);
    }
}

/* surface_forward_kernel, wcsph.cu:306-339 (cohesion; diameter = 2 r_p,
 * launch_surface_forward wcsph.cu:466-468) */
static void surface_kernel(const grid_t* g, const real* const pos[3], const real* masses, const real* rest_densities,
                           const real* tensions, const real* boundary_tension, real* const out[3]) {
    const real h = (real)g->d->support_radius;
    const real diameter = K(2.0) * (real)g->d->particle_radius;
#pragma omp parallel for schedule(static)
    for (int64_t ii = 0; ii < (int64_t)g->n; ++ii) {
        const uint32_t i = (uint32_t)ii;
        const v3 p = v3_load(pos, i);
        v3 acc = v3_make(0, 0, 0);
        cell_range r[27];
        const int nr = stencil(g, p, r);
        for (int k = 0; k < nr; ++k) {
            for (uint32_t s = r[k].first; s < r[k].last; ++s) {
                const uint32_t j = g->idx[s];
                if (j == i) continue;
                const v3 d = v3_sub(p, v3_load(pos, j));
                const real dist = v3_length(d);
                if (dist >= h) continue;
                const real w = dist > diameter ? cubic(d, h) : cubic(v3_make(diameter, K(0.0), K(0.0)), h);
                const real factor = -tensions[i] * masses[j] * w / masses[i];
                acc = v3_add(acc, v3_scale(d, factor));
            }
            for (uint32_t s = r[k].bfirst; s < r[k].blast; ++s) {
                const uint32_t b = g->bidx[s];
                const v3 d = v3_sub(p, boundary_position(g, b));
                const real dist = v3_length(d);
                if (dist >= h) continue;
                const real w = dist > diameter ? cubic(d, h) : cubic(v3_make(diameter, K(0.0), K(0.0)), h);
                const real factor = -boundary_tension[i] * rest_densities[i] * boundary_volume(g, b) * w / masses[i];
                acc = v3_add(acc, v3_scale(d, factor));
            }
        }
        v3_store(out, i, acc);
    }
}

/* SPH<WCSPH>::forward_step (sph.ixx:350-354) = build(step, x) -> density(x, x)
 * -> WCSPH::forward (sph.cpp:214-226): EOS, pressure, artificial viscosity,
 * surface, a = ((a_p + a_visc) + a_surf) + (g + u), integrate. */
void SUF(oracle_wcsph_step)(const oracle_domain* d, uint64_t step, const real* pos[3], const real* vel[3],
                            const real* ctl[3], const real* masses, const real* rest_densities,
                            const real* viscosities, const real* tensions, const real* speed, const real* exponent,
                            const real* boundary_tension, real* next_pos[3], real* next_vel[3], real* densities_out,
                            real* pressures_out) {
    const uint32_t n = d->particle_count;
    const real dt = (real)d->time_step;
    const cbox box = collision_box(d, step);
    grid_t g;
    grid_build(&g, d, step, pos[0], pos[1], pos[2]);
    real* rho = alloc_real(n);
    real* p = alloc_real(n);
    vfield ap = alloc_vfield(n), av = alloc_vfield(n), as = alloc_vfield(n), ext = alloc_vfield(n),
           total = alloc_vfield(n);
    density_kernel(&g, 0, pos, pos, masses, rest_densities, rho);
    eos_kernel(n, rho, rest_densities, speed, exponent, p);
    pressure_kernel(&g, pos, masses, rest_densities, rho, p, ap.c);
    artificial_viscosity_kernel(&g, pos, vel, masses, rest_densities, viscosities, speed, rho, av.c);
    surface_kernel(&g, pos, masses, rest_densities, tensions, boundary_tension, as.c);
    add_kernel(n, CV(ap), CV(av), total.c);
    add_kernel(n, CV(total), CV(as), total.c);
    /* external_forward_kernel, wcsph.cu:86-90 */
    const v3 gravity = v3_make((real)d->gravity[0], (real)d->gravity[1], (real)d->gravity[2]);
    for (uint32_t i = 0; i < n; ++i) v3_store(ext.c, i, v3_add(gravity, v3_load(ctl, i)));
    add_kernel(n, CV(total), CV(ext), total.c);
    sph_predict_kernel(n, dt, box, pos, vel, CV(total), NULL, next_pos, next_vel);
    if (densities_out) memcpy(densities_out, rho, sizeof(real) * n);
    if (pressures_out) memcpy(pressures_out, p, sizeof(real) * n);
    free(rho);
    free(p);
    free_vfield(&ap);
    free_vfield(&av);
    free_vfield(&as);
    free_vfield(&ext);
    free_vfield(&total);
    grid_free(&g);
}

/* ------------------------------------------------------------------ PBF */
/* lambda_forward_kernel, pbf.cu:95-127 */
static void lambda_kernel(const grid_t* g, const real* const topo[3], const real* const pos[3], const real* masses,
                          const real* rest_densities, const real* densities, const real* relaxation,
                          real* const gsum[3], real* denominators, real* lambdas) {
    const real h = (real)g->d->support_radius;
    const real sc = spiky_coefficient(h);
#pragma omp parallel for schedule(static)
    for (int64_t ii = 0; ii < (int64_t)g->n; ++ii) {
        const uint32_t i = (uint32_t)ii;
        const v3 p = v3_load(pos, i);
        const real rho0 = rest_densities[i];
        v3 sg = v3_make(0, 0, 0);
        real den = relaxation[i];
        cell_range r[27];
        const int nr = stencil(g, v3_load(topo, i), r);
        for (int k = 0; k < nr; ++k) {
            for (uint32_t s = r[k].first; s < r[k].last; ++s) {
                const uint32_t j = g->idx[s];
                if (j == i) continue;
                const v3 gr = spiky_gradient(v3_sub(p, v3_load(pos, j)), h, sc);
                const real c = masses[j] / rho0;
                sg = v3_add(sg, v3_scale(gr, c));
                den += c * c * v3_dot(gr, gr);
            }
            for (uint32_t s = r[k].bfirst; s < r[k].blast; ++s) {
                const uint32_t b = g->bidx[s];
                sg = v3_add(sg, v3_scale(spiky_gradient(v3_sub(p, boundary_position(g, b)), h, sc), boundary_volume(g, b)));
            }
        }
        den += v3_dot(sg, sg);
        if (gsum) v3_store(gsum, i, sg);
        if (denominators) denominators[i] = den;
        lambdas[i] = -(densities[i] / rho0 - K(1.0)) / den;
    }
}

/* correction_forward_kernel, pbf.cu:254-288 */
static void correction_kernel(const grid_t* g, const real* const topo[3], const real* const pos[3], const real* masses,
                              const real* rest_densities, const real* lambdas, const real* strength,
                              const real* exponent, const real* radius, real* const out[3]) {
    const real h = (real)g->d->support_radius;
    const real sc = spiky_coefficient(h);
#pragma omp parallel for schedule(static)
    for (int64_t ii = 0; ii < (int64_t)g->n; ++ii) {
        const uint32_t i = (uint32_t)ii;
        const v3 p = v3_load(pos, i);
        const real rho0 = rest_densities[i];
        v3 corr = v3_make(0, 0, 0);
        cell_range r[27];
        const int nr = stencil(g, v3_load(topo, i), r);
        for (int k = 0; k < nr; ++k) {
            for (uint32_t s = r[k].first; s < r[k].last; ++s) {
                const uint32_t j = g->idx[s];
                if (j == i) continue;
                const v3 d = v3_sub(p, v3_load(pos, j));
                const real art = artificial_value(d, h, strength[i], exponent[i], radius[i]);
                const real c = masses[j] / rho0;
                corr = v3_add(corr, v3_scale(spiky_gradient(d, h, sc), c * (lambdas[i] + lambdas[j] + art)));
            }
            for (uint32_t s = r[k].bfirst; s < r[k].blast; ++s) {
                const uint32_t b = g->bidx[s];
                const v3 d = v3_sub(p, boundary_position(g, b));
                const real art = artificial_value(d, h, strength[i], exponent[i], radius[i]);
                corr = v3_add(corr, v3_scale(spiky_gradient(d, h, sc), boundary_volume(g, b) * (lambdas[i] + art)));
            }
        }
        v3_store(out, i, corr);
    }
}

/* vorticity_forward_kernel, pbf.cu:476-497 */
static void vorticity_kernel(const grid_t* g, const real* const topo[3], const real* const pos[3],
                             const real* const vel[3], real* const out[3]) {
    const real h = (real)g->d->support_radius;
    const real sc = spiky_coefficient(h);
#pragma omp parallel for schedule(static)
    for (int64_t ii = 0; ii < (int64_t)g->n; ++ii) {
        const uint32_t i = (uint32_t)ii;
        const v3 p = v3_load(pos, i);
        const v3 v = v3_load(vel, i);
        v3 w = v3_make(0, 0, 0);
        cell_range r[27];
        const int nr = stencil(g, v3_load(topo, i), r);
        for (int k = 0; k < nr; ++k)
            for (uint32_t s = r[k].first; s < r[k].last; ++s) {
                const uint32_t j = g->idx[s];
                if (j == i) continue;
                w = v3_add(w, v3_cross(v3_sub(v3_load(vel, j), v), spiky_gradient(v3_sub(p, v3_load(pos, j)), h, sc)));
            }
        v3_store(out, i, w);
    }
}

/* magnitude_forward_kernel + normal_forward_kernel, pbf.cu:565-593 */
static void normal_kernel(const grid_t* g, const real* const topo[3], const real* const pos[3],
                          const real* magnitudes, real* const normals[3], real* normalizers) {
    const real h = (real)g->d->support_radius;
    const real sc = spiky_coefficient(h);
#pragma omp parallel for schedule(static)
    for (int64_t ii = 0; ii < (int64_t)g->n; ++ii) {
        const uint32_t i = (uint32_t)ii;
        const v3 p = v3_load(pos, i);
        v3 gr = v3_make(0, 0, 0);
        cell_range r[27];
        const int nr = stencil(g, v3_load(topo, i), r);
        for (int k = 0; k < nr; ++k)
            for (uint32_t s = r[k].first; s < r[k].last; ++s) {
                const uint32_t j = g->idx[s];
                if (j == i) continue;
                gr = v3_add(gr, v3_scale(spiky_gradient(v3_sub(p, v3_load(pos, j)), h, sc), magnitudes[j] - magnitudes[i]));
            }
        const real nrm = v3_length(gr);
        normalizers[i] = nrm;
        v3_store(normals, i, nrm > K(0.0) ? v3_scale(gr, K(1.0) / nrm) : v3_make(0, 0, 0));
    }
}

/* xsph_forward_kernel, pbf.cu:700-721 */
static void xsph_kernel(const grid_t* g, const real* const topo[3], const real* const pos[3], const real* const vel[3],
                        const real* viscosity, real* const out[3]) {
    const real h = (real)g->d->support_radius;
#pragma omp parallel for schedule(static)
    for (int64_t ii = 0; ii < (int64_t)g->n; ++ii) {
        const uint32_t i = (uint32_t)ii;
        const v3 p = v3_load(pos, i);
        const v3 v = v3_load(vel, i);
        v3 sm = v3_make(0, 0, 0);
        cell_range r[27];
        const int nr = stencil(g, v3_load(topo, i), r);
        for (int k = 0; k < nr; ++k)
            for (uint32_t s = r[k].first; s < r[k].last; ++s) {
                const uint32_t j = g->idx[s];
                if (j == i) continue;
                sm = v3_add(sm, v3_scale(v3_sub(v3_load(vel, j), v), poly6(v3_sub(p, v3_load(pos, j)), h)));
            }
        v3_store(out, i, v3_add(v, v3_scale(sm, viscosity[i])));
    }
}

void SUF(oracle_pbf_lambda)(const oracle_domain* d, uint64_t grid_step, const real* topology[3],
                            const real* positions[3], const real* masses, const real* rest_densities,
                            const real* densities, const real* relaxation, real* gradient_sums[3],
                            real* denominators, real* lambdas) {
    grid_t g;
    grid_build(&g, d, grid_step, topology[0], topology[1], topology[2]);
    lambda_kernel(&g, topology, positions, masses, rest_densities, densities, relaxation, gradient_sums, denominators,
                  lambdas);
    grid_free(&g);
}

/* PBF::forward_step, pbf.cpp:167-191 */
void SUF(oracle_pbf_step)(const oracle_domain* d, uint32_t iterations, uint64_t step, const real* pos[3],
                          const real* vel[3], const real* ctl[3], const real* masses, const real* rest_densities,
                          const real* relaxation, const real* strength, const real* exponent, const real* radius,
                          const real* xsph, const real* confinement, real* next_pos[3], real* next_vel[3],
                          real* densities_out, real* lambdas_out, real* predicted_out[3],
                          real* vorticity_magnitudes_out) {
    const uint32_t n = d->particle_count;
    const real dt = (real)d->time_step;
    const v3 gravity = v3_make((real)d->gravity[0], (real)d->gravity[1], (real)d->gravity[2]);
    vfield xs = alloc_vfield(n), pc = alloc_vfield(n), corr = alloc_vfield(n), v = alloc_vfield(n);
    vfield w = alloc_vfield(n), nrm = alloc_vfield(n), vc = alloc_vfield(n);
    real* rho = alloc_real(n);
    real* lam = alloc_real(n);
    real* mag = alloc_real(n);
    real* nz = alloc_real(n);
    /* predict_forward_kernel, pbf.cu:73-78 */
#pragma omp parallel for schedule(static)
    for (int64_t ii = 0; ii < (int64_t)n; ++ii) {
        const uint32_t i = (uint32_t)ii;
        const v3 pv = v3_add(v3_load(vel, i), v3_scale(v3_add(gravity, v3_load(ctl, i)), dt));
        v3_store(xs.c, i, v3_add(v3_load(pos, i), v3_scale(pv, dt)));
    }
    grid_t g;
    grid_build(&g, d, step + 1u, xs.c[0], xs.c[1], xs.c[2]);
    for (int c = 0; c < 3; ++c) memcpy(pc.c[c], xs.c[c], sizeof(real) * n);
    const cbox box = collision_box(d, step + 1u);
    for (uint32_t it = 0; it < iterations; ++it) {
        density_kernel(&g, 1, CV(xs), CV(pc), masses, rest_densities, rho);
        lambda_kernel(&g, CV(xs), CV(pc), masses, rest_densities, rho, relaxation, NULL, NULL, lam);
        correction_kernel(&g, CV(xs), CV(pc), masses, rest_densities, lam, strength, exponent, radius, corr.c);
        /* project_forward_kernel, pbf.cu:417-431 */
#pragma omp parallel for schedule(static)
        for (int64_t ii = 0; ii < (int64_t)n; ++ii) {
            const uint32_t i = (uint32_t)ii;
            const v3 cand = v3_add(v3_load(CV(pc), i), v3_load(CV(corr), i));
            v3_store(pc.c, i,
                     v3_make(R_FMIN(box.hi[0], R_FMAX(box.lo[0], cand.x)), R_FMIN(box.hi[1], R_FMAX(box.lo[1], cand.y)),
                             R_FMIN(box.hi[2], R_FMAX(box.lo[2], cand.z))));
        }
    }
    /* reconstruct_forward_kernel, pbf.cu:456-460 */
    const real inv_dt = K(1.0) / dt;
#pragma omp parallel for schedule(static)
    for (int64_t ii = 0; ii < (int64_t)n; ++ii) {
        const uint32_t i = (uint32_t)ii;
        v3_store(v.c, i, v3_scale(v3_sub(v3_load(CV(pc), i), v3_load(pos, i)), inv_dt));
    }
    vorticity_kernel(&g, CV(xs), CV(pc), CV(v), w.c);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (int64_t)n; ++i) mag[i] = v3_length(v3_load(CV(w), (uint32_t)i));
    normal_kernel(&g, CV(xs), CV(pc), mag, nrm.c, nz);
    /* confinement_forward_kernel, pbf.cu:674-678 */
#pragma omp parallel for schedule(static)
    for (int64_t ii = 0; ii < (int64_t)n; ++ii) {
        const uint32_t i = (uint32_t)ii;
        v3_store(vc.c, i,
                 v3_add(v3_load(CV(v), i),
                        v3_scale(v3_cross(v3_load(CV(nrm), i), v3_load(CV(w), i)), dt * confinement[i])));
    }
    xsph_kernel(&g, CV(xs), CV(pc), CV(vc), xsph, next_vel);
    for (int c = 0; c < 3; ++c) memcpy(next_pos[c], pc.c[c], sizeof(real) * n);
    if (densities_out) memcpy(densities_out, rho, sizeof(real) * n);
    if (lambdas_out) memcpy(lambdas_out, lam, sizeof(real) * n);
    if (predicted_out)
        for (int c = 0; c < 3; ++c)
            if (predicted_out[c]) memcpy(predicted_out[c], xs.c[c], sizeof(real) * n);
    if (vorticity_magnitudes_out) memcpy(vorticity_magnitudes_out, mag, sizeof(real) * n);
    free_vfield(&xs);
    free_vfield(&pc);
    free_vfield(&corr);
    free_vfield(&v);
    free_vfield(&w);
    free_vfield(&nrm);
    free_vfield(&vc);
    free(rho);
    free(lam);
    free(mag);
    free(nz);
    grid_free(&g);
}

/* ======================================================= PBF derivatives */
/* Forward-mode (JVP) and reverse-mode (VJP) derivatives of PBF::forward_step
 * with the neighbour graph, the box-collision masks and every branch frozen
 * at the primal step: PBF::jvp_step (pbf.cpp:193-215) and PBF::vjp_step
 * (pbf.cpp:217-258) with the kernels of pbf.cu and density.cu. The JVP runs
 * in the primal precision; the VJP promotes the primal values and works in
 * FP64, and every adjoint output is accumulated (+=). Pair terms are
 * differentiated in gather form: particle a collects the terms in which it
 * is the centre and those in which it is a neighbour, so no output is
 * written by two particles. Boundary particles are not supported here. */

typedef struct {
    double x, y, z;
} d3;
static inline d3 d3_make(double x, double y, double z) {
    d3 r = {x, y, z};
    return r;
}
static inline d3 d3_of(v3 a) { return d3_make((double)a.x, (double)a.y, (double)a.z); }
static inline d3 d3_add(d3 a, d3 b) { return d3_make(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline d3 d3_sub(d3 a, d3 b) { return d3_make(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline d3 d3_scale(d3 v, double f) { return d3_make(f * v.x, f * v.y, f * v.z); }
static inline double d3_dot(d3 a, d3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline d3 d3_cross(d3 a, d3 b) {
    return d3_make(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
static inline d3 d3_load(double* const f[3], uint32_t i) { return d3_make(f[0][i], f[1][i], f[2][i]); }
static inline d3 d3_cload(const double* const f[3], uint32_t i) { return d3_make(f[0][i], f[1][i], f[2][i]); }
static inline void d3_acc(double* const f[3], uint32_t i, d3 v) {
    f[0][i] += v.x;
    f[1][i] += v.y;
    f[2][i] += v.z;
}

/* Derivative of the spiky gradient G(d) = f(r) d, f = k (h - r)^2 / r, along
 * w: H(d) w = f w + (f'(r) / r) (d . w) d with f' = -k (h - r)(h + r) / r^2
 * (zero where G is zero: r = 0 or r >= h). H is symmetric and even in d. */
static d3 spiky_hessian_d(v3 d, d3 w, double h, double k) {
    const d3 dd = d3_of(d);
    const double r = sqrt(d3_dot(dd, dd));
    if (r == 0.0 || r >= h) return d3_make(0, 0, 0);
    const double f = k * (h - r) * (h - r) / r;
    const double fp = -k * (h - r) * (h + r) / (r * r);
    return d3_add(d3_scale(w, f), d3_scale(dd, fp / r * d3_dot(dd, w)));
}

/* The same product in the primal precision (the JVP's G-dot). */
static v3 spiky_hessian_r(v3 d, v3 w, real h, real k) {
    const real r = v3_length(d);
    if (r == K(0.0) || r >= h) return v3_make(0, 0, 0);
    const real f = k * (h - r) * (h - r) / r;
    const real fp = -k * (h - r) * (h + r) / (r * r);
    return v3_add(v3_scale(w, f), v3_scale(d, fp / r * v3_dot(d, w)));
}

/* Artificial pressure s(d) = -k (W(d) / W(dq))^n (pbf.cu:56-71) with its
 * gradient in d and its partial derivatives in k, n and dq. */
typedef struct {
    real value;
    v3 gradient;
    real d_strength, d_exponent, d_radius;
} artificial_t;

static artificial_t artificial_full(v3 d, real h, real strength, real exponent, real radius) {
    artificial_t a;
    memset(&a, 0, sizeof(a));
    const real reference = poly6(v3_make(radius, K(0.0), K(0.0)), h);
    const real kernel = poly6(d, h);
    if (kernel == K(0.0)) return a;
    const real ratio = kernel / reference;
    const real power = R_POW(ratio, exponent);
    a.value = -strength * power;
    a.d_strength = -power;
    a.d_exponent = a.value * R_LOG(ratio);
    /* dW(dq)/d dq = -6 alpha (h^2 - dq^2)^2 dq, alpha = 315 / (64 pi h^9) */
    const real hr = h * h;
    const real diff = hr - radius * radius;
    const real radial = radius >= h ? K(0.0)
                                    : K(-6.0) * K(315.0) / (K(64.0) * R_PI * hr * hr * hr * h * h * h) * radius * diff * diff;
    a.d_radius = -a.value * exponent * radial / reference;
    a.gradient = v3_scale(poly6_gradient(d, h), -strength * exponent * R_POW(ratio, exponent - K(1.0)) / reference);
    return a;
}

/* Every primal quantity of one PBF step that the derivatives read. */
typedef struct {
    uint32_t n, iterations;
    grid_t g;
    real h, spiky, dt, inv_dt;
    vfield xs;     /* predicted positions: topology and p[0] */
    vfield* p;     /* positions entering iteration k; p[iterations] is the result */
    real** rho;    /* per iteration */
    real** lam;
    real** den;
    vfield* gsum;
    uint8_t** mask; /* collision mask of the projection of iteration k */
    vfield vrec, w, nrm, vc;
    real *mag, *nz;
} pbf_record;

static void pbf_record_forward(pbf_record* r, const oracle_domain* d, uint32_t iterations, uint64_t step,
                               const real* pos[3], const real* vel[3], const real* ctl[3], const real* masses,
                               const real* rest_densities, const real* relaxation, const real* strength,
                               const real* exponent, const real* radius, const real* confinement) {
    const uint32_t n = d->particle_count;
    memset(r, 0, sizeof(*r));
    r->n = n;
    r->iterations = iterations;
    r->h = (real)d->support_radius;
    r->spiky = spiky_coefficient(r->h);
    r->dt = (real)d->time_step;
    r->inv_dt = K(1.0) / r->dt;
    const v3 gravity = v3_make((real)d->gravity[0], (real)d->gravity[1], (real)d->gravity[2]);
    r->xs = alloc_vfield(n);
    for (uint32_t i = 0; i < n; ++i) {
        const v3 pv = v3_add(v3_load(vel, i), v3_scale(v3_add(gravity, v3_load(ctl, i)), r->dt));
        v3_store(r->xs.c, i, v3_add(v3_load(pos, i), v3_scale(pv, r->dt)));
    }
    grid_build(&r->g, d, step + 1u, r->xs.c[0], r->xs.c[1], r->xs.c[2]);
    r->p = (vfield*)calloc(iterations + 1u, sizeof(vfield));
    r->rho = (real**)calloc(iterations + 1u, sizeof(real*));
    r->lam = (real**)calloc(iterations + 1u, sizeof(real*));
    r->den = (real**)calloc(iterations + 1u, sizeof(real*));
    r->gsum = (vfield*)calloc(iterations + 1u, sizeof(vfield));
    r->mask = (uint8_t**)calloc(iterations + 1u, sizeof(uint8_t*));
    r->p[0] = alloc_vfield(n);
    for (int c = 0; c < 3; ++c) memcpy(r->p[0].c[c], r->xs.c[c], sizeof(real) * n);
    const cbox box = collision_box(d, step + 1u);
    vfield corr = alloc_vfield(n);
    for (uint32_t it = 0; it < iterations; ++it) {
        r->rho[it] = alloc_real(n);
        r->lam[it] = alloc_real(n);
        r->den[it] = alloc_real(n);
        r->gsum[it] = alloc_vfield(n);
        r->mask[it] = (uint8_t*)malloc(n + 16u);
        r->p[it + 1] = alloc_vfield(n);
        density_kernel(&r->g, 1, CV(r->xs), CV(r->p[it]), masses, rest_densities, r->rho[it]);
        lambda_kernel(&r->g, CV(r->xs), CV(r->p[it]), masses, rest_densities, r->rho[it], relaxation, r->gsum[it].c,
                      r->den[it], r->lam[it]);
        correction_kernel(&r->g, CV(r->xs), CV(r->p[it]), masses, rest_densities, r->lam[it], strength, exponent, radius,
                          corr.c);
        for (uint32_t i = 0; i < n; ++i) {
            const v3 cand = v3_add(v3_load(CV(r->p[it]), i), v3_load(CV(corr), i));
            uint8_t m = 0;
            if (cand.x < box.lo[0] || cand.x > box.hi[0]) m |= 1u;
            if (cand.y < box.lo[1] || cand.y > box.hi[1]) m |= 2u;
            if (cand.z < box.lo[2] || cand.z > box.hi[2]) m |= 4u;
            r->mask[it][i] = m;
            v3_store(r->p[it + 1].c, i,
                     v3_make(R_FMIN(box.hi[0], R_FMAX(box.lo[0], cand.x)), R_FMIN(box.hi[1], R_FMAX(box.lo[1], cand.y)),
                             R_FMIN(box.hi[2], R_FMAX(box.lo[2], cand.z))));
        }
    }
    free_vfield(&corr);
    const vfield pc = r->p[iterations];
    r->vrec = alloc_vfield(n);
    r->w = alloc_vfield(n);
    r->nrm = alloc_vfield(n);
    r->vc = alloc_vfield(n);
    r->mag = alloc_real(n);
    r->nz = alloc_real(n);
    for (uint32_t i = 0; i < n; ++i) v3_store(r->vrec.c, i, v3_scale(v3_sub(v3_load(CV(pc), i), v3_load(pos, i)), r->inv_dt));
    vorticity_kernel(&r->g, CV(r->xs), CV(pc), CV(r->vrec), r->w.c);
    for (uint32_t i = 0; i < n; ++i) r->mag[i] = v3_length(v3_load(CV(r->w), i));
    normal_kernel(&r->g, CV(r->xs), CV(pc), r->mag, r->nrm.c, r->nz);
    for (uint32_t i = 0; i < n; ++i)
        v3_store(r->vc.c, i,
                 v3_add(v3_load(CV(r->vrec), i),
                        v3_scale(v3_cross(v3_load(CV(r->nrm), i), v3_load(CV(r->w), i)), r->dt * confinement[i])));
}

static void pbf_record_free(pbf_record* r) {
    for (uint32_t it = 0; it <= r->iterations; ++it) {
        free_vfield(&r->p[it]);
        if (it < r->iterations) {
            free(r->rho[it]);
            free(r->lam[it]);
            free(r->den[it]);
            free_vfield(&r->gsum[it]);
            free(r->mask[it]);
        }
    }
    free(r->p);
    free(r->rho);
    free(r->lam);
    free(r->den);
    free(r->gsum);
    free(r->mask);
    free_vfield(&r->xs);
    free_vfield(&r->vrec);
    free_vfield(&r->w);
    free_vfield(&r->nrm);
    free_vfield(&r->vc);
    free(r->mag);
    free(r->nz);
    grid_free(&r->g);
}

/* ---------------------------------------------------------------- PBF JVP */
void SUF(oracle_pbf_jvp)(const oracle_domain* d, uint32_t iterations, uint64_t step, const real* pos[3],
                         const real* vel[3], const real* ctl[3], const real* masses, const real* rest_densities,
                         const real* relaxation, const real* strength, const real* exponent, const real* radius,
                         const real* xsph, const real* confinement, const real* pos_t[3], const real* vel_t[3],
                         const real* ctl_t[3], const real* masses_t, const real* rest_densities_t,
                         const real* relaxation_t, const real* strength_t, const real* exponent_t,
                         const real* radius_t, const real* xsph_t, const real* confinement_t, real* next_pos_t[3],
                         real* next_vel_t[3]) {
    pbf_record r;
    pbf_record_forward(&r, d, iterations, step, pos, vel, ctl, masses, rest_densities, relaxation, strength, exponent,
                       radius, confinement);
    const uint32_t n = r.n;
    const real h = r.h, k = r.spiky, dt = r.dt;
    const grid_t* g = &r.g;
    vfield pt = alloc_vfield(n), pn = alloc_vfield(n);
    real* rho_t = alloc_real(n);
    real* lam_t = alloc_real(n);
    /* predict_jvp_kernel, pbf.cu:80-84 */
    for (uint32_t i = 0; i < n; ++i)
        v3_store(pt.c, i,
                 v3_add(v3_load(pos_t, i), v3_scale(v3_add(v3_load(vel_t, i), v3_scale(v3_load(ctl_t, i), dt)), dt)));
    for (uint32_t it = 0; it < iterations; ++it) {
        const real* const* p = CV(r.p[it]);
        /* density_jvp_kernel (density.cu:42-67, poly6) */
#pragma omp parallel for schedule(static)
        for (int64_t ii = 0; ii < (int64_t)n; ++ii) {
            const uint32_t i = (uint32_t)ii;
            const v3 pi = v3_load(p, i), ti = v3_load(CV(pt), i);
            cell_range cr[27];
            const int nr = stencil(g, v3_load(CV(r.xs), i), cr);
            real acc = K(0.0);
            for (int c = 0; c < nr; ++c) {
                for (uint32_t s = cr[c].first; s < cr[c].last; ++s) {
                    const uint32_t j = g->idx[s];
                    const v3 dd = v3_sub(pi, v3_load(p, j));
                    acc += masses_t[j] * poly6(dd, h) + masses[j] * v3_dot(poly6_gradient(dd, h), v3_sub(ti, v3_load(CV(pt), j)));
                }
                /* boundary samples (density.cu:59-63): rho0_i V_b W, the sample fixed */
                for (uint32_t s = cr[c].bfirst; s < cr[c].blast; ++s) {
                    const uint32_t b = g->bidx[s];
                    const v3 dd = v3_sub(pi, boundary_position(g, b));
                    acc += boundary_volume(g, b) *
                           (rest_densities_t[i] * poly6(dd, h) + rest_densities[i] * v3_dot(poly6_gradient(dd, h), ti));
                }
            }
            rho_t[i] = acc;
        }
        /* lambda_jvp_kernel (pbf.cu:129-176) */
#pragma omp parallel for schedule(static)
        for (int64_t ii = 0; ii < (int64_t)n; ++ii) {
            const uint32_t i = (uint32_t)ii;
            const v3 pi = v3_load(p, i), ti = v3_load(CV(pt), i);
            const real rho0 = rest_densities[i], rho0_t = rest_densities_t[i];
            v3 sg = v3_make(0, 0, 0), sg_t = v3_make(0, 0, 0);
            real den_t = relaxation_t[i];
            cell_range cr[27];
            const int nr = stencil(g, v3_load(CV(r.xs), i), cr);
            for (int c = 0; c < nr; ++c) {
                for (uint32_t s = cr[c].first; s < cr[c].last; ++s) {
                    const uint32_t j = g->idx[s];
                    if (j == i) continue;
                    const v3 dd = v3_sub(pi, v3_load(p, j));
                    const v3 gr = spiky_gradient(dd, h, k);
                    const v3 gr_t = spiky_hessian_r(dd, v3_sub(ti, v3_load(CV(pt), j)), h, k);
                    const real cf = masses[j] / rho0;
                    const real cf_t = masses_t[j] / rho0 - masses[j] * rho0_t / (rho0 * rho0);
                    const v3 ng = v3_scale(gr, cf);
                    const v3 ng_t = v3_add(v3_scale(gr, cf_t), v3_scale(gr_t, cf));
                    sg = v3_add(sg, ng);
                    sg_t = v3_add(sg_t, ng_t);
                    den_t += K(2.0) * v3_dot(ng, ng_t);
                }
                /* boundary samples (pbf.cu:164-168): V_b grad W in the gradient sum only */
                for (uint32_t s = cr[c].bfirst; s < cr[c].blast; ++s) {
                    const uint32_t b = g->bidx[s];
                    const v3 dd = v3_sub(pi, boundary_position(g, b));
                    const real vb = boundary_volume(g, b);
                    sg = v3_add(sg, v3_scale(spiky_gradient(dd, h, k), vb));
                    sg_t = v3_add(sg_t, v3_scale(spiky_hessian_r(dd, ti, h, k), vb));
                }
            }
            den_t += K(2.0) * v3_dot(sg, sg_t);
            const real den = r.den[it][i];
            const real con = r.rho[it][i] / rho0 - K(1.0);
            const real con_t = rho_t[i] / rho0 - r.rho[it][i] * rho0_t / (rho0 * rho0);
            lam_t[i] = -con_t / den + con * den_t / (den * den);
        }
        /* correction_jvp_kernel (pbf.cu:290-333) + project_jvp_kernel (pbf.cu:433-443) */
#pragma omp parallel for schedule(static)
        for (int64_t ii = 0; ii < (int64_t)n; ++ii) {
            const uint32_t i = (uint32_t)ii;
            const v3 pi = v3_load(p, i), ti = v3_load(CV(pt), i);
            const real rho0 = rest_densities[i], rho0_t = rest_densities_t[i];
            v3 res = v3_make(0, 0, 0);
            cell_range cr[27];
            const int nr = stencil(g, v3_load(CV(r.xs), i), cr);
            for (int c = 0; c < nr; ++c) {
                for (uint32_t s = cr[c].first; s < cr[c].last; ++s) {
                    const uint32_t j = g->idx[s];
                    if (j == i) continue;
                    const v3 dd = v3_sub(pi, v3_load(p, j));
                    const v3 dd_t = v3_sub(ti, v3_load(CV(pt), j));
                    const artificial_t a = artificial_full(dd, h, strength[i], exponent[i], radius[i]);
                    const real a_t = v3_dot(a.gradient, dd_t) + a.d_strength * strength_t[i] +
                                     a.d_exponent * exponent_t[i] + a.d_radius * radius_t[i];
                    const real cf = masses[j] / rho0;
                    const real cf_t = masses_t[j] / rho0 - masses[j] * rho0_t / (rho0 * rho0);
                    const real mu = r.lam[it][i] + r.lam[it][j] + a.value;
                    const real mu_t = lam_t[i] + lam_t[j] + a_t;
                    res = v3_add(res, v3_add(v3_scale(spiky_gradient(dd, h, k), cf_t * mu + cf * mu_t),
                                             v3_scale(spiky_hessian_r(dd, dd_t, h, k), cf * mu)));
                }
                /* boundary samples (pbf.cu:321-329): V_b (lambda_i + s_corr) grad W */
                for (uint32_t s = cr[c].bfirst; s < cr[c].blast; ++s) {
                    const uint32_t b = g->bidx[s];
                    const v3 dd = v3_sub(pi, boundary_position(g, b));
                    const artificial_t a = artificial_full(dd, h, strength[i], exponent[i], radius[i]);
                    const real a_t = v3_dot(a.gradient, ti) + a.d_strength * strength_t[i] +
                                     a.d_exponent * exponent_t[i] + a.d_radius * radius_t[i];
                    const real mu = r.lam[it][i] + a.value;
                    res = v3_add(res, v3_scale(v3_add(v3_scale(spiky_gradient(dd, h, k), lam_t[i] + a_t),
                                                      v3_scale(spiky_hessian_r(dd, ti, h, k), mu)),
                                               boundary_volume(g, b)));
                }
            }
            v3 out = v3_add(ti, res);
            const uint8_t m = r.mask[it][i];
            if (m & 1u) out.x = K(0.0);
            if (m & 2u) out.y = K(0.0);
            if (m & 4u) out.z = K(0.0);
            v3_store(pn.c, i, out);
        }
        vfield tmp = pt;
        pt = pn;
        pn = tmp;
    }
    const real* const* pc = CV(r.p[iterations]);
    vfield vt = alloc_vfield(n), wt = alloc_vfield(n), nt = alloc_vfield(n), vct = alloc_vfield(n);
    real* mag_t = alloc_real(n);
    /* reconstruct_jvp_kernel (pbf.cu:462-466) */
    for (uint32_t i = 0; i < n; ++i)
        v3_store(vt.c, i, v3_scale(v3_sub(v3_load(CV(pt), i), v3_load(pos_t, i)), r.inv_dt));
    /* vorticity_jvp_kernel (pbf.cu:499-527) + magnitude_jvp_kernel (pbf.cu:595-599) */
#pragma omp parallel for schedule(static)
    for (int64_t ii = 0; ii < (int64_t)n; ++ii) {
        const uint32_t i = (uint32_t)ii;
        const v3 pi = v3_load(pc, i), ti = v3_load(CV(pt), i);
        const v3 vi = v3_load(CV(r.vrec), i), vti = v3_load(CV(vt), i);
        v3 res = v3_make(0, 0, 0);
        cell_range cr[27];
        const int nr = stencil(g, v3_load(CV(r.xs), i), cr);
        for (int c = 0; c < nr; ++c)
            for (uint32_t s = cr[c].first; s < cr[c].last; ++s) {
                const uint32_t j = g->idx[s];
                if (j == i) continue;
                const v3 dd = v3_sub(pi, v3_load(pc, j));
                const v3 dd_t = v3_sub(ti, v3_load(CV(pt), j));
                const v3 diff = v3_sub(v3_load(CV(r.vrec), j), vi);
                const v3 diff_t = v3_sub(v3_load(CV(vt), j), vti);
                res = v3_add(res, v3_add(v3_cross(diff_t, spiky_gradient(dd, h, k)),
                                         v3_cross(diff, spiky_hessian_r(dd, dd_t, h, k))));
            }
        v3_store(wt.c, i, res);
        mag_t[i] = r.mag[i] > K(0.0) ? v3_dot(v3_load(CV(r.w), i), res) / r.mag[i] : K(0.0);
    }
    /* normal_jvp_kernel (pbf.cu:600-628) + confinement_jvp_kernel (pbf.cu:680-686) */
#pragma omp parallel for schedule(static)
    for (int64_t ii = 0; ii < (int64_t)n; ++ii) {
        const uint32_t i = (uint32_t)ii;
        const v3 pi = v3_load(pc, i), ti = v3_load(CV(pt), i);
        v3 gt = v3_make(0, 0, 0);
        cell_range cr[27];
        const int nr = stencil(g, v3_load(CV(r.xs), i), cr);
        for (int c = 0; c < nr; ++c)
            for (uint32_t s = cr[c].first; s < cr[c].last; ++s) {
                const uint32_t j = g->idx[s];
                if (j == i) continue;
                const v3 dd = v3_sub(pi, v3_load(pc, j));
                const v3 dd_t = v3_sub(ti, v3_load(CV(pt), j));
                gt = v3_add(gt, v3_add(v3_scale(spiky_gradient(dd, h, k), mag_t[j] - mag_t[i]),
                                       v3_scale(spiky_hessian_r(dd, dd_t, h, k), r.mag[j] - r.mag[i])));
            }
        const v3 nv = v3_load(CV(r.nrm), i);
        const v3 ntan = r.nz[i] == K(0.0) ? v3_make(0, 0, 0)
                                          : v3_scale(v3_add(gt, v3_scale(nv, -v3_dot(nv, gt))), K(1.0) / r.nz[i]);
        v3_store(nt.c, i, ntan);
        const v3 wv = v3_load(CV(r.w), i);
        const v3 force = v3_cross(nv, wv);
        const v3 force_t = v3_add(v3_cross(ntan, wv), v3_cross(nv, v3_load(CV(wt), i)));
        v3_store(vct.c, i,
                 v3_add(v3_load(CV(vt), i),
                        v3_scale(v3_add(v3_scale(force, confinement_t[i]), v3_scale(force_t, confinement[i])), dt)));
    }
    /* xsph_jvp_kernel (pbf.cu:723-754) */
#pragma omp parallel for schedule(static)
    for (int64_t ii = 0; ii < (int64_t)n; ++ii) {
        const uint32_t i = (uint32_t)ii;
        const v3 pi = v3_load(pc, i), ti = v3_load(CV(pt), i);
        const v3 vi = v3_load(CV(r.vc), i), vti = v3_load(CV(vct), i);
        v3 sm = v3_make(0, 0, 0), sm_t = v3_make(0, 0, 0);
        cell_range cr[27];
        const int nr = stencil(g, v3_load(CV(r.xs), i), cr);
        for (int c = 0; c < nr; ++c)
            for (uint32_t s = cr[c].first; s < cr[c].last; ++s) {
                const uint32_t j = g->idx[s];
                if (j == i) continue;
                const v3 dd = v3_sub(pi, v3_load(pc, j));
                const v3 dd_t = v3_sub(ti, v3_load(CV(pt), j));
                const v3 diff = v3_sub(v3_load(CV(r.vc), j), vi);
                const v3 diff_t = v3_sub(v3_load(CV(vct), j), vti);
                const real wgt = poly6(dd, h);
                const real wgt_t = v3_dot(poly6_gradient(dd, h), dd_t);
                sm = v3_add(sm, v3_scale(diff, wgt));
                sm_t = v3_add(sm_t, v3_add(v3_scale(diff_t, wgt), v3_scale(diff, wgt_t)));
            }
        v3_store(next_vel_t, i, v3_add(vti, v3_add(v3_scale(sm, xsph_t[i]), v3_scale(sm_t, xsph[i]))));
        v3_store(next_pos_t, i, ti);
    }
    free_vfield(&pt);
    free_vfield(&pn);
    free_vfield(&vt);
    free_vfield(&wt);
    free_vfield(&nt);
    free_vfield(&vct);
    free(rho_t);
    free(lam_t);
    free(mag_t);
    pbf_record_free(&r);
}

/* ---------------------------------------------------------------- PBF VJP */
static double* dzeros(uint32_t n) { return (double*)calloc((size_t)n + 2u, sizeof(double)); }
typedef struct {
    double* c[3];
} dfield;
static dfield dfield_zeros(uint32_t n) {
    dfield f;
    for (int a = 0; a < 3; ++a) f.c[a] = dzeros(n);
    return f;
}
static void dfield_free(dfield* f) {
    for (int a = 0; a < 3; ++a) free(f->c[a]);
}
#define DV(f) ((const double* const*)(f).c)

void SUF(oracle_pbf_vjp)(const oracle_domain* d, uint32_t iterations, uint64_t step, const real* pos[3],
                         const real* vel[3], const real* ctl[3], const real* masses, const real* rest_densities,
                         const real* relaxation, const real* strength, const real* exponent, const real* radius,
                         const real* xsph, const real* confinement, const double* next_pos_a[3],
                         const double* next_vel_a[3], double* pos_a[3], double* vel_a[3], double* ctl_a[3],
                         double* masses_a, double* rest_densities_a, double* relaxation_a, double* strength_a,
                         double* exponent_a, double* radius_a, double* xsph_a, double* confinement_a) {
    pbf_record r;
    pbf_record_forward(&r, d, iterations, step, pos, vel, ctl, masses, rest_densities, relaxation, strength, exponent,
                       radius, confinement);
    const uint32_t n = r.n;
    const double h = (double)r.h, k = (double)r.spiky;
    const real hr = r.h, kr = r.spiky;
    const grid_t* g = &r.g;
    const real* const* pc = CV(r.p[iterations]);
    dfield pa = dfield_zeros(n);  /* adjoint of the positions p[current] */
    dfield vca = dfield_zeros(n); /* adjoint of v_c */
    dfield va = dfield_zeros(n);  /* adjoint of v_rec */
    dfield wa = dfield_zeros(n);  /* adjoint of omega */
    dfield na = dfield_zeros(n);  /* adjoint of the normals */
    double* maga = dzeros(n);
    for (uint32_t i = 0; i < n; ++i) d3_acc(pa.c, i, d3_cload(next_pos_a, i));

    /* xsph_vjp_kernel (pbf.cu:756-794) */
#pragma omp parallel for schedule(static)
    for (int64_t ii = 0; ii < (int64_t)n; ++ii) {
        const uint32_t a = (uint32_t)ii;
        const v3 pi = v3_load(pc, a);
        const d3 va_out = d3_cload(next_vel_a, a);
        const d3 vca_ = d3_of(v3_load(CV(r.vc), a));
        d3 pcon = d3_make(0, 0, 0), vcon = va_out;
        double ccon = 0.0;
        cell_range cr[27];
        const int nr = stencil(g, v3_load(CV(r.xs), a), cr);
        for (int c = 0; c < nr; ++c)
            for (uint32_t s = cr[c].first; s < cr[c].last; ++s) {
                const uint32_t j = g->idx[s];
                if (j == a) continue;
                const v3 dd = v3_sub(pi, v3_load(pc, j));
                const double wgt = (double)poly6(dd, hr);
                const d3 gw = d3_of(poly6_gradient(dd, hr));
                const d3 diff = d3_sub(d3_of(v3_load(CV(r.vc), j)), vca_);
                const d3 vb_j = d3_cload(next_vel_a, j);
                vcon = d3_add(vcon, d3_scale(va_out, -(double)xsph[a] * wgt));
                vcon = d3_add(vcon, d3_scale(vb_j, (double)xsph[j] * wgt));
                pcon = d3_add(pcon, d3_scale(gw, (double)xsph[a] * d3_dot(va_out, diff) -
                                                    (double)xsph[j] * d3_dot(vb_j, diff)));
                ccon += wgt * d3_dot(va_out, diff);
            }
        d3_acc(pa.c, a, pcon);
        d3_acc(vca.c, a, vcon);
        xsph_a[a] += ccon;
    }
    /* confinement_vjp_kernel (pbf.cu:688-698) */
    const double dt = (double)r.dt;
    for (uint32_t i = 0; i < n; ++i) {
        const d3 out = d3_load(vca.c, i);
        const d3 wv = d3_of(v3_load(CV(r.w), i));
        const d3 nv = d3_of(v3_load(CV(r.nrm), i));
        const double f = dt * (double)confinement[i];
        d3_acc(va.c, i, out);
        d3_acc(wa.c, i, d3_scale(d3_cross(out, nv), f));
        d3_acc(na.c, i, d3_scale(d3_cross(wv, out), f));
        confinement_a[i] += dt * d3_dot(out, d3_cross(nv, wv));
    }
    /* normal_vjp_kernel (pbf.cu:635-672): gradient adjoint (n_bar - n (n . n_bar)) / N */
    dfield ga = dfield_zeros(n);
    for (uint32_t i = 0; i < n; ++i) {
        if (r.nz[i] == K(0.0)) continue;
        const d3 nb = d3_load(na.c, i);
        const d3 nv = d3_of(v3_load(CV(r.nrm), i));
        d3_acc(ga.c, i, d3_scale(d3_sub(nb, d3_scale(nv, d3_dot(nv, nb))), 1.0 / (double)r.nz[i]));
    }
#pragma omp parallel for schedule(static)
    for (int64_t ii = 0; ii < (int64_t)n; ++ii) {
        const uint32_t a = (uint32_t)ii;
        const v3 pi = v3_load(pc, a);
        const d3 gaa = d3_load(ga.c, a);
        d3 pcon = d3_make(0, 0, 0);
        double mcon = 0.0;
        cell_range cr[27];
        const int nr = stencil(g, v3_load(CV(r.xs), a), cr);
        for (int c = 0; c < nr; ++c)
            for (uint32_t s = cr[c].first; s < cr[c].last; ++s) {
                const uint32_t j = g->idx[s];
                if (j == a) continue;
                const v3 dd = v3_sub(pi, v3_load(pc, j));
                const d3 gr = d3_of(spiky_gradient(dd, hr, kr));
                const d3 both = d3_add(gaa, d3_load(ga.c, j));
                pcon = d3_add(pcon, spiky_hessian_d(dd, d3_scale(both, (double)r.mag[j] - (double)r.mag[a]), h, k));
                mcon -= d3_dot(gr, both);
            }
        d3_acc(pa.c, a, pcon);
        maga[a] += mcon;
    }
    for (uint32_t i = 0; i < n; ++i)
        if (r.mag[i] > K(0.0)) d3_acc(wa.c, i, d3_scale(d3_of(v3_load(CV(r.w), i)), maga[i] / (double)r.mag[i]));
    /* vorticity_vjp_kernel (pbf.cu:529-563) */
#pragma omp parallel for schedule(static)
    for (int64_t ii = 0; ii < (int64_t)n; ++ii) {
        const uint32_t a = (uint32_t)ii;
        const v3 pi = v3_load(pc, a);
        const d3 vi = d3_of(v3_load(CV(r.vrec), a));
        const d3 wb = d3_load(wa.c, a);
        d3 pcon = d3_make(0, 0, 0), vcon = d3_make(0, 0, 0);
        cell_range cr[27];
        const int nr = stencil(g, v3_load(CV(r.xs), a), cr);
        for (int c = 0; c < nr; ++c)
            for (uint32_t s = cr[c].first; s < cr[c].last; ++s) {
                const uint32_t j = g->idx[s];
                if (j == a) continue;
                const v3 dd = v3_sub(pi, v3_load(pc, j));
                const d3 gr = d3_of(spiky_gradient(dd, hr, kr));
                const d3 both = d3_add(wb, d3_load(wa.c, j));
                const d3 diff = d3_sub(d3_of(v3_load(CV(r.vrec), j)), vi);
                vcon = d3_sub(vcon, d3_cross(gr, both));
                pcon = d3_add(pcon, spiky_hessian_d(dd, d3_cross(both, diff), h, k));
            }
        d3_acc(pa.c, a, pcon);
        d3_acc(va.c, a, vcon);
    }
    /* reconstruct_vjp_kernel (pbf.cu:468-474) */
    const double inv_dt = (double)r.inv_dt;
    for (uint32_t i = 0; i < n; ++i) {
        const d3 cb = d3_scale(d3_load(va.c, i), inv_dt);
        d3_acc(pos_a, i, d3_scale(cb, -1.0));
        d3_acc(pa.c, i, cb);
    }
    /* constraint iterations in reverse (pbf.cpp:226-256) */
    dfield pn = dfield_zeros(n), ca = dfield_zeros(n);
    double* lama = dzeros(n);
    double* rhoa = dzeros(n);
    for (uint32_t it = iterations; it-- > 0u;) {
        const real* const* p = CV(r.p[it]);
        const real* lam = r.lam[it];
        for (uint32_t i = 0; i < n; ++i) {
            for (int c = 0; c < 3; ++c) pn.c[c][i] = 0.0;
            lama[i] = 0.0;
            rhoa[i] = 0.0;
            /* project_vjp_kernel (pbf.cu:445-454) */
            d3 b = d3_load(pa.c, i);
            const uint8_t m = r.mask[it][i];
            if (m & 1u) b.x = 0.0;
            if (m & 2u) b.y = 0.0;
            if (m & 4u) b.z = 0.0;
            ca.c[0][i] = b.x;
            ca.c[1][i] = b.y;
            ca.c[2][i] = b.z;
            d3_acc(pn.c, i, b);
        }
        /* correction_vjp_kernel (pbf.cu:335-415) */
#pragma omp parallel for schedule(static)
        for (int64_t ii = 0; ii < (int64_t)n; ++ii) {
            const uint32_t a = (uint32_t)ii;
            const v3 pi = v3_load(p, a);
            const double rho0 = (double)rest_densities[a];
            const d3 oa = d3_load(ca.c, a);
            d3 pcon = d3_make(0, 0, 0);
            double lcon = 0.0, mcon = 0.0, rcon = 0.0, scon = 0.0, econ = 0.0, qcon = 0.0;
            cell_range cr[27];
            const int nr = stencil(g, v3_load(CV(r.xs), a), cr);
            for (int c = 0; c < nr; ++c) {
                for (uint32_t s = cr[c].first; s < cr[c].last; ++s) {
                    const uint32_t j = g->idx[s];
                    if (j == a) continue;
                    const v3 dd = v3_sub(pi, v3_load(p, j));
                    const d3 gr = d3_of(spiky_gradient(dd, hr, kr));
                    /* a as the centre */
                    const artificial_t art = artificial_full(dd, hr, strength[a], exponent[a], radius[a]);
                    const double cf = (double)masses[j] / rho0;
                    const double mu = (double)lam[a] + (double)lam[j] + (double)art.value;
                    const double sa = cf * d3_dot(oa, gr);
                    pcon = d3_add(pcon, spiky_hessian_d(dd, d3_scale(oa, cf * mu), h, k));
                    pcon = d3_add(pcon, d3_scale(d3_of(art.gradient), sa));
                    lcon += sa;
                    rcon -= cf * mu * d3_dot(oa, gr) / rho0;
                    scon += sa * (double)art.d_strength;
                    econ += sa * (double)art.d_exponent;
                    qcon += sa * (double)art.d_radius;
                    /* a as a neighbour of j: d_ja = -d_aj, G(d_ja) = -G(d_aj), H even, grad s odd */
                    const v3 rd = v3_scale(dd, K(-1.0));
                    const artificial_t rart = artificial_full(rd, hr, strength[j], exponent[j], radius[j]);
                    const double rrho0 = (double)rest_densities[j];
                    const double rcf = (double)masses[a] / rrho0;
                    const double rmu = (double)lam[j] + (double)lam[a] + (double)rart.value;
                    const d3 ob = d3_load(ca.c, j);
                    const d3 rgr = d3_of(spiky_gradient(rd, hr, kr));
                    const double rsa = rcf * d3_dot(ob, rgr);
                    pcon = d3_sub(pcon, spiky_hessian_d(rd, d3_scale(ob, rcf * rmu), h, k));
                    pcon = d3_sub(pcon, d3_scale(d3_of(rart.gradient), rsa));
                    lcon += rsa;
                    mcon += rmu / rrho0 * d3_dot(ob, rgr);
                }
                /* boundary samples (pbf.cu:390-397): the sample is fixed, so
                 * the pair reaches only a's adjoints */
                for (uint32_t s = cr[c].bfirst; s < cr[c].blast; ++s) {
                    const uint32_t b = g->bidx[s];
                    const v3 dd = v3_sub(pi, boundary_position(g, b));
                    const d3 gr = d3_of(spiky_gradient(dd, hr, kr));
                    const artificial_t art = artificial_full(dd, hr, strength[a], exponent[a], radius[a]);
                    const double vol = (double)boundary_volume(g, b);
                    const double mu = (double)lam[a] + (double)art.value;
                    const double sa = vol * d3_dot(oa, gr);
                    pcon = d3_add(pcon, spiky_hessian_d(dd, d3_scale(oa, vol * mu), h, k));
                    pcon = d3_add(pcon, d3_scale(d3_of(art.gradient), sa));
                    lcon += sa;
                    scon += sa * (double)art.d_strength;
                    econ += sa * (double)art.d_exponent;
                    qcon += sa * (double)art.d_radius;
                }
            }
            d3_acc(pn.c, a, pcon);
            lama[a] += lcon;
            masses_a[a] += mcon;
            rest_densities_a[a] += rcon;
            strength_a[a] += scon;
            exponent_a[a] += econ;
            radius_a[a] += qcon;
        }
        /* lambda_vjp_kernel (pbf.cu:178-252): D_bar = lam_bar C / D^2 */
#pragma omp parallel for schedule(static)
        for (int64_t ii = 0; ii < (int64_t)n; ++ii) {
            const uint32_t a = (uint32_t)ii;
            const v3 pi = v3_load(p, a);
            const double rho0 = (double)rest_densities[a];
            const double den = (double)r.den[it][a];
            const double con = (double)r.rho[it][a] / rho0 - 1.0;
            const double dbar = lama[a] * con / (den * den);
            const d3 sgi = d3_of(v3_load(CV(r.gsum[it]), a));
            d3 pcon = d3_make(0, 0, 0);
            double mcon = 0.0, rcon = lama[a] * (double)r.rho[it][a] / (den * rho0 * rho0);
            cell_range cr[27];
            const int nr = stencil(g, v3_load(CV(r.xs), a), cr);
            for (int c = 0; c < nr; ++c) {
                for (uint32_t s = cr[c].first; s < cr[c].last; ++s) {
                    const uint32_t j = g->idx[s];
                    if (j == a) continue;
                    const v3 dd = v3_sub(pi, v3_load(p, j));
                    const d3 gr = d3_of(spiky_gradient(dd, hr, kr));
                    /* a as the centre: g_bar = 2 D_bar c (c g + S); c_bar = 2 D_bar (c |g|^2 + g . S) */
                    const double cf = (double)masses[j] / rho0;
                    const d3 gbar = d3_scale(d3_add(d3_scale(gr, cf), sgi), 2.0 * dbar * cf);
                    pcon = d3_add(pcon, spiky_hessian_d(dd, gbar, h, k));
                    const double cbar = 2.0 * dbar * (cf * d3_dot(gr, gr) + d3_dot(gr, sgi));
                    rcon -= cbar * cf / rho0;
                    /* a as a neighbour of j */
                    const double jrho0 = (double)rest_densities[j];
                    const double jden = (double)r.den[it][j];
                    const double jcon = (double)r.rho[it][j] / jrho0 - 1.0;
                    const double jdbar = lama[j] * jcon / (jden * jden);
                    const d3 sgj = d3_of(v3_load(CV(r.gsum[it]), j));
                    const v3 rd = v3_scale(dd, K(-1.0));
                    const d3 rgr = d3_of(spiky_gradient(rd, hr, kr));
                    const double rcf = (double)masses[a] / jrho0;
                    const d3 rgbar = d3_scale(d3_add(d3_scale(rgr, rcf), sgj), 2.0 * jdbar * rcf);
                    pcon = d3_sub(pcon, spiky_hessian_d(rd, rgbar, h, k));
                    const double rcbar = 2.0 * jdbar * (rcf * d3_dot(rgr, rgr) + d3_dot(rgr, sgj));
                    mcon += rcbar / jrho0;
                }
                /* boundary samples (pbf.cu:226-229): V_b grad W enters S only, so
                 * its position adjoint is H(d) (2 D_bar V_b S) */
                for (uint32_t s = cr[c].bfirst; s < cr[c].blast; ++s) {
                    const uint32_t b = g->bidx[s];
                    const v3 dd = v3_sub(pi, boundary_position(g, b));
                    pcon = d3_add(pcon, spiky_hessian_d(dd, d3_scale(sgi, 2.0 * dbar * (double)boundary_volume(g, b)), h, k));
                }
            }
            d3_acc(pn.c, a, pcon);
            rhoa[a] -= lama[a] / (den * rho0);
            masses_a[a] += mcon;
            rest_densities_a[a] += rcon;
            relaxation_a[a] += dbar;
        }
        /* density_vjp_kernel (density.cu:69-98, poly6; the self term has zero gradient) */
#pragma omp parallel for schedule(static)
        for (int64_t ii = 0; ii < (int64_t)n; ++ii) {
            const uint32_t a = (uint32_t)ii;
            const v3 pi = v3_load(p, a);
            d3 pcon = d3_make(0, 0, 0);
            double mcon = 0.0, rcon = 0.0;
            cell_range cr[27];
            const int nr = stencil(g, v3_load(CV(r.xs), a), cr);
            for (int c = 0; c < nr; ++c) {
                for (uint32_t s = cr[c].first; s < cr[c].last; ++s) {
                    const uint32_t j = g->idx[s];
                    const v3 dd = v3_sub(pi, v3_load(p, j));
                    const d3 gw = d3_of(poly6_gradient(dd, hr));
                    pcon = d3_add(pcon, d3_scale(gw, rhoa[a] * (double)masses[j] + rhoa[j] * (double)masses[a]));
                    mcon += rhoa[j] * (double)poly6(dd, hr);
                }
                /* boundary samples (density.cu:88-93): rho0_a V_b W */
                for (uint32_t s = cr[c].bfirst; s < cr[c].blast; ++s) {
                    const uint32_t b = g->bidx[s];
                    const v3 dd = v3_sub(pi, boundary_position(g, b));
                    const double vol = (double)boundary_volume(g, b);
                    pcon = d3_add(pcon, d3_scale(d3_of(poly6_gradient(dd, hr)), rhoa[a] * (double)rest_densities[a] * vol));
                    rcon += rhoa[a] * vol * (double)poly6(dd, hr);
                }
            }
            d3_acc(pn.c, a, pcon);
            masses_a[a] += mcon;
            rest_densities_a[a] += rcon;
        }
        dfield tmp = pa;
        pa = pn;
        pn = tmp;
    }
    /* predict_vjp_kernel (pbf.cu:86-93) */
    const double dt2 = (double)(r.dt * r.dt);
    for (uint32_t i = 0; i < n; ++i) {
        const d3 b = d3_load(pa.c, i);
        d3_acc(pos_a, i, b);
        d3_acc(vel_a, i, d3_scale(b, dt));
        d3_acc(ctl_a, i, d3_scale(b, dt2));
    }
    dfield_free(&pa);
    dfield_free(&pn);
    dfield_free(&ca);
    dfield_free(&vca);
    dfield_free(&va);
    dfield_free(&wa);
    dfield_free(&na);
    dfield_free(&ga);
    free(maga);
    free(lama);
    free(rhoa);
    pbf_record_free(&r);
}

/* ===================================================== DFSPH derivatives */
/* SPH<DFSPH>::jvp_step and vjp_step (sph.ixx:356-365) around DFSPH::jvp
 * (sph.cpp:612-659) and DFSPH::vjp (sph.cpp:661-723), with the derivative
 * kernels of sph/common.cu, sph/dfsph.cu and density/density.cu. The
 * neighbour graph, the box-collision masks and the active set of every
 * max(0, .) pressure update are frozen at the primal step. The JVP runs in
 * the primal precision (the reference's FP32 tangents). The VJP promotes the
 * primal values and accumulates FP64 adjoints (+=) like the reference's
 * kernels, keeping its mixed-precision operand order (products of two FP32
 * primal values are rounded to FP32 before promotion), except the warm-start
 * pressure adjoints, which are copied (sph.cpp:713). Pair terms are gathered
 * from both ends of a pair: particle i collects the terms in which it is the
 * centre and those in which it is the neighbour. */

enum { DF_MASS = 0, DF_REST, DF_VISC, DF_DIV, DF_DEN, DF_PARAMS };

/* Every primal quantity of one DFSPH step that the derivatives read. Phase
 * q = 0 is the divergence solve, q = 1 the density solve; iterate k holds
 * the pressures p[q][k] and accelerations a[q][k] after k updates, and
 * the predicted positions xs[q][k] and densities rs[q][k] of update k >= 1. */
typedef struct {
    uint32_t n, iters[2];
    int warm;
    /* density of the volume V = m / density in the pressure update: the state
     * density (DFSPH, IISPH) or, when set, the rest density (PCISPH) */
    int volume_rest;
    grid_t g;
    real h, dt, G;
    cbox box;
    const real* pos[3];
    const real* vel[3];
    const real* prm[DF_PARAMS];
    real* rho;
    vfield anp, adiv, total;
    real** p[2];
    vfield* a[2];
    vfield* xs[2];
    real** rs[2];
} dfsph_record;

/* DFSPH::forward (sph.cpp:572-610) keeping every iterate. With nd = 0, no
 * warm start and the density-phase relaxation this is also PCISPH::forward
 * (sph.cpp:295-309, volume_rest set: V = m / rho0, pcisph.cu:104-110) and
 * IISPH::forward (sph.cpp:408-421, V = m / rho, iisph.cu:104-110): the
 * divergence phase then only adds a zero acceleration (+0 + a = a) and the
 * pressures start from zero (clear_iteration, sph.cpp:298, 411). */
static void dfsph_record_forward(dfsph_record* r, const oracle_domain* d, uint32_t nd, uint32_t ne, int warm,
                                 int volume_rest, uint64_t step, const real* pos[3], const real* vel[3],
                                 const real* ctl[3], const real* const prm[DF_PARAMS], const real* const warm_p[2]) {
    const uint32_t n = d->particle_count;
    memset(r, 0, sizeof(*r));
    r->n = n;
    r->iters[0] = nd;
    r->iters[1] = ne;
    r->warm = warm;
    r->volume_rest = volume_rest;
    r->h = (real)d->support_radius;
    r->dt = (real)d->time_step;
    r->G = SUF(oracle_reference_gradient_norm)(d);
    r->box = collision_box(d, step);
    for (int c = 0; c < 3; ++c) {
        r->pos[c] = pos[c];
        r->vel[c] = vel[c];
    }
    for (int k = 0; k < DF_PARAMS; ++k) r->prm[k] = prm[k];
    grid_build(&r->g, d, step, pos[0], pos[1], pos[2]);
    r->rho = alloc_real(n);
    r->anp = alloc_vfield(n);
    r->adiv = alloc_vfield(n);
    r->total = alloc_vfield(n);
    density_kernel(&r->g, 0, pos, pos, prm[DF_MASS], prm[DF_REST], r->rho);
    non_pressure_kernel(&r->g, pos, vel, ctl, prm[DF_MASS], prm[DF_VISC], r->rho, r->anp.c);
    vfield acc = alloc_vfield(n);
    for (int q = 0; q < 2; ++q) {
        const uint32_t it = r->iters[q];
        r->p[q] = (real**)calloc(it + 1u, sizeof(real*));
        r->rs[q] = (real**)calloc(it + 1u, sizeof(real*));
        r->a[q] = (vfield*)calloc(it + 1u, sizeof(vfield));
        r->xs[q] = (vfield*)calloc(it + 1u, sizeof(vfield));
        r->p[q][0] = alloc_real(n);
        for (uint32_t i = 0; i < n; ++i) r->p[q][0][i] = warm ? warm_p[q][i] : K(0.0);
        r->a[q][0] = alloc_vfield(n);
        pressure_kernel(&r->g, pos, prm[DF_MASS], prm[DF_REST], r->rho, r->p[q][0], r->a[q][0].c);
        for (uint32_t k = 1; k <= it; ++k) {
            const real* const* ap = CV(r->a[q][k - 1u]);
            if (q == 1) {
                add_kernel(n, CV(r->adiv), ap, acc.c);
                ap = CV(acc);
            }
            r->xs[q][k] = alloc_vfield(n);
            sph_predict_kernel(n, r->dt, r->box, pos, vel, CV(r->anp), ap, r->xs[q][k].c, NULL);
            r->rs[q][k] = alloc_real(n);
            density_kernel(&r->g, 0, pos, CV(r->xs[q][k]), prm[DF_MASS], prm[DF_REST], r->rs[q][k]);
            r->p[q][k] = alloc_real(n);
            memcpy(r->p[q][k], r->p[q][k - 1u], sizeof(real) * n);
            projection_update_kernel(n, r->dt, r->G, prm[DF_MASS], volume_rest ? prm[DF_REST] : r->rho,
                                     q ? prm[DF_REST] : r->rho, r->rs[q][k], prm[q ? DF_DEN : DF_DIV], r->p[q][k]);
            r->a[q][k] = alloc_vfield(n);
            pressure_kernel(&r->g, pos, prm[DF_MASS], prm[DF_REST], r->rho, r->p[q][k], r->a[q][k].c);
        }
        if (q == 0)
            for (int c = 0; c < 3; ++c) memcpy(r->adiv.c[c], r->a[0][it].c[c], sizeof(real) * n);
    }
    add_kernel(n, CV(r->adiv), CV(r->a[1][ne]), r->total.c);
    free_vfield(&acc);
}

static void dfsph_record_free(dfsph_record* r) {
    for (int q = 0; q < 2; ++q) {
        for (uint32_t k = 0; k <= r->iters[q]; ++k) {
            free(r->p[q][k]);
            free(r->rs[q][k]);
            free_vfield(&r->a[q][k]);
            free_vfield(&r->xs[q][k]);
        }
        free(r->p[q]);
        free(r->rs[q]);
        free(r->a[q]);
        free(r->xs[q]);
    }
    free(r->rho);
    free_vfield(&r->anp);
    free_vfield(&r->adiv);
    free_vfield(&r->total);
    grid_free(&r->g);
}

/* ------------------------------------------------------------ SPH JVP */
/* cubic_gradient_tangent, density/device.cuh:45-57 */
static v3 cubic_gradient_tangent(v3 d, v3 dd, real h) {
    const real dist = v3_length(d);
    const real coef = K(8.0) / (R_PI * h * h * h);
    if (dist == K(0.0)) return v3_scale(dd, K(-12.0) * coef / (h * h));
    if (dist >= h) return v3_make(0, 0, 0);
    const real q = dist / h;
    const real deriv = q <= K(0.5) ? K(18.0) * q * q - K(12.0) * q : K(-6.0) * (K(1.0) - q) * (K(1.0) - q);
    const real second = q <= K(0.5) ? K(36.0) * q - K(12.0) : K(12.0) * (K(1.0) - q);
    const real radial = v3_dot(d, dd) / dist;
    const real rs = coef * deriv / (h * dist);
    const real rsd = coef * (second / (h * h * dist) - deriv / (h * dist * dist));
    return v3_add(v3_scale(dd, rs), v3_scale(d, rsd * radial));
}

/* viscosity_laplacian_tangent, density/device.cuh:78-82 */
static real viscosity_laplacian_tangent(v3 d, v3 dd, real h) {
    const real dist = v3_length(d);
    if (dist == K(0.0) || dist >= h) return K(0.0);
    return K(-45.0) * v3_dot(d, dd) / (R_PI * h * h * h * h * h * h * dist);
}

/* density_jvp_kernel, density/density.cu:42-67 (cubic kernel) */
/* the density kernel and its gradient: poly6 (PBF) or the cubic spline,
 * density_kernel / density_kernel_gradient of density.cu:9-16 */
static real density_w(v3 d, real h, int pbf) { return pbf ? poly6(d, h) : cubic(d, h); }
static v3 density_g(v3 d, real h, int pbf) { return pbf ? poly6_gradient(d, h) : cubic_gradient(d, h); }

static void density_jvp_kernel(const grid_t* g, int pbf, const real* const topo[3], const real* const pos[3],
                               const real* const pos_t[3], const real* masses, const real* masses_t,
                               const real* rest, const real* rest_t, real* out) {
    const real h = (real)g->d->support_radius;
#pragma omp parallel for schedule(static)
    for (int64_t ii = 0; ii < (int64_t)g->n; ++ii) {
        const uint32_t i = (uint32_t)ii;
        const v3 p = v3_load(pos, i);
        const v3 t = v3_load(pos_t, i);
        cell_range cr[27];
        const int nr = stencil(g, v3_load(topo, i), cr);
        real result = K(0.0);
        for (int k = 0; k < nr; ++k) {
            for (uint32_t s = cr[k].first; s < cr[k].last; ++s) {
                const uint32_t j = g->idx[s];
                const v3 d = v3_sub(p, v3_load(pos, j));
                result += masses_t[j] * density_w(d, h, pbf) +
                          masses[j] * v3_dot(density_g(d, h, pbf), v3_sub(t, v3_load(pos_t, j)));
            }
            for (uint32_t s = cr[k].bfirst; s < cr[k].blast; ++s) {
                const uint32_t b = g->bidx[s];
                const v3 d = v3_sub(p, boundary_position(g, b));
                result += boundary_volume(g, b) *
                          (rest_t[i] * density_w(d, h, pbf) + rest[i] * v3_dot(density_g(d, h, pbf), t));
            }
        }
        out[i] = result;
    }
}

/* non_pressure_jvp_kernel, sph/common.cu:49-96 */
static void non_pressure_jvp_kernel(const grid_t* g, const real* const pos[3], const real* const vel[3],
                                    const real* const ctl_t[3], const real* const pos_t[3],
                                    const real* const vel_t[3], const real* masses, const real* masses_t,
                                    const real* visc, const real* visc_t, const real* rho, const real* rho_t,
                                    real* const out[3]) {
    const real h = (real)g->d->support_radius;
#pragma omp parallel for schedule(static)
    for (int64_t ii = 0; ii < (int64_t)g->n; ++ii) {
        const uint32_t i = (uint32_t)ii;
        const v3 p = v3_load(pos, i);
        const v3 v = v3_load(vel, i);
        const v3 pt = v3_load(pos_t, i);
        const v3 vt = v3_load(vel_t, i);
        v3 result = v3_load(ctl_t, i);
        cell_range cr[27];
        const int nr = stencil(g, p, cr);
        for (int k = 0; k < nr; ++k) {
            for (uint32_t s = cr[k].first; s < cr[k].last; ++s) {
                const uint32_t j = g->idx[s];
                if (j == i) continue;
                const v3 d = v3_sub(p, v3_load(pos, j));
                const v3 dd = v3_sub(pt, v3_load(pos_t, j));
                const real lap = viscosity_laplacian(d, h);
                const real lap_t = viscosity_laplacian_tangent(d, dd, h);
                const real nu = K(0.5) * (visc[i] + visc[j]);
                const real nu_t = K(0.5) * (visc_t[i] + visc_t[j]);
                const real mass = masses[j];
                const real dens = rho[j];
                const real w = nu * mass * lap / dens;
                const real w_t = nu_t * mass * lap / dens + nu * masses_t[j] * lap / dens + nu * mass * lap_t / dens -
                                 nu * mass * lap * rho_t[j] / (dens * dens);
                const v3 diff = v3_sub(v3_load(vel, j), v);
                const v3 diff_t = v3_sub(v3_load(vel_t, j), vt);
                result = v3_add(result, v3_add(v3_scale(diff, w_t), v3_scale(diff_t, w)));
            }
            for (uint32_t s = cr[k].bfirst; s < cr[k].blast; ++s) {
                const uint32_t b = g->bidx[s];
                const v3 d = v3_sub(p, boundary_position(g, b));
                const real lap = viscosity_laplacian(d, h);
                const real w = visc[i] * boundary_volume(g, b) * lap;
                const real w_t = boundary_volume(g, b) * (visc_t[i] * lap + visc[i] * viscosity_laplacian_tangent(d, pt, h));
                const v3 diff = v3_sub(boundary_velocity(g, b), v);
                result = v3_add(result, v3_add(v3_scale(diff, w_t), v3_scale(vt, -w)));
            }
        }
        v3_store(out, i, result);
    }
}

/* pressure_jvp_kernel, sph/common.cu:196-237 */
static void pressure_jvp_kernel(const grid_t* g, const real* const pos[3], const real* const pos_t[3],
                                const real* masses, const real* masses_t, const real* rest, const real* rest_t,
                                const real* rho, const real* rho_t, const real* p, const real* p_t,
                                real* const out[3]) {
    const real h = (real)g->d->support_radius;
#pragma omp parallel for schedule(static)
    for (int64_t ii = 0; ii < (int64_t)g->n; ++ii) {
        const uint32_t i = (uint32_t)ii;
        const v3 x = v3_load(pos, i);
        const v3 xt = v3_load(pos_t, i);
        const real dens = rho[i];
        const real first = p[i] / (dens * dens);
        const real first_t = p_t[i] / (dens * dens) - K(2.0) * p[i] * rho_t[i] / (dens * dens * dens);
        v3 result = v3_make(0, 0, 0);
        cell_range cr[27];
        const int nr = stencil(g, x, cr);
        for (int k = 0; k < nr; ++k) {
            for (uint32_t s = cr[k].first; s < cr[k].last; ++s) {
                const uint32_t j = g->idx[s];
                if (j == i) continue;
                const real nd = rho[j];
                const real second = p[j] / (nd * nd);
                const real second_t = p_t[j] / (nd * nd) - K(2.0) * p[j] * rho_t[j] / (nd * nd * nd);
                const v3 d = v3_sub(x, v3_load(pos, j));
                const v3 dd = v3_sub(xt, v3_load(pos_t, j));
                const real factor = -masses[j] * (first + second);
                const real factor_t = -masses_t[j] * (first + second) - masses[j] * (first_t + second_t);
                result = v3_add(result, v3_add(v3_scale(cubic_gradient(d, h), factor_t),
                                               v3_scale(cubic_gradient_tangent(d, dd, h), factor)));
            }
            for (uint32_t s = cr[k].bfirst; s < cr[k].blast; ++s) {
                const uint32_t b = g->bidx[s];
                const v3 d = v3_sub(x, boundary_position(g, b));
                const real vol = boundary_volume(g, b);
                const real factor = -rest[i] * vol * first;
                const real factor_t = -vol * (rest_t[i] * first + rest[i] * first_t);
                result = v3_add(result, v3_add(v3_scale(cubic_gradient(d, h), factor_t),
                                               v3_scale(cubic_gradient_tangent(d, xt, h), factor)));
            }
        }
        v3_store(out, i, result);
    }
}

/* Box-collision bits of the clamp in predict_forward_kernel
 * (sph/dfsph.cu:26-44) for acceleration a1 + a2. */
static int predict_collisions(cbox box, real dt, v3 x, v3 v, v3 a1, v3 a2, v3* xs) {
    const v3 vs = v3_add(v, v3_scale(v3_add(a1, a2), dt));
    const v3 p = v3_add(x, v3_scale(vs, dt));
    if (xs) *xs = p;
    return (p.x < box.lo[0] || p.x > box.hi[0] ? 1 : 0) | (p.y < box.lo[1] || p.y > box.hi[1] ? 2 : 0) |
           (p.z < box.lo[2] || p.z > box.hi[2] ? 4 : 0);
}

/* predict_jvp_kernel, sph/dfsph.cu:46-66 */
static void predict_jvp_kernel(uint32_t n, real dt, cbox box, const real* const pos[3], const real* const vel[3],
                               const real* const a1[3], const real* const a2[3], const real* const pos_t[3],
                               const real* const vel_t[3], const real* const a1_t[3], const real* const a2_t[3],
                               real* const out_pos_t[3], real* const out_vel_t[3]) {
#pragma omp parallel for schedule(static)
    for (int64_t ii = 0; ii < (int64_t)n; ++ii) {
        const uint32_t i = (uint32_t)ii;
        const int col = predict_collisions(box, dt, v3_load(pos, i), v3_load(vel, i), v3_load(a1, i), v3_load(a2, i), NULL);
        v3 vd = v3_add(v3_load(vel_t, i), v3_scale(v3_add(v3_load(a1_t, i), v3_load(a2_t, i)), dt));
        v3 xd = v3_add(v3_load(pos_t, i), v3_scale(vd, dt));
        if (col & 1) xd.x = K(0.0);
        if (col & 2) xd.y = K(0.0);
        if (col & 4) xd.z = K(0.0);
        if (box.no_slip && col) {
            vd = v3_make(0, 0, 0);
        } else {
            if (col & 1) vd.x = K(0.0);
            if (col & 2) vd.y = K(0.0);
            if (col & 4) vd.z = K(0.0);
        }
        if (out_pos_t) v3_store(out_pos_t, i, xd);
        if (out_vel_t) v3_store(out_vel_t, i, vd);
    }
}

/* projection_update_jvp_kernel, sph/dfsph.cu:112-121; out may alias prev_t */
static void projection_update_jvp_kernel(uint32_t n, real dt, real G, const real* masses, const real* masses_t,
                                         const real* rho, const real* rho_t, const real* targets,
                                         const real* targets_t, const real* prev, const real* predicted,
                                         const real* relax, const real* prev_t, const real* predicted_t,
                                         const real* relax_t, real* out) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (int64_t)n; ++i) {
        const real error = predicted[i] - targets[i];
        const real volume = masses[i] / rho[i];
        const real delta = K(1.0) / (K(2.0) * dt * dt * volume * volume * G);
        const real delta_t = K(-2.0) * delta * (masses_t[i] / masses[i] - rho_t[i] / rho[i]);
        const real candidate = prev[i] + relax[i] * delta * error;
        out[i] = candidate > K(0.0) ? prev_t[i] + relax_t[i] * delta * error + relax[i] * delta_t * error +
                                          relax[i] * delta * (predicted_t[i] - targets_t[i])
                                    : K(0.0);
    }
}

void SUF(oracle_dfsph_jvp)(const oracle_domain* d, uint32_t nd, uint32_t ne, int warm_start, int volume_rest,
                           uint64_t step, const real* pos[3], const real* vel[3], const real* ctl[3],
                           const real* const prm[DF_PARAMS], const real* warm_div, const real* warm_den,
                           const real* pos_t[3], const real* vel_t[3], const real* ctl_t[3],
                           const real* const prm_t[DF_PARAMS], const real* warm_div_t, const real* warm_den_t,
                           real* next_pos_t[3], real* next_vel_t[3], real* next_warm_div_t, real* next_warm_den_t) {
    dfsph_record r;
    const real* const warm_p[2] = {warm_div, warm_den};
    dfsph_record_forward(&r, d, nd, ne, warm_start, volume_rest, step, pos, vel, ctl, prm, warm_p);
    const uint32_t n = r.n;
    const grid_t* g = &r.g;
    const real *m = prm[DF_MASS], *rest = prm[DF_REST], *m_t = prm_t[DF_MASS], *rest_t = prm_t[DF_REST];
    /* sph.ixx:357: tangent of the state densities */
    real* rho_t = alloc_real(n);
    density_jvp_kernel(g, 0, pos, pos, pos_t, m, m_t, rest, rest_t, rho_t);
    vfield anp_t = alloc_vfield(n), ap_t = alloc_vfield(n), adiv_t = alloc_vfield(n), tot_t = alloc_vfield(n);
    vfield xs_t = alloc_vfield(n), acc = alloc_vfield(n);
    non_pressure_jvp_kernel(g, pos, vel, ctl_t, pos_t, vel_t, m, m_t, prm[DF_VISC], prm_t[DF_VISC], r.rho, rho_t,
                            anp_t.c);
    real* p_t = alloc_real(n);
    real* rs_t = alloc_real(n);
    const real* const warm_t[2] = {warm_div_t, warm_den_t};
    real* const out_warm_t[2] = {next_warm_div_t, next_warm_den_t};
    for (int q = 0; q < 2; ++q) {
        for (uint32_t i = 0; i < n; ++i) p_t[i] = warm_start ? warm_t[q][i] : K(0.0);
        pressure_jvp_kernel(g, pos, pos_t, m, m_t, rest, rest_t, r.rho, rho_t, r.p[q][0], p_t, ap_t.c);
        for (uint32_t k = 1; k <= r.iters[q]; ++k) {
            const real* const* a = CV(r.a[q][k - 1u]);
            const real* const* at = CV(ap_t);
            if (q == 1) {
                add_kernel(n, CV(r.adiv), a, acc.c);
                add_kernel(n, CV(adiv_t), at, tot_t.c);
                a = CV(acc);
                at = CV(tot_t);
            }
            predict_jvp_kernel(n, r.dt, r.box, pos, vel, CV(r.anp), a, pos_t, vel_t, CV(anp_t), at, xs_t.c, NULL);
            density_jvp_kernel(g, 0, pos, CV(r.xs[q][k]), CV(xs_t), m, m_t, rest, rest_t, rs_t);
            projection_update_jvp_kernel(n, r.dt, r.G, m, m_t, volume_rest ? rest : r.rho,
                                         volume_rest ? rest_t : rho_t, q ? rest : r.rho, q ? rest_t : rho_t,
                                         r.p[q][k - 1u], r.rs[q][k], prm[q ? DF_DEN : DF_DIV], p_t, rs_t,
                                         prm_t[q ? DF_DEN : DF_DIV], p_t);
            pressure_jvp_kernel(g, pos, pos_t, m, m_t, rest, rest_t, r.rho, rho_t, r.p[q][k], p_t, ap_t.c);
        }
        if (q == 0)
            for (int c = 0; c < 3; ++c) memcpy(adiv_t.c[c], ap_t.c[c], sizeof(real) * n);
        if (out_warm_t[q]) memcpy(out_warm_t[q], p_t, sizeof(real) * n);
    }
    add_kernel(n, CV(adiv_t), CV(ap_t), tot_t.c);
    predict_jvp_kernel(n, r.dt, r.box, pos, vel, CV(r.anp), CV(r.total), pos_t, vel_t, CV(anp_t), CV(tot_t), next_pos_t,
                       next_vel_t);
    free(rho_t);
    free(p_t);
    free(rs_t);
    free_vfield(&anp_t);
    free_vfield(&ap_t);
    free_vfield(&adiv_t);
    free_vfield(&tot_t);
    free_vfield(&xs_t);
    free_vfield(&acc);
    dfsph_record_free(&r);
}

/* The primal of the recorded step, next positions and velocities (the last
 * prediction, sph.cpp:609): pins dfsph_record_forward, with the options the
 * derivatives use, to the forward steps of DFSPH, PCISPH and IISPH. */
void SUF(oracle_dfsph_record_step)(const oracle_domain* d, uint32_t nd, uint32_t ne, int warm_start, int volume_rest,
                                   uint64_t step, const real* pos[3], const real* vel[3], const real* ctl[3],
                                   const real* const prm[DF_PARAMS], const real* warm_div, const real* warm_den,
                                   real* next_pos[3], real* next_vel[3]) {
    dfsph_record r;
    const real* const warm_p[2] = {warm_div, warm_den};
    dfsph_record_forward(&r, d, nd, ne, warm_start, volume_rest, step, pos, vel, ctl, prm, warm_p);
    sph_predict_kernel(r.n, r.dt, r.box, pos, vel, CV(r.anp), CV(r.total), next_pos, next_vel);
    dfsph_record_free(&r);
}

/* ------------------------------------------------------------ SPH VJP */
/* cubic_hessian_product, density/device.cuh:59-70 */
static d3 cubic_hessian_product(v3 d, d3 w, real h) {
    const double dist = sqrt((double)v3_dot(d, d));
    const double coef = 8.0 / ((double)R_PI * h * h * h);
    if (dist == 0.0) return d3_scale(w, -12.0 * coef / (h * h));
    if (dist >= h) return d3_make(0, 0, 0);
    const double q = dist / h;
    const double deriv = q <= 0.5 ? 18.0 * q * q - 12.0 * q : -6.0 * (1.0 - q) * (1.0 - q);
    const double second = q <= 0.5 ? 36.0 * q - 12.0 : 12.0 * (1.0 - q);
    const double rs = coef * deriv / (h * dist);
    const double rsd = coef * (second / (h * h * dist) - deriv / (h * dist * dist));
    const d3 dd = d3_of(d);
    return d3_add(d3_scale(w, rs), d3_scale(dd, rsd * d3_dot(w, dd) / dist));
}

/* viscosity_laplacian_gradient, density/device.cuh:84-88 */
static d3 viscosity_laplacian_gradient(v3 d, real h, double s) {
    const double dist = sqrt((double)v3_dot(d, d));
    if (dist == 0.0 || dist >= h) return d3_make(0, 0, 0);
    return d3_scale(d3_of(d), -45.0 * s / ((double)R_PI * pow((double)h, 6.0) * dist));
}

/* pressure_vjp_kernel, sph/common.cu:239-294 */
static void pressure_vjp_kernel(const grid_t* g, const real* const pos[3], const real* masses, const real* rest,
                                const real* rho, const real* p, const double* const acc_a[3], double* const pos_a[3],
                                double* rho_a, double* p_a, double* masses_a, double* rest_a) {
    const real h = (real)g->d->support_radius;
#pragma omp parallel for schedule(static)
    for (int64_t ii = 0; ii < (int64_t)g->n; ++ii) {
        const uint32_t i = (uint32_t)ii;
        const v3 x = v3_load(pos, i);
        const d3 local = d3_cload(acc_a, i);
        const double first = (double)p[i] / (rho[i] * rho[i]);
        d3 position = d3_make(0, 0, 0);
        double first_a = 0.0, mass_c = 0.0, rest_c = 0.0;
        cell_range cr[27];
        const int nr = stencil(g, x, cr);
        for (int k = 0; k < nr; ++k) {
            for (uint32_t s = cr[k].first; s < cr[k].last; ++s) {
                const uint32_t j = g->idx[s];
                if (j == i) continue;
                const v3 d = v3_sub(x, v3_load(pos, j));
                const d3 grad = d3_of(cubic_gradient(d, h));
                const double second = (double)p[j] / (rho[j] * rho[j]);
                const double sum = first + second;
                const double local_factor = -masses[j] * sum;
                first_a -= masses[j] * d3_dot(local, grad);
                position = d3_add(position, cubic_hessian_product(d, d3_scale(local, local_factor), h));
                const d3 cross = d3_cload(acc_a, j);
                const double cross_projection = d3_dot(cross, grad);
                mass_c += sum * cross_projection;
                first_a += masses[i] * cross_projection;
                position = d3_add(position, cubic_hessian_product(d, d3_scale(cross, masses[i] * sum), h));
            }
            for (uint32_t s = cr[k].bfirst; s < cr[k].blast; ++s) {
                const uint32_t b = g->bidx[s];
                const v3 d = v3_sub(x, boundary_position(g, b));
                const d3 grad = d3_of(cubic_gradient(d, h));
                const real vol = boundary_volume(g, b);
                const double factor = -rest[i] * vol * first;
                const double factor_a = d3_dot(local, grad);
                rest_c -= factor_a * vol * first;
                first_a -= factor_a * rest[i] * vol;
                position = d3_add(position, cubic_hessian_product(d, d3_scale(local, factor), h));
            }
        }
        d3_acc(pos_a, i, position);
        masses_a[i] += mass_c;
        rest_a[i] += rest_c;
        p_a[i] += first_a / (rho[i] * rho[i]);
        rho_a[i] -= 2.0 * first_a * p[i] / (rho[i] * rho[i] * rho[i]);
    }
}

/* non_pressure_vjp_kernel, sph/common.cu:98-163 */
static void non_pressure_vjp_kernel(const grid_t* g, const real* const pos[3], const real* const vel[3],
                                    const real* masses, const real* visc, const real* rho, const double* const acc_a[3],
                                    double* const pos_a[3], double* const vel_a[3], double* const ctl_a[3],
                                    double* rho_a, double* masses_a, double* visc_a) {
    const real h = (real)g->d->support_radius;
#pragma omp parallel for schedule(static)
    for (int64_t ii = 0; ii < (int64_t)g->n; ++ii) {
        const uint32_t i = (uint32_t)ii;
        const v3 x = v3_load(pos, i);
        const v3 v = v3_load(vel, i);
        const d3 local = d3_cload(acc_a, i);
        d3 position = d3_make(0, 0, 0), velocity = d3_make(0, 0, 0);
        double dens_c = 0.0, mass_c = 0.0, visc_c = 0.0;
        cell_range cr[27];
        const int nr = stencil(g, x, cr);
        for (int k = 0; k < nr; ++k) {
            for (uint32_t s = cr[k].first; s < cr[k].last; ++s) {
                const uint32_t j = g->idx[s];
                if (j == i) continue;
                const v3 d = v3_sub(x, v3_load(pos, j));
                const v3 diff = v3_sub(v3_load(vel, j), v);
                const real lap = viscosity_laplacian(d, h);
                const real nu = K(0.5) * (visc[i] + visc[j]);
                const real mass = masses[j];
                const real dens = rho[j];
                const real w = nu * mass * lap / dens;
                const double w_a = d3_dot(local, d3_of(diff));
                velocity = d3_add(velocity, d3_scale(local, -w));
                const double nu_a = w_a * mass * lap / dens;
                visc_c += 0.5 * nu_a;
                position = d3_add(position, viscosity_laplacian_gradient(d, h, w_a * nu * mass / dens));
                const d3 cross = d3_cload(acc_a, j);
                const real cross_w = nu * masses[i] * lap / rho[i];
                const double cross_w_a = d3_dot(cross, d3_of(v3_scale(diff, K(-1.0))));
                velocity = d3_add(velocity, d3_scale(cross, cross_w));
                mass_c += cross_w_a * nu * lap / rho[i];
                dens_c -= cross_w_a * nu * masses[i] * lap / (rho[i] * rho[i]);
                visc_c += 0.5 * cross_w_a * masses[i] * lap / rho[i];
                position = d3_add(position, viscosity_laplacian_gradient(d, h, cross_w_a * nu * masses[i] / rho[i]));
            }
            for (uint32_t s = cr[k].bfirst; s < cr[k].blast; ++s) {
                const uint32_t b = g->bidx[s];
                const v3 d = v3_sub(x, boundary_position(g, b));
                const v3 diff = v3_sub(boundary_velocity(g, b), v);
                const real lap = viscosity_laplacian(d, h);
                const real vol = boundary_volume(g, b);
                const real w = visc[i] * vol * lap;
                const double w_a = d3_dot(local, d3_of(diff));
                velocity = d3_add(velocity, d3_scale(local, -w));
                visc_c += w_a * vol * lap;
                position = d3_add(position, viscosity_laplacian_gradient(d, h, w_a * visc[i] * vol));
            }
        }
        d3_acc(pos_a, i, position);
        d3_acc(vel_a, i, velocity);
        d3_acc(ctl_a, i, local);
        rho_a[i] += dens_c;
        masses_a[i] += mass_c;
        visc_a[i] += visc_c;
    }
}

/* density_vjp_kernel, density/density.cu:69-100 (cubic kernel) */
static void density_vjp_kernel(const grid_t* g, int pbf, const real* const topo[3], const real* const pos[3],
                               const real* masses, const real* rest, const double* rho_a, double* const pos_a[3],
                               double* masses_a, double* rest_a) {
    const real h = (real)g->d->support_radius;
#pragma omp parallel for schedule(static)
    for (int64_t ii = 0; ii < (int64_t)g->n; ++ii) {
        const uint32_t i = (uint32_t)ii;
        const v3 x = v3_load(pos, i);
        const double out_a = rho_a[i];
        d3 position = d3_make(0, 0, 0);
        double mass_c = 0.0, rest_c = 0.0;
        cell_range cr[27];
        const int nr = stencil(g, v3_load(topo, i), cr);
        for (int k = 0; k < nr; ++k) {
            for (uint32_t s = cr[k].first; s < cr[k].last; ++s) {
                const uint32_t j = g->idx[s];
                const v3 d = v3_sub(x, v3_load(pos, j));
                position = d3_add(position, d3_scale(d3_of(density_g(d, h, pbf)), out_a * masses[j] + rho_a[j] * masses[i]));
                mass_c += rho_a[j] * density_w(d, h, pbf);
            }
            for (uint32_t s = cr[k].bfirst; s < cr[k].blast; ++s) {
                const uint32_t b = g->bidx[s];
                const v3 d = v3_sub(x, boundary_position(g, b));
                const real vol = boundary_volume(g, b);
                position = d3_add(position, d3_scale(d3_of(density_g(d, h, pbf)), out_a * rest[i] * vol));
                rest_c += out_a * vol * density_w(d, h, pbf);
            }
        }
        d3_acc(pos_a, i, position);
        masses_a[i] += mass_c;
        rest_a[i] += rest_c;
    }
}

/* projection_update_vjp_kernel, sph/dfsph.cu:123-139 */
static void projection_update_vjp_kernel(uint32_t n, real dt, real G, const real* masses, const real* rho,
                                         const real* targets, const real* prev, const real* predicted,
                                         const real* relax, const double* p_a, double* masses_a, double* rho_a,
                                         double* targets_a, double* prev_a, double* predicted_a, double* relax_a) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (int64_t)n; ++i) {
        const double error = (double)predicted[i] - targets[i];
        const double volume = (double)masses[i] / rho[i];
        const double delta = 1.0 / (2.0 * dt * dt * volume * volume * G);
        const double candidate = prev[i] + relax[i] * delta * error;
        if (candidate > 0.0) {
            const double adj = p_a[i];
            prev_a[i] += adj;
            predicted_a[i] += relax[i] * delta * adj;
            targets_a[i] -= relax[i] * delta * adj;
            masses_a[i] -= 2.0 * relax[i] * delta * error * adj / masses[i];
            rho_a[i] += 2.0 * relax[i] * delta * error * adj / rho[i];
            relax_a[i] += delta * error * adj;
        }
    }
}

/* When set, predict_vjp_kernel restates the reference's own VJP of the
 * prediction (sph/dfsph.cu:68-102, also integrate_vjp_kernel,
 * sph/common.cu:344-366): a no-slip collision there zeroes the whole velocity
 * adjoint, including dt * xbar_c of the components that did not collide,
 * although those positions still depend on v_c and a_c. That is neither the
 * transpose of predict_jvp_kernel nor the derivative of the prediction; by
 * default the exact transpose is used (the two agree everywhere else). */
static int reference_no_slip_adjoint = 0;
void SUF(oracle_set_reference_no_slip_adjoint)(int enabled) { reference_no_slip_adjoint = enabled != 0; }

/* predict_vjp_kernel, sph/dfsph.cu:68-102 (vs_a may be NULL: zero) */
static void predict_vjp_kernel(uint32_t n, real dt, cbox box, const real* const pos[3], const real* const vel[3],
                               const real* const a1[3], const real* const a2[3], const double* const xs_a[3],
                               const double* const vs_a[3], double* const pos_a[3], double* const vel_a[3],
                               double* const a1_a[3], double* const a2_a[3]) {
#pragma omp parallel for schedule(static)
    for (int64_t ii = 0; ii < (int64_t)n; ++ii) {
        const uint32_t i = (uint32_t)ii;
        const int col = predict_collisions(box, dt, v3_load(pos, i), v3_load(vel, i), v3_load(a1, i), v3_load(a2, i), NULL);
        const d3 xa = d3_cload(xs_a, i);
        const d3 va = vs_a ? d3_cload(vs_a, i) : d3_make(0, 0, 0);
        const d3 position = d3_make(col & 1 ? 0.0 : xa.x, col & 2 ? 0.0 : xa.y, col & 4 ? 0.0 : xa.z);
        d3 velocity;
        if (box.no_slip && col) {
            /* every output velocity component is reset; the positions of the
             * components that did not collide keep x + dt (v + dt a) */
            velocity = reference_no_slip_adjoint ? d3_make(0, 0, 0) : d3_scale(position, dt);
        } else {
            velocity = d3_add(va, d3_scale(position, dt));
            if (col & 1) velocity.x = 0.0;
            if (col & 2) velocity.y = 0.0;
            if (col & 4) velocity.z = 0.0;
        }
        d3_acc(pos_a, i, position);
        d3_acc(vel_a, i, velocity);
        const d3 acc = d3_scale(velocity, dt);
        d3_acc(a1_a, i, acc);
        d3_acc(a2_a, i, acc);
    }
}

static void dfield_copy(dfield* dst, const double* const src[3], uint32_t n) {
    for (int c = 0; c < 3; ++c) memcpy(dst->c[c], src[c], sizeof(double) * n);
}
static void dfield_clear(dfield* f, uint32_t n) {
    for (int c = 0; c < 3; ++c) memset(f->c[c], 0, sizeof(double) * n);
}

void SUF(oracle_dfsph_vjp)(const oracle_domain* d, uint32_t nd, uint32_t ne, int warm_start, int volume_rest,
                           uint64_t step, const real* pos[3], const real* vel[3], const real* ctl[3],
                           const real* const prm[DF_PARAMS], const real* warm_div, const real* warm_den,
                           const double* next_pos_a[3], const double* next_vel_a[3], const double* next_warm_div_a,
                           const double* next_warm_den_a, double* pos_a[3], double* vel_a[3], double* ctl_a[3],
                           double* const prm_a[DF_PARAMS], double* warm_div_a, double* warm_den_a) {
    dfsph_record r;
    const real* const warm_p[2] = {warm_div, warm_den};
    dfsph_record_forward(&r, d, nd, ne, warm_start, volume_rest, step, pos, vel, ctl, prm, warm_p);
    const uint32_t n = r.n;
    const grid_t* g = &r.g;
    const real *m = prm[DF_MASS], *rest = prm[DF_REST];
    double *m_a = prm_a[DF_MASS], *rest_a = prm_a[DF_REST];
    double* rho_a = dzeros(n);
    double* target_a = dzeros(n);
    double* cur_p = dzeros(n);
    double* prev_p = dzeros(n);
    double* rs_a = dzeros(n);
    dfield anp_a = dfield_zeros(n), tot_a = dfield_zeros(n), adiv_a = dfield_zeros(n);
    dfield cur_ap = dfield_zeros(n), prev_ap = dfield_zeros(n), xs_a = dfield_zeros(n);
    vfield acc = alloc_vfield(n);
    /* sph.cpp:671: the last prediction gives the next state */
    predict_vjp_kernel(n, r.dt, r.box, pos, vel, CV(r.anp), CV(r.total), next_pos_a, next_vel_a, pos_a, vel_a, anp_a.c,
                       tot_a.c);
    dfield_copy(&adiv_a, DV(tot_a), n);
    dfield_copy(&cur_ap, DV(tot_a), n);
    for (uint32_t i = 0; i < n; ++i) cur_p[i] = next_warm_den_a ? next_warm_den_a[i] : 0.0;
    double* const warm_a[2] = {warm_div_a, warm_den_a};
    /* reverse_phase (sph.cpp:675-714): density phase, then divergence phase */
    for (int q = 1; q >= 0; --q) {
        if (q == 0) {
            dfield_copy(&cur_ap, DV(adiv_a), n);
            for (uint32_t i = 0; i < n; ++i) cur_p[i] = next_warm_div_a ? next_warm_div_a[i] : 0.0;
        }
        for (uint32_t k = r.iters[q]; k >= 1u; --k) {
            memset(prev_p, 0, sizeof(double) * n);
            memset(rs_a, 0, sizeof(double) * n);
            dfield_clear(&prev_ap, n);
            dfield_clear(&xs_a, n);
            pressure_vjp_kernel(g, pos, m, rest, r.rho, r.p[q][k], DV(cur_ap), pos_a, rho_a, cur_p, m_a, rest_a);
            projection_update_vjp_kernel(n, r.dt, r.G, m, volume_rest ? rest : r.rho, q ? rest : r.rho,
                                         r.p[q][k - 1u], r.rs[q][k], prm[q ? DF_DEN : DF_DIV], cur_p, m_a,
                                         volume_rest ? rest_a : rho_a, q ? rest_a : target_a, prev_p, rs_a,
                                         prm_a[q ? DF_DEN : DF_DIV]);
            density_vjp_kernel(g, 0, pos, CV(r.xs[q][k]), m, rest, rs_a, xs_a.c, m_a, rest_a);
            if (q == 0) {
                predict_vjp_kernel(n, r.dt, r.box, pos, vel, CV(r.anp), CV(r.a[0][k - 1u]), DV(xs_a), NULL, pos_a,
                                   vel_a, anp_a.c, prev_ap.c);
            } else {
                add_kernel(n, CV(r.adiv), CV(r.a[1][k - 1u]), acc.c);
                dfield_clear(&tot_a, n);
                predict_vjp_kernel(n, r.dt, r.box, pos, vel, CV(r.anp), CV(acc), DV(xs_a), NULL, pos_a, vel_a,
                                   anp_a.c, tot_a.c);
                /* add_adjoint_kernel, sph/common.cu:376-388 */
                for (int c = 0; c < 3; ++c)
                    for (uint32_t i = 0; i < n; ++i) {
                        adiv_a.c[c][i] += tot_a.c[c][i];
                        prev_ap.c[c][i] += tot_a.c[c][i];
                    }
            }
            double* t = cur_p;
            cur_p = prev_p;
            prev_p = t;
            dfield tf = cur_ap;
            cur_ap = prev_ap;
            prev_ap = tf;
        }
        pressure_vjp_kernel(g, pos, m, rest, r.rho, r.p[q][0], DV(cur_ap), pos_a, rho_a, cur_p, m_a, rest_a);
        if (warm_start && warm_a[q]) memcpy(warm_a[q], cur_p, sizeof(double) * n);
    }
    for (uint32_t i = 0; i < n; ++i) rho_a[i] += target_a[i];
    non_pressure_vjp_kernel(g, pos, vel, m, prm[DF_VISC], r.rho, DV(anp_a), pos_a, vel_a, ctl_a, rho_a, m_a,
                            prm_a[DF_VISC]);
    /* sph.ixx:364: the state densities */
    density_vjp_kernel(g, 0, pos, pos, m, rest, rho_a, pos_a, m_a, rest_a);
    free(rho_a);
    free(target_a);
    free(cur_p);
    free(prev_p);
    free(rs_a);
    dfield_free(&anp_a);
    dfield_free(&tot_a);
    dfield_free(&adiv_a);
    dfield_free(&cur_ap);
    dfield_free(&prev_ap);
    dfield_free(&xs_a);
    free_vfield(&acc);
    dfsph_record_free(&r);
}

/* ------------------------------------------------------ WCSPH derivatives */
/* SPH<WCSPH>::jvp_step / vjp_step (sph.ixx:356-365) -> WCSPH::jvp / vjp
 * (sph.cpp:228-249): EOS, pressure, artificial viscosity, surface cohesion,
 * external acceleration and integration, differentiated with the neighbour
 * graph, the compression test of every pair and the box-collision masks
 * frozen at the primal step. Parameters: masses, rest densities, viscosities,
 * surface tensions, speed of sound, Tait exponent, boundary surface tension. */
enum { WS_MASS = 0, WS_REST, WS_VISC, WS_TENSION, WS_SPEED, WS_EXP, WS_BTENSION, WS_PARAMS };

/* eos_jvp_kernel, wcsph.cu:102-117 */
static void eos_jvp_kernel(uint32_t n, const real* rho, const real* rho_t, const real* rest, const real* rest_t,
                           const real* speed, const real* speed_t, const real* exponent, const real* exponent_t,
                           real* out) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (int64_t)n; ++i) {
        const real r0 = rest[i];
        const real c = speed[i];
        const real g = exponent[i];
        const real ratio = rho[i] / r0;
        const real powered = R_POW(ratio, g);
        const real difference = powered - K(1.0);
        const real d_rho = c * c * R_POW(ratio, g - K(1.0));
        const real d_rest = c * c * difference / g - c * c * powered;
        const real d_speed = K(2.0) * r0 * c * difference / g;
        const real d_exponent = r0 * c * c * (g * powered * R_LOG(ratio) - difference) / (g * g);
        out[i] = d_rho * rho_t[i] + d_rest * rest_t[i] + d_speed * speed_t[i] + d_exponent * exponent_t[i];
    }
}

/* eos_vjp_kernel, wcsph.cu:119-134 (FP64 from the FP32 primal) */
static void eos_vjp_kernel(uint32_t n, const real* rho, const real* rest, const real* speed, const real* exponent,
                           const double* p_a, double* rho_a, double* rest_a, double* speed_a, double* exponent_a) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (int64_t)n; ++i) {
        const double r0 = rest[i];
        const double c = speed[i];
        const double g = exponent[i];
        const double ratio = (double)rho[i] / r0;
        const double powered = pow(ratio, g);
        const double difference = powered - 1.0;
        const double a = p_a[i];
        rho_a[i] += a * c * c * pow(ratio, g - 1.0);
        rest_a[i] += a * (c * c * difference / g - c * c * powered);
        speed_a[i] += a * 2.0 * r0 * c * difference / g;
        exponent_a[i] += a * r0 * c * c * (g * powered * log(ratio) - difference) / (g * g);
    }
}

/* The tangent of one artificial-viscosity pair term -m pi W'(d) given the
 * tangents of its inputs (viscosity_jvp_kernel, wcsph.cu:197-217 and, for a
 * boundary sample, :226-234). nu = k (alpha c / rho_sum) with k = 2 h (fluid)
 * or h (boundary, where alpha c / rho are particle i's own). */
static v3 viscosity_pair_tangent(v3 d, v3 dd, v3 dv, v3 dvd, real nu, real nu_t, real mass, real mass_t, real h) {
    const real compression = v3_dot(dv, d);
    const real compression_t = v3_dot(dvd, d) + v3_dot(dv, dd);
    const real denominator = v3_dot(d, d) + K(0.01) * h * h;
    const real denominator_t = K(2.0) * v3_dot(d, dd);
    const real pi = -nu * compression / denominator;
    const real pi_t = -(nu_t * compression + nu * compression_t) / denominator +
                      nu * compression * denominator_t / (denominator * denominator);
    const v3 gradient = cubic_gradient(d, h);
    const v3 gradient_t = cubic_gradient_tangent(d, dd, h);
    return v3_add(v3_scale(gradient, -(mass_t * pi + mass * pi_t)), v3_scale(gradient_t, -mass * pi));
}

/* viscosity_jvp_kernel, wcsph.cu:176-237 */
static void artificial_viscosity_jvp_kernel(const grid_t* g, const real* const pos[3], const real* const vel[3],
                                            const real* const pos_t[3], const real* const vel_t[3],
                                            const real* const prm[WS_PARAMS], const real* const prm_t[WS_PARAMS],
                                            const real* rho, const real* rho_t, real* const out[3]) {
    const real h = (real)g->d->support_radius;
    const real *m = prm[WS_MASS], *m_t = prm_t[WS_MASS], *rest = prm[WS_REST], *rest_t = prm_t[WS_REST];
    const real *mu = prm[WS_VISC], *mu_t = prm_t[WS_VISC], *c = prm[WS_SPEED], *c_t = prm_t[WS_SPEED];
#pragma omp parallel for schedule(static)
    for (int64_t ii = 0; ii < (int64_t)g->n; ++ii) {
        const uint32_t i = (uint32_t)ii;
        const v3 p = v3_load(pos, i);
        const v3 v = v3_load(vel, i);
        const v3 pt = v3_load(pos_t, i);
        const v3 vt = v3_load(vel_t, i);
        v3 result = v3_make(0, 0, 0);
        cell_range cr[27];
        const int nr = stencil(g, p, cr);
        for (int k = 0; k < nr; ++k) {
            for (uint32_t s = cr[k].first; s < cr[k].last; ++s) {
                const uint32_t j = g->idx[s];
                if (j == i) continue;
                const v3 d = v3_sub(p, v3_load(pos, j));
                const v3 dv = v3_sub(v, v3_load(vel, j));
                if (v3_dot(dv, d) >= K(0.0)) continue;
                const real alpha = K(0.5) * (mu[i] + mu[j]);
                const real alpha_t = K(0.5) * (mu_t[i] + mu_t[j]);
                const real speed = K(0.5) * (c[i] + c[j]);
                const real speed_t = K(0.5) * (c_t[i] + c_t[j]);
                const real rho_sum = rho[i] + rho[j];
                const real rho_sum_t = rho_t[i] + rho_t[j];
                const real nu = K(2.0) * alpha * h * speed / rho_sum;
                const real nu_t = K(2.0) * h * (alpha_t * speed + alpha * speed_t) / rho_sum - nu * rho_sum_t / rho_sum;
                result = v3_add(result, viscosity_pair_tangent(d, v3_sub(pt, v3_load(pos_t, j)), dv,
                                                               v3_sub(vt, v3_load(vel_t, j)), nu, nu_t, m[j], m_t[j],
                                                               h));
            }
            for (uint32_t s = cr[k].bfirst; s < cr[k].blast; ++s) {
                const uint32_t b = g->bidx[s];
                const v3 d = v3_sub(p, boundary_position(g, b));
                const v3 dv = v3_sub(v, boundary_velocity(g, b));
                if (v3_dot(dv, d) >= K(0.0)) continue;
                const real nu = mu[i] * h * c[i] / rho[i];
                const real nu_t = h * (mu_t[i] * c[i] + mu[i] * c_t[i]) / rho[i] - nu * rho_t[i] / rho[i];
                const real vol = boundary_volume(g, b);
                result = v3_add(result, viscosity_pair_tangent(d, pt, dv, vt, nu, nu_t, rest[i] * vol,
                                                               rest_t[i] * vol, h));
            }
        }
        v3_store(out, i, result);
    }
}

/* The weight of a cohesion pair (surface_forward_kernel, wcsph.cu:324): the
 * kernel value at the distance, clamped below at the particle diameter. */
static real cohesion_weight(v3 d, real h, real diameter) {
    return v3_length(d) > diameter ? cubic(d, h) : cubic(v3_make(diameter, K(0.0), K(0.0)), h);
}

/* surface_jvp_kernel, wcsph.cu:341-386; a pair term is -s m_b w / m_i d with
 * (s, m_b) = (sigma_i, m_j) for a fluid neighbour and (sigma_b,i,
 * rho0_i V_b) for a boundary sample */
static void surface_jvp_kernel(const grid_t* g, const real* const pos[3], const real* const pos_t[3],
                               const real* const prm[WS_PARAMS], const real* const prm_t[WS_PARAMS],
                               real* const out[3]) {
    const real h = (real)g->d->support_radius;
    const real diameter = K(2.0) * (real)g->d->particle_radius;
    const real *m = prm[WS_MASS], *m_t = prm_t[WS_MASS], *rest = prm[WS_REST], *rest_t = prm_t[WS_REST];
    const real *st = prm[WS_TENSION], *st_t = prm_t[WS_TENSION], *bt = prm[WS_BTENSION], *bt_t = prm_t[WS_BTENSION];
#pragma omp parallel for schedule(static)
    for (int64_t ii = 0; ii < (int64_t)g->n; ++ii) {
        const uint32_t i = (uint32_t)ii;
        const v3 p = v3_load(pos, i);
        const v3 pt = v3_load(pos_t, i);
        const real mi = m[i];
        v3 result = v3_make(0, 0, 0);
        cell_range cr[27];
        const int nr = stencil(g, p, cr);
        for (int k = 0; k < nr; ++k) {
            for (uint32_t s = cr[k].first; s < cr[k].last; ++s) {
                const uint32_t j = g->idx[s];
                if (j == i) continue;
                const v3 d = v3_sub(p, v3_load(pos, j));
                const v3 dd = v3_sub(pt, v3_load(pos_t, j));
                const real dist = v3_length(d);
                if (dist >= h) continue;
                const real w = cohesion_weight(d, h, diameter);
                const real w_t = dist > diameter ? v3_dot(cubic_gradient(d, h), dd) : K(0.0);
                const real mj = m[j];
                const real factor = -st[i] * mj * w / mi;
                const real factor_t = -st_t[i] * mj * w / mi - st[i] * m_t[j] * w / mi - st[i] * mj * w_t / mi +
                                      st[i] * mj * w * m_t[i] / (mi * mi);
                result = v3_add(result, v3_add(v3_scale(d, factor_t), v3_scale(dd, factor)));
            }
            for (uint32_t s = cr[k].bfirst; s < cr[k].blast; ++s) {
                const uint32_t b = g->bidx[s];
                const v3 d = v3_sub(p, boundary_position(g, b));
                const real dist = v3_length(d);
                if (dist >= h) continue;
                const real w = cohesion_weight(d, h, diameter);
                const real w_t = dist > diameter ? v3_dot(cubic_gradient(d, h), pt) : K(0.0);
                const real vol = boundary_volume(g, b);
                const real mb = rest[i] * vol;
                const real mb_t = rest_t[i] * vol;
                const real factor = -bt[i] * mb * w / mi;
                const real factor_t = -bt_t[i] * mb * w / mi - bt[i] * mb_t * w / mi - bt[i] * mb * w_t / mi +
                                      bt[i] * mb * w * m_t[i] / (mi * mi);
                result = v3_add(result, v3_add(v3_scale(d, factor_t), v3_scale(pt, factor)));
            }
        }
        v3_store(out, i, result);
    }
}

void SUF(oracle_wcsph_jvp)(const oracle_domain* d, uint64_t step, const real* pos[3], const real* vel[3],
                           const real* ctl[3], const real* const prm[WS_PARAMS], const real* pos_t[3],
                           const real* vel_t[3], const real* ctl_t[3], const real* const prm_t[WS_PARAMS],
                           real* next_pos_t[3], real* next_vel_t[3]) {
    const uint32_t n = d->particle_count;
    const real dt = (real)d->time_step;
    const cbox box = collision_box(d, step);
    grid_t g;
    grid_build(&g, d, step, pos[0], pos[1], pos[2]);
    const real *m = prm[WS_MASS], *rest = prm[WS_REST];
    real* rho = alloc_real(n);
    real* p = alloc_real(n);
    vfield ap = alloc_vfield(n), av = alloc_vfield(n), as = alloc_vfield(n), total = alloc_vfield(n);
    density_kernel(&g, 0, pos, pos, m, rest, rho);
    eos_kernel(n, rho, rest, prm[WS_SPEED], prm[WS_EXP], p);
    pressure_kernel(&g, pos, m, rest, rho, p, ap.c);
    artificial_viscosity_kernel(&g, pos, vel, m, rest, prm[WS_VISC], prm[WS_SPEED], rho, av.c);
    surface_kernel(&g, pos, m, rest, prm[WS_TENSION], prm[WS_BTENSION], as.c);
    add_kernel(n, CV(ap), CV(av), total.c);
    add_kernel(n, CV(total), CV(as), total.c);
    const v3 gravity = v3_make((real)d->gravity[0], (real)d->gravity[1], (real)d->gravity[2]);
    for (uint32_t i = 0; i < n; ++i) v3_store(ap.c, i, v3_add(gravity, v3_load(ctl, i)));
    add_kernel(n, CV(total), CV(ap), total.c);
    /* tangents (sph.cpp:228-237; the state densities' tangent, sph.ixx:357) */
    real* rho_t = alloc_real(n);
    real* p_t = alloc_real(n);
    vfield apt = alloc_vfield(n), avt = alloc_vfield(n), ast = alloc_vfield(n), tt = alloc_vfield(n),
           zero = alloc_vfield(n);
    for (int c = 0; c < 3; ++c) memset(zero.c[c], 0, sizeof(real) * n);
    density_jvp_kernel(&g, 0, pos, pos, pos_t, m, prm_t[WS_MASS], rest, prm_t[WS_REST], rho_t);
    eos_jvp_kernel(n, rho, rho_t, rest, prm_t[WS_REST], prm[WS_SPEED], prm_t[WS_SPEED], prm[WS_EXP], prm_t[WS_EXP],
                   p_t);
    pressure_jvp_kernel(&g, pos, pos_t, m, prm_t[WS_MASS], rest, prm_t[WS_REST], rho, rho_t, p, p_t, apt.c);
    artificial_viscosity_jvp_kernel(&g, pos, vel, pos_t, vel_t, prm, prm_t, rho, rho_t, avt.c);
    surface_jvp_kernel(&g, pos, pos_t, prm, prm_t, ast.c);
    add_kernel(n, CV(apt), CV(avt), tt.c);
    add_kernel(n, CV(tt), CV(ast), tt.c);
    add_kernel(n, CV(tt), ctl_t, tt.c);
    predict_jvp_kernel(n, dt, box, pos, vel, CV(total), CV(zero), pos_t, vel_t, CV(tt), CV(zero), next_pos_t,
                       next_vel_t);
    free(rho);
    free(p);
    free(rho_t);
    free(p_t);
    free_vfield(&ap);
    free_vfield(&av);
    free_vfield(&as);
    free_vfield(&total);
    free_vfield(&apt);
    free_vfield(&avt);
    free_vfield(&ast);
    free_vfield(&tt);
    free_vfield(&zero);
    grid_free(&g);
}

/* viscosity_reverse, wcsph.cu:40-72: the adjoint of one pair term
 * -m_b pi(d, dv) W'(d) with pi = -nu (dv . d) / (|d|^2 + 0.01 h^2),
 * nu = 2 h alpha c / (rho_a + rho_b), for the acceleration adjoint w of its
 * centre; zero for a separating pair. */
typedef struct {
    d3 displacement, velocity;
    double mass_b, rho_a, rho_b, mu_a, mu_b, c_a, c_b;
} viscosity_adjoint_t;

static viscosity_adjoint_t viscosity_reverse(v3 d, v3 dv, real rho_a, real rho_b, real mu_a, real mu_b, real c_a,
                                             real c_b, real mass_b, real h, d3 w) {
    viscosity_adjoint_t r;
    memset(&r, 0, sizeof(r));
    const double compression = v3_dot(dv, d);
    if (compression >= 0.0) return r;
    const double denominator = (double)v3_dot(d, d) + 0.01 * h * h;
    /* the sums of two FP32 primal values are FP32 sums, as in the
     * reference's float arguments */
    const double alpha = 0.5 * (double)(mu_a + mu_b);
    const double speed = 0.5 * (double)(c_a + c_b);
    const double rho_sum = (double)(rho_a + rho_b);
    const double nu = 2.0 * alpha * h * speed / rho_sum;
    const double pi = -nu * compression / denominator;
    const d3 gradient = d3_of(cubic_gradient(d, h));
    r.mass_b = -pi * d3_dot(w, gradient);
    const double pi_a = -(double)mass_b * d3_dot(w, gradient);
    r.displacement = cubic_hessian_product(d, d3_scale(w, -(double)mass_b * pi), h);
    const double nu_a = -compression * pi_a / denominator;
    const double compression_a = -nu * pi_a / denominator;
    const double denominator_a = nu * compression * pi_a / (denominator * denominator);
    r.velocity = d3_scale(d3_of(d), compression_a);
    r.displacement = d3_add(r.displacement, d3_add(d3_scale(d3_of(dv), compression_a),
                                                   d3_scale(d3_of(d), 2.0 * denominator_a)));
    const double alpha_a = 2.0 * h * speed * nu_a / rho_sum;
    const double speed_a = 2.0 * alpha * h * nu_a / rho_sum;
    const double rho_sum_a = -nu * nu_a / rho_sum;
    r.mu_a = r.mu_b = 0.5 * alpha_a;
    r.c_a = r.c_b = 0.5 * speed_a;
    r.rho_a = r.rho_b = rho_sum_a;
    return r;
}

/* viscosity_vjp_kernel, wcsph.cu:239-304, gather form: particle i collects
 * the terms where it is the centre (w_i) and those where it is the
 * neighbour (w_j, with d and dv reversed) */
static void artificial_viscosity_vjp_kernel(const grid_t* g, const real* const pos[3], const real* const vel[3],
                                            const real* const prm[WS_PARAMS], const real* rho,
                                            const double* const acc_a[3], double* const pos_a[3],
                                            double* const vel_a[3], double* rho_a, double* const prm_a[WS_PARAMS]) {
    const real h = (real)g->d->support_radius;
    const real *m = prm[WS_MASS], *rest = prm[WS_REST], *mu = prm[WS_VISC], *c = prm[WS_SPEED];
#pragma omp parallel for schedule(static)
    for (int64_t ii = 0; ii < (int64_t)g->n; ++ii) {
        const uint32_t i = (uint32_t)ii;
        const v3 p = v3_load(pos, i);
        const v3 v = v3_load(vel, i);
        const d3 wi = d3_cload(acc_a, i);
        d3 position = d3_make(0, 0, 0), velocity = d3_make(0, 0, 0);
        double mass_c = 0.0, rest_c = 0.0, rho_c = 0.0, mu_c = 0.0, c_c = 0.0;
        cell_range cr[27];
        const int nr = stencil(g, p, cr);
        for (int k = 0; k < nr; ++k) {
            for (uint32_t s = cr[k].first; s < cr[k].last; ++s) {
                const uint32_t j = g->idx[s];
                if (j == i) continue;
                const v3 d = v3_sub(p, v3_load(pos, j));
                const v3 dv = v3_sub(v, v3_load(vel, j));
                const viscosity_adjoint_t own =
                    viscosity_reverse(d, dv, rho[i], rho[j], mu[i], mu[j], c[i], c[j], m[j], h, wi);
                position = d3_add(position, own.displacement);
                velocity = d3_add(velocity, own.velocity);
                rho_c += own.rho_a;
                mu_c += own.mu_a;
                c_c += own.c_a;
                const viscosity_adjoint_t other =
                    viscosity_reverse(v3_scale(d, K(-1.0)), v3_scale(dv, K(-1.0)), rho[j], rho[i], mu[j], mu[i], c[j],
                                      c[i], m[i], h, d3_cload(acc_a, j));
                position = d3_sub(position, other.displacement);
                velocity = d3_sub(velocity, other.velocity);
                mass_c += other.mass_b;
                rho_c += other.rho_b;
                mu_c += other.mu_b;
                c_c += other.c_b;
            }
            for (uint32_t s = cr[k].bfirst; s < cr[k].blast; ++s) {
                const uint32_t b = g->bidx[s];
                const v3 d = v3_sub(p, boundary_position(g, b));
                const v3 dv = v3_sub(v, boundary_velocity(g, b));
                const real vol = boundary_volume(g, b);
                /* the boundary term's nu = mu_i h c_i / rho_i is the fluid
                 * form with both ends equal to particle i (wcsph.cu:290) */
                const viscosity_adjoint_t own =
                    viscosity_reverse(d, dv, rho[i], rho[i], mu[i], mu[i], c[i], c[i], rest[i] * vol, h, wi);
                position = d3_add(position, own.displacement);
                velocity = d3_add(velocity, own.velocity);
                rest_c += own.mass_b * vol;
                rho_c += own.rho_a + own.rho_b;
                mu_c += own.mu_a + own.mu_b;
                c_c += own.c_a + own.c_b;
            }
        }
        d3_acc(pos_a, i, position);
        d3_acc(vel_a, i, velocity);
        rho_a[i] += rho_c;
        prm_a[WS_MASS][i] += mass_c;
        prm_a[WS_REST][i] += rest_c;
        prm_a[WS_VISC][i] += mu_c;
        prm_a[WS_SPEED][i] += c_c;
    }
}

/* surface_reverse, wcsph.cu:74-88: the adjoint of one cohesion term
 * -s m_b w(d) / m_a d for the acceleration adjoint w of its centre */
typedef struct {
    d3 displacement;
    double mass_a, mass_b, tension;
} surface_adjoint_t;

static surface_adjoint_t surface_reverse(v3 d, real mass_a, real mass_b, real tension, real h, real diameter, d3 w) {
    surface_adjoint_t r;
    memset(&r, 0, sizeof(r));
    const real dist = v3_length(d);
    if (dist >= h) return r;
    const real weight = cohesion_weight(d, h, diameter);
    const double factor = -(double)tension * mass_b * weight / mass_a;
    const double factor_a = d3_dot(w, d3_of(d));
    r.displacement = d3_scale(w, factor);
    if (dist > diameter)
        r.displacement = d3_add(r.displacement, d3_scale(d3_of(cubic_gradient(d, h)),
                                                         -(double)tension * mass_b * factor_a / mass_a));
    r.tension = -(double)mass_b * weight * factor_a / mass_a;
    r.mass_b = -(double)tension * weight * factor_a / mass_a;
    r.mass_a = (double)tension * mass_b * weight * factor_a / ((double)mass_a * mass_a);
    return r;
}

/* surface_vjp_kernel, wcsph.cu:388-435, gather form */
static void surface_vjp_kernel(const grid_t* g, const real* const pos[3], const real* const prm[WS_PARAMS],
                               const double* const acc_a[3], double* const pos_a[3],
                               double* const prm_a[WS_PARAMS]) {
    const real h = (real)g->d->support_radius;
    const real diameter = K(2.0) * (real)g->d->particle_radius;
    const real *m = prm[WS_MASS], *rest = prm[WS_REST], *st = prm[WS_TENSION], *bt = prm[WS_BTENSION];
#pragma omp parallel for schedule(static)
    for (int64_t ii = 0; ii < (int64_t)g->n; ++ii) {
        const uint32_t i = (uint32_t)ii;
        const v3 p = v3_load(pos, i);
        const d3 wi = d3_cload(acc_a, i);
        d3 position = d3_make(0, 0, 0);
        double mass_c = 0.0, rest_c = 0.0, tension_c = 0.0, boundary_c = 0.0;
        cell_range cr[27];
        const int nr = stencil(g, p, cr);
        for (int k = 0; k < nr; ++k) {
            for (uint32_t s = cr[k].first; s < cr[k].last; ++s) {
                const uint32_t j = g->idx[s];
                if (j == i) continue;
                const v3 d = v3_sub(p, v3_load(pos, j));
                const surface_adjoint_t own = surface_reverse(d, m[i], m[j], st[i], h, diameter, wi);
                position = d3_add(position, own.displacement);
                mass_c += own.mass_a;
                tension_c += own.tension;
                const surface_adjoint_t other =
                    surface_reverse(v3_scale(d, K(-1.0)), m[j], m[i], st[j], h, diameter, d3_cload(acc_a, j));
                position = d3_sub(position, other.displacement);
                mass_c += other.mass_b;
            }
            for (uint32_t s = cr[k].bfirst; s < cr[k].blast; ++s) {
                const uint32_t b = g->bidx[s];
                const v3 d = v3_sub(p, boundary_position(g, b));
                const real vol = boundary_volume(g, b);
                const surface_adjoint_t own = surface_reverse(d, m[i], rest[i] * vol, bt[i], h, diameter, wi);
                position = d3_add(position, own.displacement);
                mass_c += own.mass_a;
                rest_c += own.mass_b * vol;
                boundary_c += own.tension;
            }
        }
        d3_acc(pos_a, i, position);
        prm_a[WS_MASS][i] += mass_c;
        prm_a[WS_REST][i] += rest_c;
        prm_a[WS_TENSION][i] += tension_c;
        prm_a[WS_BTENSION][i] += boundary_c;
    }
}

void SUF(oracle_wcsph_vjp)(const oracle_domain* d, uint64_t step, const real* pos[3], const real* vel[3],
                           const real* ctl[3], const real* const prm[WS_PARAMS], const double* next_pos_a[3],
                           const double* next_vel_a[3], double* pos_a[3], double* vel_a[3], double* ctl_a[3],
                           double* const prm_a[WS_PARAMS]) {
    const uint32_t n = d->particle_count;
    const real dt = (real)d->time_step;
    const cbox box = collision_box(d, step);
    grid_t g;
    grid_build(&g, d, step, pos[0], pos[1], pos[2]);
    const real *m = prm[WS_MASS], *rest = prm[WS_REST];
    real* rho = alloc_real(n);
    real* p = alloc_real(n);
    vfield ap = alloc_vfield(n), av = alloc_vfield(n), as = alloc_vfield(n), total = alloc_vfield(n),
           zero = alloc_vfield(n);
    for (int c = 0; c < 3; ++c) memset(zero.c[c], 0, sizeof(real) * n);
    density_kernel(&g, 0, pos, pos, m, rest, rho);
    eos_kernel(n, rho, rest, prm[WS_SPEED], prm[WS_EXP], p);
    pressure_kernel(&g, pos, m, rest, rho, p, ap.c);
    artificial_viscosity_kernel(&g, pos, vel, m, rest, prm[WS_VISC], prm[WS_SPEED], rho, av.c);
    surface_kernel(&g, pos, m, rest, prm[WS_TENSION], prm[WS_BTENSION], as.c);
    add_kernel(n, CV(ap), CV(av), total.c);
    add_kernel(n, CV(total), CV(as), total.c);
    const v3 gravity = v3_make((real)d->gravity[0], (real)d->gravity[1], (real)d->gravity[2]);
    for (uint32_t i = 0; i < n; ++i) v3_store(ap.c, i, v3_add(gravity, v3_load(ctl, i)));
    add_kernel(n, CV(total), CV(ap), total.c);
    /* sph.cpp:241-248 */
    dfield total_a = dfield_zeros(n), scratch = dfield_zeros(n);
    double* rho_a = dzeros(n);
    double* p_a = dzeros(n);
    predict_vjp_kernel(n, dt, box, pos, vel, CV(total), CV(zero), next_pos_a, next_vel_a, pos_a, vel_a, total_a.c,
                       scratch.c);
    for (int c = 0; c < 3; ++c)
        for (uint32_t i = 0; i < n; ++i) ctl_a[c][i] += total_a.c[c][i];
    surface_vjp_kernel(&g, pos, prm, DV(total_a), pos_a, prm_a);
    artificial_viscosity_vjp_kernel(&g, pos, vel, prm, rho, DV(total_a), pos_a, vel_a, rho_a, prm_a);
    pressure_vjp_kernel(&g, pos, m, rest, rho, p, DV(total_a), pos_a, rho_a, p_a, prm_a[WS_MASS], prm_a[WS_REST]);
    eos_vjp_kernel(n, rho, rest, prm[WS_SPEED], prm[WS_EXP], p_a, rho_a, prm_a[WS_REST], prm_a[WS_SPEED],
                   prm_a[WS_EXP]);
    /* sph.ixx:364: the state densities */
    density_vjp_kernel(&g, 0, pos, pos, m, rest, rho_a, pos_a, prm_a[WS_MASS], prm_a[WS_REST]);
    free(rho);
    free(p);
    free(rho_a);
    free(p_a);
    free_vfield(&ap);
    free_vfield(&av);
    free_vfield(&as);
    free_vfield(&total);
    free_vfield(&zero);
    dfield_free(&total_a);
    dfield_free(&scratch);
    grid_free(&g);
}

#ifndef ORACLE_F64
int oracle_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
void oracle_set_threads(int threads) {
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#else
    (void)threads;
#endif
}
#endif
