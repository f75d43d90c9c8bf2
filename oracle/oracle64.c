/*
 * oracle64.c -- FP64 textbook physics oracle and the threaded FP32 CPU
 * baseline of the reference's per-step path.
 *
 * TEST INFRASTRUCTURE ONLY (checker and CPU baseline; never shipped).
 *
 * The reference's force code (src/Forces.jl:6-74) is not valid physics: it
 * broadcasts a scalar distance to all three components, hard-codes
 * eps=-1e10, sigma=1e-4 and subtracts accumulated forces.  Only its entry
 * point shapes (force_lennardjones!, force_coulomb!, sum_forces!) are kept;
 * the maths here is the textbook form:
 *   LJ:      U = 4 eps [(s/r)^12 - (s/r)^6],  F_i = 24 eps/r^2 [2(s/r)^12 - (s/r)^6] r_ij
 *   Coulomb: U = k q_i q_j / r,               F_i = k q_i q_j r_ij / r^3
 * with r_ij = x_i - x_j, each term cut off at its own radius (optionally
 * energy-shifted), virial W = sum_{i<j} r_ij . F_ij.
 * The integrator is kick-drift-force-kick velocity Verlet; the reference's
 * stale-force variant (src/Simulator.jl:198-225) is replaced by the standard
 * scheme.  Reflecting walls follow boundary_reflect! (src/Simulator.jl:81-111):
 * a coordinate outside [lo, hi] is clamped and that velocity component negated.
 */
#define _GNU_SOURCE
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "refcpu.h"

static inline void o64_pair(const o64_params *p, double qi, double qj, double dx, double dy,
                            double dz, double r2, double *f, double *u) {
    double fs = 0.0, e = 0.0;
    double r = sqrt(r2);
    if (p->rc_lj > 0.0 && r < p->rc_lj) {
        double sr2 = (p->sigma * p->sigma) / r2;
        double sr6 = sr2 * sr2 * sr2;
        double sr12 = sr6 * sr6;
        fs += 24.0 * p->eps * (2.0 * sr12 - sr6) / r2;
        e += 4.0 * p->eps * (sr12 - sr6);
        if (p->shift_lj) {
            double c2 = (p->sigma * p->sigma) / (p->rc_lj * p->rc_lj);
            double c6 = c2 * c2 * c2;
            e -= 4.0 * p->eps * (c6 * c6 - c6);
        }
    }
    if (p->rc_coul > 0.0 && r < p->rc_coul) {
        double kqq = p->k_coul * qi * qj;
        fs += kqq / (r2 * r);
        e += kqq / r;
        if (p->shift_coul) e -= kqq / p->rc_coul;
    }
    f[0] = fs * dx;
    f[1] = fs * dy;
    f[2] = fs * dz;
    *u = e;
}

void o64_forces_pairs(const double *x, const double *q, int32_t N, const int32_t *pi,
                      const int32_t *pj, int64_t npairs, const o64_params *p, double *F,
                      double *U, double *W) {
    memset(F, 0, sizeof(double) * 3 * (size_t)N);
    double u = 0.0, w = 0.0;
    double rc = fmax(p->rc_lj, p->rc_coul);
    for (int64_t k = 0; k < npairs; ++k) {
        int32_t i = pi[k], j = pj[k];
        double dx = x[3 * i] - x[3 * j], dy = x[3 * i + 1] - x[3 * j + 1],
               dz = x[3 * i + 2] - x[3 * j + 2];
        double r2 = dx * dx + dy * dy + dz * dz;
        if (!(r2 < rc * rc)) continue;
        double f[3], e;
        o64_pair(p, q ? q[i] : 0.0, q ? q[j] : 0.0, dx, dy, dz, r2, f, &e);
        for (int d = 0; d < 3; ++d) { F[3 * i + d] += f[d]; F[3 * j + d] -= f[d]; }
        u += e;
        w += f[0] * dx + f[1] * dy + f[2] * dz;
    }
    if (U) *U = u;
    if (W) *W = w;
}

/* Uniform cell list over the bounding box; cell edge >= rc. */
typedef struct {
    int32_t nc[3];
    double lo[3], inv[3];
    int32_t *head, *next;
} cells;

static void cells_build(cells *c, const double *x, int32_t N, double rc) {
    double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int32_t i = 0; i < N; ++i)
        for (int d = 0; d < 3; ++d) {
            lo[d] = fmin(lo[d], x[3 * i + d]);
            hi[d] = fmax(hi[d], x[3 * i + d]);
        }
    int64_t total = 1;
    for (int d = 0; d < 3; ++d) {
        double ext = hi[d] - lo[d];
        int32_t n = (int32_t)floor(ext / rc);
        if (n < 1) n = 1;
        if (n > 512) n = 512;
        c->nc[d] = n;
        c->lo[d] = lo[d];
        c->inv[d] = ext > 0 ? n / ext : 0.0;
        total *= n;
    }
    c->head = (int32_t *)malloc(sizeof(int32_t) * (size_t)total);
    c->next = (int32_t *)malloc(sizeof(int32_t) * (size_t)N);
    for (int64_t k = 0; k < total; ++k) c->head[k] = -1;
    for (int32_t i = N - 1; i >= 0; --i) {
        int32_t ci[3];
        for (int d = 0; d < 3; ++d) {
            int32_t t = (int32_t)floor((x[3 * i + d] - c->lo[d]) * c->inv[d]);
            ci[d] = t < 0 ? 0 : (t >= c->nc[d] ? c->nc[d] - 1 : t);
        }
        int64_t k = ((int64_t)ci[2] * c->nc[1] + ci[1]) * c->nc[0] + ci[0];
        c->next[i] = c->head[k];
        c->head[k] = i;
    }
}

static inline int32_t cell_of(const cells *c, const double *xi, int d) {
    int32_t t = (int32_t)floor((xi[d] - c->lo[d]) * c->inv[d]);
    return t < 0 ? 0 : (t >= c->nc[d] ? c->nc[d] - 1 : t);
}

typedef struct {
    const double *x, *q;
    int32_t N, lo_i, hi_i;
    const o64_params *p;
    const cells *c;
    double *F; /* private N x 3 */
    double u, w;
    int64_t npairs;
} o64_task;

static void *o64_worker(void *arg) {
    o64_task *t = (o64_task *)arg;
    const double *x = t->x;
    const cells *c = t->c;
    double rc = fmax(t->p->rc_lj, t->p->rc_coul);
    double rc2 = rc * rc;
    for (int32_t i = t->lo_i; i < t->hi_i; ++i) {
        int32_t ci[3];
        for (int d = 0; d < 3; ++d) ci[d] = cell_of(c, x + 3 * (int64_t)i, d);
        for (int32_t dz = -1; dz <= 1; ++dz)
            for (int32_t dy = -1; dy <= 1; ++dy)
                for (int32_t dx = -1; dx <= 1; ++dx) {
                    int32_t cx = ci[0] + dx, cy = ci[1] + dy, cz = ci[2] + dz;
                    if (cx < 0 || cy < 0 || cz < 0 || cx >= c->nc[0] || cy >= c->nc[1] ||
                        cz >= c->nc[2])
                        continue;
                    int64_t k = ((int64_t)cz * c->nc[1] + cy) * c->nc[0] + cx;
                    for (int32_t j = c->head[k]; j >= 0; j = c->next[j]) {
                        if (j <= i) continue;
                        double rx = x[3 * i] - x[3 * j], ry = x[3 * i + 1] - x[3 * j + 1],
                               rz = x[3 * i + 2] - x[3 * j + 2];
                        double r2 = rx * rx + ry * ry + rz * rz;
                        if (!(r2 < rc2)) continue;
                        double f[3], e;
                        o64_pair(t->p, t->q ? t->q[i] : 0.0, t->q ? t->q[j] : 0.0, rx, ry, rz,
                                 r2, f, &e);
                        for (int d = 0; d < 3; ++d) {
                            t->F[3 * i + d] += f[d];
                            t->F[3 * j + d] -= f[d];
                        }
                        t->u += e;
                        t->w += f[0] * rx + f[1] * ry + f[2] * rz;
                        t->npairs++;
                    }
                }
    }
    return NULL;
}

#define O64_THREADS 8

int64_t o64_forces(const double *x, const double *q, int32_t N, const o64_params *p,
                   double *F, double *U, double *W) {
    double rc = fmax(p->rc_lj, p->rc_coul);
    cells c;
    cells_build(&c, x, N, rc);
    int nt = N < 4096 ? 1 : O64_THREADS;
    o64_task tk[O64_THREADS];
    pthread_t th[O64_THREADS];
    int32_t per = (N + nt - 1) / nt;
    for (int t = 0; t < nt; ++t) {
        int32_t lo = t * per, hi = lo + per > N ? N : lo + per;
        tk[t] = (o64_task){x, q, N, lo, hi, p, &c, NULL, 0.0, 0.0, 0};
        tk[t].F = (double *)calloc(3 * (size_t)N, sizeof(double));
        if (nt == 1) o64_worker(&tk[t]);
        else pthread_create(&th[t], NULL, o64_worker, &tk[t]);
    }
    if (nt > 1)
        for (int t = 0; t < nt; ++t) pthread_join(th[t], NULL);
    memset(F, 0, sizeof(double) * 3 * (size_t)N);
    double u = 0.0, w = 0.0;
    int64_t np = 0;
    for (int t = 0; t < nt; ++t) {
        for (int64_t k = 0; k < 3 * (int64_t)N; ++k) F[k] += tk[t].F[k];
        u += tk[t].u;
        w += tk[t].w;
        np += tk[t].npairs;
        free(tk[t].F);
    }
    free(c.head);
    free(c.next);
    if (U) *U = u;
    if (W) *W = w;
    return np;
}

/* Velocity rescaling of rescale_velocity! (src/Simulator.jl:119-144), called
 * every thermostat_every steps (src/Simulator.jl:241-243), with the
 * temperature from the kinetic energy, T = 2 K / (3 N) (k_B = 1; the
 * reference sums |v| instead of v^2, Simulator.jl:129-132):
 *   beta = sqrt(1 + gamma (T0 / T - 1)),  v <- beta v.
 * It acts on the full-step velocity v_s after step s's force, before the
 * kick-drift that starts step s + 1 (s > 0 and s a multiple of the period),
 * so a rescale at the last step of a run is not applied within that run. */
static void o64_rescale(double *v, int32_t N, double K, double t0, double gamma) {
    const double T = 2.0 * K / (3.0 * (double)N);
    double beta = 1.0;
    if (T > 0.0) {
        const double b2 = 1.0 + gamma * (t0 / T - 1.0);
        beta = sqrt(b2 > 0.0 ? b2 : 0.0);
    }
    for (int64_t k = 0; k < 3 * (int64_t)N; ++k) v[k] *= beta;
}

int o64_vv_run_thermo(double *x, double *v, const double *q, const double *m, int32_t N,
                      const o64_params *p, double dt, int32_t nsteps, int32_t walls, double lo,
                      double hi, int32_t thermostat_every, double t0, double gamma, double *F,
                      double *energy) {
    double U, W;
    o64_forces(x, q, N, p, F, &U, &W);
    double K = 0.0;
    for (int32_t i = 0; i < N; ++i)
        K += 0.5 * m[i] * (v[3 * i] * v[3 * i] + v[3 * i + 1] * v[3 * i + 1] + v[3 * i + 2] * v[3 * i + 2]);
    if (energy) { energy[0] = K; energy[1] = U; energy[2] = W; }
    for (int32_t s = 1; s <= nsteps; ++s) {
        if (thermostat_every > 0 && s - 1 > 0 && (s - 1) % thermostat_every == 0)
            o64_rescale(v, N, K, t0, gamma);
        for (int32_t i = 0; i < N; ++i) {
            double h = 0.5 * dt / m[i];
            for (int d = 0; d < 3; ++d) {
                v[3 * i + d] += h * F[3 * i + d];
                x[3 * i + d] += dt * v[3 * i + d];
                if (walls) {
                    if (x[3 * i + d] < lo) { x[3 * i + d] = lo; v[3 * i + d] = -v[3 * i + d]; }
                    if (x[3 * i + d] > hi) { x[3 * i + d] = hi; v[3 * i + d] = -v[3 * i + d]; }
                }
            }
        }
        o64_forces(x, q, N, p, F, &U, &W);
        K = 0.0;
        for (int32_t i = 0; i < N; ++i) {
            double h = 0.5 * dt / m[i];
            for (int d = 0; d < 3; ++d) v[3 * i + d] += h * F[3 * i + d];
            K += 0.5 * m[i] * (v[3 * i] * v[3 * i] + v[3 * i + 1] * v[3 * i + 1] + v[3 * i + 2] * v[3 * i + 2]);
        }
        if (energy) { energy[3 * s] = K; energy[3 * s + 1] = U; energy[3 * s + 2] = W; }
    }
    return REF_OK;
}

int o64_vv_run(double *x, double *v, const double *q, const double *m, int32_t N,
               const o64_params *p, double dt, int32_t nsteps, int32_t walls, double lo,
               double hi, double *F, double *energy) {
    return o64_vv_run_thermo(x, v, q, m, N, p, dt, nsteps, walls, lo, hi, 0, 0.0, 0.0, F, energy);
}

/* ------------------------------------------------------------------------- */
/* Threaded FP32 CPU baseline: the reference's neighbour path                 */
/* (leafbuild_traverse_bvh with Polyester-style strided threads), textbook    */
/* LJ + Coulomb from the pair list with per-thread force arrays reduced at   */
/* the end (docs/src/devdiary.md:1236), and FP32 velocity Verlet.             */
/* ------------------------------------------------------------------------- */
static double now_s(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

typedef struct {
    const float *x, *q;
    const ref_pair *pairs;
    int64_t lo, hi;
    const base_params *p;
    float *F;
    int64_t inrange;
} bf_task;

static void *bf_worker(void *arg) {
    bf_task *t = (bf_task *)arg;
    const base_params *p = t->p;
    const float rc2 = p->rc * p->rc, rcc2 = p->rc_coul * p->rc_coul, s2 = p->sigma * p->sigma;
    for (int64_t k = t->lo; k < t->hi; ++k) {
        int32_t i = t->pairs[k].i, j = t->pairs[k].j;
        float dx = t->x[3 * i] - t->x[3 * j], dy = t->x[3 * i + 1] - t->x[3 * j + 1],
              dz = t->x[3 * i + 2] - t->x[3 * j + 2];
        float r2 = dx * dx + dy * dy + dz * dz;
        float fs = 0.0f;
        int in = 0;
        if (r2 < rc2) {
            float ir2 = 1.0f / r2, sr2 = s2 * ir2, sr6 = sr2 * sr2 * sr2;
            fs += 24.0f * p->eps * ir2 * sr6 * (2.0f * sr6 - 1.0f);
            in = 1;
        }
        if (t->q && r2 < rcc2) {
            float ir = 1.0f / sqrtf(r2);
            fs += p->k_coul * t->q[i] * t->q[j] * ir * ir * ir;
            in = 1;
        }
        t->inrange += in;
        t->F[3 * i] += fs * dx; t->F[3 * i + 1] += fs * dy; t->F[3 * i + 2] += fs * dz;
        t->F[3 * j] -= fs * dx; t->F[3 * j + 1] -= fs * dy; t->F[3 * j + 2] -= fs * dz;
    }
    return NULL;
}

static int64_t base_forces(const float *x, const float *q, int32_t N, const ref_pair *pairs,
                           int64_t np, const base_params *p, float *F, float *scratch) {
    int nt = p->nthreads < 1 ? 1 : (p->nthreads > 256 ? 256 : p->nthreads);
    bf_task tk[256];
    pthread_t th[256];
    int64_t per = (np + nt - 1) / nt;
    for (int t = 0; t < nt; ++t) {
        float *Ft = scratch + (size_t)t * 3 * (size_t)N;
        memset(Ft, 0, sizeof(float) * 3 * (size_t)N);
        int64_t lo = t * per, hi = lo + per > np ? np : lo + per;
        if (lo > hi) lo = hi;
        tk[t] = (bf_task){x, q, pairs, lo, hi, p, Ft, 0};
        if (nt == 1) bf_worker(&tk[t]);
        else pthread_create(&th[t], NULL, bf_worker, &tk[t]);
    }
    if (nt > 1)
        for (int t = 0; t < nt; ++t) pthread_join(th[t], NULL);
    int64_t in = 0;
    memset(F, 0, sizeof(float) * 3 * (size_t)N);
    for (int t = 0; t < nt; ++t) {
        const float *Ft = scratch + (size_t)t * 3 * (size_t)N;
        for (int64_t k = 0; k < 3 * (int64_t)N; ++k) F[k] += Ft[k];
        in += tk[t].inrange;
    }
    return in;
}

int base_run(float *x, float *v, const float *q, int32_t N, const base_params *p,
             int32_t nsteps, double *times, int64_t *pairs_eval) {
    if (p->apl < 1 || N % p->apl != 0 || N / p->apl < 2) return REF_ERR_ARG;
    int nt = p->nthreads < 1 ? 1 : (p->nthreads > 256 ? 256 : p->nthreads);
    float *F = (float *)calloc(3 * (size_t)N, sizeof(float));
    float *scratch = (float *)malloc(sizeof(float) * 3 * (size_t)N * (size_t)nt);
    ref_pair *pairs = NULL;
    int64_t np = 0, cap = 0;
    float lo3[3] = {p->lo, p->lo, p->lo};
    float sc = 1024.0f / (p->hi - p->lo);
    float sc3[3] = {sc, sc, sc};
    double tm[4] = {0, 0, 0, 0};
    int64_t pe = 0;
    const float rlist = p->rc + p->skin;
    const int32_t every = p->rebuild_every < 1 ? 1 : p->rebuild_every;
    for (int32_t s = 0; s <= nsteps; ++s) {
        /* s == 0 computes the initial forces; steps 1..nsteps integrate */
        if (s < nsteps && s % every == 0) {
            double tt[3];
            for (;;) {
                int64_t got = ref_leafbuild_traverse_bvh(x, N, p->apl, rlist, p->predicate,
                                                         p->morton_mode, lo3, sc3, pairs, cap,
                                                         nt, NULL, tt);
                if (got < 0) { free(F); free(scratch); free(pairs); return (int)got; }
                tm[0] += tt[0]; tm[1] += tt[1]; tm[2] += tt[2];
                if (got <= cap) { np = got; break; }
                cap = got + got / 8 + 1024;
                pairs = (ref_pair *)realloc(pairs, sizeof(ref_pair) * (size_t)cap);
            }
        }
        double t0 = now_s();
        int64_t in = base_forces(x, q, N, pairs, np, p, F, scratch);
        if (s > 0) pe += in;
        if (s == nsteps) {
            for (int64_t k = 0; k < 3 * (int64_t)N; ++k) v[k] += 0.5f * p->dt * F[k];
            tm[3] += now_s() - t0;
            break;
        }
        /* finish step s (second half kick), start step s+1 (kick + drift) */
        for (int32_t i = 0; i < N; ++i)
            for (int d = 0; d < 3; ++d) {
                float vv = v[3 * i + d];
                if (s > 0) vv += 0.5f * p->dt * F[3 * i + d];
                vv += 0.5f * p->dt * F[3 * i + d];
                float xx = x[3 * i + d] + p->dt * vv;
                if (p->walls) {
                    if (xx < p->lo) { xx = p->lo; vv = -vv; }
                    if (xx > p->hi) { xx = p->hi; vv = -vv; }
                }
                v[3 * i + d] = vv;
                x[3 * i + d] = xx;
            }
        tm[3] += now_s() - t0;
    }
    if (times) memcpy(times, tm, sizeof(tm));
    if (pairs_eval) *pairs_eval = pe;
    free(F); free(scratch); free(pairs);
    return REF_OK;
}
