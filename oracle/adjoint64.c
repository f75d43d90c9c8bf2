/* adjoint64.c -- FP64 CPU oracle of the differentiable simulation (TEST
 * INFRASTRUCTURE ONLY: the checker of nd_adjoint.hip, used by tests/).
 *
 * Same model as the library's adjoint: unit masses, no walls, velocity
 * Verlet (kick-drift-kick), shifted-force potentials
 *   U_sf(r) = U(r) - U(rc) - (r - rc) U'(rc),  U_LJ = 4 eps ((s/r)^12 - (s/r)^6),
 *   U_C = k q_i q_j / r,
 * and the loss L = sum_b (g_b - t_b)^2 of a smoothed radial distribution of
 * the final frame (kernel (1 - (d/h)^2)^3 on |d| < h).  It is written
 * independently of the GPU code: all pairs by brute force, the radial
 * Hessian grad grad U = U'' r^ r^T + (U'/r)(I - r^ r^T) instead of the
 * s(r) form, the whole trajectory stored, and the reverse sweep taken step by
 * step with two Jacobian products per step (no merging).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define ADJ64_PI 3.14159265358979323846

typedef struct {
    double eps, sigma, rc, k_coul, rc_coul, dt, rdf_rmax, rdf_h, volume;
    int nbins, nsteps;
} adj64_params;

/* radial derivatives of the shifted-force pair potential: U', U'' and the
 * parameter derivatives of U' (eps, sigma, and for Coulomb per unit q_i q_j) */
typedef struct {
    double d1, d2, d1_eps, d1_sig, d1_qq;
} radial;

static double lj_d1(double r, double e, double s) {  /* dU_LJ/dr */
    double sr6 = pow(s / r, 6.0);
    return 4.0 * e * (-12.0 * sr6 * sr6 + 6.0 * sr6) / r;
}
static double lj_d2(double r, double e, double s) {
    double sr6 = pow(s / r, 6.0);
    return 4.0 * e * (156.0 * sr6 * sr6 - 42.0 * sr6) / (r * r);
}
static double lj_d1_sig(double r, double e, double s) {  /* d/dsigma of dU_LJ/dr */
    double sr6 = pow(s / r, 6.0);
    return 4.0 * e * (-144.0 * sr6 * sr6 + 36.0 * sr6) / (r * s);
}

static radial pair_radial(double r, double qq, const adj64_params *p) {
    radial o = {0, 0, 0, 0, 0};
    if (r < p->rc) {
        o.d1 += lj_d1(r, p->eps, p->sigma) - lj_d1(p->rc, p->eps, p->sigma);
        o.d2 += lj_d2(r, p->eps, p->sigma);
        o.d1_eps = (lj_d1(r, p->eps, p->sigma) - lj_d1(p->rc, p->eps, p->sigma)) / p->eps;
        o.d1_sig = lj_d1_sig(r, p->eps, p->sigma) - lj_d1_sig(p->rc, p->eps, p->sigma);
    }
    if (p->rc_coul > 0.0 && r < p->rc_coul) {
        double u1 = -p->k_coul / (r * r) + p->k_coul / (p->rc_coul * p->rc_coul);  /* per qq */
        o.d1 += qq * u1;
        o.d2 += qq * 2.0 * p->k_coul / (r * r * r);
        o.d1_qq = u1;
    }
    return o;
}

static double reach(const adj64_params *p) {
    double r = p->rc;
    if (p->rc_coul > r) r = p->rc_coul;
    return r;
}

/* F = -grad U */
static void forces(const double *x, const double *q, int n, const adj64_params *p, double *F) {
    memset(F, 0, sizeof(double) * 3 * (size_t)n);
    double R = reach(p);
    for (int i = 0; i < n; ++i)
        for (int j = i + 1; j < n; ++j) {
            double d[3], r2 = 0.0;
            for (int c = 0; c < 3; ++c) { d[c] = x[3 * i + c] - x[3 * j + c]; r2 += d[c] * d[c]; }
            if (r2 >= R * R) continue;
            double r = sqrt(r2);
            radial u = pair_radial(r, q ? q[i] * q[j] : 0.0, p);
            for (int c = 0; c < 3; ++c) {
                double f = -u.d1 * d[c] / r;
                F[3 * i + c] += f;
                F[3 * j + c] -= f;
            }
        }
}

/* The shifted-force potential energy from its definition,
 * U_sf(r) = U(r) - U(rc) - (r - rc) U'(rc) for r < rc, summed over pairs:
 * U_LJ = 4 eps ((s/r)^12 - (s/r)^6) and U'_LJ(rc) from its own closed form
 * here, U_C = k q_i q_j / r, U'_C = -k q_i q_j / r^2.  It is the function the
 * forces above must be minus the gradient of. */
static double lj_u(double r, double e, double s) {
    double sr6 = pow(s / r, 6.0);
    return 4.0 * e * (sr6 * sr6 - sr6);
}
static double energy(const double *x, const double *q, int n, const adj64_params *p) {
    double U = 0.0, R = reach(p);
    double dlj_rc = 4.0 * p->eps * (-12.0 * pow(p->sigma / p->rc, 12.0) + 6.0 * pow(p->sigma / p->rc, 6.0)) / p->rc;
    for (int i = 0; i < n; ++i)
        for (int j = i + 1; j < n; ++j) {
            double r2 = 0.0;
            for (int c = 0; c < 3; ++c) { double d = x[3 * i + c] - x[3 * j + c]; r2 += d * d; }
            if (r2 >= R * R) continue;
            double r = sqrt(r2);
            if (r < p->rc)
                U += lj_u(r, p->eps, p->sigma) - lj_u(p->rc, p->eps, p->sigma) - (r - p->rc) * dlj_rc;
            if (q && p->rc_coul > 0.0 && r < p->rc_coul) {
                double kqq = p->k_coul * q[i] * q[j];
                U += kqq / r - kqq / p->rc_coul + (r - p->rc_coul) * kqq / (p->rc_coul * p->rc_coul);
            }
        }
    return U;
}

/* The oracle's forward-model forces F = -grad U_sf (n x 3) and U_sf, for tests. */
void adj64_forces(const double *x, const double *q, int n, const adj64_params *p, double *F, double *U) {
    forces(x, q, n, p, F);
    if (U) *U = energy(x, q, n, p);
}

/* lam += J(x)^T a (J = dF/dx = -Hess U, symmetric); grads += (dF/dtheta)^T a */
static void jt_product(const double *x, const double *q, int n, const adj64_params *p,
                       const double *a, double *lam, double *dq, double *dpar) {
    double R = reach(p);
    for (int i = 0; i < n; ++i)
        for (int j = i + 1; j < n; ++j) {
            double d[3], r2 = 0.0;
            for (int c = 0; c < 3; ++c) { d[c] = x[3 * i + c] - x[3 * j + c]; r2 += d[c] * d[c]; }
            if (r2 >= R * R) continue;
            double r = sqrt(r2), e[3];
            for (int c = 0; c < 3; ++c) e[c] = d[c] / r;
            radial u = pair_radial(r, q ? q[i] * q[j] : 0.0, p);
            double da[3], eda = 0.0;
            for (int c = 0; c < 3; ++c) { da[c] = a[3 * i + c] - a[3 * j + c]; eda += e[c] * da[c]; }
            /* Hess_pair * da = U'' e (e.da) + (U'/r)(da - e (e.da)); F_i = -grad_i U */
            for (int c = 0; c < 3; ++c) {
                double h = u.d2 * e[c] * eda + (u.d1 / r) * (da[c] - e[c] * eda);
                lam[3 * i + c] -= h;
                lam[3 * j + c] += h;
            }
            /* dF_i/dtheta = -(dU'/dtheta) e, dF_j/dtheta = +(dU'/dtheta) e */
            dpar[0] += -u.d1_eps * eda;
            dpar[1] += -u.d1_sig * eda;
            if (q && u.d1_qq != 0.0) {
                dq[i] += -u.d1_qq * q[j] * eda;
                dq[j] += -u.d1_qq * q[i] * eda;
            }
        }
}

static double kern(double d, double h) {
    double u = d / h, t = 1.0 - u * u;
    return fabs(u) < 1.0 ? t * t * t : 0.0;
}
static double dkern(double d, double h) {
    double u = d / h, t = 1.0 - u * u;
    return fabs(u) < 1.0 ? -6.0 * u * t * t / h : 0.0;
}

static void rdf(const double *x, int n, const adj64_params *p, double *g) {
    double dr = p->rdf_rmax / p->nbins, npair = 0.5 * n * (n - 1.0);
    for (int b = 0; b < p->nbins; ++b) g[b] = 0.0;
    for (int i = 0; i < n; ++i)
        for (int j = i + 1; j < n; ++j) {
            double r2 = 0.0;
            for (int c = 0; c < 3; ++c) { double d = x[3 * i + c] - x[3 * j + c]; r2 += d * d; }
            double r = sqrt(r2);
            for (int b = 0; b < p->nbins; ++b) g[b] += kern(r - (b + 0.5) * dr, p->rdf_h);
        }
    for (int b = 0; b < p->nbins; ++b) {
        double rb = (b + 0.5) * dr;
        g[b] *= p->volume / (npair * 4.0 * ADJ64_PI * rb * rb * dr);
    }
}

/* Forward run (and, if dx0 != NULL, the gradients).  Returns 0 on success. */
int adj64_run(const double *x0, const double *v0, const double *q, int n, const adj64_params *p,
              const double *target, double *rdf_out, double *loss, double *dx0, double *dv0,
              double *dq, double *dpar) {
    int T = p->nsteps;
    size_t m = 3 * (size_t)n;
    double *X = malloc(sizeof(double) * m * (size_t)(T + 1));
    double *v = malloc(sizeof(double) * m), *F = malloc(sizeof(double) * m);
    double *g = malloc(sizeof(double) * (size_t)p->nbins);
    if (!X || !v || !F || !g) { free(X); free(v); free(F); free(g); return -1; }
    double c = 0.5 * p->dt;
    memcpy(X, x0, sizeof(double) * m);
    memcpy(v, v0, sizeof(double) * m);
    forces(X, q, n, p, F);
    for (int k = 0; k < T; ++k) {
        double *xk = X + m * (size_t)k, *x1 = X + m * (size_t)(k + 1);
        for (size_t t = 0; t < m; ++t) {
            v[t] += c * F[t];
            x1[t] = xk[t] + p->dt * v[t];
        }
        forces(x1, q, n, p, F);
        for (size_t t = 0; t < m; ++t) v[t] += c * F[t];
    }
    const double *xT = X + m * (size_t)T;
    rdf(xT, n, p, g);
    double L = 0.0;
    if (target)
        for (int b = 0; b < p->nbins; ++b) L += (g[b] - target[b]) * (g[b] - target[b]);
    if (rdf_out) memcpy(rdf_out, g, sizeof(double) * (size_t)p->nbins);
    if (loss) *loss = L;
    if (dx0 && target) {
        double *lx = calloc(m, sizeof(double)), *lv = calloc(m, sizeof(double));
        double *a = malloc(sizeof(double) * m);
        double dpar_[2] = {0.0, 0.0};
        double *dq_ = calloc((size_t)n, sizeof(double));
        /* dL/dx_T */
        double dr = p->rdf_rmax / p->nbins, npair = 0.5 * n * (n - 1.0);
        for (int i = 0; i < n; ++i)
            for (int j = i + 1; j < n; ++j) {
                double d[3], r2 = 0.0;
                for (int cc = 0; cc < 3; ++cc) { d[cc] = xT[3 * i + cc] - xT[3 * j + cc]; r2 += d[cc] * d[cc]; }
                double r = sqrt(r2), w = 0.0;
                for (int b = 0; b < p->nbins; ++b) {
                    double rb = (b + 0.5) * dr;
                    double nb = p->volume / (npair * 4.0 * ADJ64_PI * rb * rb * dr);
                    w += 2.0 * (g[b] - target[b]) * nb * dkern(r - rb, p->rdf_h);
                }
                for (int cc = 0; cc < 3; ++cc) {
                    lx[3 * i + cc] += w * d[cc] / r;
                    lx[3 * j + cc] -= w * d[cc] / r;
                }
            }
        /* reverse sweep, one step at a time */
        for (int k = T - 1; k >= 0; --k) {
            const double *xk = X + m * (size_t)k, *x1 = X + m * (size_t)(k + 1);
            /* v_{k+1} = v_h + c F(x_{k+1}) */
            for (size_t t = 0; t < m; ++t) a[t] = c * lv[t];
            jt_product(x1, q, n, p, a, lx, dq_, dpar_);
            /* x_{k+1} = x_k + dt v_h: lambda_h = lambda_v + dt lambda_x */
            for (size_t t = 0; t < m; ++t) lv[t] += p->dt * lx[t];
            /* v_h = v_k + c F(x_k) */
            for (size_t t = 0; t < m; ++t) a[t] = c * lv[t];
            jt_product(xk, q, n, p, a, lx, dq_, dpar_);
        }
        memcpy(dx0, lx, sizeof(double) * m);
        if (dv0) memcpy(dv0, lv, sizeof(double) * m);
        if (dq) memcpy(dq, dq_, sizeof(double) * (size_t)n);
        if (dpar) { dpar[0] = dpar_[0]; dpar[1] = dpar_[1]; }
        free(lx); free(lv); free(a); free(dq_);
    }
    free(X); free(v); free(F); free(g);
    return 0;
}
