// nd_adjoint.hip -- differentiable velocity Verlet with a hand-written adjoint.
//
// Forward model (one system, unit masses, no walls):
//   v_h = v_k + c F(x_k),  x_{k+1} = x_k + dt v_h,  v_{k+1} = v_h + c F(x_{k+1}),  c = dt / 2
// with shifted-force Lennard-Jones and shifted-force Coulomb pair forces
//   F_ij = s(r) r_ij,  s(r) = (F(r) - F(rc)) / r   (F(r) = -dU/dr)
// so every force goes to zero continuously at its cutoff.  The loss is
//   L = sum_b (g_b - t_b)^2,  g_b = norm_b sum_{i<j} K(r_ij - r_b),
//   K(d) = (1 - (d/h)^2)^3 for |d| < h (zero outside), r_b = (b + 1/2) dr,
// a smoothed radial distribution of the final frame against a target t.
//
// Reverse sweep.  With lambda_x, lambda_v the adjoints of x_{k+1}, v_{k+1}:
//   lambda_h = lambda_v + dt lambda_x;  lambda_v <- lambda_h
//   a = w_k lambda_h  (w_k = 2c for 0 < k, c for k = 0)
//   lambda_x <- lambda_x + J(x_k) a,   theta_bar += (dF(x_k)/dtheta)^T a
// J = dF/dx is symmetric, so the two Jacobian terms of a step that meet at
// x_k (from v_{k+1} = v_h + c F(x_{k+1}) of step k-1 and v_h = v_k + c F(x_k)
// of step k) merge into one Hessian-vector product per step.  Per pair,
//   M_ij = s I + (s'(r)/r) r_ij r_ij^T,  (J a)_i = sum_j M_ij (a_i - a_j).
// The forward pass keeps (x, v) every checkpoint_every steps; the sweep
// recomputes each segment forward from its checkpoint, storing its positions,
// then walks it backwards.  Pair lists (radius max(cutoffs, rdf range) +
// skin) come from the library's LBVH pair-list path and are rebuilt at the
// same steps (multiples of rebuild_every) in the forward and the recompute.
//
// Kernels run one lane per atom over the atom's row of the pair list (every
// j paired with i; rows are rebuilt with the list), so nothing is added with
// atomics.  Forces are summed in 2^-40 fixed point: the sum does not depend
// on the order within a row, which the list build leaves unspecified, so the
// recompute reproduces the forward trajectory bit for bit in both
// precisions.  The reverse sweep sums in FP64.
//
// No reference counterpart: NaiveDynamics.jl has no gradients.  This is the
// differentiation form of the simulate! hot path (Simulator.jl:154-256).
#include <algorithm>
#include <cmath>
#include <vector>

#include "nd_internal.hpp"

namespace {

constexpr double kFix = 1099511627776.0;  // 2^40
constexpr double kInvFix = 1.0 / 1099511627776.0;

struct AdjP {
    int32_t n, nbins;
    double eps, sigma, rc, rc2, kc, rcc, rcc2, dt, c;
    double flj_rc;   // F_LJ(rc)
    double glj_rc;   // dF_LJ/dsigma at rc
    double rmax, h, dr;
};

// F_LJ(r) = 24 eps (2 sigma^12 / r^13 - sigma^6 / r^7)
__host__ __device__ inline double lj_F(double r, double eps, double sigma) {
    const double sr = sigma / r, sr6 = sr * sr * sr * sr * sr * sr;
    return 24.0 * eps * (2.0 * sr6 * sr6 - sr6) / r;
}
// dF_LJ/dsigma = 24 eps (24 sigma^11 / r^13 - 6 sigma^5 / r^7)
__host__ __device__ inline double lj_G(double r, double eps, double sigma) {
    const double sr = sigma / r, sr6 = sr * sr * sr * sr * sr * sr;
    return 24.0 * eps * (24.0 * sr6 * sr6 - 6.0 * sr6) / (r * sigma);
}

// s(r) of the forward force in precision T (LJ + Coulomb, both shifted).
template <class T>
__device__ __forceinline__ T pair_s(T r2, T qq, const AdjP &P) {
    T s = (T)0;
    const T r = sqrt(r2), ir = (T)1 / r;
    if (r2 < (T)P.rc2) {
        const T sr2 = (T)(P.sigma * P.sigma) / r2, sr6 = sr2 * sr2 * sr2;
        const T F = (T)(24.0 * P.eps) * (sr6 * sr6 + sr6 * sr6 - sr6) * ir;
        s += (F - (T)P.flj_rc) * ir;
    }
    if (r2 < (T)P.rcc2) {
        const T F = (T)P.kc * qq * ((T)1 / r2 - (T)(1.0 / P.rcc2));
        s += F * ir;
    }
    return s;
}

// FP64 pair terms of the reverse sweep: s, ds = s'(r)/r, ds/deps, ds/dsigma
// and H with ds_coulomb/dq_i = kc q_j H.
struct PairD {
    double s, ds, s_eps, s_sig, H;
};
__device__ __forceinline__ PairD pair_d(double r2, double qi, double qj, const AdjP &P) {
    PairD o{0.0, 0.0, 0.0, 0.0, 0.0};
    const double r = sqrt(r2), ir = 1.0 / r, ir3 = ir * ir * ir;
    if (r2 < P.rc2) {
        const double sr2 = P.sigma * P.sigma / r2, sr6 = sr2 * sr2 * sr2, sr12 = sr6 * sr6;
        const double F = 24.0 * P.eps * (2.0 * sr12 - sr6) * ir;
        const double Fp = 24.0 * P.eps * (-26.0 * sr12 + 7.0 * sr6) * ir * ir;
        const double s = (F - P.flj_rc) * ir;
        o.s += s;
        o.ds += (Fp * r - (F - P.flj_rc)) * ir3;
        o.s_eps = s / P.eps;
        const double G = 24.0 * P.eps * (24.0 * sr12 - 6.0 * sr6) * ir / P.sigma;
        o.s_sig = (G - P.glj_rc) * ir;
    }
    if (r2 < P.rcc2) {
        const double qq = qi * qj;
        const double Hf = ir * ir - 1.0 / P.rcc2;  // F_c = kc qq Hf
        const double F = P.kc * qq * Hf;
        const double Fp = -2.0 * P.kc * qq * ir3;
        o.s += F * ir;
        o.ds += (Fp * r - F) * ir3;
        o.H = Hf * ir;
    }
    return o;
}

// Sorted FP32 positions (w = 0) for the list walk.
__global__ void __launch_bounds__(nd::kBlock)
k_adj_pos4(const float *__restrict__ xf, const int32_t *__restrict__ perm, int32_t n,
           float4 *__restrict__ out) {
    const int32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const int64_t o = 3 * (int64_t)perm[s];
    out[s] = make_float4(xf[o], xf[o + 1], xf[o + 2], 0.0f);
}

// Neighbour rows in the list walk's layout: sorted slot s (atom perm[s])
// keeps its entries in column s & 63 of the block of its 64 slots (stride
// 64), as original atom ids once k_adj_map_ids has run; lcnt[s] entries.
struct Rows {
    const int32_t *lists, *lcnt, *perm;
    int32_t cap;
    __device__ __forceinline__ const int32_t *col(int32_t s) const {
        return lists + ((int64_t)(s >> 6) * cap) * 64 + (s & 63);
    }
    __device__ __forceinline__ int32_t count(int32_t s) const { return lcnt[s] < cap ? lcnt[s] : cap; }
};

// The walk writes sorted indices; the adjoint's kernels want original ids.
// One lane per slot and list row (blockIdx.y = row): no per-lane loop.
__global__ void __launch_bounds__(nd::kBlock)
k_adj_map_ids(const int32_t *__restrict__ perm, int32_t n, int32_t cap, int32_t *__restrict__ lists,
              const int32_t *__restrict__ lcnt) {
    const int32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    const int32_t k = blockIdx.y;
    if (s >= n || k >= (lcnt[s] < cap ? lcnt[s] : cap)) return;
    int32_t *e = lists + ((int64_t)(s >> 6) * cap + k) * 64 + (s & 63);
    *e = perm[*e];
}

// Rows are split over kRowLanes consecutive lanes of a wave (lane s of an
// atom takes entries s, s + kRowLanes, ...): at N = 32,768 one lane per atom
// leaves 2 waves per CU and every row a long chain of dependent loads.
constexpr int kRowLanes = 4;

// F_i = sum_j s(r_ij) r_ij over row i, summed in 2^-40 fixed point: the sum
// is the same whatever the order of the row (and however it is split).
template <class T>
__global__ void __launch_bounds__(nd::kBlock)
k_adj_force(Rows R, const T *__restrict__ x, const double *__restrict__ q, AdjP P,
            unsigned long long *__restrict__ F) {
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    const int32_t sl = (int32_t)(t / kRowLanes), sub = (int32_t)(t % kRowLanes);
    const int32_t i = sl < P.n ? R.perm[sl] : 0;
    long long fx = 0, fy = 0, fz = 0;
    if (sl < P.n) {
    const int32_t *col = R.col(sl);
    const int32_t c = R.count(sl);
    const T xi = x[3 * (int64_t)i], yi = x[3 * (int64_t)i + 1], zi = x[3 * (int64_t)i + 2];
    const double qi = q[i];
    for (int32_t k = sub; k < c; k += kRowLanes) {
        const int32_t j = col[(int64_t)k * 64];
        const T dx = xi - x[3 * (int64_t)j], dy = yi - x[3 * (int64_t)j + 1], dz = zi - x[3 * (int64_t)j + 2];
        const T r2 = dx * dx + dy * dy + dz * dz;
        if (!(r2 < (T)P.rc2 || r2 < (T)P.rcc2)) continue;
        const T s = pair_s<T>(r2, (T)(qi * q[j]), P);
        fx += llrint((double)(s * dx) * kFix);
        fy += llrint((double)(s * dy) * kFix);
        fz += llrint((double)(s * dz) * kFix);
    }
    }
    // exact integer sums over the atom's lanes (all lanes take part)
#pragma unroll
    for (int off = 1; off < kRowLanes; off <<= 1) {
        fx += __shfl_xor(fx, off, 64);
        fy += __shfl_xor(fy, off, 64);
        fz += __shfl_xor(fz, off, 64);
    }
    if (sl < P.n && sub == 0) {
        F[3 * (int64_t)i] = (unsigned long long)fx;
        F[3 * (int64_t)i + 1] = (unsigned long long)fy;
        F[3 * (int64_t)i + 2] = (unsigned long long)fz;
    }
}

__device__ __forceinline__ double fix_val(unsigned long long v) { return (double)(long long)v * kInvFix; }

// v += c F; x += dt v
template <class T>
__global__ void __launch_bounds__(nd::kBlock)
k_adj_kick_drift(T *__restrict__ x, T *__restrict__ v, const unsigned long long *__restrict__ F,
                 AdjP P) {
    const int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (k >= 3 * (int64_t)P.n) return;
    T vv = v[k] + (T)P.c * (T)fix_val(F[k]);
    v[k] = vv;
    x[k] = x[k] + (T)P.dt * vv;
}

// v += c F
template <class T>
__global__ void __launch_bounds__(nd::kBlock)
k_adj_kick(T *__restrict__ v, const unsigned long long *__restrict__ F, AdjP P) {
    const int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (k >= 3 * (int64_t)P.n) return;
    v[k] = v[k] + (T)P.c * (T)fix_val(F[k]);
}

template <class T>
__global__ void __launch_bounds__(nd::kBlock)
k_adj_to_f32(const T *__restrict__ x, int64_t m, float *__restrict__ out) {
    const int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (k < m) out[k] = (float)x[k];
}

// RDF kernel weight of pair distance r for bin b
__device__ __forceinline__ double rdf_K(double d, double h) {
    const double u = d / h, t = 1.0 - u * u;
    return fabs(u) < 1.0 ? t * t * t : 0.0;
}
__device__ __forceinline__ double rdf_dK(double d, double h) {
    const double u = d / h, t = 1.0 - u * u;
    return fabs(u) < 1.0 ? -6.0 * u * t * t / h : 0.0;
}

// raw histogram sum_{i<j} K(r_ij - r_b) (normalisation on the host); bins
// are summed per workgroup in LDS, then added once per workgroup
constexpr int kMaxBins = 1024;

template <class T>
__global__ void __launch_bounds__(nd::kBlock)
k_rdf_hist(Rows R, const T *__restrict__ x, AdjP P, double *__restrict__ hist) {
    __shared__ double hb[kMaxBins];
    for (int b = threadIdx.x; b < P.nbins; b += blockDim.x) hb[b] = 0.0;
    __syncthreads();
    const int32_t sl = blockIdx.x * blockDim.x + threadIdx.x;
    if (sl < P.n) {
        const int32_t i = R.perm[sl];
        const double xi = (double)x[3 * (int64_t)i], yi = (double)x[3 * (int64_t)i + 1],
                     zi = (double)x[3 * (int64_t)i + 2];
        const int32_t *col = R.col(sl);
        const int32_t c = R.count(sl);
        for (int32_t k = 0; k < c; ++k) {
            const int32_t j = col[(int64_t)k * 64];
            if (j < i) continue;  // each unordered pair once
            const double dx = xi - (double)x[3 * (int64_t)j], dy = yi - (double)x[3 * (int64_t)j + 1],
                         dz = zi - (double)x[3 * (int64_t)j + 2];
            const double r = sqrt(dx * dx + dy * dy + dz * dz);
            if (!(r < P.rmax + P.h)) continue;
            const int32_t b0 = max(0, (int32_t)floor((r - P.h) / P.dr - 0.5));
            const int32_t b1 = min(P.nbins - 1, (int32_t)ceil((r + P.h) / P.dr - 0.5));
            for (int32_t b = b0; b <= b1; ++b) {
                const double w = rdf_K(r - (b + 0.5) * P.dr, P.h);
                if (w != 0.0) atomicAdd(&hb[b], w);
            }
        }
    }
    __syncthreads();
    for (int b = threadIdx.x; b < P.nbins; b += blockDim.x)
        if (hb[b] != 0.0) unsafeAtomicAdd(&hist[b], hb[b]);
}

// lambda_x += dL/dx with L = sum_b (g_b - t_b)^2: wb[b] = 2 (g_b - t_b) norm_b
template <class T>
__global__ void __launch_bounds__(nd::kBlock)
k_rdf_grad(Rows R, const T *__restrict__ x, AdjP P, const double *__restrict__ wb,
           double *__restrict__ lam) {
    const int32_t sl = blockIdx.x * blockDim.x + threadIdx.x;
    if (sl >= P.n) return;
    const int32_t i = R.perm[sl];
    const int32_t *col = R.col(sl);
    const int32_t c = R.count(sl);
    const double xi = (double)x[3 * (int64_t)i], yi = (double)x[3 * (int64_t)i + 1],
                 zi = (double)x[3 * (int64_t)i + 2];
    double gx = 0.0, gy = 0.0, gz = 0.0;
    for (int32_t k = 0; k < c; ++k) {
        const int32_t j = col[(int64_t)k * 64];
        const double dx = xi - (double)x[3 * (int64_t)j], dy = yi - (double)x[3 * (int64_t)j + 1],
                     dz = zi - (double)x[3 * (int64_t)j + 2];
        const double r = sqrt(dx * dx + dy * dy + dz * dz);
        if (!(r < P.rmax + P.h)) continue;
        const int32_t b0 = max(0, (int32_t)floor((r - P.h) / P.dr - 0.5));
        const int32_t b1 = min(P.nbins - 1, (int32_t)ceil((r + P.h) / P.dr - 0.5));
        double g = 0.0;
        for (int32_t b = b0; b <= b1; ++b) g += wb[b] * rdf_dK(r - (b + 0.5) * P.dr, P.h);
        const double s = g / r;
        gx += s * dx;
        gy += s * dy;
        gz += s * dz;
    }
    lam[3 * (int64_t)i] += gx;
    lam[3 * (int64_t)i + 1] += gy;
    lam[3 * (int64_t)i + 2] += gz;
}

// lambda_h = lambda_v + dt lambda_x; lambda_v = lambda_h; a = w lambda_h
__global__ void __launch_bounds__(nd::kBlock)
k_adj_vec(const double *__restrict__ lx, double *__restrict__ lv, double *__restrict__ a, int64_t m,
          double dt, double w) {
    const int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (k >= m) return;
    const double h = lv[k] + dt * lx[k];
    lv[k] = h;
    a[k] = w * h;
}

// lambda_x += J(x) a; dq, (eps, sigma) += (dF/dtheta)^T a.  Row i gives atom
// i's own terms; each unordered pair sits in two rows, so its eps / sigma
// term is taken with weight 1/2 from each.
template <class T>
__global__ void __launch_bounds__(nd::kBlock)
k_adj_hvp(Rows R, const T *__restrict__ x, const double *__restrict__ q, AdjP P,
          const double *__restrict__ a, double *__restrict__ lam, double *__restrict__ dq,
          double *__restrict__ dpar) {
    __shared__ double sh[2][nd::kBlock / 64];
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    const int32_t sl = (int32_t)(t / kRowLanes), sub = (int32_t)(t % kRowLanes);
    const int32_t i = sl < P.n ? R.perm[sl] : 0;
    double ge = 0.0, gs = 0.0;
    double mx = 0.0, my = 0.0, mz = 0.0, gq = 0.0;
    if (sl < P.n) {
        const double xi = (double)x[3 * (int64_t)i], yi = (double)x[3 * (int64_t)i + 1],
                     zi = (double)x[3 * (int64_t)i + 2];
        const double ai = a[3 * (int64_t)i], bi = a[3 * (int64_t)i + 1], ci = a[3 * (int64_t)i + 2];
        const double qi = q[i];
        const int32_t *col = R.col(sl);
        const int32_t c = R.count(sl);
        for (int32_t k = sub; k < c; k += kRowLanes) {
            const int32_t j = col[(int64_t)k * 64];
            const double dx = xi - (double)x[3 * (int64_t)j], dy = yi - (double)x[3 * (int64_t)j + 1],
                         dz = zi - (double)x[3 * (int64_t)j + 2];
            const double r2 = dx * dx + dy * dy + dz * dz;
            if (!(r2 < P.rc2 || r2 < P.rcc2)) continue;
            const PairD d = pair_d(r2, qi, q[j], P);
            const double ax = ai - a[3 * (int64_t)j], ay = bi - a[3 * (int64_t)j + 1],
                         az = ci - a[3 * (int64_t)j + 2];
            const double rda = dx * ax + dy * ay + dz * az;
            mx += d.s * ax + d.ds * rda * dx;
            my += d.s * ay + d.ds * rda * dy;
            mz += d.s * az + d.ds * rda * dz;
            gq += P.kc * q[j] * d.H * rda;
            ge += 0.5 * d.s_eps * rda;
            gs += 0.5 * d.s_sig * rda;
        }
    }
    // the atom's lanes combine their partial sums in a fixed order
#pragma unroll
    for (int off = 1; off < kRowLanes; off <<= 1) {
        mx += __shfl_xor(mx, off, 64);
        my += __shfl_xor(my, off, 64);
        mz += __shfl_xor(mz, off, 64);
        gq += __shfl_xor(gq, off, 64);
    }
    if (sl < P.n && sub == 0) {
        lam[3 * (int64_t)i] += mx;
        lam[3 * (int64_t)i + 1] += my;
        lam[3 * (int64_t)i + 2] += mz;
        dq[i] += gq;
    }
    for (int o = 32; o > 0; o >>= 1) {
        ge += __shfl_down(ge, o, 64);
        gs += __shfl_down(gs, o, 64);
    }
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (lane == 0) { sh[0][w] = ge; sh[1][w] = gs; }
    __syncthreads();
    if (threadIdx.x == 0) {
        double e = 0.0, s = 0.0;
        for (int t = 0; t < (int)(blockDim.x >> 6); ++t) { e += sh[0][t]; s += sh[1][t]; }
        unsafeAtomicAdd(&dpar[0], e);
        unsafeAtomicAdd(&dpar[1], s);
    }
}

__global__ void k_zero_f64(double *p, int64_t m) {
    const int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (k < m) p[k] = 0.0;
}

}  // namespace

struct nd_adj {
    nd_ctx *ctx = nullptr;
    nd_adj_params p{};
    AdjP P{};
    bool dbl = true;
    size_t esz = 8;  // bytes per state scalar
    int32_t apl = 1;
    float rlist = 0.0f;
    // state and scratch
    nd::DevBuf x, v, F, xf, xs, codes, perm, nodes, store, flags, q, ckx, ckv, seg, lx, lv, a, dq,
        dpar, hist, wb, rec, lists, lcnt;
    nd::SortBufs sort;
    int32_t cap = 0;  // row capacity
    int32_t nck = 0;
    bool have_forward = false;
    std::vector<double> target, rdf, norm;
    double loss = 0.0;
};

namespace {

size_t state_bytes(const nd_adj *s) { return s->esz * 3 * (size_t)s->p.n; }

template <class T> T *at(nd::DevBuf &b, size_t off_elems = 0) { return b.as<T>() + off_elems; }

// neighbour rows (radius rlist, Hilbert key LBVH) at positions xp (device, T)
// the walk's blocks of [cap][64] slots per 64 sorted atoms
void reserve_rows(nd_adj *s) {
    const size_t n64 = ((size_t)s->p.n + 63) & ~(size_t)63;
    s->lists.reserve(sizeof(int32_t) * (size_t)s->cap * n64);
}

Rows rows_of(nd_adj *s) {
    return Rows{s->lists.as<int32_t>(), s->lcnt.as<int32_t>(), s->perm.as<int32_t>(), s->cap};
}

template <class T>
void build_list(nd_adj *s, const T *xp) {
    nd_ctx *ctx = s->ctx;
    hipStream_t st = ctx->stream;
    const int32_t n = s->p.n, nL = n / s->apl;
    const int64_t m = 3 * (int64_t)n;
    int32_t *codes = s->codes.as<int32_t>(), *perm = s->perm.as<int32_t>(), *flags = s->flags.as<int32_t>();
    k_adj_to_f32<T><<<nd::grid_for(m), nd::kBlock, 0, st>>>(xp, m, s->xf.as<float>());
    nd::launch_fill_i32(st, flags, 8, 0);
    nd::launch_morton(ctx, s->xf.as<float>(), n, ND_MORTON_HILBERT, nullptr, nullptr, codes, perm, flags + 2);
    nd::radix_sort_pairs(st, s->sort, codes, perm, n, 0, 30);  // Hilbert keys are < 2^30
    k_adj_pos4<<<nd::grid_for(n), nd::kBlock, 0, st>>>(s->xf.as<float>(), perm, n, s->xs.as<float4>());
    nd::launch_lbvh(st, s->xs.as<float>(), 4, codes, nL, s->apl, s->rlist, 0, s->nodes.as<nd_node>(),
                    s->store.as<int32_t>(), flags + 1);
    for (int attempt = 0; attempt < 4; ++attempt) {
        // the simulation's list walk over the sorted atoms, then original ids
        nd::launch_list_walk(st, s->nodes.as<nd_node>(), s->rec.as<float4>(), s->xs.as<float4>(), n, nL,
                             s->apl, s->rlist, s->cap, s->lists.as<int32_t>(), s->lcnt.as<int32_t>(), flags + 4);
        k_adj_map_ids<<<dim3(nd::grid_for(n), (unsigned)s->cap), nd::kBlock, 0, st>>>(
            perm, n, s->cap, s->lists.as<int32_t>(), s->lcnt.as<int32_t>());
        ND_HIP(hipGetLastError());
        int32_t h[8];
        ND_HIP(hipMemcpyAsync(h, flags, sizeof(h), hipMemcpyDeviceToHost, st));
        ND_HIP(hipStreamSynchronize(st));
        if (h[2]) throw nd::Error(ND_ERR_DOMAIN, "adjoint: non-finite coordinate");
        if (h[1] || h[6]) throw nd::Error(ND_ERR_STATE, "adjoint: malformed neighbour tree");
        if (h[4] <= s->cap) return;
        s->cap = h[4] + h[4] / 4 + 8;
        // k_adj_map_ids takes one grid row per list row
        if (s->cap > 65535) throw nd::Error(ND_ERR_CAPACITY, "adjoint: more than 65,535 neighbours per atom");
        reserve_rows(s);
        nd::launch_fill_i32(st, flags + 4, 3, 0);
    }
    throw nd::Error(ND_ERR_CAPACITY, "adjoint: neighbour rows kept overflowing");
}

template <class T>
void compute_force(nd_adj *s, const T *xs) {
    hipStream_t st = s->ctx->stream;
    const int64_t m = 3 * (int64_t)s->p.n;
    (void)m;
    k_adj_force<T><<<nd::grid_for((int64_t)s->p.n * kRowLanes), nd::kBlock, 0, st>>>(rows_of(s),
                                                                 xs, s->q.as<double>(), s->P,
                                                                 s->F.as<unsigned long long>());
    ND_HIP(hipGetLastError());
}

// Advance steps [k0, k1) from (x, v) with F = F(x_k0) already in s->F and the
// list valid for x_k0.  If seg != null, x_k (k0 <= k < k1) is stored there.
template <class T>
void advance(nd_adj *s, int32_t k0, int32_t k1, T *seg) {
    hipStream_t st = s->ctx->stream;
    const int64_t m = 3 * (int64_t)s->p.n;
    T *x = s->x.as<T>(), *v = s->v.as<T>();
    for (int32_t k = k0; k < k1; ++k) {
        if (seg)
            ND_HIP(hipMemcpyAsync(seg + (size_t)(k - k0) * m, x, state_bytes(s), hipMemcpyDeviceToDevice, st));
        k_adj_kick_drift<T><<<nd::grid_for(m), nd::kBlock, 0, st>>>(x, v, s->F.as<unsigned long long>(), s->P);
        if ((k + 1) % s->p.rebuild_every == 0) build_list<T>(s, x);
        compute_force<T>(s, x);
        k_adj_kick<T><<<nd::grid_for(m), nd::kBlock, 0, st>>>(v, s->F.as<unsigned long long>(), s->P);
        ND_HIP(hipGetLastError());
    }
}

template <class T>
void forward_impl(nd_adj *s) {
    hipStream_t st = s->ctx->stream;
    const int32_t K = s->p.checkpoint_every, T_ = s->p.nsteps;
    T *x = s->x.as<T>(), *v = s->v.as<T>();
    build_list<T>(s, x);
    compute_force<T>(s, x);
    for (int32_t c = 0; c * K < T_; ++c) {
        ND_HIP(hipMemcpyAsync(s->ckx.as<T>() + (size_t)c * 3 * s->p.n, x, state_bytes(s), hipMemcpyDeviceToDevice, st));
        ND_HIP(hipMemcpyAsync(s->ckv.as<T>() + (size_t)c * 3 * s->p.n, v, state_bytes(s), hipMemcpyDeviceToDevice, st));
        const int32_t k1 = (c + 1) * K < T_ ? (c + 1) * K : T_;
        advance<T>(s, c * K, k1, (T *)nullptr);
    }
    // final frame: radial distribution (pairs of the list valid at x_T)
    if (T_ % s->p.rebuild_every != 0) build_list<T>(s, x);
    const int32_t nb = s->p.nbins;
    k_zero_f64<<<nd::grid_for(nb), nd::kBlock, 0, st>>>(s->hist.as<double>(), nb);
    k_rdf_hist<T><<<nd::grid_for(s->p.n), nd::kBlock, 0, st>>>(rows_of(s),
                                                                x, s->P, s->hist.as<double>());
    ND_HIP(hipGetLastError());
    std::vector<double> h(nb);
    ND_HIP(hipMemcpyAsync(h.data(), s->hist.p, sizeof(double) * nb, hipMemcpyDeviceToHost, st));
    ND_HIP(hipStreamSynchronize(st));
    s->rdf.resize(nb);
    s->loss = 0.0;
    for (int32_t b = 0; b < nb; ++b) {
        s->rdf[b] = s->norm[b] * h[b];
        if (!s->target.empty()) s->loss += (s->rdf[b] - s->target[b]) * (s->rdf[b] - s->target[b]);
    }
    s->have_forward = true;
}

template <class T>
void backward_impl(nd_adj *s) {
    hipStream_t st = s->ctx->stream;
    const int32_t K = s->p.checkpoint_every, T_ = s->p.nsteps, W = s->p.rebuild_every, n = s->p.n;
    const int64_t m = 3 * (int64_t)n;
    double *lx = s->lx.as<double>(), *lv = s->lv.as<double>(), *a = s->a.as<double>();
    k_zero_f64<<<nd::grid_for(m), nd::kBlock, 0, st>>>(lx, m);
    k_zero_f64<<<nd::grid_for(m), nd::kBlock, 0, st>>>(lv, m);
    k_zero_f64<<<nd::grid_for(n), nd::kBlock, 0, st>>>(s->dq.as<double>(), n);
    k_zero_f64<<<1, 64, 0, st>>>(s->dpar.as<double>(), 2);
    // dL/dx_T: x still holds x_T and the list is valid there
    std::vector<double> wb(s->p.nbins);
    for (int32_t b = 0; b < s->p.nbins; ++b) wb[b] = 2.0 * (s->rdf[b] - s->target[b]) * s->norm[b];
    ND_HIP(hipMemcpyAsync(s->wb.p, wb.data(), sizeof(double) * wb.size(), hipMemcpyHostToDevice, st));
    k_rdf_grad<T><<<nd::grid_for(n), nd::kBlock, 0, st>>>(rows_of(s),
                                                           s->x.as<T>(), s->P, s->wb.as<double>(), lx);
    ND_HIP(hipGetLastError());
    T *seg = s->seg.as<T>();
    for (int32_t c = (T_ - 1) / K; c >= 0; --c) {
        const int32_t k0 = c * K, k1 = (c + 1) * K < T_ ? (c + 1) * K : T_;
        // recompute the segment from its checkpoint, storing x_k0 .. x_{k1-1}
        ND_HIP(hipMemcpyAsync(s->x.p, s->ckx.as<T>() + (size_t)c * m, state_bytes(s), hipMemcpyDeviceToDevice, st));
        ND_HIP(hipMemcpyAsync(s->v.p, s->ckv.as<T>() + (size_t)c * m, state_bytes(s), hipMemcpyDeviceToDevice, st));
        build_list<T>(s, s->x.as<T>());  // k0 is a multiple of W
        compute_force<T>(s, s->x.as<T>());
        advance<T>(s, k0, k1, seg);
        int32_t list_at = -1;
        for (int32_t k = k1 - 1; k >= k0; --k) {
            const int32_t w0 = (k / W) * W;
            const T *xk = seg + (size_t)(k - k0) * m;
            if (list_at != w0) {
                build_list<T>(s, seg + (size_t)(w0 - k0) * m);
                list_at = w0;
            }
            k_adj_vec<<<nd::grid_for(m), nd::kBlock, 0, st>>>(lx, lv, a, m, s->P.dt,
                                                                k == 0 ? s->P.c : 2.0 * s->P.c);
            k_adj_hvp<T><<<nd::grid_for((int64_t)n * kRowLanes), nd::kBlock, 0, st>>>(rows_of(s),
                                                                  xk, s->q.as<double>(), s->P, a, lx,
                                                                  s->dq.as<double>(), s->dpar.as<double>());
            ND_HIP(hipGetLastError());
        }
    }
}

}  // namespace

extern "C" {

int nd_adj_create(nd_ctx *ctx, const nd_adj_params *p, nd_adj **out) {
    ND_API_BEGIN
    ND_REQUIRE(ctx && p && out, "nd_adj_create: null argument");
    *out = nullptr;
    ND_REQUIRE(p->n >= 8 && p->n % 4 == 0 && p->nsteps >= 1 && p->checkpoint_every >= 1 &&
                   p->rebuild_every >= 1,
               "nd_adj_create: n >= 8 and a multiple of 4 (list leaves of 4, 8 or 16 atoms); nsteps, "
               "checkpoint_every, rebuild_every >= 1");
    ND_REQUIRE(p->checkpoint_every % p->rebuild_every == 0,
               "nd_adj_create: checkpoint_every must be a multiple of rebuild_every");
    ND_REQUIRE(p->precision == ND_PREC_FP32 || p->precision == ND_PREC_FP64,
               "nd_adj_create: precision must be FP32 or FP64");
    ND_REQUIRE(p->nbins >= 1 && p->nbins <= kMaxBins && p->rdf_rmax > 0.0 && p->rdf_h > 0.0 &&
                   p->volume > 0.0,
               "nd_adj_create: 1..1024 RDF bins; range, kernel width and volume must be positive");
    ND_REQUIRE(p->dt > 0.0 && p->eps > 0.0 && p->sigma > 0.0 && p->rc > 0.0 && p->skin >= 0.0,
               "nd_adj_create: dt, eps, sigma, rc > 0 and skin >= 0");
    ND_HIP(hipSetDevice(ctx->device));
    nd_adj *s = new nd_adj();
    try {
        s->ctx = ctx;
        s->p = *p;
        s->dbl = p->precision == ND_PREC_FP64;
        s->esz = s->dbl ? 8 : 4;
        s->apl = 4;
        for (int32_t a : {16, 8, 4})
            if (p->n % a == 0 && p->n / a >= 2) { s->apl = a; break; }
        const double reach = std::max(std::max(p->rc, p->rc_coul), p->rdf_rmax + p->rdf_h);
        // FP32 copies of the coordinates feed the list: pad by their rounding
        s->rlist = (float)((reach + p->skin) * (1.0 + 2.0e-6) + 1e-6);
        const size_t n = (size_t)p->n, m = 3 * n;
        const int32_t nck = (p->nsteps + p->checkpoint_every - 1) / p->checkpoint_every;
        s->nck = nck;
        s->x.reserve(s->esz * m);
        s->v.reserve(s->esz * m);
        s->F.reserve(8 * m);
        s->xf.reserve(4 * m);
        s->q.reserve(8 * n);
        s->ckx.reserve(s->esz * m * (size_t)nck);
        s->ckv.reserve(s->esz * m * (size_t)nck);
        s->seg.reserve(s->esz * m * (size_t)p->checkpoint_every);
        s->lx.reserve(8 * m);
        s->lv.reserve(8 * m);
        s->a.reserve(8 * m);
        s->dq.reserve(8 * n);
        s->dpar.reserve(16);
        s->hist.reserve(8 * (size_t)p->nbins);
        s->xs.reserve(sizeof(float4) * n);
        s->lcnt.reserve(4 * n);
        s->rec.reserve(8 * sizeof(float4) * (size_t)(p->n / s->apl - 1));
        s->codes.reserve(4 * n);
        s->perm.reserve(4 * n);
        s->nodes.reserve(sizeof(nd_node) * (size_t)(2 * (p->n / s->apl) - 1));
        s->store.reserve(sizeof(int32_t) * nd::lbvh_scratch_ints(p->n / s->apl));
        s->flags.reserve(64);  // [0..3] tree build status words, [4..6] the walk's counters
        s->wb.reserve(8 * (size_t)p->nbins);
        s->cap = 96;
        reserve_rows(s);
        // RDF normalisation: ideal-gas pair count in each bin's shell
        const double dr = p->rdf_rmax / p->nbins, npair = 0.5 * (double)p->n * (p->n - 1);
        s->norm.resize(p->nbins);
        for (int32_t b = 0; b < p->nbins; ++b) {
            const double rb = (b + 0.5) * dr;
            s->norm[b] = p->volume / (npair * 4.0 * M_PI * rb * rb * dr);
        }
        nd_adj_set_potential(s, p->eps, p->sigma);
    } catch (...) {
        nd_adj_destroy(s);
        throw;
    }
    *out = s;
    ND_API_END(ctx)
}

int nd_adj_destroy(nd_adj *s) {
    if (!s) return ND_OK;
    if (s->ctx) {
        (void)hipSetDevice(s->ctx->device);
        (void)hipStreamSynchronize(s->ctx->stream);
    }
    nd::DevBuf *bufs[] = {&s->rec, &s->lists, &s->lcnt, &s->x, &s->v, &s->F, &s->xf, &s->xs, &s->codes, &s->perm, &s->nodes,
                          &s->store, &s->flags, &s->q, &s->ckx, &s->ckv, &s->seg, &s->lx, &s->lv,
                          &s->a, &s->dq, &s->dpar, &s->hist, &s->wb,
                          &s->sort.keys_alt, &s->sort.vals_alt, &s->sort.hist, &s->sort.flags};
    for (auto *b : bufs) b->release();
    delete s;
    return ND_OK;
}

int nd_adj_set_potential(nd_adj *s, double eps, double sigma) {
    nd_ctx *ctx = s ? s->ctx : nullptr;
    ND_API_BEGIN
    ND_REQUIRE(s && eps > 0.0 && sigma > 0.0, "nd_adj_set_potential: eps, sigma > 0");
    const nd_adj_params &p = s->p;
    AdjP &P = s->P;
    P.n = p.n;
    P.nbins = p.nbins;
    P.eps = eps;
    P.sigma = sigma;
    P.rc = p.rc;
    P.rc2 = p.rc * p.rc;
    P.kc = p.k_coul;
    P.rcc = p.rc_coul;
    P.rcc2 = p.rc_coul > 0.0 ? p.rc_coul * p.rc_coul : 0.0;
    P.dt = p.dt;
    P.c = 0.5 * p.dt;
    P.flj_rc = lj_F(p.rc, eps, sigma);
    P.glj_rc = lj_G(p.rc, eps, sigma);
    P.rmax = p.rdf_rmax;
    P.h = p.rdf_h;
    P.dr = p.rdf_rmax / p.nbins;
    s->have_forward = false;
    ND_API_END(ctx)
}

int nd_adj_forward(nd_adj *s, const double *h_x0, const double *h_v0, const double *h_q,
                   const double *h_target, double *h_rdf, double *h_loss) {
    nd_ctx *ctx = s ? s->ctx : nullptr;
    ND_API_BEGIN
    ND_REQUIRE(s && h_x0 && h_v0, "nd_adj_forward: null argument");
    ND_HIP(hipSetDevice(ctx->device));
    const size_t n = (size_t)s->p.n, m = 3 * n;
    hipStream_t st = ctx->stream;
    std::vector<double> q(n, 0.0);
    if (h_q) q.assign(h_q, h_q + n);
    ND_HIP(hipMemcpyAsync(s->q.p, q.data(), 8 * n, hipMemcpyHostToDevice, st));
    if (s->dbl) {
        ND_HIP(hipMemcpyAsync(s->x.p, h_x0, 8 * m, hipMemcpyHostToDevice, st));
        ND_HIP(hipMemcpyAsync(s->v.p, h_v0, 8 * m, hipMemcpyHostToDevice, st));
    } else {
        std::vector<float> xf(h_x0, h_x0 + m), vf(h_v0, h_v0 + m);
        ND_HIP(hipMemcpyAsync(s->x.p, xf.data(), 4 * m, hipMemcpyHostToDevice, st));
        ND_HIP(hipMemcpyAsync(s->v.p, vf.data(), 4 * m, hipMemcpyHostToDevice, st));
        ND_HIP(hipStreamSynchronize(st));
    }
    s->target.clear();
    if (h_target) s->target.assign(h_target, h_target + s->p.nbins);
    if (s->dbl) forward_impl<double>(s);
    else forward_impl<float>(s);
    if (h_rdf) std::copy(s->rdf.begin(), s->rdf.end(), h_rdf);
    if (h_loss) *h_loss = s->loss;
    ND_API_END(ctx)
}

int nd_adj_final_state(nd_adj *s, double *h_x, double *h_v) {
    nd_ctx *ctx = s ? s->ctx : nullptr;
    ND_API_BEGIN
    ND_REQUIRE(s && s->have_forward, "nd_adj_final_state: run nd_adj_forward first");
    ND_HIP(hipSetDevice(ctx->device));
    const size_t m = 3 * (size_t)s->p.n;
    hipStream_t st = ctx->stream;
    if (s->dbl) {
        if (h_x) ND_HIP(hipMemcpyAsync(h_x, s->x.p, 8 * m, hipMemcpyDeviceToHost, st));
        if (h_v) ND_HIP(hipMemcpyAsync(h_v, s->v.p, 8 * m, hipMemcpyDeviceToHost, st));
        ND_HIP(hipStreamSynchronize(st));
    } else {
        std::vector<float> xf(m), vf(m);
        ND_HIP(hipMemcpyAsync(xf.data(), s->x.p, 4 * m, hipMemcpyDeviceToHost, st));
        ND_HIP(hipMemcpyAsync(vf.data(), s->v.p, 4 * m, hipMemcpyDeviceToHost, st));
        ND_HIP(hipStreamSynchronize(st));
        for (size_t k = 0; k < m; ++k) {
            if (h_x) h_x[k] = xf[k];
            if (h_v) h_v[k] = vf[k];
        }
    }
    ND_API_END(ctx)
}

int nd_adj_backward(nd_adj *s, double *h_dx0, double *h_dv0, double *h_dq, double *h_dparams) {
    nd_ctx *ctx = s ? s->ctx : nullptr;
    ND_API_BEGIN
    ND_REQUIRE(s && s->have_forward && !s->target.empty(),
               "nd_adj_backward: run nd_adj_forward with a target first");
    ND_HIP(hipSetDevice(ctx->device));
    if (s->dbl) backward_impl<double>(s);
    else backward_impl<float>(s);
    hipStream_t st = ctx->stream;
    const size_t n = (size_t)s->p.n;
    if (h_dx0) ND_HIP(hipMemcpyAsync(h_dx0, s->lx.p, 8 * 3 * n, hipMemcpyDeviceToHost, st));
    if (h_dv0) ND_HIP(hipMemcpyAsync(h_dv0, s->lv.p, 8 * 3 * n, hipMemcpyDeviceToHost, st));
    if (h_dq) ND_HIP(hipMemcpyAsync(h_dq, s->dq.p, 8 * n, hipMemcpyDeviceToHost, st));
    if (h_dparams) ND_HIP(hipMemcpyAsync(h_dparams, s->dpar.p, 16, hipMemcpyDeviceToHost, st));
    ND_HIP(hipStreamSynchronize(st));
    // the recompute overwrote the final state: the forward must be rerun before another backward
    s->have_forward = false;
    ND_API_END(ctx)
}

}  // extern "C"
