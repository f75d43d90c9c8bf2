// nd_force.hip -- forces over an explicit pair list: force_lennardjones!,
// force_coulomb!, sum_forces! (src/Forces.jl:15-74).
//
// The reference's arithmetic in Forces.jl is not physical (scalar distance
// broadcast to x/y/z, hard-coded eps=-1e10, sigma=1e-4, no Newton's third
// law); these kernels keep the entry points and compute the textbook forces
//   LJ:      F_i = 24 eps / r^2 [2 (s/r)^12 - (s/r)^6] r_ij,  F_j = -F_i
//   Coulomb: F_i = k q_i q_j r_ij / r^3,                       F_j = -F_i
// with r_ij = x_i - x_j for every listed pair inside the cutoff.  One lane per
// pair.  Both atoms' force components are summed as 64-bit integers in units
// of 2^-32 (integer addition is associative, so the result does not depend on
// the order in which lanes add, and the unit is finer than FP32's at every
// force above 2^-9); the pair energy is summed per workgroup in FP64 and the
// workgroup sums are added in index order by one workgroup.  Every pair index
// is checked against [0, n): a pair outside it is skipped and reported.
#include "nd_internal.hpp"

namespace {

constexpr double kFixScale = 4294967296.0;          // 2^32 units per force unit
constexpr double kFixLimit = 1073741824.0;          // |per-pair component| < 2^30
enum : int32_t { kBadIndex = 1, kFixRange = 2 };

__device__ __forceinline__ double block_sum(double v, double *sh) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (lane == 0) sh[w] = v;
    __syncthreads();
    double s = 0.0;
    if (threadIdx.x == 0)
        for (int k = 0; k < (int)(blockDim.x >> 6); ++k) s += sh[k];
    return s;
}

__device__ __forceinline__ void add_fix(unsigned long long *acc, double f, int32_t *status) {
    if (!(fabs(f) < kFixLimit)) {  // also catches NaN
        atomicOr(status, kFixRange);
        return;
    }
    const long long q = __double2ll_rn(f * kFixScale);
    atomicAdd(acc, (unsigned long long)q);
}

template <int KIND>  // 0: Lennard-Jones, 1: Coulomb
__global__ void __launch_bounds__(nd::kBlock)
k_pair_force(unsigned long long *__restrict__ acc, const nd_pair *__restrict__ pairs, int64_t npairs,
             const float *__restrict__ xyz, const float *__restrict__ charge, int32_t n, float eps,
             float sigma2, float rc2, float kc, float shift, double *__restrict__ partials,
             int32_t *__restrict__ status) {
    __shared__ double sh[nd::kBlock / 64];
    const int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    double e = 0.0;
    if (k < npairs) {
        const int32_t i = pairs[k].i, j = pairs[k].j;
        if ((uint32_t)i >= (uint32_t)n || (uint32_t)j >= (uint32_t)n) {
            atomicOr(status, kBadIndex);
        } else {
            const float dx = xyz[3 * (int64_t)i] - xyz[3 * (int64_t)j];
            const float dy = xyz[3 * (int64_t)i + 1] - xyz[3 * (int64_t)j + 1];
            const float dz = xyz[3 * (int64_t)i + 2] - xyz[3 * (int64_t)j + 2];
            const float r2 = dx * dx + dy * dy + dz * dz;
            if (r2 < rc2) {
                float fs;
                const float ir2 = 1.0f / r2;
                if (KIND == 0) {
                    const float sr2 = sigma2 * ir2;
                    const float sr6 = sr2 * sr2 * sr2;
                    fs = 24.0f * eps * ir2 * sr6 * (2.0f * sr6 - 1.0f);
                    e = (double)(4.0f * eps * sr6 * (sr6 - 1.0f)) - (double)shift;
                } else {
                    const float qq = kc * charge[i] * charge[j];
                    const float ir = sqrtf(ir2);
                    fs = qq * ir * ir2;
                    e = (double)(qq * ir) - (double)(qq * shift);
                }
                const double fx = (double)(fs * dx), fy = (double)(fs * dy), fz = (double)(fs * dz);
                add_fix(&acc[3 * (int64_t)i], fx, status);
                add_fix(&acc[3 * (int64_t)i + 1], fy, status);
                add_fix(&acc[3 * (int64_t)i + 2], fz, status);
                add_fix(&acc[3 * (int64_t)j], -fx, status);
                add_fix(&acc[3 * (int64_t)j + 1], -fy, status);
                add_fix(&acc[3 * (int64_t)j + 2], -fz, status);
            }
        }
    }
    const double s = block_sum(e, sh);
    if (threadIdx.x == 0) partials[blockIdx.x] = s;
}

__global__ void __launch_bounds__(nd::kBlock)
k_fix_to_float(float *__restrict__ force, const unsigned long long *__restrict__ acc, int64_t n3) {
    const int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (k < n3) force[k] = (float)((double)(long long)acc[k] * (1.0 / kFixScale));
}

// Sum of the workgroup energies in a fixed order (one workgroup).
__global__ void __launch_bounds__(nd::kBlock)
k_sum_partials(const double *__restrict__ partials, int64_t m, double *__restrict__ out) {
    __shared__ double sh[nd::kBlock / 64];
    double v = 0.0;
    for (int64_t k = threadIdx.x; k < m; k += blockDim.x) v += partials[k];
    const double s = block_sum(v, sh);
    if (threadIdx.x == 0) out[0] = s;
}

__global__ void __launch_bounds__(nd::kBlock)
k_sum3(float *__restrict__ out, const float *__restrict__ a, const float *__restrict__ b,
       int64_t n3) {
    const int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (k < n3) out[k] = a[k] + b[k];
}

template <int KIND>
void run_pair_force(nd_ctx *ctx, float *d_force, const nd_pair *d_pairs, int64_t npairs,
                    const float *d_xyz, const float *d_charge, int32_t n,
                    const nd_force_params *p, double *h_energy) {
    ND_HIP(hipSetDevice(ctx->device));
    const int64_t n3 = 3 * (int64_t)n;
    const int64_t nblk = npairs > 0 ? (npairs + nd::kBlock - 1) / nd::kBlock : 0;
    ctx->force_acc.reserve(sizeof(unsigned long long) * (size_t)(n3 > 0 ? n3 : 1));
    ctx->force_part.reserve(sizeof(double) * (size_t)(nblk + 1));
    unsigned long long *acc = ctx->force_acc.as<unsigned long long>();
    double *part = ctx->force_part.as<double>();
    nd::clear_status(ctx);
    if (n3 > 0) nd::launch_fill_i32(ctx->stream, reinterpret_cast<int32_t *>(acc), 2 * n3, 0);
    const float rc = KIND == 0 ? p->rc : p->rc_coul;
    float shift = 0.0f;
    if (KIND == 0 && p->shift_lj) {
        const float c2 = p->sigma * p->sigma / (rc * rc), c6 = c2 * c2 * c2;
        shift = 4.0f * p->eps * c6 * (c6 - 1.0f);
    }
    if (KIND == 1 && p->shift_coul) shift = 1.0f / rc;
    if (npairs > 0)
        k_pair_force<KIND><<<(unsigned)nblk, nd::kBlock, 0, ctx->stream>>>(
            acc, d_pairs, npairs, d_xyz, d_charge, n, p->eps, p->sigma * p->sigma, rc * rc, p->k_coul,
            shift, part, ctx->status.as<int32_t>());
    k_sum_partials<<<1, nd::kBlock, 0, ctx->stream>>>(part, nblk, part + nblk);
    if (n3 > 0) k_fix_to_float<<<nd::grid_for(n3), nd::kBlock, 0, ctx->stream>>>(d_force, acc, n3);
    ND_HIP(hipGetLastError());
    double e = 0.0;
    ND_HIP(hipMemcpyAsync(&e, part + nblk, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    const int32_t st = nd::read_status(ctx);  // synchronises
    if (st & kBadIndex) throw nd::Error(ND_ERR_ARG, "pair index outside [0, n)");
    if (st & kFixRange) throw nd::Error(ND_ERR_DOMAIN, "pair force outside the summation range (|F| >= 2^30)");
    if (h_energy) *h_energy = e;
}

}  // namespace

extern "C" {

int nd_force_lennardjones(nd_ctx *ctx, float *d_force, const nd_pair *d_pairs, int64_t npairs,
                          const float *d_xyz, int32_t n, const nd_force_params *p,
                          double *h_energy) {
    ND_API_BEGIN
    ND_REQUIRE(ctx && d_force && d_xyz && p && n >= 0 && npairs >= 0 && (d_pairs || npairs == 0),
               "force_lennardjones!: bad argument");
    run_pair_force<0>(ctx, d_force, d_pairs, npairs, d_xyz, nullptr, n, p, h_energy);
    ND_API_END(ctx)
}

int nd_force_coulomb(nd_ctx *ctx, float *d_force, const nd_pair *d_pairs, int64_t npairs,
                     const float *d_xyz, const float *d_charge, int32_t n,
                     const nd_force_params *p, double *h_energy) {
    ND_API_BEGIN
    ND_REQUIRE(ctx && d_force && d_xyz && d_charge && p && n >= 0 && npairs >= 0 &&
                   (d_pairs || npairs == 0),
               "force_coulomb!: bad argument");
    run_pair_force<1>(ctx, d_force, d_pairs, npairs, d_xyz, d_charge, n, p, h_energy);
    ND_API_END(ctx)
}

int nd_sum_forces(nd_ctx *ctx, float *d_force, const float *d_f1, const float *d_f2, int32_t n) {
    ND_API_BEGIN
    ND_REQUIRE(ctx && d_force && d_f1 && d_f2 && n >= 0, "sum_forces!: bad argument");
    ND_HIP(hipSetDevice(ctx->device));
    const int64_t n3 = 3 * (int64_t)n;
    if (n3 > 0) {
        k_sum3<<<nd::grid_for(n3), nd::kBlock, 0, ctx->stream>>>(d_force, d_f1, d_f2, n3);
        ND_HIP(hipGetLastError());
    }
    ND_HIP(hipStreamSynchronize(ctx->stream));
    ND_API_END(ctx)
}

}  // extern "C"
