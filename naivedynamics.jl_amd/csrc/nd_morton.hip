// nd_morton.hip -- K1: sort keys for the LBVH (masked reference key and the
// interleaved key), bounding box, small fill/gather kernels.
#include "nd_internal.hpp"

namespace {

using nd::cell10;
using nd::floor_i32;
using nd::kBinwidth;
using nd::kMagic0;
using nd::kMagic1;
using nd::kMagic2;
using nd::spread10;

__global__ void __launch_bounds__(nd::kBlock)
k_morton_masked(const float *__restrict__ xyz, int32_t n, int32_t *__restrict__ codes,
                int32_t *__restrict__ idx, int32_t *__restrict__ status) {
    int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float x = xyz[3 * (int64_t)i + 0];
    const float y = xyz[3 * (int64_t)i + 1];
    const float z = xyz[3 * (int64_t)i + 2];
    int32_t qx = 0, qy = 0, qz = 0;
    bool ok = floor_i32(x / kBinwidth, qx);
    ok &= floor_i32(y / kBinwidth, qy);
    ok &= floor_i32(z / kBinwidth, qz);
    codes[i] = (qx & kMagic0) | (qy & kMagic1) | (qz & kMagic2);
    if (idx) idx[i] = i;
    if (!ok) atomicOr(status, nd::kStatusDomain);
}

template <bool HILBERT>
__global__ void __launch_bounds__(nd::kBlock)
k_morton_interleave(const float *__restrict__ xyz, int32_t n, const float *__restrict__ box,
                    int32_t *__restrict__ codes, int32_t *__restrict__ idx,
                    int32_t *__restrict__ status) {
    int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float x = xyz[3 * (int64_t)i + 0];
    const float y = xyz[3 * (int64_t)i + 1];
    const float z = xyz[3 * (int64_t)i + 2];
    bool ok = __builtin_isfinite(x) && __builtin_isfinite(y) && __builtin_isfinite(z);
    const uint32_t cx = cell10(x, box[0], box[3]), cy = cell10(y, box[1], box[4]),
                   cz = cell10(z, box[2], box[5]);
    const uint32_t c = HILBERT ? nd::hilbert30(cx, cy, cz) : (spread10(cx) | (spread10(cy) << 1) | (spread10(cz) << 2));
    codes[i] = (int32_t)c;
    if (idx) idx[i] = i;
    if (!ok) atomicOr(status, nd::kStatusDomain);
}

// Order-preserving float <-> int map for atomic min/max.
__device__ __forceinline__ int32_t f2ord(float f) {
    int32_t i = __float_as_int(f);
    return i >= 0 ? i : i ^ 0x7FFFFFFF;
}
__device__ __forceinline__ float ord2f(int32_t i) {
    return __int_as_float(i >= 0 ? i : i ^ 0x7FFFFFFF);
}

__global__ void k_bbox_init(int32_t *bb) {
    if (threadIdx.x < 3) bb[threadIdx.x] = 0x7FFFFFFF;
    else if (threadIdx.x < 6) bb[threadIdx.x] = (int32_t)0x80000000;
}

__global__ void __launch_bounds__(nd::kBlock)
k_bbox(const float *__restrict__ xyz, int32_t n, int32_t *__restrict__ bb) {
    __shared__ int32_t s[6];
    if (threadIdx.x < 6) s[threadIdx.x] = threadIdx.x < 3 ? 0x7FFFFFFF : (int32_t)0x80000000;
    __syncthreads();
    int32_t mn[3] = {0x7FFFFFFF, 0x7FFFFFFF, 0x7FFFFFFF};
    int32_t mx[3] = {(int32_t)0x80000000, (int32_t)0x80000000, (int32_t)0x80000000};
    for (int32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        for (int d = 0; d < 3; ++d) {
            float v = xyz[3 * (int64_t)i + d];
            if (!__builtin_isfinite(v)) continue;
            int32_t o = f2ord(v);
            mn[d] = min(mn[d], o);
            mx[d] = max(mx[d], o);
        }
    for (int d = 0; d < 3; ++d) {
        atomicMin(&s[d], mn[d]);
        atomicMax(&s[3 + d], mx[d]);
    }
    __syncthreads();
    if (threadIdx.x < 3) atomicMin(&bb[threadIdx.x], s[threadIdx.x]);
    else if (threadIdx.x < 6) atomicMax(&bb[threadIdx.x], s[threadIdx.x]);
}

// bbox (ordered ints) -> (lo[3], scale[3]) floats in place, scale = 1024/(hi-lo).
__global__ void k_bbox_finalize(int32_t *bb) {
    if (threadIdx.x != 0) return;
    float lo[3], sc[3];
    for (int d = 0; d < 3; ++d) {
        lo[d] = ord2f(bb[d]);
        float hi = ord2f(bb[3 + d]);
        float ext = hi - lo[d];
        sc[d] = ext > 0.0f ? 1024.0f / ext : 0.0f;
    }
    float *f = reinterpret_cast<float *>(bb);
    for (int d = 0; d < 3; ++d) { f[d] = lo[d]; f[3 + d] = sc[d]; }
}

__global__ void k_fill_i32(int32_t *p, int64_t n, int32_t v) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

__global__ void __launch_bounds__(nd::kBlock)
k_gather_xyz(const float *__restrict__ xyz, const int32_t *__restrict__ perm, int32_t n,
             float *__restrict__ out) {
    int32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    int64_t o = perm[s];
    out[3 * (int64_t)s + 0] = xyz[3 * o + 0];
    out[3 * (int64_t)s + 1] = xyz[3 * o + 1];
    out[3 * (int64_t)s + 2] = xyz[3 * o + 2];
}

}  // namespace

namespace nd {

void launch_fill_i32(hipStream_t s, int32_t *p, int64_t n, int32_t v) {
    if (n <= 0) return;
    unsigned g = (unsigned)std::min<int64_t>((n + kBlock - 1) / kBlock, 4096);
    k_fill_i32<<<g, kBlock, 0, s>>>(p, n, v);
    ND_HIP(hipGetLastError());
}

void launch_gather_xyz(hipStream_t s, const float *xyz, const int32_t *perm, int32_t n,
                       float *out) {
    if (n <= 0) return;
    k_gather_xyz<<<grid_for(n), kBlock, 0, s>>>(xyz, perm, n, out);
    ND_HIP(hipGetLastError());
}

void launch_bbox(hipStream_t st, const float *xyz, int32_t n, float *bbox6) {
    int32_t *bb = reinterpret_cast<int32_t *>(bbox6);
    k_bbox_init<<<1, 64, 0, st>>>(bb);
    // few workgroups: their atomics on the six words serialise in L2
    unsigned g = (unsigned)std::max<int64_t>(1, std::min<int64_t>(grid_for(n), 64));
    k_bbox<<<g, kBlock, 0, st>>>(xyz, n, bb);
    k_bbox_finalize<<<1, 64, 0, st>>>(bb);
    ND_HIP(hipGetLastError());
}

void launch_morton_box(hipStream_t st, float *box, const float *xyz, int32_t n, int32_t mode,
                       int32_t *codes, int32_t *idx, int32_t *status) {
    if (n <= 0) return;
    if (mode == ND_MORTON_MASKED) {
        k_morton_masked<<<grid_for(n), kBlock, 0, st>>>(xyz, n, codes, idx, status);
    } else {
        launch_bbox(st, xyz, n, box);
        if (mode == ND_MORTON_HILBERT)
            k_morton_interleave<true><<<grid_for(n), kBlock, 0, st>>>(xyz, n, box, codes, idx, status);
        else
            k_morton_interleave<false><<<grid_for(n), kBlock, 0, st>>>(xyz, n, box, codes, idx, status);
    }
    ND_HIP(hipGetLastError());
}

void launch_morton(nd_ctx *ctx, const float *xyz, int32_t n, int32_t mode, const float lo[3],
                   const float scale[3], int32_t *codes, int32_t *idx, int32_t *status) {
    if (n <= 0) return;
    ctx->bbox.reserve(64);
    float *box = ctx->bbox.as<float>();
    if (mode != ND_MORTON_MASKED && lo && scale) {
        float h[6] = {lo[0], lo[1], lo[2], scale[0], scale[1], scale[2]};
        ND_HIP(hipMemcpyAsync(box, h, sizeof(h), hipMemcpyHostToDevice, ctx->stream));
        if (mode == ND_MORTON_HILBERT)
            k_morton_interleave<true><<<grid_for(n), kBlock, 0, ctx->stream>>>(xyz, n, box, codes, idx, status);
        else
            k_morton_interleave<false><<<grid_for(n), kBlock, 0, ctx->stream>>>(xyz, n, box, codes, idx, status);
        ND_HIP(hipGetLastError());
        return;
    }
    launch_morton_box(ctx->stream, box, xyz, n, mode, codes, idx, status);
}

}  // namespace nd

extern "C" int nd_mortoncodes(nd_ctx *ctx, const float *d_xyz, int32_t n, int32_t mode,
                              const float *h_lo3, const float *h_scale3, int32_t *d_codes) {
    ND_API_BEGIN
    ND_REQUIRE(ctx && d_xyz && d_codes && n >= 0, "nd_mortoncodes: bad argument");
    ND_REQUIRE(mode == ND_MORTON_MASKED || mode == ND_MORTON_INTERLEAVE || mode == ND_MORTON_HILBERT,
               "nd_mortoncodes: unknown key mode");
    ND_REQUIRE((h_lo3 == nullptr) == (h_scale3 == nullptr),
               "nd_mortoncodes: pass both lo and scale, or neither");
    ND_HIP(hipSetDevice(ctx->device));
    nd::launch_morton(ctx, d_xyz, n, mode, h_lo3, h_scale3, d_codes, nullptr,
                      ctx->status.as<int32_t>());
    if (nd::read_status(ctx) & nd::kStatusDomain)
        throw nd::Error(ND_ERR_DOMAIN,
                        "mortoncodes!: coordinate outside the Int32 quantisation domain "
                        "(reference throws InexactError)");
    ND_API_END(ctx)
}
