// nd_sort.hip -- K2: stable LSD radix sort of (int32 key, int32 value) pairs in
// signed-key order, for sortperm (BVHTraverse.jl:570-573): 8-bit digits in
// general, three 10-bit digits for the simulation's 30-bit keys.
//
// Per pass: (1) per-tile digit histograms, (2) one exclusive scan over the
// digit-major [digit][tile] table, (3) stable scatter.  The scatter ranks
// each 64-key round of a wave with one ballot per digit bit (lanes holding
// the same digit form a peer mask; rank = popcount of lower peers), carries
// running per-wave digit counters in LDS, stages the tile in LDS in sorted
// order and writes every digit run to global memory contiguously.
#include "nd_internal.hpp"

namespace {

constexpr int kWaves = nd::kBlock / 64;     // 4
constexpr int kRounds = 4;                  // 64-key rounds per wave
constexpr int kTile = nd::kBlock * kRounds; // 1024 keys per workgroup

template <int BITS>
__device__ __forceinline__ uint32_t digit_of(int32_t key, int shift) {
    return (((uint32_t)key ^ 0x80000000u) >> shift) & ((1u << BITS) - 1u);
}

// Per-tile histogram of one digit (2^BITS bins).
template <int BITS>
__global__ void __launch_bounds__(nd::kBlock)
k_radix_hist(const int32_t *__restrict__ keys, int32_t n, int shift, int32_t ntiles,
             int32_t *__restrict__ hist) {
    constexpr int R = 1 << BITS, DPT = R / nd::kBlock;
    __shared__ int32_t cnt[R];
    const int t = threadIdx.x;
#pragma unroll
    for (int c = 0; c < DPT; ++c) cnt[t * DPT + c] = 0;
    __syncthreads();
    const int64_t base = (int64_t)blockIdx.x * kTile;
#pragma unroll
    for (int r = 0; r < kRounds; ++r) {
        int64_t i = base + r * nd::kBlock + t;
        if (i < n) atomicAdd(&cnt[digit_of<BITS>(keys[i], shift)], 1);
    }
    __syncthreads();
#pragma unroll
    for (int c = 0; c < DPT; ++c) hist[(int64_t)(t * DPT + c) * ntiles + blockIdx.x] = cnt[t * DPT + c];
}

// Exclusive scan of v over the 256 lanes of a workgroup; returns the prefix
// and leaves the workgroup total in *total (shared).
__device__ __forceinline__ int32_t block_exclusive_scan256(int32_t v, int32_t *wsum, int32_t *total) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    int32_t inc = v;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const int32_t u = __shfl_up(inc, off, 64);
        if (lane >= off) inc += u;
    }
    if (lane == 63) wsum[w] = inc;
    __syncthreads();
    int32_t base = 0;
    for (int k = 0; k < w; ++k) base += wsum[k];
    if (threadIdx.x == nd::kBlock - 1) *total = base + inc;
    __syncthreads();
    return base + inc - v;
}

// One workgroup per digit: exclusive scan of that digit's row of tile counts
// (hist[d][0..ntiles)) in place; the row total goes to totals[d].
__global__ void __launch_bounds__(nd::kBlock)
k_radix_scan_rows(int32_t *__restrict__ hist, int32_t ntiles, int32_t *__restrict__ totals) {
    __shared__ int32_t wsum[kWaves];
    __shared__ int32_t tot;
    int32_t *row = hist + (int64_t)blockIdx.x * ntiles;
    int32_t carry = 0;
    for (int32_t c0 = 0; c0 < ntiles; c0 += nd::kBlock) {
        const int32_t k = c0 + threadIdx.x;
        const int32_t v = k < ntiles ? row[k] : 0;
        const int32_t ex = block_exclusive_scan256(v, wsum, &tot);
        if (k < ntiles) row[k] = carry + ex;
        carry += tot;
        __syncthreads();
    }
    if (threadIdx.x == 0) totals[blockIdx.x] = carry;
}

// Stable scatter of one tile for one digit.  Thread t owns digits
// t * DPT .. t * DPT + DPT - 1 in the per-digit bookkeeping.
template <int BITS>
__global__ void __launch_bounds__(nd::kBlock)
k_radix_scatter(const int32_t *__restrict__ keys_in, const int32_t *__restrict__ vals_in,
                int32_t n, int shift, int32_t ntiles, const int32_t *__restrict__ hist,
                const int32_t *__restrict__ totals, int32_t *__restrict__ keys_out,
                int32_t *__restrict__ vals_out) {
    constexpr int R = 1 << BITS, DPT = R / nd::kBlock;
    __shared__ int32_t wcnt[kWaves][R];
    __shared__ int32_t bpre[R];
    __shared__ int32_t gofs[R];
    __shared__ int32_t wsum[kWaves];
    __shared__ int32_t tot0;
    __shared__ int32_t sk[kTile];
    __shared__ int32_t sv[kTile];

    const int t = threadIdx.x;
    const int w = t >> 6;
    const int lane = t & 63;
    const int64_t base = (int64_t)blockIdx.x * kTile;
    const int32_t nvalid = (int32_t)((n - base) < kTile ? (n - base) : kTile);

    // global start of digit d in this tile = (keys with smaller digits) +
    // (keys with digit d in earlier tiles)
    {
        int32_t tv[DPT], sum = 0;
#pragma unroll
        for (int c = 0; c < DPT; ++c) { tv[c] = totals[t * DPT + c]; sum += tv[c]; }
        int32_t run = block_exclusive_scan256(sum, wsum, &tot0);
#pragma unroll
        for (int c = 0; c < DPT; ++c) {
            const int d = t * DPT + c;
            gofs[d] = run + hist[(int64_t)d * ntiles + blockIdx.x];
            run += tv[c];
        }
    }
#pragma unroll
    for (int k = 0; k < kWaves; ++k)
#pragma unroll
        for (int c = 0; c < DPT; ++c) wcnt[k][t * DPT + c] = 0;
    __syncthreads();

    // rank each 64-key round of a wave: lanes with the same digit (found with
    // one ballot per digit bit) form a peer mask; rank = running count of the
    // digit in this wave + number of lower peers
    const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    int32_t key[kRounds], val[kRounds], rank[kRounds];
    uint32_t dig[kRounds];
#pragma unroll
    for (int r = 0; r < kRounds; ++r) {
        const int32_t li = w * (64 * kRounds) + r * 64 + lane;  // position within the tile
        const bool valid = li < nvalid;
        key[r] = valid ? keys_in[base + li] : 0;
        val[r] = valid ? vals_in[base + li] : 0;
        const uint32_t d = digit_of<BITS>(key[r], shift);
        dig[r] = d;
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < BITS; ++b) {
            const bool bit = (d >> b) & 1u;
            const uint64_t m = __ballot(bit);
            peers &= bit ? m : ~m;
        }
        int32_t before = 0;
        if (valid) before = wcnt[w][d];
        rank[r] = before + __popcll(peers & lt);
        if (valid && (peers & lt) == 0) wcnt[w][d] = before + __popcll(peers);
    }
    __syncthreads();

    // per digit: offsets of each wave's run, and the block-level digit prefix
    {
        int32_t tv[DPT], sum = 0;
#pragma unroll
        for (int c = 0; c < DPT; ++c) {
            const int d = t * DPT + c;
            int32_t tot = 0;
#pragma unroll
            for (int k = 0; k < kWaves; ++k) {
                const int32_t cw = wcnt[k][d];
                wcnt[k][d] = tot;
                tot += cw;
            }
            tv[c] = tot;
            sum += tot;
        }
        int32_t run = block_exclusive_scan256(sum, wsum, &tot0);
#pragma unroll
        for (int c = 0; c < DPT; ++c) {
            bpre[t * DPT + c] = run;
            run += tv[c];
        }
    }
    __syncthreads();

#pragma unroll
    for (int r = 0; r < kRounds; ++r) {
        const int32_t li = w * (64 * kRounds) + r * 64 + lane;
        if (li < nvalid) {
            const int32_t pos = bpre[dig[r]] + wcnt[w][dig[r]] + rank[r];
            sk[pos] = key[r];
            sv[pos] = val[r];
        }
    }
    __syncthreads();

    // every digit run leaves contiguously
#pragma unroll
    for (int r = 0; r < kRounds; ++r) {
        const int32_t k = r * nd::kBlock + t;
        if (k < nvalid) {
            const int32_t kk = sk[k];
            const uint32_t d = digit_of<BITS>(kk, shift);
            const int64_t g = (int64_t)gofs[d] + (k - bpre[d]);
            if (g >= 0 && g < n) {
                keys_out[g] = kk;
                vals_out[g] = sv[k];
            }
        }
    }
}

__global__ void __launch_bounds__(nd::kBlock)
k_keybits(const int32_t *__restrict__ keys, int32_t n, uint32_t *__restrict__ out) {
    uint32_t o = 0u, a = 0xFFFFFFFFu;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        uint32_t k = (uint32_t)keys[i] ^ 0x80000000u;
        o |= k;
        a &= k;
    }
    // one atomic pair per workgroup: atomics on one address serialise in
    // one L2 channel, so the grid is kept small (see the launch)
    __shared__ uint32_t so, sa;
    if (threadIdx.x == 0) { so = 0u; sa = 0xFFFFFFFFu; }
    __syncthreads();
    atomicOr(&so, o);
    atomicAnd(&sa, a);
    __syncthreads();
    if (threadIdx.x == 0) {
        atomicOr(&out[0], so);
        atomicAnd(&out[1], sa);
    }
}

__global__ void k_keybits_init(uint32_t *out) {
    if (threadIdx.x == 0) { out[0] = 0u; out[1] = 0xFFFFFFFFu; }
}

__global__ void __launch_bounds__(nd::kBlock) k_iota(int32_t *p, int32_t n) {
    int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = i;
}

// quantized_positions! (BVHTraverse.jl:290-314), one axis: the coordinate of
// atom order[i] (i itself on the first axis) as a signed key in Julia's
// isless order on Float32 (-0 before +0, every NaN after +Inf and tied).
__global__ void __launch_bounds__(nd::kBlock)
k_rank_keys(const float *__restrict__ xyz, int32_t n, int32_t axis, int32_t first, int32_t *__restrict__ keys,
            int32_t *__restrict__ order) {
    const int32_t i = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= n) return;
    const int32_t a = first ? i : order[i];
    const float v = xyz[3 * (int64_t)a + axis];
    const int32_t b = __float_as_int(v);
    keys[i] = (v != v) ? INT32_MAX : (b < 0 ? (b ^ 0x7FFFFFFF) : b);
    order[i] = a;
}

// q[atom][axis] = 1-based place of the atom in the axis' stable sort
__global__ void __launch_bounds__(nd::kBlock)
k_rank_scatter(const int32_t *__restrict__ order, int32_t n, int32_t axis, int32_t *__restrict__ q) {
    const int32_t i = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i < n) q[3 * (int64_t)order[i] + axis] = i + 1;
}

}  // namespace

namespace nd {

size_t radix_scratch_blocks(int32_t n) { return (size_t)((n + kTile - 1) / kTile); }

namespace {

// One LSD pass per BITS-bit digit from first_bit up to end_bit; returns the
// number of passes (odd: the result is in the alternate buffers).
template <int BITS>
int radix_passes(hipStream_t s, SortBufs &b, int32_t *keys, int32_t *vals, int32_t n,
                 int32_t first_bit, int32_t end_bit) {
    constexpr int R = 1 << BITS;
    const int32_t ntiles = (int32_t)radix_scratch_blocks(n);
    b.keys_alt.reserve(sizeof(int32_t) * (size_t)n);
    b.vals_alt.reserve(sizeof(int32_t) * (size_t)n);
    b.hist.reserve(sizeof(int32_t) * ((size_t)R * (size_t)ntiles + R));
    int32_t *hist = b.hist.as<int32_t>(), *totals = hist + (size_t)R * (size_t)ntiles;
    int32_t *ki = keys, *vi = vals, *ko = b.keys_alt.as<int32_t>(), *vo = b.vals_alt.as<int32_t>();
    int passes = 0;
    for (int shift = first_bit; shift < end_bit; shift += BITS) {
        k_radix_hist<BITS><<<ntiles, kBlock, 0, s>>>(ki, n, shift, ntiles, hist);
        k_radix_scan_rows<<<R, kBlock, 0, s>>>(hist, ntiles, totals);
        k_radix_scatter<BITS><<<ntiles, kBlock, 0, s>>>(ki, vi, n, shift, ntiles, hist, totals, ko, vo);
        std::swap(ki, ko);
        std::swap(vi, vo);
        ++passes;
    }
    ND_HIP(hipGetLastError());
    return passes;
}

}  // namespace

void radix_sort_pairs(hipStream_t s, SortBufs &b, int32_t *keys, int32_t *vals, int32_t n,
                      int32_t first_bit, int32_t end_bit) {
    if (n <= 1 || end_bit <= first_bit) return;
    const int passes = radix_passes<8>(s, b, keys, vals, n, first_bit, end_bit);
    if (passes & 1) {
        ND_HIP(hipMemcpyAsync(keys, b.keys_alt.p, sizeof(int32_t) * (size_t)n, hipMemcpyDeviceToDevice, s));
        ND_HIP(hipMemcpyAsync(vals, b.vals_alt.p, sizeof(int32_t) * (size_t)n, hipMemcpyDeviceToDevice, s));
    }
}

void radix_sort_keys10(hipStream_t s, SortBufs &b, int32_t *keys, int32_t *vals, int32_t n,
                       int32_t key_bits, int32_t **keys_sorted, int32_t **vals_sorted) {
    *keys_sorted = keys;
    *vals_sorted = vals;
    if (n <= 1) return;
    const int passes = radix_passes<10>(s, b, keys, vals, n, 0, key_bits);
    if (passes & 1) {
        *keys_sorted = b.keys_alt.as<int32_t>();
        *vals_sorted = b.vals_alt.as<int32_t>();
    }
}

// Sorts keys/vals in place over only the 8-bit digit windows whose bits vary
// (synchronises once to read the varying-bit mask).
void radix_sort_pairs_auto(nd_ctx *ctx, int32_t *keys, int32_t *vals, int32_t n) {
    if (n <= 1) return;
    ctx->sort.flags.reserve(64);
    uint32_t *fl = ctx->sort.flags.as<uint32_t>();
    k_keybits_init<<<1, 64, 0, ctx->stream>>>(fl);
    unsigned g = (unsigned)std::min<int64_t>(grid_for(n), 64);
    k_keybits<<<g, kBlock, 0, ctx->stream>>>(keys, n, fl);
    uint32_t h[2];
    ND_HIP(hipMemcpyAsync(h, fl, sizeof(h), hipMemcpyDeviceToHost, ctx->stream));
    ND_HIP(hipStreamSynchronize(ctx->stream));
    const uint32_t vary = h[0] ^ h[1];
    for (int shift = 0; shift < 32; shift += 8)
        if ((vary >> shift) & 0xFFu) radix_sort_pairs(ctx->stream, ctx->sort, keys, vals, n, shift, shift + 8);
}

void launch_iota(hipStream_t s, int32_t *p, int32_t n) {
    if (n > 0) k_iota<<<grid_for(n), kBlock, 0, s>>>(p, n);
    ND_HIP(hipGetLastError());
}

}  // namespace nd

extern "C" int nd_sortperm(nd_ctx *ctx, const int32_t *d_codes, int32_t n,
                           int32_t *d_codes_sorted, int32_t *d_perm) {
    ND_API_BEGIN
    ND_REQUIRE(ctx && d_codes && d_codes_sorted && d_perm && n >= 0, "nd_sortperm: bad argument");
    ND_HIP(hipSetDevice(ctx->device));
    if (d_codes_sorted != d_codes)
        ND_HIP(hipMemcpyAsync(d_codes_sorted, d_codes, sizeof(int32_t) * (size_t)n,
                              hipMemcpyDeviceToDevice, ctx->stream));
    nd::launch_iota(ctx->stream, d_perm, n);
    nd::radix_sort_pairs_auto(ctx, d_codes_sorted, d_perm, n);
    ND_HIP(hipStreamSynchronize(ctx->stream));
    ND_API_END(ctx)
}

extern "C" int nd_quantized_positions(nd_ctx *ctx, const float *d_xyz, int32_t n, int32_t *d_quantized) {
    ND_API_BEGIN
    ND_REQUIRE(ctx && n >= 0 && (n == 0 || (d_xyz && d_quantized)), "nd_quantized_positions: bad argument");
    ND_HIP(hipSetDevice(ctx->device));
    if (n == 0) return ND_OK;
    ctx->rank_keys.reserve(2 * sizeof(int32_t) * (size_t)n);
    int32_t *keys = ctx->rank_keys.as<int32_t>(), *order = keys + n;
    // x, then y and z, each a stable sort of the order the previous axis left
    for (int32_t axis = 0; axis < 3; ++axis) {
        k_rank_keys<<<nd::grid_for(n), nd::kBlock, 0, ctx->stream>>>(d_xyz, n, axis, axis == 0, keys, order);
        ND_HIP(hipGetLastError());
        nd::radix_sort_pairs_auto(ctx, keys, order, n);
        k_rank_scatter<<<nd::grid_for(n), nd::kBlock, 0, ctx->stream>>>(order, n, axis, d_quantized);
        ND_HIP(hipGetLastError());
    }
    ND_HIP(hipStreamSynchronize(ctx->stream));
    ND_API_END(ctx)
}
