// nd_cells.hip -- uniform cell-list neighbour search: the comparison point the
// reference benchmarks its LBVH against (CellListMap's neighborlist(x, cutoff),
// assets/BVHBenchSuite.jl:307; also called at src/MDInput.jl:290).
//
// Cells of side >= cutoff on a grid over the bounding box of the finite
// coordinates; atoms stably sorted by cell number (x fastest), so the three
// cells cx-1 .. cx+1 of one (cy, cz) row are one contiguous run of the sorted
// copy.  Every unordered pair is found once from its earlier atom in cell
// order through a half shell of five runs: the rest of its own cell plus cell
// cx+1, the row (cy+1, cz), and the three rows (cy-1 .. cy+1, cz+1).  One lane
// per sorted atom; a counting pass, an exclusive scan and an emitting pass
// place each atom's pairs at a fixed offset, so the output order is
// deterministic.  Distances use the reference's exact fold and both of its
// predicates, so the pair set and every distance equal the brute-force
// list's bit for bit.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "nd_internal.hpp"

namespace {

// Order-preserving int image of a finite float (for atomic min / max).
__device__ __forceinline__ int32_t f2ord(float v) {
    const int32_t b = __float_as_int(v);
    return b >= 0 ? b : (b ^ 0x7FFFFFFF);
}
inline float ord2f_host(int32_t o) {
    const int32_t b = o >= 0 ? o : (o ^ 0x7FFFFFFF);
    float f;
    std::memcpy(&f, &b, sizeof f);
    return f;
}

__global__ void k_cell_bbox_init(int32_t *bb) {
    if (threadIdx.x < 3) bb[threadIdx.x] = 0x7FFFFFFF;
    else if (threadIdx.x < 6) bb[threadIdx.x] = (int32_t)0x80000000;
}

// Per-axis min / max over the atoms whose three coordinates are all finite.
__global__ void __launch_bounds__(nd::kBlock)
k_cell_bbox(const float *__restrict__ xyz, int32_t n, int32_t *__restrict__ bb) {
    __shared__ int32_t s[6];
    if (threadIdx.x < 6) s[threadIdx.x] = threadIdx.x < 3 ? 0x7FFFFFFF : (int32_t)0x80000000;
    __syncthreads();
    int32_t mn[3] = {0x7FFFFFFF, 0x7FFFFFFF, 0x7FFFFFFF};
    int32_t mx[3] = {(int32_t)0x80000000, (int32_t)0x80000000, (int32_t)0x80000000};
    for (int32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const float v[3] = {xyz[3 * (int64_t)i], xyz[3 * (int64_t)i + 1], xyz[3 * (int64_t)i + 2]};
        if (!(__builtin_isfinite(v[0]) && __builtin_isfinite(v[1]) && __builtin_isfinite(v[2])))
            continue;
        for (int d = 0; d < 3; ++d) {
            mn[d] = min(mn[d], f2ord(v[d]));
            mx[d] = max(mx[d], f2ord(v[d]));
        }
    }
    for (int d = 0; d < 3; ++d) {
        atomicMin(&s[d], mn[d]);
        atomicMax(&s[3 + d], mx[d]);
    }
    __syncthreads();
    if (threadIdx.x < 3) atomicMin(&bb[threadIdx.x], s[threadIdx.x]);
    else if (threadIdx.x < 6) atomicMax(&bb[threadIdx.x], s[threadIdx.x]);
}

template <int STRIDE>
__global__ void __launch_bounds__(nd::kBlock)
k_cell_keys(const float *__restrict__ xyz, int32_t n, nd::CellGrid cg, int32_t *__restrict__ keys,
            int32_t *__restrict__ vals) {
    const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float *p = xyz + (int64_t)STRIDE * i;
    const float x = p[0], y = p[1], z = p[2];
    int32_t k = cg.ncell;
    if (__builtin_isfinite(x) && __builtin_isfinite(y) && __builtin_isfinite(z))
        k = (nd::cell_coord(cg, z, 2) * cg.g[1] + nd::cell_coord(cg, y, 1)) * cg.g[0] +
            nd::cell_coord(cg, x, 0);
    keys[i] = k;
    vals[i] = i;
}

template <int STRIDE>
__global__ void __launch_bounds__(nd::kBlock)
k_cell_gather(const float *__restrict__ xyz, const int32_t *__restrict__ vals, int32_t n,
              float4 *__restrict__ cpos) {
    const int32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int32_t i = vals[t];
    const float *p = xyz + (int64_t)STRIDE * i;
    cpos[t] = make_float4(p[0], p[1], p[2], __int_as_float(i));
}

// cstart[k] = lower_bound(keys, k) for k = 0..ncell.
__global__ void __launch_bounds__(nd::kBlock)
k_cell_start(const int32_t *__restrict__ keys, int32_t n, int32_t ncell, int32_t *__restrict__ cstart) {
    const int32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k > ncell) return;
    int32_t lo = 0, hi = n;
    while (lo < hi) {
        const int32_t mid = (int32_t)(((uint32_t)lo + (uint32_t)hi) >> 1);
        if (keys[mid] < k) lo = mid + 1;
        else hi = mid;
    }
    cstart[k] = lo;
}

// Half-shell scan of sorted atom t (see the file comment); pairs are
// (min, max) of the two original indices and the distance.
template <bool EMIT>
__global__ void __launch_bounds__(nd::kBlock)
k_cell_pairs(const float4 *__restrict__ cpos, const int32_t *__restrict__ keys, int32_t n,
             nd::CellGrid cg, const int32_t *__restrict__ cstart, float r, float r2, int32_t pred,
             int64_t *__restrict__ counts, const int64_t *__restrict__ offsets,
             nd_pair *__restrict__ out) {
    const int32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int32_t key = keys[t];
    int64_t c = 0;
    nd_pair *o = nullptr;
    int64_t limit = 0;
    if (EMIT) {
        o = out + offsets[t];
        limit = offsets[t + 1] - offsets[t];
    }
    if (key < cg.ncell) {
        const float4 p = cpos[t];
        const int32_t oi = __float_as_int(p.w);
        const int32_t g0 = cg.g[0], g1 = cg.g[1], g2 = cg.g[2];
        const int32_t cx = key % g0, cyz = key / g0, cy = cyz % g1, cz = cyz / g1;
        auto run = [&](int32_t b, int32_t e) {
            for (int32_t u = b; u < e; ++u) {
                const float4 q = cpos[u];
                float d;
                if (nd::accept(nd::sqdist(p.x, p.y, p.z, q.x, q.y, q.z), r, r2, pred, d)) {
                    if (EMIT && c < limit) {
                        const int32_t oj = __float_as_int(q.w);
                        nd_pair pr;
                        pr.i = oi < oj ? oi : oj;
                        pr.j = oi < oj ? oj : oi;
                        pr.d = d;
                        o[c] = pr;
                    }
                    ++c;
                }
            }
        };
        const int32_t x0 = cx > 0 ? cx - 1 : 0, x1 = cx + 1 < g0 ? cx + 1 : g0 - 1;
        run(t + 1, cstart[key + (cx + 1 < g0 ? 2 : 1)]);
        if (cy + 1 < g1) {
            const int32_t b = (cz * g1 + cy + 1) * g0;
            run(cstart[b + x0], cstart[b + x1 + 1]);
        }
        if (cz + 1 < g2)
            for (int32_t y = cy > 0 ? cy - 1 : 0; y <= cy + 1 && y < g1; ++y) {
                const int32_t b = ((cz + 1) * g1 + y) * g0;
                run(cstart[b + x0], cstart[b + x1 + 1]);
            }
    }
    if (!EMIT) counts[t] = c;
}

}  // namespace

namespace nd {

CellGrid make_cell_grid(const double lo[3], const double ext[3], double r, int64_t max_cells) {
    CellGrid cg{};
    const double side = r * (1.0 + 1e-4);
    int64_t g[3];
    for (int d = 0; d < 3; ++d) {
        const double q = ext[d] > 0.0 ? std::floor(ext[d] / side) : 1.0;
        g[d] = q < 1.0 ? 1 : (q > 1024.0 ? 1024 : (int64_t)q);
    }
    if (max_cells < 1) max_cells = 1;
    while (g[0] * g[1] * g[2] > max_cells) {
        const double f = std::cbrt((double)max_cells / (double)(g[0] * g[1] * g[2]));
        for (int d = 0; d < 3; ++d) g[d] = std::max<int64_t>(1, (int64_t)std::floor((double)g[d] * f));
    }
    for (int d = 0; d < 3; ++d) {
        cg.lo[d] = lo[d];
        cg.g[d] = (int32_t)g[d];
        cg.inv[d] = ext[d] > 0.0 ? (double)g[d] / ext[d] : 0.0;
    }
    cg.ncell = (int32_t)(g[0] * g[1] * g[2]);
    return cg;
}

const int32_t *launch_cell_index(hipStream_t st, SortBufs &sb, const float *xyz, int32_t stride,
                                 int32_t n, const CellGrid &cg, int32_t *keys, int32_t *vals,
                                 float4 *cpos, int32_t *cstart) {
    if (stride != 3 && stride != 4) throw Error(ND_ERR_ARG, "cell index: stride must be 3 or 4");
    const unsigned g = grid_for(n);
    if (stride == 3) k_cell_keys<3><<<g, kBlock, 0, st>>>(xyz, n, cg, keys, vals);
    else k_cell_keys<4><<<g, kBlock, 0, st>>>(xyz, n, cg, keys, vals);
    // keys <= ncell < 2^key_bits
    int32_t key_bits = 1;
    while (key_bits < 31 && ((int64_t)1 << key_bits) <= (int64_t)cg.ncell) ++key_bits;
    int32_t *ks = keys, *vs = vals;
    radix_sort_keys10(st, sb, keys, vals, n, key_bits, &ks, &vs);
    if (stride == 3) k_cell_gather<3><<<g, kBlock, 0, st>>>(xyz, vs, n, cpos);
    else k_cell_gather<4><<<g, kBlock, 0, st>>>(xyz, vs, n, cpos);
    k_cell_start<<<grid_for((int64_t)cg.ncell + 1), kBlock, 0, st>>>(ks, n, cg.ncell, cstart);
    ND_HIP(hipGetLastError());
    return ks;
}

}  // namespace nd

extern "C" {

int nd_cell_neighborlist(nd_ctx *ctx, const float *d_xyz, int32_t n, float cutoff,
                         int32_t predicate, nd_pair *d_out, int64_t cap, int64_t *count) {
    ND_API_BEGIN
    ND_REQUIRE(ctx && count && n >= 0 && (d_xyz || n == 0), "nd_cell_neighborlist: bad argument");
    ND_REQUIRE(std::isfinite(cutoff) && cutoff > 0.0f, "nd_cell_neighborlist: cutoff must be finite and > 0");
    ND_REQUIRE(predicate == ND_PRED_SCALAR || predicate == ND_PRED_SIMD, "unknown pair predicate");
    ND_HIP(hipSetDevice(ctx->device));
    *count = 0;
    if (n < 2) return ND_OK;
    hipStream_t st = ctx->stream;
    ctx->bbox.reserve(64);
    int32_t *bb = ctx->bbox.as<int32_t>();
    k_cell_bbox_init<<<1, 64, 0, st>>>(bb);
    // few workgroups: their atomics on the six words serialise in L2
    k_cell_bbox<<<(unsigned)std::min<int64_t>(nd::grid_for(n), 64), nd::kBlock, 0, st>>>(d_xyz, n, bb);
    ND_HIP(hipGetLastError());
    int32_t hb[6];
    ND_HIP(hipMemcpyAsync(hb, bb, sizeof hb, hipMemcpyDeviceToHost, st));
    ND_HIP(hipStreamSynchronize(st));
    double lo[3], ext[3];
    for (int d = 0; d < 3; ++d) {
        if (hb[d] > hb[3 + d]) return ND_OK;  // no atom with finite coordinates
        lo[d] = (double)ord2f_host(hb[d]);
        ext[d] = (double)ord2f_host(hb[3 + d]) - lo[d];
    }
    const nd::CellGrid cg =
        nd::make_cell_grid(lo, ext, (double)cutoff, std::min<int64_t>(1 << 22, 2 * (int64_t)n + 64));
    ctx->cell_keys.reserve(sizeof(int32_t) * (size_t)n);
    ctx->cell_vals.reserve(sizeof(int32_t) * (size_t)n);
    ctx->cell_pos.reserve(sizeof(float4) * (size_t)n);
    ctx->cell_start.reserve(sizeof(int32_t) * ((size_t)cg.ncell + 1));
    float4 *cpos = ctx->cell_pos.as<float4>();
    int32_t *cstart = ctx->cell_start.as<int32_t>();
    const int32_t *keys = nd::launch_cell_index(st, ctx->sort, d_xyz, 3, n, cg, ctx->cell_keys.as<int32_t>(),
                                                ctx->cell_vals.as<int32_t>(), cpos, cstart);
    ctx->counts.reserve(sizeof(int64_t) * ((size_t)n + 1));
    ctx->offsets.reserve(sizeof(int64_t) * ((size_t)n + 1));
    const float r2 = cutoff * cutoff;
    k_cell_pairs<false><<<nd::grid_for(n), nd::kBlock, 0, st>>>(
        cpos, keys, n, cg, cstart, cutoff, r2, predicate, ctx->counts.as<int64_t>(), nullptr, nullptr);
    ND_HIP(hipGetLastError());
    const int64_t total = nd::scan_counts(ctx, n);
    *count = total;
    if (total > cap || (total > 0 && d_out == nullptr))
        throw nd::Error(ND_ERR_CAPACITY, "pair buffer too small; *count = required");
    if (total == 0) return ND_OK;
    k_cell_pairs<true><<<nd::grid_for(n), nd::kBlock, 0, st>>>(
        cpos, keys, n, cg, cstart, cutoff, r2, predicate, nullptr, ctx->offsets.as<int64_t>(), d_out);
    ND_HIP(hipGetLastError());
    ND_HIP(hipStreamSynchronize(st));
    ND_API_END(ctx)
}

int nd_cell_neighborlist_host(nd_ctx *ctx, const float *h_xyz, int32_t n, float cutoff,
                              int32_t predicate, nd_pair *h_out, int64_t cap, int64_t *count) {
    ND_API_BEGIN
    ND_REQUIRE(ctx && count && n >= 0 && (h_xyz || n == 0), "nd_cell_neighborlist_host: bad argument");
    ND_HIP(hipSetDevice(ctx->device));
    ctx->xyz_in.reserve(sizeof(float) * 3 * (size_t)n);
    if (n > 0)
        ND_HIP(hipMemcpyAsync(ctx->xyz_in.p, h_xyz, sizeof(float) * 3 * (size_t)n, hipMemcpyHostToDevice,
                              ctx->stream));
    if (cap > 0) ctx->pairs_out.reserve(sizeof(nd_pair) * (size_t)cap);
    const int rc = nd_cell_neighborlist(ctx, ctx->xyz_in.as<float>(), n, cutoff, predicate,
                                        cap > 0 ? ctx->pairs_out.as<nd_pair>() : nullptr, cap, count);
    if (rc != ND_OK) return rc;
    if (*count > 0 && h_out)
        ND_HIP(hipMemcpyAsync(h_out, ctx->pairs_out.p, sizeof(nd_pair) * (size_t)*count, hipMemcpyDeviceToHost,
                              ctx->stream));
    ND_HIP(hipStreamSynchronize(ctx->stream));
    ND_API_END(ctx)
}

}  // extern "C"
