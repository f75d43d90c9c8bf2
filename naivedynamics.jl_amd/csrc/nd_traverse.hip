// nd_traverse.hip -- K4 (reference form): leaf-pair stackless traversal that
// emits a compacted (i, j, d) pair list, plus the one-call neighbour list,
// tree export and the brute-force pair list.
//
// leafneighbor_traverse (BVHTraverse.jl:1236-1323): for every query leaf q,
// same-leaf pairs first (onecluster_proximitytest!, :1021-1040), then a rope
// walk starting at q's skip: overlapping internal node -> go left,
// overlapping leaf -> apl x apl tile test (twocluster_proximitytest!,
// :1042-1055) and follow its skip, no overlap -> follow skip.  One lane per
// query leaf; a counting pass, an exclusive scan and an emitting pass place
// every leaf's pairs at a fixed offset, so the output is in exactly the order
// of a single-threaded reference run.
#include "nd_internal.hpp"

namespace {

using nd::accept;
using nd::sqdist;

struct Box { float mn[3], mx[3]; int32_t left, skip; };

__device__ __forceinline__ Box load_node(const nd_node *__restrict__ nodes, int32_t k) {
    const float4 *p = reinterpret_cast<const float4 *>(nodes + (k - 1));
    const float4 a = p[0];
    const float4 b = p[1];
    Box n;
    n.mn[0] = a.x; n.mn[1] = a.y; n.mn[2] = a.z;
    n.mx[0] = a.w; n.mx[1] = b.x; n.mx[2] = b.y;
    n.left = __float_as_int(b.z);
    n.skip = __float_as_int(b.w);
    return n;
}

// aabb_overlap_test (BVHTraverse.jl:1094-1098), strict on all three axes.
__device__ __forceinline__ bool overlap(const Box &a, const Box &b) {
    return (a.mn[0] < b.mx[0]) & (a.mx[0] > b.mn[0]) & (a.mn[1] < b.mx[1]) &
           (a.mx[1] > b.mn[1]) & (a.mn[2] < b.mx[2]) & (a.mx[2] > b.mn[2]);
}

template <bool EMIT>
struct Sink {
    int64_t n = 0, limit = 0;
    nd_pair *out = nullptr;
    __device__ __forceinline__ void push(int32_t i, int32_t j, float d) {
        if (EMIT && n < limit) {  // limit = this leaf's count from the counting pass
            nd_pair p; p.i = i; p.j = j; p.d = d;
            out[n] = p;
        }
        ++n;
    }
};

template <bool EMIT>
__global__ void __launch_bounds__(nd::kBlock)
k_leaf_traverse(const nd_node *__restrict__ nodes, const float *__restrict__ xyz,
                const int32_t *__restrict__ idx, int32_t nL, int32_t apl, float r, float r2,
                int32_t pred, int64_t *__restrict__ counts, const int64_t *__restrict__ offsets,
                nd_pair *__restrict__ out, int32_t *__restrict__ status) {
    const int32_t q = 1 + (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
    if (q > nL) return;
    Sink<EMIT> sink;
    if (EMIT) {
        sink.out = out + offsets[q - 1];
        sink.limit = offsets[q] - offsets[q - 1];
    }
    const Box ql = load_node(nodes, q);
    const int64_t qlo = (int64_t)(q - 1) * apl;
    float d;
    for (int32_t a = 0; a < apl - 1; ++a) {
        const float ax = xyz[3 * (qlo + a)], ay = xyz[3 * (qlo + a) + 1], az = xyz[3 * (qlo + a) + 2];
        for (int32_t b = a + 1; b < apl; ++b) {
            const float s = sqdist(ax, ay, az, xyz[3 * (qlo + b)], xyz[3 * (qlo + b) + 1],
                                   xyz[3 * (qlo + b) + 2]);
            if (accept(s, r, r2, pred, d)) sink.push(idx[qlo + a], idx[qlo + b], d);
        }
    }
    int32_t t = ql.skip;
    // A DFS walk visits each node at most once: more than 2nL visits or an
    // index outside [1, 2nL-1] means a malformed tree; stop instead of looping.
    int32_t visits = 0;
    while (t != 0) {
        if (++visits > 2 * nL || t < 1 || t > 2 * nL - 1) {
            atomicOr(status, nd::kStatusTree);
            break;
        }
        const Box tn = load_node(nodes, t);
        if (overlap(ql, tn)) {
            if (tn.left == 0) {
                const int64_t tlo = (int64_t)(t - 1) * apl;
                for (int32_t a = 0; a < apl; ++a) {
                    const float ax = xyz[3 * (qlo + a)], ay = xyz[3 * (qlo + a) + 1],
                                az = xyz[3 * (qlo + a) + 2];
                    for (int32_t b = 0; b < apl; ++b) {
                        const float s = sqdist(ax, ay, az, xyz[3 * (tlo + b)],
                                               xyz[3 * (tlo + b) + 1], xyz[3 * (tlo + b) + 2]);
                        if (accept(s, r, r2, pred, d)) sink.push(idx[qlo + a], idx[tlo + b], d);
                    }
                }
                t = tn.skip;
            } else {
                t = tn.left;
            }
        } else {
            t = tn.skip;
        }
    }
    if (!EMIT) counts[q - 1] = sink.n;
}

// apl == 4 specialisation: the query leaf's atoms stay in registers.
template <bool EMIT>
__global__ void __launch_bounds__(nd::kBlock)
k_leaf_traverse4(const nd_node *__restrict__ nodes, const float *__restrict__ xyz,
                 const int32_t *__restrict__ idx, int32_t nL, float r, float r2, int32_t pred,
                 int64_t *__restrict__ counts, const int64_t *__restrict__ offsets,
                 nd_pair *__restrict__ out, int32_t *__restrict__ status) {
    const int32_t q = 1 + (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
    if (q > nL) return;
    Sink<EMIT> sink;
    if (EMIT) {
        sink.out = out + offsets[q - 1];
        sink.limit = offsets[q] - offsets[q - 1];
    }
    const Box ql = load_node(nodes, q);
    const int64_t qlo = (int64_t)(q - 1) * 4;
    float qx[4], qy[4], qz[4];
    int32_t qi[4];
#pragma unroll
    for (int a = 0; a < 4; ++a) {
        qx[a] = xyz[3 * (qlo + a)];
        qy[a] = xyz[3 * (qlo + a) + 1];
        qz[a] = xyz[3 * (qlo + a) + 2];
        qi[a] = EMIT ? idx[qlo + a] : 0;
    }
    float d;
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int b = a + 1; b < 4; ++b)
            if (accept(sqdist(qx[a], qy[a], qz[a], qx[b], qy[b], qz[b]), r, r2, pred, d))
                sink.push(qi[a], qi[b], d);
    int32_t t = ql.skip;
    // A DFS walk visits each node at most once: more than 2nL visits or an
    // index outside [1, 2nL-1] means a malformed tree; stop instead of looping.
    int32_t visits = 0;
    while (t != 0) {
        if (++visits > 2 * nL || t < 1 || t > 2 * nL - 1) {
            atomicOr(status, nd::kStatusTree);
            break;
        }
        const Box tn = load_node(nodes, t);
        if (overlap(ql, tn)) {
            if (tn.left == 0) {
                const int64_t tlo = (int64_t)(t - 1) * 4;
                float tx[4], ty[4], tz[4];
#pragma unroll
                for (int b = 0; b < 4; ++b) {
                    tx[b] = xyz[3 * (tlo + b)];
                    ty[b] = xyz[3 * (tlo + b) + 1];
                    tz[b] = xyz[3 * (tlo + b) + 2];
                }
#pragma unroll
                for (int a = 0; a < 4; ++a)
#pragma unroll
                    for (int b = 0; b < 4; ++b)
                        if (accept(sqdist(qx[a], qy[a], qz[a], tx[b], ty[b], tz[b]), r, r2, pred, d))
                            sink.push(qi[a], EMIT ? idx[tlo + b] : 0, d);
                t = tn.skip;
            } else {
                t = tn.left;
            }
        } else {
            t = tn.skip;
        }
    }
    if (!EMIT) counts[q - 1] = sink.n;
}

// Exclusive scan of n int64 counts into offsets[0..n]; offsets[n] = total.
__global__ void __launch_bounds__(1024)
k_scan_i64(const int64_t *__restrict__ in, int64_t n, int64_t *__restrict__ out) {
    __shared__ int64_t part[1024];
    const int t = threadIdx.x;
    const int64_t per = (n + 1023) / 1024;
    const int64_t lo = t * per, hi = lo + per < n ? lo + per : n;
    int64_t s = 0;
    for (int64_t k = lo; k < hi; ++k) s += in[k];
    part[t] = s;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        int64_t v = t >= off ? part[t - off] : 0;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    int64_t run = part[t] - s;
    for (int64_t k = lo; k < hi; ++k) {
        int64_t c = in[k];
        out[k] = run;
        run += c;
    }
    if (t == 1023) out[n] = part[1023];
}

// threshold_pairs(unique_pairs(x), r) (AllToAll.jl:65-88): lane i scans j > i.
template <bool EMIT>
__global__ void __launch_bounds__(nd::kBlock)
k_bruteforce(const float *__restrict__ xyz, int32_t n, float r, int64_t *__restrict__ counts,
             const int64_t *__restrict__ offsets, nd_pair *__restrict__ out) {
    const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n - 1) {
        if (!EMIT && i < n) counts[i] = 0;
        return;
    }
    const float ax = xyz[3 * (int64_t)i], ay = xyz[3 * (int64_t)i + 1], az = xyz[3 * (int64_t)i + 2];
    int64_t c = 0;
    nd_pair *o = EMIT ? out + offsets[i] : nullptr;
    for (int32_t j = i + 1; j < n; ++j) {
        const float d = sqrtf(sqdist(ax, ay, az, xyz[3 * (int64_t)j], xyz[3 * (int64_t)j + 1],
                                     xyz[3 * (int64_t)j + 2]));
        if (d < r) {
            if (EMIT) { nd_pair p; p.i = i; p.j = j; p.d = d; o[c] = p; }
            ++c;
        }
    }
    if (!EMIT) counts[i] = c;
}

}  // namespace

namespace nd {

int64_t scan_counts(nd_ctx *ctx, int64_t n) {
    // counts in ctx->counts (n), offsets to ctx->offsets (n+1); returns total.
    k_scan_i64<<<1, 1024, 0, ctx->stream>>>(ctx->counts.as<int64_t>(), n, ctx->offsets.as<int64_t>());
    ND_HIP(hipGetLastError());
    int64_t total = 0;
    ND_HIP(hipMemcpyAsync(&total, ctx->offsets.as<int64_t>() + n, sizeof(int64_t),
                          hipMemcpyDeviceToHost, ctx->stream));
    ND_HIP(hipStreamSynchronize(ctx->stream));
    return total;
}

int64_t leaf_traverse(nd_ctx *ctx, const nd_node *nodes, const float *xyz_sorted,
                      const int32_t *idx_sorted, int32_t n, int32_t apl, float r, int32_t pred,
                      nd_pair *out, int64_t cap) {
    const int32_t nL = n / apl;
    const float r2 = r * r;  // (neighbor_distance)^2 in Float32
    ctx->counts.reserve(sizeof(int64_t) * (size_t)(nL + 1));
    ctx->offsets.reserve(sizeof(int64_t) * (size_t)(nL + 1));
    int64_t *cnt = ctx->counts.as<int64_t>();
    const int64_t *off = ctx->offsets.as<int64_t>();
    int32_t *st = ctx->status.as<int32_t>();
    if (apl == 4)
        k_leaf_traverse4<false><<<grid_for(nL), kBlock, 0, ctx->stream>>>(
            nodes, xyz_sorted, idx_sorted, nL, r, r2, pred, cnt, off, nullptr, st);
    else
        k_leaf_traverse<false><<<grid_for(nL), kBlock, 0, ctx->stream>>>(
            nodes, xyz_sorted, idx_sorted, nL, apl, r, r2, pred, cnt, off, nullptr, st);
    ND_HIP(hipGetLastError());
    const int64_t total = scan_counts(ctx, nL);
    if (nd::read_status(ctx) & nd::kStatusTree)
        throw nd::Error(ND_ERR_STATE, "leafneighbor_traverse: malformed tree (rope cycle or bad index)");
    if (total > cap || out == nullptr) return total;
    if (apl == 4)
        k_leaf_traverse4<true><<<grid_for(nL), kBlock, 0, ctx->stream>>>(
            nodes, xyz_sorted, idx_sorted, nL, r, r2, pred, cnt, off, out, st);
    else
        k_leaf_traverse<true><<<grid_for(nL), kBlock, 0, ctx->stream>>>(
            nodes, xyz_sorted, idx_sorted, nL, apl, r, r2, pred, cnt, off, out, st);
    ND_HIP(hipGetLastError());
    ND_HIP(hipStreamSynchronize(ctx->stream));
    return total;
}

}  // namespace nd

namespace {

void check_spec(const nd_bvh_spec *spec, nd_bvh_spec &chk) {
    ND_REQUIRE(spec, "null SpheresBVHSpecs");
    if (nd_bvh_spec_init(&chk, spec->neighbor_distance, spec->atom_count, spec->atomsperleaf) !=
        ND_OK)
        throw nd::Error(ND_ERR_ARG,
                        "SpheresBVHSpecs: atomsperleaf must divide atom_count and give >= 2 leaves");
}

void check_pred(int32_t pred) {
    ND_REQUIRE(pred == ND_PRED_SCALAR || pred == ND_PRED_SIMD, "unknown pair predicate");
}

}  // namespace

extern "C" {

int nd_leafneighbor_traverse(nd_ctx *ctx, const nd_node *d_nodes, const float *d_xyz_sorted,
                             const int32_t *d_index_sorted, const nd_bvh_spec *spec,
                             int32_t predicate, nd_pair *d_out, int64_t cap, int64_t *count) {
    ND_API_BEGIN
    ND_REQUIRE(ctx && d_nodes && d_xyz_sorted && d_index_sorted && count,
               "nd_leafneighbor_traverse: null argument");
    nd_bvh_spec chk;
    check_spec(spec, chk);
    check_pred(predicate);
    ND_HIP(hipSetDevice(ctx->device));
    const int64_t total = nd::leaf_traverse(ctx, d_nodes, d_xyz_sorted, d_index_sorted,
                                            chk.atom_count, chk.atomsperleaf,
                                            chk.neighbor_distance, predicate, d_out, cap);
    *count = total;
    if (total > cap) throw nd::Error(ND_ERR_CAPACITY, "pair buffer too small; *count = required");
    ND_API_END(ctx)
}

int nd_leafbuild_traverse_bvh(nd_ctx *ctx, const float *d_xyz, const nd_bvh_spec *spec,
                              int32_t predicate, int32_t morton_mode, nd_pair *d_out,
                              int64_t cap, int64_t *count) {
    ND_API_BEGIN
    ND_REQUIRE(ctx && d_xyz && count, "nd_leafbuild_traverse_bvh: null argument");
    nd_bvh_spec chk;
    check_spec(spec, chk);
    check_pred(predicate);
    ND_REQUIRE(morton_mode == ND_MORTON_MASKED || morton_mode == ND_MORTON_INTERLEAVE,
               "unknown key mode");
    ND_HIP(hipSetDevice(ctx->device));
    const int32_t n = chk.atom_count, apl = chk.atomsperleaf, nL = chk.leaves_count;
    ctx->has_tree = false;
    ctx->codes.reserve(sizeof(int32_t) * (size_t)n);
    ctx->codes_sorted.reserve(sizeof(int32_t) * (size_t)n);
    ctx->perm.reserve(sizeof(int32_t) * (size_t)n);
    ctx->xyz_sorted.reserve(sizeof(float) * 3 * (size_t)n);
    ctx->nodes.reserve(sizeof(nd_node) * (size_t)(2 * nL - 1));
    ctx->store.reserve(sizeof(int32_t) * nd::lbvh_scratch_ints(nL));
    // mortoncodes! -> codes (and perm = 1:N)
    nd::launch_morton(ctx, d_xyz, n, morton_mode, nullptr, nullptr, ctx->codes_sorted.as<int32_t>(),
                      ctx->perm.as<int32_t>(), ctx->status.as<int32_t>());
    if (nd::read_status(ctx) & nd::kStatusDomain)
        throw nd::Error(ND_ERR_DOMAIN,
                        "mortoncodes!: coordinate outside the Int32 quantisation domain "
                        "(reference throws InexactError)");
    // sortperm + permute!
    nd::radix_sort_pairs_auto(ctx, ctx->codes_sorted.as<int32_t>(), ctx->perm.as<int32_t>(), n);
    nd::launch_gather_xyz(ctx->stream, d_xyz, ctx->perm.as<int32_t>(), n, ctx->xyz_sorted.as<float>());
    // leafcluster_primitives + bounding_volume_hierarchy!
    nd::launch_lbvh(ctx->stream, ctx->xyz_sorted.as<float>(), 3, ctx->codes_sorted.as<int32_t>(), nL,
                    apl, chk.neighbor_distance, 1, ctx->nodes.as<nd_node>(), ctx->store.as<int32_t>(),
                    ctx->status.as<int32_t>());
    ctx->has_tree = true;
    ctx->tree_n = n;
    ctx->tree_apl = apl;
    const int64_t total = nd::leaf_traverse(ctx, ctx->nodes.as<nd_node>(), ctx->xyz_sorted.as<float>(),
                                            ctx->perm.as<int32_t>(), n, apl, chk.neighbor_distance,
                                            predicate, d_out, cap);
    *count = total;
    if (total > cap) throw nd::Error(ND_ERR_CAPACITY, "pair buffer too small; *count = required");
    ND_API_END(ctx)
}

int nd_leafbuild_traverse_bvh_host(nd_ctx *ctx, const float *h_xyz, const nd_bvh_spec *spec,
                                   int32_t predicate, int32_t morton_mode, nd_pair *h_out,
                                   int64_t cap, int64_t *count) {
    ND_API_BEGIN
    ND_REQUIRE(ctx && h_xyz && spec && count, "nd_leafbuild_traverse_bvh_host: null argument");
    nd_bvh_spec chk;
    check_spec(spec, chk);
    ND_HIP(hipSetDevice(ctx->device));
    const size_t n = (size_t)chk.atom_count;
    ctx->xyz_in.reserve(sizeof(float) * 3 * n);
    ND_HIP(hipMemcpyAsync(ctx->xyz_in.p, h_xyz, sizeof(float) * 3 * n, hipMemcpyHostToDevice,
                          ctx->stream));
    if (cap > 0) ctx->pairs_out.reserve(sizeof(nd_pair) * (size_t)cap);
    int rc = nd_leafbuild_traverse_bvh(ctx, ctx->xyz_in.as<float>(), spec, predicate, morton_mode,
                                       cap > 0 ? ctx->pairs_out.as<nd_pair>() : nullptr, cap, count);
    if (rc != ND_OK) return rc;
    if (*count > 0 && h_out)
        ND_HIP(hipMemcpyAsync(h_out, ctx->pairs_out.p, sizeof(nd_pair) * (size_t)*count,
                              hipMemcpyDeviceToHost, ctx->stream));
    ND_HIP(hipStreamSynchronize(ctx->stream));
    ND_API_END(ctx)
}

int nd_tree_export(nd_ctx *ctx, nd_node *h_nodes, int32_t *h_perm, int32_t *h_codes_sorted) {
    ND_API_BEGIN
    ND_REQUIRE(ctx, "nd_tree_export: null context");
    if (!ctx->has_tree) throw nd::Error(ND_ERR_STATE, "nd_tree_export: no tree has been built");
    const size_t n = (size_t)ctx->tree_n, nL = n / (size_t)ctx->tree_apl;
    if (h_nodes)
        ND_HIP(hipMemcpyAsync(h_nodes, ctx->nodes.p, sizeof(nd_node) * (2 * nL - 1),
                              hipMemcpyDeviceToHost, ctx->stream));
    if (h_perm)
        ND_HIP(hipMemcpyAsync(h_perm, ctx->perm.p, sizeof(int32_t) * n, hipMemcpyDeviceToHost,
                              ctx->stream));
    if (h_codes_sorted)
        ND_HIP(hipMemcpyAsync(h_codes_sorted, ctx->codes_sorted.p, sizeof(int32_t) * n,
                              hipMemcpyDeviceToHost, ctx->stream));
    ND_HIP(hipStreamSynchronize(ctx->stream));
    ND_API_END(ctx)
}

int nd_unique_threshold_pairs(nd_ctx *ctx, const float *d_xyz, int32_t n, float threshold,
                              nd_pair *d_out, int64_t cap, int64_t *count) {
    ND_API_BEGIN
    ND_REQUIRE(ctx && d_xyz && count && n >= 0, "nd_unique_threshold_pairs: bad argument");
    ND_HIP(hipSetDevice(ctx->device));
    *count = 0;
    if (n < 2) return ND_OK;
    ctx->counts.reserve(sizeof(int64_t) * (size_t)(n + 1));
    ctx->offsets.reserve(sizeof(int64_t) * (size_t)(n + 1));
    k_bruteforce<false><<<nd::grid_for(n), nd::kBlock, 0, ctx->stream>>>(
        d_xyz, n, threshold, ctx->counts.as<int64_t>(), nullptr, nullptr);
    ND_HIP(hipGetLastError());
    const int64_t total = nd::scan_counts(ctx, n);
    *count = total;
    if (total > cap) throw nd::Error(ND_ERR_CAPACITY, "pair buffer too small; *count = required");
    k_bruteforce<true><<<nd::grid_for(n), nd::kBlock, 0, ctx->stream>>>(
        d_xyz, n, threshold, nullptr, ctx->offsets.as<int64_t>(), d_out);
    ND_HIP(hipGetLastError());
    ND_HIP(hipStreamSynchronize(ctx->stream));
    ND_API_END(ctx)
}

}  // extern "C"
