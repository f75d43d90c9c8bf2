// nd_lbvh.hip -- K3: leaf boxes fused with the Apetrei bottom-up LBVH build
// (leafcluster_primitives, BVHTraverse.jl:376-410; exptdelta :676-696;
// bvh_interior! :704-846; bounding_volume_hierarchy! :849-860).
//
// One lane per leaf.  A lane writes its leaf node, then climbs: at every
// internal node the first lane to arrive parks its range end in store[] with
// an agent-scope compare-and-swap and stops; the second lane (which sees the
// first lane's node through the CAS's acquire half) merges the sibling box,
// writes the parent node and continues.  The tree topology depends only on
// the sorted codes, so it is identical to a serial build.
#include "nd_internal.hpp"

namespace {

__device__ __forceinline__ int32_t delta(int32_t i, const int32_t *__restrict__ codes, int32_t nL,
                                         int32_t apl) {
    if (i >= nL || i < 1) return INT32_MAX;
    const int32_t x = codes[(int64_t)(i - 1) * apl] ^ codes[(int64_t)i * apl];
    uint32_t t = (uint32_t)x;
    if (x == 0) t += 0x80000000u + (uint32_t)(i ^ (i + 1));  // typemin + (i xor i+1)
    t -= 1u;
    return (int32_t)t;
}

__device__ __forceinline__ int32_t rope(int32_t r, int32_t delr, const int32_t *__restrict__ codes,
                                        int32_t nL, int32_t apl) {
    // Next node after a subtree whose range ends at r (BVHTraverse.jl:717-741, 815-832).
    if (r == nL) return 0;
    const int32_t n1 = r + 1;
    return (delr < delta(n1, codes, nL, apl)) ? n1 : n1 + nL;
}

__device__ __forceinline__ void store_node(nd_node *nodes, int32_t k, const float mn[3],
                                           const float mx[3], int32_t left, int32_t skip) {
    float4 *p = reinterpret_cast<float4 *>(nodes + (k - 1));
    p[0] = make_float4(mn[0], mn[1], mn[2], mx[0]);
    p[1] = make_float4(mx[1], mx[2], __int_as_float(left), __int_as_float(skip));
}

__device__ __forceinline__ void merge_node(const nd_node *nodes, int32_t k, float mn[3],
                                           float mx[3]) {
    const float4 *p = reinterpret_cast<const float4 *>(nodes + (k - 1));
    const float4 a = p[0];
    const float4 b = p[1];
    // max.(keys[child].max, bounding) / min.(keys[child].min, bounding)
    mx[0] = jl_max(a.w, mx[0]);
    mx[1] = jl_max(b.x, mx[1]);
    mx[2] = jl_max(b.y, mx[2]);
    mn[0] = jl_min(a.x, mn[0]);
    mn[1] = jl_min(a.y, mn[1]);
    mn[2] = jl_min(a.z, mn[2]);
}

__device__ __forceinline__ int32_t cas_acq_rel(int32_t *p, int32_t desired) {
    int32_t expected = 0;
    __hip_atomic_compare_exchange_strong(p, &expected, desired, __ATOMIC_ACQ_REL,
                                         __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
    return expected;  // old value: 0 if this lane arrived first
}

__global__ void __launch_bounds__(nd::kBlock)
k_lbvh(const float *__restrict__ xyz, int32_t stride, const int32_t *__restrict__ codes,
       int32_t nL, int32_t apl, float h, int32_t override_root, nd_node *nodes,
       int32_t *store, int32_t *status) {
    int32_t i = 1 + (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i > nL) return;

    // ---- leaf box: min(x - r/2), max(x + r/2) over the leaf's atoms ----
    float mn[3], mx[3];
    const float *p0 = xyz + (int64_t)stride * (int64_t)(i - 1) * apl;
#pragma unroll
    for (int d = 0; d < 3; ++d) { mn[d] = p0[d] - h; mx[d] = p0[d] + h; }
    for (int32_t a = 0; a < apl; ++a) {
        const float *p = p0 + (int64_t)stride * a;
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            mn[d] = jl_min(mn[d], p[d] - h);
            mx[d] = jl_max(mx[d], p[d] + h);
        }
    }

    int32_t rangel = i, ranger = i;
    int32_t dell = delta(rangel - 1, codes, nL, apl);
    int32_t delr = delta(ranger, codes, nL, apl);
    store_node(nodes, i, mn, mx, 0, rope(i, delr, codes, nL, apl));

    // Every climb step strictly widens [rangel, ranger], so a walker takes at
    // most nL steps; the bound and the index checks only matter for codes
    // that violate delta's range (bit 31 set), where the reference is UB.
    for (int32_t steps = 0; steps <= nL; ++steps) {
        int32_t leftChild;
        if (delr < dell) {  // this subtree is a left child; sibling to the right
            leftChild = i;
            const int32_t split = ranger;
            if (split < 1 || split > nL - 1) break;
            const int32_t other = cas_acq_rel(&store[split - 1], rangel);
            if (other == 0) return;
            if (other < 1 || other > nL) break;
            ranger = other;
            delr = delta(ranger, codes, nL, apl);
            int32_t rightChild = split + 1;
            if (rightChild != ranger) rightChild += nL;
            merge_node(nodes, rightChild, mn, mx);
        } else {  // right child; sibling to the left
            const int32_t split = rangel - 1;
            if (split < 1 || split > nL - 1) break;
            const int32_t other = cas_acq_rel(&store[split - 1], ranger);
            if (other == 0) return;
            if (other < 1 || other > nL) break;
            rangel = other;
            dell = delta(rangel - 1, codes, nL, apl);
            leftChild = split;
            if (leftChild != rangel) leftChild += nL;
            merge_node(nodes, leftChild, mn, mx);
        }
        const int32_t q = (delr < dell) ? ranger : rangel;
        if (q < 1 || q > nL - 1) break;
        const int32_t parent = q + nL;
        const int32_t skip = rope(ranger, delr, codes, nL, apl);
        if (parent == nL + 1 && override_root) {
            const float z[3] = {0.0f, 0.0f, 0.0f}, o[3] = {1.0f, 1.0f, 1.0f};
            store_node(nodes, parent, z, o, leftChild, skip);
        } else {
            store_node(nodes, parent, mn, mx, leftChild, skip);
        }
        i = parent;
        if (i == nL + 1) return;
    }
    if (status) atomicOr(status, nd::kStatusTree);
}

}  // namespace

namespace nd {

void launch_lbvh(hipStream_t s, const float *xyz_sorted, int32_t xyz_stride,
                 const int32_t *codes_sorted, int32_t nL, int32_t apl, float r, int32_t override_root,
                 nd_node *nodes, int32_t *store, int32_t *status) {
    launch_fill_i32(s, store, nL - 1, 0);
    const float h = r / 2.0f;  // neighbor_distance/2 in Float32
    k_lbvh<<<grid_for(nL), kBlock, 0, s>>>(xyz_sorted, xyz_stride, codes_sorted, nL, apl, h,
                                          override_root, nodes, store, status);
    ND_HIP(hipGetLastError());
}

}  // namespace nd

extern "C" int nd_bounding_volume_hierarchy(nd_ctx *ctx, const float *d_xyz_sorted,
                                            const int32_t *d_codes_sorted,
                                            const nd_bvh_spec *spec, int32_t override_root,
                                            nd_node *d_nodes) {
    ND_API_BEGIN
    ND_REQUIRE(ctx && d_xyz_sorted && d_codes_sorted && spec && d_nodes,
               "nd_bounding_volume_hierarchy: null argument");
    nd_bvh_spec chk;
    ND_REQUIRE(nd_bvh_spec_init(&chk, spec->neighbor_distance, spec->atom_count,
                                spec->atomsperleaf) == ND_OK,
               "nd_bounding_volume_hierarchy: invalid SpheresBVHSpecs");
    ND_HIP(hipSetDevice(ctx->device));
    const int32_t nL = chk.leaves_count;
    ctx->store.reserve(sizeof(int32_t) * (size_t)nL);
    nd::launch_lbvh(ctx->stream, d_xyz_sorted, 3, d_codes_sorted, nL, chk.atomsperleaf,
                    chk.neighbor_distance, override_root, d_nodes, ctx->store.as<int32_t>(),
                    ctx->status.as<int32_t>());
    if (nd::read_status(ctx) & nd::kStatusTree)
        throw nd::Error(ND_ERR_DOMAIN, "bounding_volume_hierarchy!: codes outside the range the "
                                       "build supports (bit 31 set)");
    ND_API_END(ctx)
}
