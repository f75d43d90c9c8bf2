// nd_lbvh.hip -- K3: leaf boxes + Apetrei bottom-up LBVH build + box refit
// (leafcluster_primitives, BVHTraverse.jl:376-410; exptdelta :676-696;
// bvh_interior! :704-846; bounding_volume_hierarchy! :849-860).
//
// Topology pass, one lane per leaf: the lane writes its leaf node (box of
// its atoms +- r/2, rope), then climbs.  At every internal node the first
// lane to arrive parks its range end in store[] with a compare-and-swap and
// stops; the second lane receives the sibling's range end from the CAS,
// writes the parent's left child and rope, records the parent's leaf range
// and continues.  The range end carried by the CAS is the only datum passed
// between lanes, so a relaxed agent-scope CAS suffices: no lane reads memory
// another lane wrote in this kernel, and no cache writeback/invalidate is
// needed per level.  The topology depends only on the sorted codes, so it is
// identical to the reference's serial or threaded build.
//
// Refit pass, one wave per internal node: the node's box is the min/max of
// the leaf boxes over its contiguous leaf range.  That equals the reference's
// pairwise bottom-up merge bit for bit (Julia's min/max on non-NaN floats,
// -0 ordered below +0, is associative and commutative), and needs no
// inter-node ordering.
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

__device__ __forceinline__ int32_t cas_relaxed(int32_t *p, int32_t desired) {
    int32_t expected = 0;
    __hip_atomic_compare_exchange_strong(p, &expected, desired, __ATOMIC_RELAXED,
                                         __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return expected;  // old value: 0 if this lane arrived first
}

__global__ void __launch_bounds__(nd::kBlock)
k_lbvh_topology(const int32_t *__restrict__ codes, int32_t nL, int32_t apl, nd_node *__restrict__ nodes,
                int32_t *__restrict__ store, int2 *__restrict__ ranges,
                int32_t *__restrict__ status) {
    int32_t i = 1 + (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i > nL) return;

    // the leaf's box is written by launch_leaf_boxes; here its left (0) and rope
    int32_t rangel = i, ranger = i;
    int32_t dell = delta(rangel - 1, codes, nL, apl);
    int32_t delr = delta(ranger, codes, nL, apl);
    *reinterpret_cast<int2 *>(&nodes[i - 1].left) = make_int2(0, rope(i, delr, codes, nL, apl));

    // Every climb step strictly widens [rangel, ranger], so a walker takes at
    // most nL steps; the bound and the index checks only matter for codes
    // that violate delta's range (bit 31 set), where the reference is UB.
    for (int32_t steps = 0; steps <= nL; ++steps) {
        int32_t leftChild;
        if (delr < dell) {  // this subtree is a left child; sibling to the right
            leftChild = i;
            const int32_t split = ranger;
            if (split < 1 || split > nL - 1) break;
            const int32_t other = cas_relaxed(&store[split - 1], rangel);
            if (other == 0) return;
            if (other < 1 || other > nL) break;
            ranger = other;
            delr = delta(ranger, codes, nL, apl);
        } else {  // right child; sibling to the left
            const int32_t split = rangel - 1;
            if (split < 1 || split > nL - 1) break;
            const int32_t other = cas_relaxed(&store[split - 1], ranger);
            if (other == 0) return;
            if (other < 1 || other > nL) break;
            rangel = other;
            dell = delta(rangel - 1, codes, nL, apl);
            leftChild = split;
            if (leftChild != rangel) leftChild += nL;
        }
        const int32_t q = (delr < dell) ? ranger : rangel;
        if (q < 1 || q > nL - 1) break;
        const int32_t parent = q + nL;
        *reinterpret_cast<int2 *>(&nodes[parent - 1].left) =
            make_int2(leftChild, rope(ranger, delr, codes, nL, apl));
        ranges[q - 1] = make_int2(rangel, ranger);
        i = parent;
        if (i == nL + 1) return;
    }
    if (status) atomicOr(status, nd::kStatusTree);
}

__device__ __forceinline__ float wave_min(float v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = jl_min(v, __shfl_xor(v, off, 64));
    return v;
}
__device__ __forceinline__ float wave_max(float v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = jl_max(v, __shfl_xor(v, off, 64));
    return v;
}

struct Box6 {
    float mn0, mn1, mn2, mx0, mx1, mx2;
    __device__ void init() {
        const float inf = __int_as_float(0x7F800000);
        mn0 = mn1 = mn2 = inf;
        mx0 = mx1 = mx2 = -inf;
    }
    // node/chunk box stored as (min xyz, max x) + (max y, max z, ...)
    __device__ void merge(const float4 a, const float4 b) {
        mn0 = jl_min(mn0, a.x); mn1 = jl_min(mn1, a.y); mn2 = jl_min(mn2, a.z);
        mx0 = jl_max(mx0, a.w); mx1 = jl_max(mx1, b.x); mx2 = jl_max(mx2, b.y);
    }
    __device__ void wave_reduce() {
        mn0 = wave_min(mn0); mn1 = wave_min(mn1); mn2 = wave_min(mn2);
        mx0 = wave_max(mx0); mx1 = wave_max(mx1); mx2 = wave_max(mx2);
    }
};

constexpr int kChunk = 64;  // leaves per refit chunk

// Box of every aligned chunk of 64 leaves (one wave per chunk), stored in the
// same 32-byte layout as a node.
__global__ void __launch_bounds__(nd::kBlock)
k_lbvh_chunks(const nd_node *__restrict__ nodes, int32_t nL, float4 *__restrict__ chunks) {
    const int32_t c = (int32_t)((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
    const int lane = threadIdx.x & 63;
    if ((int64_t)c * kChunk >= nL) return;
    Box6 b;
    b.init();
    const int32_t l = c * kChunk + lane;  // 0-based leaf
    if (l < nL) {
        const float4 *p = reinterpret_cast<const float4 *>(nodes + l);
        b.merge(p[0], p[1]);
    }
    b.wave_reduce();
    if (lane == 0) {
        chunks[2 * c] = make_float4(b.mn0, b.mn1, b.mn2, b.mx0);
        chunks[2 * c + 1] = make_float4(b.mx1, b.mx2, 0.0f, 0.0f);
    }
}

constexpr int kGroup = 8;  // lanes per internal node in the refit

__device__ __forceinline__ float group_min(float v) {
#pragma unroll
    for (int off = kGroup / 2; off > 0; off >>= 1) v = jl_min(v, __shfl_xor(v, off, 64));
    return v;
}
__device__ __forceinline__ float group_max(float v) {
#pragma unroll
    for (int off = kGroup / 2; off > 0; off >>= 1) v = jl_max(v, __shfl_xor(v, off, 64));
    return v;
}

// Box of the leaves of rg (1-based, inclusive) over the kGroup lanes of an
// aligned group (all of them call it): leaves before the first whole chunk,
// whole chunks, leaves after the last whole chunk (at most 63 + nL/64 + 63
// reads); every lane of the group returns the box.
__device__ __forceinline__ Box6 range_box(const nd_node *__restrict__ leaves, const float4 *__restrict__ chunks,
                                          int2 rg, int lane) {
    const int32_t l0 = rg.x - 1, r0 = rg.y - 1;           // 0-based inclusive leaf range
    const int32_t c0 = (l0 + kChunk - 1) / kChunk;         // first whole chunk
    const int32_t c1 = (r0 + 1) / kChunk;                  // one past the last whole chunk
    Box6 b;
    b.init();
    if (c0 < c1) {
        for (int32_t l = l0 + lane; l < c0 * kChunk; l += kGroup) {
            const float4 *p = reinterpret_cast<const float4 *>(leaves + l);
            b.merge(p[0], p[1]);
        }
        for (int32_t c = c0 + lane; c < c1; c += kGroup) b.merge(chunks[2 * c], chunks[2 * c + 1]);
        for (int32_t l = c1 * kChunk + lane; l <= r0; l += kGroup) {
            const float4 *p = reinterpret_cast<const float4 *>(leaves + l);
            b.merge(p[0], p[1]);
        }
    } else {
        for (int32_t l = l0 + lane; l <= r0; l += kGroup) {
            const float4 *p = reinterpret_cast<const float4 *>(leaves + l);
            b.merge(p[0], p[1]);
        }
    }
    b.mn0 = group_min(b.mn0); b.mn1 = group_min(b.mn1); b.mn2 = group_min(b.mn2);
    b.mx0 = group_max(b.mx0); b.mx1 = group_max(b.mx1); b.mx2 = group_max(b.mx2);
    return b;
}

// Eight lanes per internal node.
__global__ void __launch_bounds__(nd::kBlock)
k_lbvh_refit(nd_node *__restrict__ nodes, const int2 *__restrict__ ranges,
             const float4 *__restrict__ chunks, int32_t nL, int32_t override_root) {
    const int32_t q = 1 + (int32_t)((blockIdx.x * blockDim.x + threadIdx.x) / kGroup);
    const int lane = threadIdx.x & (kGroup - 1);
    if (q > nL - 1) return;  // the whole 8-lane group exits together
    const int2 rg = ranges[q - 1];
    // only reachable with an incomplete topology (status already flagged)
    if (rg.x < 1 || rg.y > nL || rg.x > rg.y) return;
    const Box6 b = range_box(nodes, chunks, rg, lane);
    float mn0 = b.mn0, mn1 = b.mn1, mn2 = b.mn2, mx0 = b.mx0, mx1 = b.mx1, mx2 = b.mx2;
    if (lane == 0) {
        if (q == 1 && override_root) {  // BVHTraverse.jl:856-858
            mn0 = mn1 = mn2 = 0.0f;
            mx0 = mx1 = mx2 = 1.0f;
        }
        float *p = reinterpret_cast<float *>(nodes + (q + nL - 1));
        *reinterpret_cast<float4 *>(p) = make_float4(mn0, mn1, mn2, mx0);
        *reinterpret_cast<float2 *>(p + 4) = make_float2(mx1, mx2);
    }
}

// Leaf boxes: box of each leaf's atoms +- r/2 (leafcluster_primitives,
// BVHTraverse.jl:376-410), written into the leaf node without touching
// left/skip.  One lane per atom and a butterfly over the leaf's SEG lanes
// (SEG = atoms per leaf, a power of two up to 64: each leaf's lanes are an
// aligned segment of one wave); Julia's min/max are associative and
// commutative on non-NaN floats, so the box is the serial loop's bit for bit.
template <int SEG>
__global__ void __launch_bounds__(nd::kBlock)
k_leaf_boxes_seg(const float *__restrict__ xyz, int32_t stride, int32_t natoms, float h,
                 nd_node *__restrict__ nodes) {
    const int32_t a = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
    const bool act = a < natoms;
    const float inf = __int_as_float(0x7F800000);
    float mn0 = inf, mn1 = inf, mn2 = inf, mx0 = -inf, mx1 = -inf, mx2 = -inf;
    if (act) {
        const float *p = xyz + (int64_t)stride * a;
        mn0 = p[0] - h; mn1 = p[1] - h; mn2 = p[2] - h;
        mx0 = p[0] + h; mx1 = p[1] + h; mx2 = p[2] + h;
    }
#pragma unroll
    for (int off = 1; off < SEG; off <<= 1) {
        mn0 = jl_min(mn0, __shfl_xor(mn0, off, 64));
        mn1 = jl_min(mn1, __shfl_xor(mn1, off, 64));
        mn2 = jl_min(mn2, __shfl_xor(mn2, off, 64));
        mx0 = jl_max(mx0, __shfl_xor(mx0, off, 64));
        mx1 = jl_max(mx1, __shfl_xor(mx1, off, 64));
        mx2 = jl_max(mx2, __shfl_xor(mx2, off, 64));
    }
    if (act && (a & (SEG - 1)) == 0) {
        float *q = reinterpret_cast<float *>(nodes + a / SEG);
        *reinterpret_cast<float4 *>(q) = make_float4(mn0, mn1, mn2, mx0);
        *reinterpret_cast<float2 *>(q + 4) = make_float2(mx1, mx2);
    }
}

// Leaf boxes for any leaf size: one lane per leaf, serial over its atoms.
__global__ void __launch_bounds__(nd::kBlock)
k_leaf_boxes(const float *__restrict__ xyz, int32_t stride, int32_t nL, int32_t apl, float h,
             nd_node *__restrict__ nodes) {
    const int32_t l = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
    if (l >= nL) return;
    const float *p0 = xyz + (int64_t)stride * (int64_t)l * apl;
    float mn[3], mx[3];
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
    float *q = reinterpret_cast<float *>(nodes + l);
    *reinterpret_cast<float4 *>(q) = make_float4(mn[0], mn[1], mn[2], mx[0]);
    *reinterpret_cast<float2 *>(q + 4) = make_float2(mx[1], mx[2]);
}

// Child-box form of the internal nodes for stack-based traversal: wide[k] for
// internal node nL+1+k holds both children's boxes and numbers
// (float4 x3: min_L xyz max_L x | max_L yz min_R xy | min_R z max_R xyz;
// int4: left, right, 0, 0).  The right child is the left child's rope.
__global__ void __launch_bounds__(nd::kBlock)
k_lbvh_wide(const nd_node *__restrict__ nodes, int32_t nL, float4 *__restrict__ wide) {
    const int32_t k = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
    if (k >= nL - 1) return;
    const int32_t L = nodes[nL + k].left;
    if (L < 1 || L > 2 * nL - 1) return;
    const float4 *pl = reinterpret_cast<const float4 *>(nodes + (L - 1));
    const float4 la = pl[0], lb = pl[1];
    const int32_t R = __float_as_int(lb.w);  // rope of the left child = right sibling
    if (R < 1 || R > 2 * nL - 1) return;
    const float4 *pr = reinterpret_cast<const float4 *>(nodes + (R - 1));
    const float4 ra = pr[0], rb = pr[1];
    float4 *w = wide + 4 * (int64_t)k;
    w[0] = make_float4(la.x, la.y, la.z, la.w);
    w[1] = make_float4(lb.x, lb.y, ra.x, ra.y);
    w[2] = make_float4(ra.z, ra.w, rb.x, rb.y);
    w[3] = make_float4(__int_as_float(L), __int_as_float(R), 0.0f, 0.0f);
}

// Four-wide records: rec[8k .. 8k+7] for internal node nL+1+k hold the
// node's grandchildren (a child that is a leaf stands for itself), up to 4
// entries in the order left child's children, right child's children, as
// structure-of-arrays float4s (min x, min y, min z, max x, max y, max z; one
// lane per entry), then their node numbers (int4; 0 = empty slot, whose box
// is empty so it overlaps nothing), then padding.
__global__ void __launch_bounds__(nd::kBlock)
k_lbvh_wide4(const nd_node *__restrict__ nodes, int32_t nL, float4 *__restrict__ rec) {
    const int32_t k = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
    if (k >= nL - 1) return;
    int32_t id[4] = {0, 0, 0, 0};
    int32_t m = 0;
    const int32_t nn = 2 * nL - 1;
    const int32_t L = nodes[nL + k].left;
    const int32_t R = (L >= 1 && L <= nn) ? nodes[L - 1].skip : 0;
    const int32_t ch[2] = {L, R};
    for (int c = 0; c < 2; ++c) {
        const int32_t x = ch[c];
        if (x < 1 || x > nn) continue;  // malformed tree: the walk flags the empty record
        if (x <= nL) {
            id[m++] = x;
        } else {
            const int32_t g0 = nodes[x - 1].left;
            const int32_t g1 = (g0 >= 1 && g0 <= nn) ? nodes[g0 - 1].skip : 0;
            if (g0 >= 1 && g0 <= nn) id[m++] = g0;
            if (g1 >= 1 && g1 <= nn) id[m++] = g1;
        }
    }
    float b[6][4];
    for (int e = 0; e < 4; ++e) {
        if (id[e] != 0) {
            const nd_node &q = nodes[id[e] - 1];
            b[0][e] = q.min[0]; b[1][e] = q.min[1]; b[2][e] = q.min[2];
            b[3][e] = q.max[0]; b[4][e] = q.max[1]; b[5][e] = q.max[2];
        } else {
            b[0][e] = b[1][e] = b[2][e] = INFINITY;
            b[3][e] = b[4][e] = b[5][e] = -INFINITY;
        }
    }
    float4 *r = rec + 8 * (int64_t)k;
    for (int d = 0; d < 6; ++d) r[d] = make_float4(b[d][0], b[d][1], b[d][2], b[d][3]);
    r[6] = make_float4(__int_as_float(id[0]), __int_as_float(id[1]), __int_as_float(id[2]),
                       __int_as_float(id[3]));
    r[7] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
}

// Leaf boxes and the box of every aligned chunk of 64 leaves in one pass,
// for a tree whose topology is kept: one workgroup per chunk (64 SEG atoms,
// SEG = atoms per leaf, a power of two from 4 to 64), each lane reducing
// SEG / 4 consecutive atoms, four lanes a leaf, the workgroup the chunk
// (shuffles within the wave, then LDS).  Leaf and chunk boxes equal those of
// k_leaf_boxes_seg and k_lbvh_chunks bit for bit (Julia's min/max are
// associative and commutative).
template <int SEG>
__global__ void __launch_bounds__(nd::kBlock)
k_leaf_chunk_boxes(const float *__restrict__ xyz, int32_t stride, int32_t natoms, float h,
                   nd_node *__restrict__ nodes, float4 *__restrict__ chunks, float4 *__restrict__ tq,
                   float4 *__restrict__ lc) {
    static_assert(SEG >= 4 && SEG <= 64 && (SEG & (SEG - 1)) == 0, "leaf size");
    static_assert(kChunk * 4 == nd::kBlock, "four lanes per leaf, one workgroup per chunk");
    constexpr int kA = SEG / 4;  // atoms per lane
    __shared__ float s_box[nd::kBlock / 64][6];
    const int32_t c = blockIdx.x;
    const int tid = threadIdx.x;
    const float inf = __int_as_float(0x7F800000);
    const int32_t a0 = c * (kChunk * SEG) + tid * kA;
    float px[kA], py[kA], pz[kA];
#pragma unroll
    for (int u = 0; u < kA; ++u) {  // every load in flight before the first use
        // lanes past the end read the last atom and are masked below
        const int32_t a = a0 + u < natoms ? a0 + u : natoms - 1;
        const float *p = xyz + (int64_t)stride * a;
        px[u] = p[0]; py[u] = p[1]; pz[u] = p[2];
    }
    Box6 b;
    b.init();
#pragma unroll
    for (int u = 0; u < kA; ++u) {
        if (a0 + u < natoms) {
            b.mn0 = jl_min(b.mn0, px[u] - h); b.mn1 = jl_min(b.mn1, py[u] - h); b.mn2 = jl_min(b.mn2, pz[u] - h);
            b.mx0 = jl_max(b.mx0, px[u] + h); b.mx1 = jl_max(b.mx1, py[u] + h); b.mx2 = jl_max(b.mx2, pz[u] + h);
        }
    }
    // the leaf: four lanes
#pragma unroll
    for (int off = 1; off < 4; off <<= 1) {
        b.mn0 = jl_min(b.mn0, __shfl_xor(b.mn0, off, 64));
        b.mn1 = jl_min(b.mn1, __shfl_xor(b.mn1, off, 64));
        b.mn2 = jl_min(b.mn2, __shfl_xor(b.mn2, off, 64));
        b.mx0 = jl_max(b.mx0, __shfl_xor(b.mx0, off, 64));
        b.mx1 = jl_max(b.mx1, __shfl_xor(b.mx1, off, 64));
        b.mx2 = jl_max(b.mx2, __shfl_xor(b.mx2, off, 64));
    }
    if ((tid & 3) == 0 && a0 < natoms) {
        float *q = reinterpret_cast<float *>(nodes + a0 / SEG);
        *reinterpret_cast<float4 *>(q) = make_float4(b.mn0, b.mn1, b.mn2, b.mx0);
        *reinterpret_cast<float2 *>(q + 4) = make_float2(b.mx1, b.mx2);
    }
    if (tq != nullptr) {
        // leaf-local copies (uniform over the launch): per atom (-2 q', |q'|^2)
        // with q' = q - c about the leaf box's centre c, per leaf (c, max |q'|^2)
        const float cx = 0.5f * (b.mn0 + b.mx0), cy = 0.5f * (b.mn1 + b.mx1), cz = 0.5f * (b.mn2 + b.mx2);
        float r2 = 0.0f;
#pragma unroll
        for (int u = 0; u < kA; ++u) {
            const float qx = px[u] - cx, qy = py[u] - cy, qz = pz[u] - cz;
            const float q2 = __builtin_fmaf(qz, qz, __builtin_fmaf(qy, qy, qx * qx));
            if (a0 + u < natoms) {
                tq[a0 + u] = make_float4(-2.0f * qx, -2.0f * qy, -2.0f * qz, q2);
                r2 = fmaxf(r2, q2);
            }
        }
#pragma unroll
        for (int off = 1; off < 4; off <<= 1) r2 = fmaxf(r2, __shfl_xor(r2, off, 64));
        if ((tid & 3) == 0 && a0 < natoms) lc[a0 / SEG] = make_float4(cx, cy, cz, r2);
    }
    // the chunk: the rest of the wave, then the four waves
#pragma unroll
    for (int off = 4; off < 64; off <<= 1) {
        b.mn0 = jl_min(b.mn0, __shfl_xor(b.mn0, off, 64));
        b.mn1 = jl_min(b.mn1, __shfl_xor(b.mn1, off, 64));
        b.mn2 = jl_min(b.mn2, __shfl_xor(b.mn2, off, 64));
        b.mx0 = jl_max(b.mx0, __shfl_xor(b.mx0, off, 64));
        b.mx1 = jl_max(b.mx1, __shfl_xor(b.mx1, off, 64));
        b.mx2 = jl_max(b.mx2, __shfl_xor(b.mx2, off, 64));
    }
    const int w = tid >> 6;
    if ((tid & 63) == 0) {
        s_box[w][0] = b.mn0; s_box[w][1] = b.mn1; s_box[w][2] = b.mn2;
        s_box[w][3] = b.mx0; s_box[w][4] = b.mx1; s_box[w][5] = b.mx2;
    }
    __syncthreads();
    if (tid == 0) {
        Box6 cb;
        cb.init();
#pragma unroll
        for (int k = 0; k < nd::kBlock / 64; ++k)
            cb.merge(make_float4(s_box[k][0], s_box[k][1], s_box[k][2], s_box[k][3]),
                     make_float4(s_box[k][4], s_box[k][5], 0.0f, 0.0f));
        chunks[2 * c] = make_float4(cb.mn0, cb.mn1, cb.mn2, cb.mx0);
        chunks[2 * c + 1] = make_float4(cb.mx1, cb.mx2, 0.0f, 0.0f);
    }
}

// The refit fused with the four-wide records, for a tree whose topology is
// kept: 32 lanes per record, 8 per entry; an internal entry's box is the box
// of its leaf range (range_box, as k_lbvh_refit computes it), a leaf entry's
// its leaf box, an empty entry's the empty box.  The records equal those
// k_lbvh_wide4 makes after k_lbvh_refit bit for bit; the internal nodes'
// boxes in nodes are left as they were (the records are what the walk reads).
__global__ void __launch_bounds__(nd::kBlock)
k_refit_wide4(const nd_node *__restrict__ nodes, const int2 *__restrict__ ranges,
              const float4 *__restrict__ chunks, int32_t nL, float4 *__restrict__ rec) {
    const int32_t k = (int32_t)((blockIdx.x * blockDim.x + threadIdx.x) >> 5);
    const int e = (threadIdx.x >> 3) & 3, lane = threadIdx.x & (kGroup - 1);
    if (k >= nL - 1) return;  // whole 32-lane groups exit together
    int32_t id[4] = {0, 0, 0, 0};
    int32_t m = 0;
    const int32_t nn = 2 * nL - 1;
    const int32_t L = nodes[nL + k].left;
    const int32_t R = (L >= 1 && L <= nn) ? nodes[L - 1].skip : 0;
    const int32_t ch[2] = {L, R};
    for (int c = 0; c < 2; ++c) {
        const int32_t x = ch[c];
        if (x < 1 || x > nn) continue;  // malformed tree: the walk flags the empty record
        if (x <= nL) {
            id[m++] = x;
        } else {
            const int32_t g0 = nodes[x - 1].left;
            const int32_t g1 = (g0 >= 1 && g0 <= nn) ? nodes[g0 - 1].skip : 0;
            if (g0 >= 1 && g0 <= nn) id[m++] = g0;
            if (g1 >= 1 && g1 <= nn) id[m++] = g1;
        }
    }
    const int32_t x = id[e];
    Box6 b;
    b.init();
    if (x > nL) {
        const int2 rg = ranges[x - nL - 1];
        // an invalid range only follows an incomplete topology (status flagged)
        if (rg.x >= 1 && rg.y <= nL && rg.x <= rg.y) b = range_box(nodes, chunks, rg, lane);
    } else if (x >= 1) {
        const nd_node &q = nodes[x - 1];
        b.mn0 = q.min[0]; b.mn1 = q.min[1]; b.mn2 = q.min[2];
        b.mx0 = q.max[0]; b.mx1 = q.max[1]; b.mx2 = q.max[2];
    }
    if (lane == 0) {
        float *r = reinterpret_cast<float *>(rec + 8 * (int64_t)k);
        r[e] = b.mn0; r[4 + e] = b.mn1; r[8 + e] = b.mn2;
        r[12 + e] = b.mx0; r[16 + e] = b.mx1; r[20 + e] = b.mx2;
        r[24 + e] = __int_as_float(x);
        r[28 + e] = 0.0f;
    }
}

}  // namespace

namespace nd {

void launch_lbvh_wide4(hipStream_t s, const nd_node *nodes, int32_t nL, float4 *rec) {
    k_lbvh_wide4<<<grid_for(nL - 1), kBlock, 0, s>>>(nodes, nL, rec);
    ND_HIP(hipGetLastError());
}

// Box-only update of an existing tree (same leaf order, same topology):
// Leaf boxes: the one-lane-per-atom butterfly for power-of-two leaf sizes up
// to 64, the serial loop otherwise.
void launch_leaf_boxes(hipStream_t s, const float *xyz, int32_t stride, int32_t nL, int32_t apl,
                       float h, nd_node *nodes) {
    const int32_t na = nL * apl;
    switch (apl) {
        case 2: k_leaf_boxes_seg<2><<<grid_for(na), kBlock, 0, s>>>(xyz, stride, na, h, nodes); break;
        case 4: k_leaf_boxes_seg<4><<<grid_for(na), kBlock, 0, s>>>(xyz, stride, na, h, nodes); break;
        case 8: k_leaf_boxes_seg<8><<<grid_for(na), kBlock, 0, s>>>(xyz, stride, na, h, nodes); break;
        case 16: k_leaf_boxes_seg<16><<<grid_for(na), kBlock, 0, s>>>(xyz, stride, na, h, nodes); break;
        case 32: k_leaf_boxes_seg<32><<<grid_for(na), kBlock, 0, s>>>(xyz, stride, na, h, nodes); break;
        case 64: k_leaf_boxes_seg<64><<<grid_for(na), kBlock, 0, s>>>(xyz, stride, na, h, nodes); break;
        default: k_leaf_boxes<<<grid_for(nL), kBlock, 0, s>>>(xyz, stride, nL, apl, h, nodes);
    }
}

// leaf boxes from the current positions, chunk boxes, internal boxes.
void launch_lbvh_refit(hipStream_t s, const float *xyz_sorted, int32_t xyz_stride, int32_t nL,
                       int32_t apl, float r, nd_node *nodes, int32_t *store) {
    const size_t off_r = ((size_t)nL + 3) & ~(size_t)3;
    const size_t off_c = (off_r + 2 * (size_t)nL + 3) & ~(size_t)3;
    const int2 *ranges = reinterpret_cast<const int2 *>(store + off_r);
    float4 *chunks = reinterpret_cast<float4 *>(store + off_c);
    launch_leaf_boxes(s, xyz_sorted, xyz_stride, nL, apl, r / 2.0f, nodes);
    const int64_t nchunks = (nL + kChunk - 1) / kChunk;
    k_lbvh_chunks<<<grid_for(nchunks * 64), kBlock, 0, s>>>(nodes, nL, chunks);
    k_lbvh_refit<<<grid_for((int64_t)(nL - 1) * kGroup), kBlock, 0, s>>>(nodes, ranges, chunks, nL, 0);
    ND_HIP(hipGetLastError());
}

void launch_lbvh_refit_wide4(hipStream_t s, const float *xyz_sorted, int32_t xyz_stride, int32_t nL,
                             int32_t apl, float r, nd_node *nodes, int32_t *store, float4 *rec, float4 *tq,
                             float4 *lc) {
    const size_t off_r = ((size_t)nL + 3) & ~(size_t)3;
    const size_t off_c = (off_r + 2 * (size_t)nL + 3) & ~(size_t)3;
    const int2 *ranges = reinterpret_cast<const int2 *>(store + off_r);
    float4 *chunks = reinterpret_cast<float4 *>(store + off_c);
    const float h = r / 2.0f;
    const unsigned nchunks = (unsigned)((nL + kChunk - 1) / kChunk);
    const int32_t na = nL * apl;
    switch (apl) {
        case 4: k_leaf_chunk_boxes<4><<<nchunks, kBlock, 0, s>>>(xyz_sorted, xyz_stride, na, h, nodes, chunks, tq, lc); break;
        case 8: k_leaf_chunk_boxes<8><<<nchunks, kBlock, 0, s>>>(xyz_sorted, xyz_stride, na, h, nodes, chunks, tq, lc); break;
        case 16: k_leaf_chunk_boxes<16><<<nchunks, kBlock, 0, s>>>(xyz_sorted, xyz_stride, na, h, nodes, chunks, tq, lc); break;
        case 32: k_leaf_chunk_boxes<32><<<nchunks, kBlock, 0, s>>>(xyz_sorted, xyz_stride, na, h, nodes, chunks, tq, lc); break;
        case 64: k_leaf_chunk_boxes<64><<<nchunks, kBlock, 0, s>>>(xyz_sorted, xyz_stride, na, h, nodes, chunks, tq, lc); break;
        default:
            if (tq != nullptr) throw Error(ND_ERR_ARG, "refit: leaf-local copies need a leaf size of 4 to 64");
            launch_leaf_boxes(s, xyz_sorted, xyz_stride, nL, apl, h, nodes);
            k_lbvh_chunks<<<grid_for((int64_t)nchunks * 64), kBlock, 0, s>>>(nodes, nL, chunks);
    }
    k_refit_wide4<<<grid_for((int64_t)(nL - 1) * 32), kBlock, 0, s>>>(nodes, ranges, chunks, nL, rec);
    ND_HIP(hipGetLastError());
}

void launch_lbvh_wide(hipStream_t s, const nd_node *nodes, int32_t nL, float4 *wide) {
    k_lbvh_wide<<<grid_for(nL - 1), kBlock, 0, s>>>(nodes, nL, wide);
    ND_HIP(hipGetLastError());
}

void launch_lbvh(hipStream_t s, const float *xyz_sorted, int32_t xyz_stride,
                 const int32_t *codes_sorted, int32_t nL, int32_t apl, float r, int32_t override_root,
                 nd_node *nodes, int32_t *store, int32_t *status, bool slots_cleared,
                 bool leaves_done) {
    // store (lbvh_scratch_ints(nL) ints): nL-1 CAS slots | nL-1 int2 leaf
    // ranges | one 32-byte box per 64-leaf chunk
    if (!slots_cleared) launch_fill_i32(s, store, nL - 1, 0);
    const size_t off_r = ((size_t)nL + 3) & ~(size_t)3;
    const size_t off_c = (off_r + 2 * (size_t)nL + 3) & ~(size_t)3;
    int2 *ranges = reinterpret_cast<int2 *>(store + off_r);
    float4 *chunks = reinterpret_cast<float4 *>(store + off_c);
    const float h = r / 2.0f;  // neighbor_distance/2 in Float32
    if (!leaves_done) launch_leaf_boxes(s, xyz_sorted, xyz_stride, nL, apl, h, nodes);
    k_lbvh_topology<<<grid_for(nL), kBlock, 0, s>>>(codes_sorted, nL, apl, nodes, store, ranges, status);
    const int64_t nchunks = (nL + kChunk - 1) / kChunk;
    k_lbvh_chunks<<<grid_for(nchunks * 64), kBlock, 0, s>>>(nodes, nL, chunks);
    k_lbvh_refit<<<grid_for((int64_t)(nL - 1) * kGroup), kBlock, 0, s>>>(nodes, ranges, chunks, nL,
                                                                          override_root);
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
    // a buffer of its own: the one-call lists' store stays valid for export
    ctx->stage_store.reserve(sizeof(int32_t) * nd::lbvh_scratch_ints(nL));
    nd::launch_lbvh(ctx->stream, d_xyz_sorted, 3, d_codes_sorted, nL, chk.atomsperleaf,
                    chk.neighbor_distance, override_root, d_nodes, ctx->stage_store.as<int32_t>(),
                    ctx->status.as<int32_t>());
    if (nd::read_status(ctx) & nd::kStatusTree)
        throw nd::Error(ND_ERR_DOMAIN, "bounding_volume_hierarchy!: codes outside the range the "
                                       "build supports (bit 31 set)");
    ND_API_END(ctx)
}
