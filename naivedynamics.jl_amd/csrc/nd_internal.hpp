// nd_internal.hpp -- shared declarations of libnaivehip (gfx950 HIP).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>

#include "naivedynamics.h"

namespace nd {

constexpr int kBlock = 256;  // 4 wave64s per workgroup

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

#define ND_HIP(call)                                                                    \
    do {                                                                                \
        hipError_t e_ = (call);                                                         \
        if (e_ != hipSuccess)                                                           \
            throw ::nd::Error(ND_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

#define ND_REQUIRE(cond, msg)                                          \
    do {                                                               \
        if (!(cond)) throw ::nd::Error(ND_ERR_ARG, std::string(msg)); \
    } while (0)

// Grow-only device buffer owned by a context or simulation.
struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
    void reserve(size_t n) {
        if (n <= bytes && p) return;
        if (p) ND_HIP(hipFree(p));
        p = nullptr;
        bytes = 0;
        if (n == 0) n = 256;
        ND_HIP(hipMalloc(&p, n));
        bytes = n;
    }
    void release() {
        if (p) (void)hipFree(p);  // teardown: nothing to report to
        p = nullptr;
        bytes = 0;
    }
    template <class T> T *as() const { return static_cast<T *>(p); }
};

// Status word bits written by kernels (atomicOr).
enum : int32_t { kStatusDomain = 1, kStatusCapacity = 2, kStatusTree = 4 };

// Scratch for one radix sort of n (key, value) pairs.
struct SortBufs {
    DevBuf keys_alt, vals_alt, hist, flags;
};

}  // namespace nd

struct nd_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    std::string err;
    nd::DevBuf status;  // int32 status word + int64 counters
    nd::DevBuf codes, codes_sorted, perm, xyz_sorted, nodes, store, counts, offsets, xyz_in,
        pairs_out, bbox;
    nd::DevBuf cell_keys, cell_vals, cell_pos, cell_start;  // nd_cell_neighborlist
    nd::DevBuf trav_rec;  // four-wide records of the one-call traversal
    nd::DevBuf scan_tmp;  // tile sums of scan_counts
    nd::DevBuf force_acc, force_part;  // pair-list forces: fixed-point sums, energy partials
    nd::DevBuf async_ws;  // work buffers of nd_leafbuild_traverse_bvh_async (resized by it alone)
    nd::DevBuf stage_store;  // CAS slots etc. of the stage-level nd_bounding_volume_hierarchy
    nd::DevBuf rank_keys;    // nd_quantized_positions: keys and atom order (2N int32)
    nd::SortBufs sort;
    // last tree built by a one-call neighbour list (for nd_tree_export)
    int32_t tree_n = 0, tree_apl = 0;
    bool has_tree = false;
    const nd_node *tree_nodes = nullptr;
    const int32_t *tree_perm = nullptr, *tree_codes = nullptr;
    const int32_t *tree_store = nullptr;  // the build's CAS slots (nL - 1)
};

namespace nd {

// ---- launch helpers implemented in the kernel translation units ----
// All are asynchronous on ctx->stream.

// nd_morton.hip
void launch_morton(nd_ctx *ctx, const float *xyz, int32_t n, int32_t mode, const float lo[3],
                   const float scale[3], int32_t *codes, int32_t *idx, int32_t *status);
// the same with the box of the interleaved keys computed into box (6 floats,
// device memory) on stream st: kernels only
void launch_morton_box(hipStream_t st, float *box, const float *xyz, int32_t n, int32_t mode,
                       int32_t *codes, int32_t *idx, int32_t *status);
void launch_bbox(hipStream_t st, const float *xyz, int32_t n, float *bbox6);
void launch_fill_i32(hipStream_t s, int32_t *p, int64_t n, int32_t v);
void launch_iota(hipStream_t s, int32_t *p, int32_t n);
void radix_sort_pairs_auto(nd_ctx *ctx, int32_t *keys, int32_t *vals, int32_t n);
void launch_gather_xyz(hipStream_t s, const float *xyz, const int32_t *perm, int32_t n,
                       float *out);

// nd_sort.hip: stable LSD radix sort of int32 keys (signed order) with int32
// values; bit range [0, key_bits) of the sign-flipped key is sorted (pass
// key_bits = 32 for a full sort).  Results land in keys/vals.
void radix_sort_pairs(hipStream_t s, SortBufs &b, int32_t *keys, int32_t *vals, int32_t n,
                      int32_t first_bit, int32_t end_bit);
// Keys below 2^key_bits in 10-bit passes; the sorted pairs are left in
// *keys_sorted / *vals_sorted (keys/vals after an even number of passes,
// b's alternate buffers after an odd one).
void radix_sort_keys10(hipStream_t s, SortBufs &b, int32_t *keys, int32_t *vals, int32_t n,
                       int32_t key_bits, int32_t **keys_sorted, int32_t **vals_sorted);
size_t radix_scratch_blocks(int32_t n);

// nd_lbvh.hip
// status (may be null) receives kStatusTree if the codes break the build's
// index invariants (only possible for caller-supplied codes with bit 31 set).
// store must hold lbvh_scratch_ints(nL) int32.
inline size_t lbvh_scratch_ints(int32_t nL) { return 4 * ((size_t)nL + 16); }
// Box-only update of a tree built by launch_lbvh on the same store.
void launch_lbvh_refit(hipStream_t s, const float *xyz_sorted, int32_t xyz_stride, int32_t nL,
                       int32_t apl, float r, nd_node *nodes, int32_t *store);
// Child-box form of the internal nodes (4 float4 per internal node).
void launch_lbvh_wide(hipStream_t s, const nd_node *nodes, int32_t nL, float4 *wide);
// Four-wide records of the internal nodes (8 float4 per internal node).
void launch_lbvh_wide4(hipStream_t s, const nd_node *nodes, int32_t nL, float4 *rec);
// launch_lbvh_refit followed by launch_lbvh_wide4, fused: leaf boxes, chunk
// boxes and the records (internal node boxes in nodes are not updated).
// tq, lc (optional, leaf sizes 4 to 64): per atom (-2 q', |q'|^2) with q' the
// atom's position less its leaf box's centre c, per leaf (c, max |q'|^2).
void launch_lbvh_refit_wide4(hipStream_t s, const float *xyz_sorted, int32_t xyz_stride, int32_t nL,
                             int32_t apl, float r, nd_node *nodes, int32_t *store, float4 *rec,
                             float4 *tq = nullptr, float4 *lc = nullptr);
// slots_cleared: the caller has already zeroed store[0 .. nL-2] (the CAS slots);
// leaves_done: the caller has already written the leaf boxes (+- r/2).
void launch_lbvh(hipStream_t s, const float *xyz_sorted, int32_t xyz_stride,
                 const int32_t *codes_sorted, int32_t nL, int32_t apl, float r, int32_t override_root,
                 nd_node *nodes, int32_t *store, int32_t *status, bool slots_cleared = false,
                 bool leaves_done = false);

// Largest per-atom list capacity: a wave's [cap][64] block of int32 rows is
// addressed through a buffer resource whose byte range (cap * 256) is int32.
constexpr int32_t kMaxListRows = (1 << 23) - 1;

// nd_sim.hip: the two-phase list walk over an LBVH of float4 positions
// (sorted; leaf boxes +- rlist/2) with 4, 8, 16 or 32 atoms per leaf.  It
// writes the four-wide records into rec (8 float4 per internal node), one
// block of [cap][64] int32 neighbour slots per 64 sorted atoms into nbr (the
// neighbours' sorted indices), each atom's count into ncount, and into ctr:
// [0] the largest count (atomic max), [1] 1 if any count exceeds cap,
// [2] kStatusTree if the tree was malformed.
void launch_list_walk(hipStream_t st, const nd_node *nodes, float4 *rec, const float4 *pos, int32_t n,
                      int32_t nL, int32_t apl, float rlist, int32_t cap, int32_t *nbr, int32_t *ncount,
                      int32_t *ctr);

// nd_traverse.hip
// own_tree: the tree was built by launch_lbvh (well formed: left pointers and
// ropes describe one depth-first order), which allows the wave-cooperative walk.
// count_dev != null: stream-ordered (no host synchronisation): the required
// count goes to *count_dev, status bits to status_dev, pairs are written only
// if the count is <= cap; returns -1.
int64_t leaf_traverse(nd_ctx *ctx, const nd_node *nodes, const float *xyz_sorted,
                      const int32_t *idx_sorted, int32_t n, int32_t apl, float r, int32_t predicate,
                      nd_pair *out, int64_t cap, bool own_tree = false, int32_t *status_dev = nullptr,
                      int64_t *count_dev = nullptr);
// Exclusive scan of ctx->counts (n int64) into ctx->offsets (n + 1): launch only,
// or launch, synchronise and return the total.
void launch_scan(nd_ctx *ctx, int64_t n);
// Point-query walk (query_mode ND_QUERY_OWN_LEAF / ND_QUERY_ROOT) with the
// counting/scan/emit protocol of leaf_traverse; synchronous.
int64_t point_traverse(nd_ctx *ctx, const nd_node *nodes, const float *xyz_sorted, const int32_t *idx_sorted,
                       int32_t n, int32_t apl, float r, int32_t mode, nd_pair *out, int64_t cap);
int64_t scan_counts(nd_ctx *ctx, int64_t n);

// Work buffers of the stream-ordered one-call neighbour list, carved from one
// block of device memory of list_work_bytes(n, apl) bytes.
struct ListWork {
    int32_t *codes, *perm, *keys_alt, *vals_alt, *hist, *store;
    float *xyz_sorted, *box;
    nd_node *nodes;
    int64_t *counts, *offsets, *scan_tmp;
    float4 *rec;
};
size_t list_work_bytes(int32_t n, int32_t apl);
ListWork carve_list_work(void *base, int32_t n, int32_t apl);
// keys, 32-bit sort, permutation, tree and traversal on st without host
// synchronisation or allocation: the required pair count goes to *d_count,
// status bits to *d_status, pairs to out when the count is <= cap.
void launch_list_async(hipStream_t st, const ListWork &w, const float *d_xyz, int32_t n, int32_t apl,
                       float r, int32_t pred, int32_t morton, nd_pair *out, int64_t cap, int64_t *d_count,
                       int32_t *d_status);
// Device address of the library's static scratch array on the current device
// (looked up once per device).
void *static_scratch(int device);

// Uniform cell grid: cell d-coordinate of x is clamp(floor((x - lo[d]) * inv[d]),
// 0, g[d] - 1), computed in double so that it is exact to far below a cell;
// NaN maps to 0.  Cells are numbered (cz * g[1] + cy) * g[0] + cx (x fastest).
struct CellGrid {
    double lo[3], inv[3];
    int32_t g[3], ncell;
};
// Grid over [lo, lo + ext] (per axis) with cells of side >= r (1 + 1e-4), at most
// 1024 per axis and max_cells in total (larger cells when capped).
CellGrid make_cell_grid(const double lo[3], const double ext[3], double r, int64_t max_cells);

// nd_cells.hip: cell index of n positions (stride 3 or 4 floats): keys (cell
// number; ncell for a non-finite coordinate, so such atoms sort last and lie
// in no cell), a stable sort by key with vals = 0..n-1, cpos[t] = (x, y, z,
// bits of vals[t]) for sorted position t, and cstart[k] = first sorted
// position with key >= k for k = 0..ncell.  keys/vals must hold n int32;
// cstart ncell + 1.  Returns the sorted keys (keys, or sb's alternate buffer).
const int32_t *launch_cell_index(hipStream_t st, SortBufs &sb, const float *xyz, int32_t stride,
                                 int32_t n, const CellGrid &cg, int32_t *keys, int32_t *vals,
                                 float4 *cpos, int32_t *cstart);

__device__ __forceinline__ int32_t cell_coord(const CellGrid &cg, float x, int d) {
    const double t = ((double)x - cg.lo[d]) * cg.inv[d];
    if (!(t > 0.0)) return 0;
    if (t >= (double)(cg.g[d] - 1)) return cg.g[d] - 1;
    return (int32_t)t;
}

void set_error(nd_ctx *ctx, const std::string &m);
int32_t read_status(nd_ctx *ctx);  // synchronises; returns and clears the status word
void clear_status(nd_ctx *ctx);

inline unsigned grid_for(int64_t n, int block = kBlock) {
    return (unsigned)((n + block - 1) / block);
}

}  // namespace nd

namespace nd {

// Float32(1/1023) (BVHTraverse.jl:184): the quotient rounded once to binary32.
constexpr float kBinwidth = (float)(1.0 / 1023.0);
// mortoncodes! masks (BVHTraverse.jl:241).
constexpr int32_t kMagic0 = 0x09249249, kMagic1 = 0x12492492, kMagic2 = 0x24924924;

// Spread the low 10 bits of x to bit positions 0, 3, 6, ..., 27.
__device__ __forceinline__ uint32_t spread10(uint32_t x) {
    x &= 0x3FFu;
    x = (x | (x << 16)) & 0x030000FFu;
    x = (x | (x << 8)) & 0x0300F00Fu;
    x = (x | (x << 4)) & 0x030C30C3u;
    x = (x | (x << 2)) & 0x09249249u;
    return x;
}

// 30-bit Hilbert index of a cell (x, y, z), 10 bits per axis, by Skilling's
// transform (J. Skilling, "Programming the Hilbert curve", AIP Conf. Proc.
// 707, 2004): undo the excess work of the inverse transform from the top bit
// down, Gray-encode, then interleave the transposed coordinates with x's bit
// the most significant of each triple.  Consecutive indices are face-adjacent
// cells, and every 3k-bit prefix is a cube of the level-k octree, so an LBVH
// over these keys is an octree hierarchy like the interleaved key's but its
// fixed-size leaves are compact runs.
__device__ __forceinline__ uint32_t hilbert30(uint32_t x, uint32_t y, uint32_t z) {
#pragma unroll
    for (uint32_t q = 1u << 9; q > 1u; q >>= 1) {
        const uint32_t p = q - 1u;
        x = (x & q) ? (x ^ p) : x;
        {
            const uint32_t t = (x ^ y) & p;
            const bool inv = (y & q) != 0u;
            x = inv ? (x ^ p) : (x ^ t);
            y = inv ? y : (y ^ t);
        }
        {
            const uint32_t t = (x ^ z) & p;
            const bool inv = (z & q) != 0u;
            x = inv ? (x ^ p) : (x ^ t);
            z = inv ? z : (z ^ t);
        }
    }
    y ^= x;
    z ^= y;
    uint32_t t = 0u;
#pragma unroll
    for (uint32_t q = 1u << 9; q > 1u; q >>= 1) t = (z & q) ? (t ^ (q - 1u)) : t;
    x ^= t;
    y ^= t;
    z ^= t;
    return (spread10(x) << 2) | (spread10(y) << 1) | spread10(z);
}

// s = sum((a - b).^2) as StaticArrays folds it: ((dx*dx + dy*dy) + dz*dz).
// Exact only in files compiled with -ffp-contract=off (no multiply-add fused).
__device__ __forceinline__ float sqdist(float ax, float ay, float az, float bx, float by,
                                        float bz) {
    const float dx = ax - bx, dy = ay - by, dz = az - bz;
    const float sx = dx * dx, sy = dy * dy, sz = dz * dz;
    float s = sx + sy;
    s = s + sz;
    return s;
}

// ND_PRED_SCALAR: s < fl(r*r) then d = sqrt(s) (BVHTraverse.jl:1027,1048).
// ND_PRED_SIMD:   d = sqrt(s), d < r            (NaiveSIMD.jl:529,589).
__device__ __forceinline__ bool accept(float s, float r, float r2, int32_t pred, float &d) {
    if (pred == ND_PRED_SCALAR) {
        if (s < r2) { d = sqrtf(s); return true; }
        return false;
    }
    d = sqrtf(s);
    return d < r;
}

// Order this wave's LDS writes before its later LDS reads (other lanes' data).
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// q = round(Int32, v, RoundDown): defined for -2^31 <= v < 2^31 only.
__device__ __forceinline__ bool floor_i32(float v, int32_t &q) {
    if (!(v >= -2147483648.0f && v < 2147483648.0f)) return false;
    q = (int32_t)floorf(v);
    return true;
}

// 10-bit cell coordinate of x in the interleaved key: clamp(floor((x-lo)*s), 0, 1023).
__device__ __forceinline__ uint32_t cell10(float x, float lo, float s) {
    float t = (x - lo) * s;
    if (!(t > 0.0f)) return 0u;  // also maps NaN to 0 (NaN input is flagged)
    if (t >= 1023.0f) return 1023u;
    return (uint32_t)floorf(t);
}

}  // namespace nd

// Julia's Float32 min/max (ordering with -0 < +0; NaN propagates), branch
// free: for non-NaN floats the integer key bits ^ ((bits >> 31) & 0x7FFFFFFF)
// orders like the floats with -0 below +0, so one signed compare and a select
// decide; an unordered pair gives x - y (a NaN), as Julia's min/max do.
__device__ __forceinline__ int32_t jl_order_key(float x) {
    const int32_t b = __float_as_int(x);
    return b ^ ((b >> 31) & 0x7FFFFFFF);
}
__device__ __forceinline__ float jl_min(float x, float y) {
    const float m = jl_order_key(y) < jl_order_key(x) ? y : x;
    return __builtin_isunordered(x, y) ? x - y : m;
}
__device__ __forceinline__ float jl_max(float x, float y) {
    const float m = jl_order_key(y) > jl_order_key(x) ? y : x;
    return __builtin_isunordered(x, y) ? x - y : m;
}

#define ND_API_BEGIN try {
#define ND_API_END(ctx)                                    \
    }                                                      \
    catch (const ::nd::Error &e) {                         \
        if (ctx) ::nd::set_error(ctx, e.what());           \
        return e.code;                                     \
    }                                                      \
    catch (const std::exception &e) {                      \
        if (ctx) ::nd::set_error(ctx, e.what());           \
        return ND_ERR_HIP;                                 \
    }                                                      \
    return ND_OK;
