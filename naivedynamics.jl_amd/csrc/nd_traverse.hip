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
#include <cfloat>
#include <mutex>

#include "nd_internal.hpp"

namespace {

using nd::accept;
using nd::sqdist;
using nd::wave_lds_sync;

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
                nd_pair *__restrict__ out, int32_t *__restrict__ status, int64_t cap) {
    const int32_t q = 1 + (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
    if (q > nL) return;
    if (EMIT && offsets[nL] > cap) return;  // too small a buffer: write nothing
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
                 nd_pair *__restrict__ out, int32_t *__restrict__ status, int64_t cap) {
    const int32_t q = 1 + (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
    if (q > nL) return;
    if (EMIT && offsets[nL] > cap) return;  // too small a buffer: write nothing
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

// ---- point-query walks ----
//
// neighbor_traverse (BVHTraverse.jl:1125-1234; proximity_test! :1057-1087,
// overlap_test :1089-1092): one lane per query atom q (sorted, q < N-1),
// starting at q's own leaf; a node is entered iff the point lies strictly
// inside its box on all three axes; a leaf tests its atoms s > q with
// s_qs < fl(r*r) and follows its rope.  ROOT: gpuneighbor_traverse!
// (ext/NaiveKA.jl:352-452), every atom queries from the root and keeps the
// subjects whose original index exceeds its own.  Counting pass, exclusive
// scan over queries, emitting pass: pairs land query-major in walk order,
// the order of a single-threaded reference run.
template <bool EMIT, bool ROOT>
__global__ void __launch_bounds__(nd::kBlock)
k_point_traverse(const nd_node *__restrict__ nodes, const float *__restrict__ xyz,
                 const int32_t *__restrict__ idx, int32_t n, int32_t nL, int32_t apl, float r2,
                 int64_t *__restrict__ counts, const int64_t *__restrict__ offsets,
                 nd_pair *__restrict__ out, int32_t *__restrict__ status, int64_t cap) {
    const int32_t nq = ROOT ? n : n - 1;
    const int32_t q = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
    if (q >= nq) return;
    if (EMIT && offsets[nq] > cap) return;  // too small a buffer: write nothing
    Sink<EMIT> sink;
    if (EMIT) {
        sink.out = out + offsets[q];
        sink.limit = offsets[q + 1] - offsets[q];
    }
    const float px = xyz[3 * (int64_t)q], py = xyz[3 * (int64_t)q + 1], pz = xyz[3 * (int64_t)q + 2];
    const int32_t qi = (ROOT || EMIT) ? idx[q] : 0;
    int32_t t = ROOT ? nL + 1 : q / apl + 1;
    // each node at most once: more visits or an index outside [1, 2nL-1]
    // means a malformed tree
    int32_t visits = 0;
    while (t != 0) {
        if (++visits > 2 * nL || t < 1 || t > 2 * nL - 1) {
            atomicOr(status, nd::kStatusTree);
            break;
        }
        const Box b = load_node(nodes, t);
        const bool inside = (b.mn[0] < px) & (px < b.mx[0]) & (b.mn[1] < py) & (py < b.mx[1]) &
                            (b.mn[2] < pz) & (pz < b.mx[2]);
        if (!inside) {
            t = b.skip;
            continue;
        }
        if (b.left != 0) {
            t = b.left;
            continue;
        }
        const int64_t lo = (int64_t)(t - 1) * apl;
        // own-leaf form: subjects after the query only
        const int32_t e0 = ROOT ? 0 : (int32_t)max((int64_t)0, (int64_t)q + 1 - lo);
        for (int32_t e = e0; e < apl; ++e) {
            const int64_t s = lo + e;
            const int32_t si = (ROOT || EMIT) ? idx[s] : 0;
            if (ROOT && !(qi < si)) continue;
            const float d2 = sqdist(px, py, pz, xyz[3 * s], xyz[3 * s + 1], xyz[3 * s + 2]);
            if (d2 < r2) sink.push(qi, si, sqrtf(d2));
        }
        t = b.skip;
    }
    if (!EMIT) counts[q] = sink.n;
}

// ---- wave-cooperative form (trees built by this library) ----
//
// For a well-formed tree the rope walk of query leaf q visits, in increasing
// leaf order, exactly the leaves t > q whose boxes strictly overlap q's: it
// starts at q's rope, every later node is reached unless an ancestor after q
// fails the test, and an ancestor's box (the exact union of its leaves')
// overlaps whenever one of its leaves' does.  The root override never
// matters (no rope leads to the root).  So one wave of 64 query atoms (64/A
// leaves, lane = row a of leaf q) walks the tree once, depth first and left
// to right over four-wide records, each stack entry carrying the lanes whose
// leaf box overlaps it and whose leaf precedes the entry's last leaf.  Popped
// leaves come in increasing order; a popped leaf t is tile-tested by the
// lanes of every q < t in its mask, and each lane places its row of the tile
// after the rows of the lanes before it in its leaf.  The output equals the
// rope walk's pair for pair, in order.

constexpr int kTravStack = 128;  // >= 3 * 32 record levels + 4 (depth <= 63)

// rec[8k .. 8k+7] for internal node nL+1+k: the node's grandchildren in left
// to right order (a leaf child stands for itself) as structure-of-arrays boxes
// (min x, y, z, max x, y, z), their node numbers (0 = empty slot), and the
// last leaf of each entry's subtree (the leaf before the first leaf of its
// rope target; nL for a rope to 0).
__global__ void __launch_bounds__(nd::kBlock)
k_trav_records(const nd_node *__restrict__ nodes, int32_t nL, float4 *__restrict__ rec) {
    const int32_t k = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
    if (k >= nL - 1) return;
    const int32_t nn = 2 * nL - 1;
    int32_t id[4] = {0, 0, 0, 0}, last[4] = {0, 0, 0, 0};
    int32_t m = 0;
    const int32_t L = nodes[nL + k].left;
    const int32_t R = (L >= 1 && L <= nn) ? nodes[L - 1].skip : 0;
    const int32_t ch[2] = {L, R};
    for (int c = 0; c < 2; ++c) {
        const int32_t x = ch[c];
        if (x < 1 || x > nn) continue;
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
        const int32_t x = id[e];
        if (x != 0) {
            const nd_node &q = nodes[x - 1];
            b[0][e] = q.min[0]; b[1][e] = q.min[1]; b[2][e] = q.min[2];
            b[3][e] = q.max[0]; b[4][e] = q.max[1]; b[5][e] = q.max[2];
            if (x <= nL) {
                last[e] = x;
            } else {
                // leftmost leaf of the rope target, minus one (nL if the
                // descent does not end in a leaf: no pruning, still exact)
                int32_t y = q.skip, steps = 0;
                while (y > nL && y <= nn && steps++ < 64) y = nodes[y - 1].left;
                last[e] = q.skip == 0 ? nL : ((y >= 1 && y <= nL) ? y - 1 : nL);
            }
        } else {
            b[0][e] = b[1][e] = b[2][e] = INFINITY;
            b[3][e] = b[4][e] = b[5][e] = -INFINITY;
        }
    }
    float4 *r = rec + 8 * (int64_t)k;
    for (int d = 0; d < 6; ++d) r[d] = make_float4(b[d][0], b[d][1], b[d][2], b[d][3]);
    r[6] = make_float4(__int_as_float(id[0]), __int_as_float(id[1]), __int_as_float(id[2]),
                       __int_as_float(id[3]));
    r[7] = make_float4(__int_as_float(last[0]), __int_as_float(last[1]), __int_as_float(last[2]),
                       __int_as_float(last[3]));
}

template <int PRED>
__device__ __forceinline__ bool accept_t(float s, float r, float r2, float &d) {
    if constexpr (PRED == ND_PRED_SCALAR) {
        if (s < r2) { d = sqrtf(s); return true; }
        return false;
    } else {
        d = sqrtf(s);
        return d < r;
    }
}

template <bool EMIT, int A, int PRED>
__global__ void __launch_bounds__(nd::kBlock)
k_leaf_traverse_w(const nd_node *__restrict__ nodes, const float4 *__restrict__ rec,
                  const float *__restrict__ xyz, const int32_t *__restrict__ idx, int32_t nL,
                  float r, float r2, int32_t nslice, int64_t *__restrict__ counts,
                  const int64_t *__restrict__ offsets, nd_pair *__restrict__ out,
                  int32_t *__restrict__ status, int64_t cap) {
    static_assert(A >= 1 && A <= 64 && (64 % A) == 0, "leaf size must divide 64");
    // too small a buffer: write nothing (uniform over the launch)
    if (EMIT && offsets[(int64_t)nL * nslice] > cap) return;
    // query atoms per wave: 32 (the upper half-wave idles) keeps twice the
    // walks in flight and a smaller union of leaf boxes per walk
    constexpr int kApw = A <= 32 ? 32 : 64;
    constexpr int kW = nd::kBlock / 64;
    __shared__ int32_t s_node[kW][kTravStack], s_first[kW][kTravStack];
    __shared__ uint64_t s_mask[kW][kTravStack];
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    // wave g walks slice g % nslice (the leaves t in [tlo, thi]) for the
    // query atoms of wave group g / nslice: a leaf whose box overlaps many
    // leaves spreads its walk over nslice waves
    const int32_t gw = (int32_t)((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
    const int32_t slice = gw % nslice;
    const int32_t tlo = 1 + (int32_t)((int64_t)slice * nL / nslice);
    const int32_t thi = (int32_t)((int64_t)(slice + 1) * nL / nslice);
    const int32_t s = (gw / nslice) * kApw + lane;
    const bool active = lane < kApw && s < nL * A;
    const int32_t Q = s / A + 1;  // query leaf (1-based)
    const int a = s % A;          // row of this lane in its leaf
    const int gbase = lane - a;   // first lane of the leaf's group
    Box ql;
    if (active) {
        ql = load_node(nodes, Q);
    } else {
        ql.mn[0] = ql.mn[1] = ql.mn[2] = INFINITY;
        ql.mx[0] = ql.mx[1] = ql.mx[2] = -INFINITY;
    }
    float ax = 0.0f, ay = 0.0f, az = 0.0f;
    int32_t oi = 0;
    if (active) {
        ax = xyz[3 * (int64_t)s];
        ay = xyz[3 * (int64_t)s + 1];
        az = xyz[3 * (int64_t)s + 2];
        if (EMIT) oi = idx[s];
    }
    // pairs of (leaf Q, this slice) placed so far (the same in the group)
    int64_t run = 0, limit = 0;
    const int64_t cslot = (int64_t)(Q - 1) * nslice + slice;
    nd_pair *o = nullptr;
    if (EMIT && active) {
        o = out + offsets[cslot];
        limit = offsets[cslot + 1] - offsets[cslot];
    }
    // Tile rows: bit b of hits = pair (a, b) of this lane's row against the
    // atoms tlo .. tlo+A-1 accepted (b > a only on the own leaf).  Called
    // by the whole wave (the group scan shuffles); lanes with part == false
    // contribute nothing.
    // Places this lane's row of a tile (bit b of hits = pair (a, b) against
    // atoms tlo .. tlo+A-1) after the rows of the lanes before it in its
    // leaf, and emits it.  Called by the whole wave (the group scan shuffles).
    auto place = [&](uint64_t hits, int64_t tlo) {
        const int32_t c = __builtin_popcountll(hits);
        int32_t x = c;
#pragma unroll
        for (int off = 1; off < A; off <<= 1) {
            const int32_t v = __shfl_up(x, off, 64);
            if (a >= off) x += v;
        }
        const int32_t total = __shfl(x, gbase + A - 1, 64);
        if (EMIT && hits) {
            int64_t pos = run + (x - c);
            while (hits) {
                const int b = __builtin_ctzll(hits);
                hits &= hits - 1;
                float d;
                accept_t<PRED>(sqdist(ax, ay, az, xyz[3 * (tlo + b)], xyz[3 * (tlo + b) + 1],
                                      xyz[3 * (tlo + b) + 2]),
                               r, r2, d);
                if (pos < limit) {
                    nd_pair p;
                    p.i = oi;
                    p.j = idx[tlo + b];
                    p.d = d;
                    o[pos] = p;
                }
                ++pos;
            }
        }
        run += total;
    };
    // same-leaf pairs (a, b > a), rows in order, in the first slice
    if (slice == 0) {
        uint64_t hits = 0;
        const int64_t qlo = (int64_t)(Q - 1) * A;
        if (active)
            for (int b = a + 1; b < A; ++b) {
                float d;
                if (accept_t<PRED>(sqdist(ax, ay, az, xyz[3 * (qlo + b)], xyz[3 * (qlo + b) + 1],
                                          xyz[3 * (qlo + b) + 2]),
                                   r, r2, d))
                    hits |= (uint64_t)1 << b;
            }
        place(hits, qlo);
    }
    // a popped leaf t (wave-uniform): its atoms are fetched together, in
    // chunks of 4, then tested by the participating lanes
    auto tile = [&](bool part, int32_t t) {
        const int64_t t0 = (int64_t)(t - 1) * A;
        constexpr int CH = A < 4 ? A : 4;
        uint64_t hits = 0;
#pragma unroll
        for (int c0 = 0; c0 < A; c0 += CH) {
            float tx[CH], ty[CH], tz[CH];
#pragma unroll
            for (int k = 0; k < CH; ++k) {
                tx[k] = xyz[3 * (t0 + c0 + k)];
                ty[k] = xyz[3 * (t0 + c0 + k) + 1];
                tz[k] = xyz[3 * (t0 + c0 + k) + 2];
            }
            if (part) {
#pragma unroll
                for (int k = 0; k < CH; ++k) {
                    float d;
                    if (accept_t<PRED>(sqdist(ax, ay, az, tx[k], ty[k], tz[k]), r, r2, d))
                        hits |= (uint64_t)1 << (c0 + k);
                }
            }
        }
        place(hits, t0);
    };

    const uint64_t act = __builtin_amdgcn_ballot_w64(active);
    if (lane == 0) {
        s_node[w][0] = nL + 1;  // the root is internal (nL >= 2)
        s_first[w][0] = 1;
        s_mask[w][0] = act;
    }
    int32_t sp = act ? 1 : 0, pops = 0;
    bool bad = false;
    const uint32_t nn = (uint32_t)(2 * nL - 1);
    wave_lds_sync();
    while (true) {
        sp = __builtin_amdgcn_readfirstlane(sp);  // wave-uniform: scalar loop control
        if (sp <= 0) break;
        --sp;
        const int32_t node = __builtin_amdgcn_readfirstlane(s_node[w][sp]);
        int32_t first = __builtin_amdgcn_readfirstlane(s_first[w][sp]);  // first leaf of node
        const uint64_t mask = s_mask[w][sp];
        wave_lds_sync();  // the popped slot is read before a push reuses it
        if (++pops > 2 * nL) { bad = true; break; }
        const bool mine = (mask >> lane) & 1u;
        if (node > nL) {
            const float4 *rc = rec + 8 * (int64_t)(node - nL - 1);
            const float4 mnx = rc[0], mny = rc[1], mnz = rc[2], mxx = rc[3], mxy = rc[4], mxz = rc[5];
            const float4 ids = rc[6], lst = rc[7];
            // entries cover consecutive leaf ranges: entry e spans
            // [first(e), last(e)], first(e+1) = last(e) + 1
            int32_t efirst[4];
            efirst[0] = first;
            efirst[1] = __float_as_int(lst.x) + 1;
            efirst[2] = __float_as_int(lst.y) + 1;
            efirst[3] = __float_as_int(lst.z) + 1;
            auto push = [&](float mx0, float my0, float mz0, float mx1, float my1, float mz1, float idf,
                            float lf, int32_t ef) {
                const int32_t c = __float_as_int(idf);
                const int32_t el = __float_as_int(lf);
                if (c == 0) return;
                if ((uint32_t)(c - 1) >= nn || sp >= kTravStack) { bad = true; return; }
                if (el < tlo || ef > thi) return;  // no leaf of this slice below it
                const bool ov = mine & (el > Q) & (ql.mn[0] < mx1) & (ql.mx[0] > mx0) &
                                (ql.mn[1] < my1) & (ql.mx[1] > my0) & (ql.mn[2] < mz1) & (ql.mx[2] > mz0);
                const uint64_t mk = __builtin_amdgcn_ballot_w64(ov);
                if (mk) {
                    if (lane == 0) {
                        s_node[w][sp] = c;
                        s_first[w][sp] = ef;
                        s_mask[w][sp] = mk;
                    }
                    ++sp;
                }
            };
            // an internal node has two children, so a record has >= 2 entries
            if (__float_as_int(ids.y) == 0) { bad = true; break; }
            // last entry first: entry 0 (leftmost) ends on top
            push(mnx.w, mny.w, mnz.w, mxx.w, mxy.w, mxz.w, ids.w, lst.w, efirst[3]);
            push(mnx.z, mny.z, mnz.z, mxx.z, mxy.z, mxz.z, ids.z, lst.z, efirst[2]);
            push(mnx.y, mny.y, mnz.y, mxx.y, mxy.y, mxz.y, ids.y, lst.y, efirst[1]);
            push(mnx.x, mny.x, mnz.x, mxx.x, mxy.x, mxz.x, ids.x, lst.x, efirst[0]);
            if (bad) break;
            wave_lds_sync();
        } else {
            tile(mine & (Q < node), node);
        }
    }
    if (bad && lane == 0) atomicOr(status, nd::kStatusTree);
    if (!EMIT && active && a == 0) counts[cslot] = run;
}

// Exclusive scan of n int64 counts in tiles of kScanTile: tile sums, a scan
// of the sums (k_scan_i64), then each tile's scan plus its offset.
constexpr int kScanPer = 8, kScanTile = nd::kBlock * kScanPer;

__global__ void __launch_bounds__(nd::kBlock)
k_scan_tile_sums(const int64_t *__restrict__ in, int64_t n, int64_t *__restrict__ sums) {
    __shared__ int64_t part[nd::kBlock / 64];
    const int64_t base = (int64_t)blockIdx.x * kScanTile;
    int64_t s = 0;
#pragma unroll
    for (int k = 0; k < kScanPer; ++k) {
        const int64_t i = base + k * nd::kBlock + threadIdx.x;  // coalesced
        if (i < n) s += in[i];
    }
    for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        int64_t t = 0;
        for (int w = 0; w < nd::kBlock / 64; ++w) t += part[w];
        sums[blockIdx.x] = t;
    }
}

__global__ void __launch_bounds__(nd::kBlock)
k_scan_tile_apply(const int64_t *__restrict__ in, int64_t n, const int64_t *__restrict__ tile_off,
                  int64_t ntiles, int64_t *__restrict__ out) {
    __shared__ int64_t part[nd::kBlock];
    const int t = threadIdx.x;
    const int64_t lo = (int64_t)blockIdx.x * kScanTile + (int64_t)t * kScanPer;
    int64_t v[kScanPer];
    int64_t s = 0;
#pragma unroll
    for (int k = 0; k < kScanPer; ++k) {
        v[k] = lo + k < n ? in[lo + k] : 0;
        s += v[k];
    }
    part[t] = s;
    __syncthreads();
    for (int off = 1; off < nd::kBlock; off <<= 1) {
        const int64_t x = t >= off ? part[t - off] : 0;
        __syncthreads();
        part[t] += x;
        __syncthreads();
    }
    int64_t run = tile_off[blockIdx.x] + part[t] - s;
#pragma unroll
    for (int k = 0; k < kScanPer; ++k) {
        if (lo + k < n) out[lo + k] = run;
        run += v[k];
    }
    if (blockIdx.x == ntiles - 1 && t == 0) out[n] = tile_off[ntiles];
}

__global__ void k_copy_total(const int64_t *__restrict__ src, int64_t *__restrict__ dst) {
    if (threadIdx.x == 0) *dst = *src;
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

// int64 entries of the tile sums a scan of n counts needs (none when one
// workgroup scans them all)
inline int64_t scan_tmp_entries(int64_t n) {
    return n <= 4 * kScanTile ? 0 : 2 * ((n + kScanTile - 1) / kScanTile) + 1;
}

// counts (n) -> exclusive offsets (n + 1) on st; tmp holds scan_tmp_entries(n)
void launch_scan_buf(hipStream_t st, const int64_t *counts, int64_t n, int64_t *offsets, int64_t *tmp) {
    if (n <= 4 * kScanTile) {
        k_scan_i64<<<1, 1024, 0, st>>>(counts, n, offsets);
    } else {
        const int64_t nt = (n + kScanTile - 1) / kScanTile;
        int64_t *sums = tmp, *off = sums + nt;
        k_scan_tile_sums<<<(unsigned)nt, kBlock, 0, st>>>(counts, n, sums);
        k_scan_i64<<<1, 1024, 0, st>>>(sums, nt, off);
        k_scan_tile_apply<<<(unsigned)nt, kBlock, 0, st>>>(counts, n, off, nt, offsets);
    }
    ND_HIP(hipGetLastError());
}

void launch_scan(nd_ctx *ctx, int64_t n) {
    // counts in ctx->counts (n), offsets to ctx->offsets (n+1), no host sync
    ctx->scan_tmp.reserve(sizeof(int64_t) * (size_t)(scan_tmp_entries(n) + 1));
    launch_scan_buf(ctx->stream, ctx->counts.as<int64_t>(), n, ctx->offsets.as<int64_t>(),
                    ctx->scan_tmp.as<int64_t>());
}

int64_t scan_counts(nd_ctx *ctx, int64_t n) {
    launch_scan(ctx, n);
    int64_t total = 0;
    ND_HIP(hipMemcpyAsync(&total, ctx->offsets.as<int64_t>() + n, sizeof(int64_t),
                          hipMemcpyDeviceToHost, ctx->stream));
    ND_HIP(hipStreamSynchronize(ctx->stream));
    return total;
}

namespace {

// leaf-range slices per wave group of the cooperative walk
inline int32_t trav_slices(int32_t nL) { return nL >= 64 ? 8 : 1; }

template <bool EMIT>
void launch_traverse_w(hipStream_t st, int32_t apl, const nd_node *nodes, const float4 *rec,
                       const float *xyz, const int32_t *idx, int32_t nL, float r, float r2, int32_t pred,
                       int64_t *cnt, const int64_t *off, nd_pair *out, int32_t *status, int64_t cap) {
    const int32_t ns = trav_slices(nL);
    const int64_t waves = ((int64_t)nL * apl + (apl <= 32 ? 31 : 63)) / (apl <= 32 ? 32 : 64);
    const unsigned g = grid_for(waves * ns * 64);
#define ND_TRAV_W(A)                                                                              \
    case A:                                                                                       \
        if (pred == ND_PRED_SCALAR)                                                               \
            k_leaf_traverse_w<EMIT, A, ND_PRED_SCALAR><<<g, kBlock, 0, st>>>(nodes, rec, xyz, idx, nL, \
                                                                            r, r2, ns, cnt, off, out, status, cap); \
        else                                                                                      \
            k_leaf_traverse_w<EMIT, A, ND_PRED_SIMD><<<g, kBlock, 0, st>>>(nodes, rec, xyz, idx, nL,   \
                                                                          r, r2, ns, cnt, off, out, status, cap);   \
        break;
    switch (apl) {
        ND_TRAV_W(1) ND_TRAV_W(2) ND_TRAV_W(4) ND_TRAV_W(8) ND_TRAV_W(16) ND_TRAV_W(32) ND_TRAV_W(64)
        default: throw Error(ND_ERR_ARG, "cooperative traversal: leaf size must divide 64");
    }
#undef ND_TRAV_W
}

}  // namespace

int64_t leaf_traverse(nd_ctx *ctx, const nd_node *nodes, const float *xyz_sorted,
                      const int32_t *idx_sorted, int32_t n, int32_t apl, float r, int32_t pred,
                      nd_pair *out, int64_t cap, bool own_tree, int32_t *status_dev, int64_t *count_dev) {
    const int32_t nL = n / apl;
    const float r2 = r * r;  // (neighbor_distance)^2 in Float32
    const bool async = count_dev != nullptr;
    int32_t *st = async ? status_dev : ctx->status.as<int32_t>();
    const bool wave_walk = own_tree && apl <= 64 && 64 % apl == 0;
    // one depth-first walk per wave and leaf-range slice (same pairs, same
    // order: counts per (leaf, slice), scanned leaf-major), or the rope walk
    // per query leaf
    const int64_t nc = wave_walk ? (int64_t)nL * trav_slices(nL) : (int64_t)nL;
    ctx->counts.reserve(sizeof(int64_t) * (size_t)(nc + 1));
    ctx->offsets.reserve(sizeof(int64_t) * (size_t)(nc + 1));
    int64_t *cnt = ctx->counts.as<int64_t>();
    const int64_t *off = ctx->offsets.as<int64_t>();
    float4 *rec = nullptr;
    auto pass = [&](bool emit) {
        if (wave_walk) {
            if (emit)
                launch_traverse_w<true>(ctx->stream, apl, nodes, rec, xyz_sorted, idx_sorted, nL, r, r2, pred,
                                        cnt, off, out, st, cap);
            else
                launch_traverse_w<false>(ctx->stream, apl, nodes, rec, xyz_sorted, idx_sorted, nL, r, r2, pred,
                                         cnt, off, nullptr, st, cap);
        } else if (apl == 4) {
            if (emit)
                k_leaf_traverse4<true><<<grid_for(nL), kBlock, 0, ctx->stream>>>(
                    nodes, xyz_sorted, idx_sorted, nL, r, r2, pred, cnt, off, out, st, cap);
            else
                k_leaf_traverse4<false><<<grid_for(nL), kBlock, 0, ctx->stream>>>(
                    nodes, xyz_sorted, idx_sorted, nL, r, r2, pred, cnt, off, nullptr, st, cap);
        } else {
            if (emit)
                k_leaf_traverse<true><<<grid_for(nL), kBlock, 0, ctx->stream>>>(
                    nodes, xyz_sorted, idx_sorted, nL, apl, r, r2, pred, cnt, off, out, st, cap);
            else
                k_leaf_traverse<false><<<grid_for(nL), kBlock, 0, ctx->stream>>>(
                    nodes, xyz_sorted, idx_sorted, nL, apl, r, r2, pred, cnt, off, nullptr, st, cap);
        }
        ND_HIP(hipGetLastError());
    };
    if (wave_walk) {
        ctx->trav_rec.reserve(8 * sizeof(float4) * (size_t)(nL - 1));
        rec = ctx->trav_rec.as<float4>();
        k_trav_records<<<grid_for(nL - 1), kBlock, 0, ctx->stream>>>(nodes, nL, rec);
    }
    pass(false);
    if (async) {
        // stream-ordered: the emitting pass checks the total against cap itself
        launch_scan(ctx, nc);
        if (out) pass(true);
        k_copy_total<<<1, 64, 0, ctx->stream>>>(off + nc, count_dev);
        ND_HIP(hipGetLastError());
        return -1;
    }
    const int64_t total = scan_counts(ctx, nc);
    if (nd::read_status(ctx) & nd::kStatusTree)
        throw nd::Error(ND_ERR_STATE, "leafneighbor_traverse: malformed tree (rope cycle or bad index)");
    if (total > cap || out == nullptr) return total;
    pass(true);
    ND_HIP(hipStreamSynchronize(ctx->stream));
    return total;
}

int64_t point_traverse(nd_ctx *ctx, const nd_node *nodes, const float *xyz_sorted, const int32_t *idx_sorted,
                       int32_t n, int32_t apl, float r, int32_t mode, nd_pair *out, int64_t cap) {
    const int32_t nL = n / apl;
    const float r2 = r * r;  // (neighbor_distance)^2 in Float32
    const bool root = mode == ND_QUERY_ROOT;
    const int64_t nq = root ? n : n - 1;
    ctx->counts.reserve(sizeof(int64_t) * (size_t)(nq + 1));
    ctx->offsets.reserve(sizeof(int64_t) * (size_t)(nq + 1));
    int64_t *cnt = ctx->counts.as<int64_t>();
    const int64_t *off = ctx->offsets.as<int64_t>();
    int32_t *st = ctx->status.as<int32_t>();
    auto pass = [&](bool emit) {
        const unsigned g = grid_for(nq);
        if (root) {
            if (emit)
                k_point_traverse<true, true><<<g, kBlock, 0, ctx->stream>>>(nodes, xyz_sorted, idx_sorted, n, nL,
                                                                           apl, r2, cnt, off, out, st, cap);
            else
                k_point_traverse<false, true><<<g, kBlock, 0, ctx->stream>>>(nodes, xyz_sorted, idx_sorted, n, nL,
                                                                            apl, r2, cnt, off, nullptr, st, cap);
        } else {
            if (emit)
                k_point_traverse<true, false><<<g, kBlock, 0, ctx->stream>>>(nodes, xyz_sorted, idx_sorted, n, nL,
                                                                            apl, r2, cnt, off, out, st, cap);
            else
                k_point_traverse<false, false><<<g, kBlock, 0, ctx->stream>>>(nodes, xyz_sorted, idx_sorted, n, nL,
                                                                             apl, r2, cnt, off, nullptr, st, cap);
        }
        ND_HIP(hipGetLastError());
    };
    pass(false);
    const int64_t total = scan_counts(ctx, nq);
    if (nd::read_status(ctx) & nd::kStatusTree)
        throw nd::Error(ND_ERR_STATE, "neighbor_traverse: malformed tree (rope cycle or bad index)");
    if (total > cap || out == nullptr) return total;
    pass(true);
    ND_HIP(hipStreamSynchronize(ctx->stream));
    return total;
}

namespace {

inline size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

// counts of the traversal: per (leaf, slice) for leaf sizes dividing 64 (the
// one-walk-per-wave traversal of a library-built tree), else per query leaf
int64_t list_count_slots(int32_t n, int32_t apl) {
    const int32_t nL = n / apl;
    return (apl <= 64 && 64 % apl == 0) ? (int64_t)nL * trav_slices(nL) : (int64_t)nL;
}

}  // namespace

size_t list_work_bytes(int32_t n, int32_t apl) {
    const int32_t nL = n / apl;
    const size_t nn = (size_t)n, ntiles = radix_scratch_blocks(n);
    const int64_t nc = list_count_slots(n, apl);
    size_t b = 0;
    b += 4 * align256(sizeof(int32_t) * nn);                           // codes, perm, keys_alt, vals_alt
    b += align256(sizeof(int32_t) * (256 * ntiles + 256));             // 8-bit digit histogram
    b += align256(sizeof(int32_t) * lbvh_scratch_ints(nL));            // CAS slots, ranges, chunks
    b += align256(sizeof(float) * 3 * nn) + 256;                       // sorted positions, box
    b += align256(sizeof(nd_node) * (size_t)(2 * nL - 1));             // tree
    b += 2 * align256(sizeof(int64_t) * (size_t)(nc + 1));             // counts, offsets
    b += align256(sizeof(int64_t) * (size_t)(scan_tmp_entries(nc + 1) + 1));
    b += align256(8 * sizeof(float4) * (size_t)(nL > 1 ? nL - 1 : 1));  // four-wide records
    return b;
}

ListWork carve_list_work(void *base, int32_t n, int32_t apl) {
    const int32_t nL = n / apl;
    const size_t nn = (size_t)n, ntiles = radix_scratch_blocks(n);
    const int64_t nc = list_count_slots(n, apl);
    char *p = static_cast<char *>(base);
    auto take = [&](size_t bytes) {
        char *q = p;
        p += align256(bytes);
        return q;
    };
    ListWork w;
    w.codes = reinterpret_cast<int32_t *>(take(sizeof(int32_t) * nn));
    w.perm = reinterpret_cast<int32_t *>(take(sizeof(int32_t) * nn));
    w.keys_alt = reinterpret_cast<int32_t *>(take(sizeof(int32_t) * nn));
    w.vals_alt = reinterpret_cast<int32_t *>(take(sizeof(int32_t) * nn));
    w.hist = reinterpret_cast<int32_t *>(take(sizeof(int32_t) * (256 * ntiles + 256)));
    w.store = reinterpret_cast<int32_t *>(take(sizeof(int32_t) * lbvh_scratch_ints(nL)));
    w.xyz_sorted = reinterpret_cast<float *>(take(sizeof(float) * 3 * nn));
    w.box = reinterpret_cast<float *>(take(256));
    w.nodes = reinterpret_cast<nd_node *>(take(sizeof(nd_node) * (size_t)(2 * nL - 1)));
    w.counts = reinterpret_cast<int64_t *>(take(sizeof(int64_t) * (size_t)(nc + 1)));
    w.offsets = reinterpret_cast<int64_t *>(take(sizeof(int64_t) * (size_t)(nc + 1)));
    w.scan_tmp = reinterpret_cast<int64_t *>(take(sizeof(int64_t) * (size_t)(scan_tmp_entries(nc + 1) + 1)));
    w.rec = reinterpret_cast<float4 *>(take(8 * sizeof(float4) * (size_t)(nL > 1 ? nL - 1 : 1)));
    return w;
}

void launch_list_async(hipStream_t st, const ListWork &w, const float *d_xyz, int32_t n, int32_t apl,
                       float r, int32_t pred, int32_t morton, nd_pair *out, int64_t cap, int64_t *d_count,
                       int32_t *d_status) {
    const int32_t nL = n / apl;
    launch_fill_i32(st, d_status, 1, 0);
    launch_morton_box(st, w.box, d_xyz, n, morton, w.codes, w.perm, d_status);
    // every key bit (no host round trip to find the varying digits); the
    // sort's buffers are views of the work block, sized so nothing grows
    SortBufs sb;
    const size_t ntiles = radix_scratch_blocks(n);
    sb.keys_alt = DevBuf{w.keys_alt, sizeof(int32_t) * (size_t)n};
    sb.vals_alt = DevBuf{w.vals_alt, sizeof(int32_t) * (size_t)n};
    sb.hist = DevBuf{w.hist, sizeof(int32_t) * (256 * ntiles + 256)};
    radix_sort_pairs(st, sb, w.codes, w.perm, n, 0, 32);
    launch_gather_xyz(st, d_xyz, w.perm, n, w.xyz_sorted);
    launch_lbvh(st, w.xyz_sorted, 3, w.codes, nL, apl, r, 1, w.nodes, w.store, d_status);
    // traversal: count, scan, emit (the emitting pass checks the total
    // against cap itself), total to *d_count
    const float r2 = r * r;
    const int64_t nc = list_count_slots(n, apl);
    const bool wave_walk = apl <= 64 && 64 % apl == 0;
    auto pass = [&](bool emit) {
        if (wave_walk) {
            if (emit)
                launch_traverse_w<true>(st, apl, w.nodes, w.rec, w.xyz_sorted, w.perm, nL, r, r2, pred, w.counts,
                                        w.offsets, out, d_status, cap);
            else
                launch_traverse_w<false>(st, apl, w.nodes, w.rec, w.xyz_sorted, w.perm, nL, r, r2, pred, w.counts,
                                         w.offsets, nullptr, d_status, cap);
        } else if (apl == 4) {
            if (emit)
                k_leaf_traverse4<true><<<grid_for(nL), kBlock, 0, st>>>(w.nodes, w.xyz_sorted, w.perm, nL, r, r2,
                                                                        pred, w.counts, w.offsets, out, d_status, cap);
            else
                k_leaf_traverse4<false><<<grid_for(nL), kBlock, 0, st>>>(w.nodes, w.xyz_sorted, w.perm, nL, r, r2,
                                                                         pred, w.counts, w.offsets, nullptr, d_status, cap);
        } else {
            if (emit)
                k_leaf_traverse<true><<<grid_for(nL), kBlock, 0, st>>>(w.nodes, w.xyz_sorted, w.perm, nL, apl, r,
                                                                       r2, pred, w.counts, w.offsets, out, d_status, cap);
            else
                k_leaf_traverse<false><<<grid_for(nL), kBlock, 0, st>>>(w.nodes, w.xyz_sorted, w.perm, nL, apl, r,
                                                                        r2, pred, w.counts, w.offsets, nullptr, d_status, cap);
        }
        ND_HIP(hipGetLastError());
    };
    if (wave_walk) {
        k_trav_records<<<grid_for(nL - 1), kBlock, 0, st>>>(w.nodes, nL, w.rec);
        ND_HIP(hipGetLastError());
    }
    pass(false);
    launch_scan_buf(st, w.counts, nc, w.offsets, w.scan_tmp);
    if (out) pass(true);
    k_copy_total<<<1, 64, 0, st>>>(w.offsets + nc, d_count);
    ND_HIP(hipGetLastError());
}

}  // namespace nd

// ---- static scratch of the allocation-free one-call list ----
namespace {

// Work buffers of nd_leafbuild_traverse_bvh_noalloc: one array per GPU, held
// by the library until the process ends.
constexpr size_t kScratchBytes = (size_t)32 << 20;
__device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
void *g_scratch[kMaxDevices] = {};
std::mutex g_scratch_mu;

}  // namespace

namespace nd {

void *static_scratch(int device) {
    if (device < 0 || device >= kMaxDevices) throw Error(ND_ERR_ARG, "device index beyond the scratch table");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch[device]) {
        // the device's own copy of the array (host code never uses its name
        // as an address); the caller has made `device` current
        void *p = nullptr;
        ND_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        g_scratch[device] = p;
    }
    return g_scratch[device];
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

void check_query(int32_t mode) {
    ND_REQUIRE(mode == ND_QUERY_OWN_LEAF || mode == ND_QUERY_ROOT, "unknown query mode");
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
    ctx->tree_nodes = ctx->nodes.as<nd_node>();
    ctx->tree_perm = ctx->perm.as<int32_t>();
    ctx->tree_codes = ctx->codes_sorted.as<int32_t>();
    ctx->tree_store = ctx->store.as<int32_t>();
    const int64_t total = nd::leaf_traverse(ctx, ctx->nodes.as<nd_node>(), ctx->xyz_sorted.as<float>(),
                                            ctx->perm.as<int32_t>(), n, apl, chk.neighbor_distance,
                                            predicate, d_out, cap, true);
    *count = total;
    if (total > cap) throw nd::Error(ND_ERR_CAPACITY, "pair buffer too small; *count = required");
    ND_API_END(ctx)
}

int nd_leafbuild_traverse_bvh_async(nd_ctx *ctx, const float *d_xyz, const nd_bvh_spec *spec,
                                    int32_t predicate, int32_t morton_mode, nd_pair *d_out, int64_t cap,
                                    int64_t *d_count, int32_t *d_status) {
    ND_API_BEGIN
    ND_REQUIRE(ctx && d_xyz && d_count && d_status, "nd_leafbuild_traverse_bvh_async: null argument");
    nd_bvh_spec chk;
    check_spec(spec, chk);
    check_pred(predicate);
    ND_REQUIRE(morton_mode == ND_MORTON_MASKED || morton_mode == ND_MORTON_INTERLEAVE,
               "unknown key mode");
    ND_HIP(hipSetDevice(ctx->device));
    const int32_t n = chk.atom_count, apl = chk.atomsperleaf;
    // this form's own work block: no other entry point resizes it, so a graph
    // captured after a first call of this size stays valid
    ctx->async_ws.reserve(nd::list_work_bytes(n, apl));
    const nd::ListWork w = nd::carve_list_work(ctx->async_ws.p, n, apl);
    nd::launch_list_async(ctx->stream, w, d_xyz, n, apl, chk.neighbor_distance, predicate, morton_mode, d_out,
                          cap, d_count, d_status);
    ctx->has_tree = true;
    ctx->tree_n = n;
    ctx->tree_apl = apl;
    ctx->tree_nodes = w.nodes;
    ctx->tree_perm = w.perm;
    ctx->tree_codes = w.codes;
    ctx->tree_store = w.store;
    ND_API_END(ctx)
}

int nd_leafbuild_traverse_bvh_noalloc(nd_ctx *ctx, const float *d_xyz, const nd_bvh_spec *spec,
                                      int32_t predicate, int32_t morton_mode, nd_pair *d_out, int64_t cap,
                                      int64_t *d_count, int32_t *d_status) {
    ND_API_BEGIN
    ND_REQUIRE(ctx && d_xyz && d_count && d_status, "nd_leafbuild_traverse_bvh_noalloc: null argument");
    nd_bvh_spec chk;
    check_spec(spec, chk);
    check_pred(predicate);
    ND_REQUIRE(morton_mode == ND_MORTON_MASKED || morton_mode == ND_MORTON_INTERLEAVE,
               "unknown key mode");
    const int32_t n = chk.atom_count, apl = chk.atomsperleaf;
    if (nd::list_work_bytes(n, apl) > kScratchBytes)
        throw nd::Error(ND_ERR_CAPACITY, "nd_leafbuild_traverse_bvh_noalloc: work buffers for " + std::to_string(n) +
                                             " atoms (" + std::to_string(nd::list_work_bytes(n, apl)) +
                                             " bytes) exceed the static scratch of " +
                                             std::to_string(kScratchBytes) + " bytes");
    ND_HIP(hipSetDevice(ctx->device));
    const nd::ListWork w = nd::carve_list_work(nd::static_scratch(ctx->device), n, apl);
    nd::launch_list_async(ctx->stream, w, d_xyz, n, apl, chk.neighbor_distance, predicate, morton_mode, d_out,
                          cap, d_count, d_status);
    ctx->has_tree = true;
    ctx->tree_n = n;
    ctx->tree_apl = apl;
    ctx->tree_nodes = w.nodes;
    ctx->tree_perm = w.perm;
    ctx->tree_codes = w.codes;
    ctx->tree_store = w.store;
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
        ND_HIP(hipMemcpyAsync(h_nodes, ctx->tree_nodes, sizeof(nd_node) * (2 * nL - 1),
                              hipMemcpyDeviceToHost, ctx->stream));
    if (h_perm)
        ND_HIP(hipMemcpyAsync(h_perm, ctx->tree_perm, sizeof(int32_t) * n, hipMemcpyDeviceToHost,
                              ctx->stream));
    if (h_codes_sorted)
        ND_HIP(hipMemcpyAsync(h_codes_sorted, ctx->tree_codes, sizeof(int32_t) * n,
                              hipMemcpyDeviceToHost, ctx->stream));
    ND_HIP(hipStreamSynchronize(ctx->stream));
    ND_API_END(ctx)
}

int nd_neighbor_traverse(nd_ctx *ctx, const nd_node *d_nodes, const float *d_xyz_sorted,
                         const int32_t *d_index_sorted, const nd_bvh_spec *spec, int32_t query_mode,
                         nd_pair *d_out, int64_t cap, int64_t *count) {
    ND_API_BEGIN
    ND_REQUIRE(ctx && d_nodes && d_xyz_sorted && d_index_sorted && count, "nd_neighbor_traverse: null argument");
    nd_bvh_spec chk;
    check_spec(spec, chk);
    check_query(query_mode);
    ND_HIP(hipSetDevice(ctx->device));
    const int64_t total = nd::point_traverse(ctx, d_nodes, d_xyz_sorted, d_index_sorted, chk.atom_count,
                                             chk.atomsperleaf, chk.neighbor_distance, query_mode, d_out, cap);
    *count = total;
    if (total > cap) throw nd::Error(ND_ERR_CAPACITY, "pair buffer too small; *count = required");
    ND_API_END(ctx)
}

int nd_build_traverse_bvh(nd_ctx *ctx, const float *d_xyz, const nd_bvh_spec *spec, int32_t query_mode,
                          int32_t morton_mode, nd_pair *d_out, int64_t cap, int64_t *count) {
    ND_API_BEGIN
    ND_REQUIRE(ctx && d_xyz && count, "nd_build_traverse_bvh: null argument");
    nd_bvh_spec chk;
    check_spec(spec, chk);
    check_query(query_mode);
    ND_REQUIRE(morton_mode == ND_MORTON_MASKED || morton_mode == ND_MORTON_INTERLEAVE, "unknown key mode");
    // the build halves its radius argument for the leaf half-width; 2r is
    // exact in Float32 below this bound, so the leaves get +- r exactly
    ND_REQUIRE(chk.neighbor_distance <= 0.5f * FLT_MAX, "nd_build_traverse_bvh: neighbor_distance too large");
    ND_HIP(hipSetDevice(ctx->device));
    const int32_t n = chk.atom_count, apl = chk.atomsperleaf, nL = chk.leaves_count;
    ctx->has_tree = false;
    ctx->codes_sorted.reserve(sizeof(int32_t) * (size_t)n);
    ctx->perm.reserve(sizeof(int32_t) * (size_t)n);
    ctx->xyz_sorted.reserve(sizeof(float) * 3 * (size_t)n);
    ctx->nodes.reserve(sizeof(nd_node) * (size_t)(2 * nL - 1));
    ctx->store.reserve(sizeof(int32_t) * nd::lbvh_scratch_ints(nL));
    // mortoncodes! + sortperm + permute! (BVHTraverse.jl:511-516)
    nd::launch_morton(ctx, d_xyz, n, morton_mode, nullptr, nullptr, ctx->codes_sorted.as<int32_t>(),
                      ctx->perm.as<int32_t>(), ctx->status.as<int32_t>());
    if (nd::read_status(ctx) & nd::kStatusDomain)
        throw nd::Error(ND_ERR_DOMAIN,
                        "mortoncodes!: coordinate outside the Int32 quantisation domain "
                        "(reference throws InexactError)");
    nd::radix_sort_pairs_auto(ctx, ctx->codes_sorted.as<int32_t>(), ctx->perm.as<int32_t>(), n);
    nd::launch_gather_xyz(ctx->stream, d_xyz, ctx->perm.as<int32_t>(), n, ctx->xyz_sorted.as<float>());
    // cluster_primitives (+- r leaves) + bounding_volume_hierarchy!; TreeData
    // overwrites the root box (BVHTraverse.jl:856-858), the root-started form's
    // build kernel keeps it (NaiveKA.jl:301-302)
    nd::launch_lbvh(ctx->stream, ctx->xyz_sorted.as<float>(), 3, ctx->codes_sorted.as<int32_t>(), nL, apl,
                    2.0f * chk.neighbor_distance, query_mode == ND_QUERY_OWN_LEAF ? 1 : 0,
                    ctx->nodes.as<nd_node>(), ctx->store.as<int32_t>(), ctx->status.as<int32_t>());
    ctx->has_tree = true;
    ctx->tree_n = n;
    ctx->tree_apl = apl;
    ctx->tree_nodes = ctx->nodes.as<nd_node>();
    ctx->tree_perm = ctx->perm.as<int32_t>();
    ctx->tree_codes = ctx->codes_sorted.as<int32_t>();
    ctx->tree_store = ctx->store.as<int32_t>();
    const int64_t total = nd::point_traverse(ctx, ctx->nodes.as<nd_node>(), ctx->xyz_sorted.as<float>(),
                                             ctx->perm.as<int32_t>(), n, apl, chk.neighbor_distance, query_mode,
                                             d_out, cap);
    *count = total;
    if (total > cap) throw nd::Error(ND_ERR_CAPACITY, "pair buffer too small; *count = required");
    ND_API_END(ctx)
}

int nd_tree_store_export(nd_ctx *ctx, int32_t *h_store) {
    ND_API_BEGIN
    ND_REQUIRE(ctx && h_store, "nd_tree_store_export: null argument");
    if (!ctx->has_tree) throw nd::Error(ND_ERR_STATE, "nd_tree_store_export: no tree has been built");
    const size_t nL = (size_t)ctx->tree_n / (size_t)ctx->tree_apl;
    ND_HIP(hipMemcpyAsync(h_store, ctx->tree_store, sizeof(int32_t) * (nL - 1), hipMemcpyDeviceToHost,
                          ctx->stream));
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
    if (total > 0 && !d_out) throw nd::Error(ND_ERR_ARG, "nd_unique_threshold_pairs: null output buffer");
    k_bruteforce<true><<<nd::grid_for(n), nd::kBlock, 0, ctx->stream>>>(
        d_xyz, n, threshold, nullptr, ctx->offsets.as<int64_t>(), d_out);
    ND_HIP(hipGetLastError());
    ND_HIP(hipStreamSynchronize(ctx->stream));
    ND_API_END(ctx)
}

}  // extern "C"
