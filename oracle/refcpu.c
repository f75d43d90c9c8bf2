/*
 * refcpu.c -- CPU restatement of NaiveDynamics.jl's LBVH neighbour-search path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the oracle the HIP library is checked
 * against and the threaded CPU baseline that bench.py times.  Nothing in the
 * product library links, loads or calls it.
 *
 * Every function restates the algorithm of the Julia reference (file:line
 * citations below, relative to the reference repository root).  Arithmetic is
 * plain IEEE-754 binary32 in the same operation order as the Julia code; build
 * with -ffp-contract=off and without -ffast-math so no multiply-add is fused.
 *
 * Index conventions: positions, codes and permutations are 0-based.  Tree nodes
 * keep the reference numbering: node k (1-based; leaves 1..nL, internal nodes
 * nL+1..2nL-1, root nL+1, 0 = sentinel) lives in array slot k-1, and the
 * left/skip fields hold 1-based node numbers exactly as GridKey does
 * (src/Neighbors/BVHTraverse.jl:102-109).  Pair records hold 0-based original
 * atom indices (the Julia reference is 1-based; add 1 to compare).
 */
#define _GNU_SOURCE
#include <math.h>
#include <pthread.h>
#include <stdatomic.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "refcpu.h"

/* ------------------------------------------------------------------------- */
/* A4/A5: masked "Morton" code.  BVHTraverse.jl:184 (binwidth) and 237-288.   */
/* ------------------------------------------------------------------------- */

/* Float32(1/1023): the quotient is formed in Float64 and rounded once. */
#define REF_BINWIDTH ((float)(1.0 / 1023.0))

static const int32_t REF_MAGIC[3] = {0x09249249, 0x12492492, 0x24924924};

float ref_binwidth(void) { return REF_BINWIDTH; }

/* round(Int32, v, RoundDown) throws InexactError unless floor(v) fits Int32;
 * for binary32 that is exactly -2^31 <= v < 2^31 (NaN fails both tests). */
static inline int ref_floor_i32(float v, int32_t *q) {
    if (!(v >= -2147483648.0f && v < 2147483648.0f)) return 0;
    *q = (int32_t)floorf(v);
    return 1;
}

/* mortoncodes! (BVHTraverse.jl:257-288): yin = q & M, yang = q | ~M and
 * yin & yang == q & M, OR-ed over the three axes. */
int ref_morton_masked(const float *xyz, int64_t n, int32_t *codes) {
    for (int64_t i = 0; i < n; ++i) {
        int32_t code = 0;
        for (int d = 0; d < 3; ++d) {
            float v = xyz[3 * i + d] / REF_BINWIDTH;
            int32_t q;
            if (!ref_floor_i32(v, &q)) return REF_ERR_DOMAIN;
            int32_t yin = q & REF_MAGIC[d];
            int32_t yang = q | ~REF_MAGIC[d];
            code |= (yin & yang);
        }
        codes[i] = code;
    }
    return REF_OK;
}

/* Performance-mode key (not in the reference): a true 30-bit Morton
 * interleave of 10-bit cell coordinates q_d = clamp(floor((x_d-lo_d)*s_d)),
 * s_d = fl32(1024 / (hi_d - lo_d)).  Bit 3k+d of the code is bit k of q_d. */
static inline uint32_t ref_spread10(uint32_t v) {
    uint32_t r = 0;
    for (int k = 0; k < 10; ++k) r |= ((v >> k) & 1u) << (3 * k);
    return r;
}

int ref_morton_interleave(const float *xyz, int64_t n, const float lo[3],
                          const float scale[3], int32_t *codes) {
    for (int64_t i = 0; i < n; ++i) {
        uint32_t code = 0;
        for (int d = 0; d < 3; ++d) {
            float x = xyz[3 * i + d];
            if (!isfinite(x)) return REF_ERR_DOMAIN;
            float t = (x - lo[d]) * scale[d];
            int32_t q = (t <= 0.0f) ? 0 : (t >= 1023.0f ? 1023 : (int32_t)floorf(t));
            code |= ref_spread10((uint32_t)q) << d;
        }
        codes[i] = (int32_t)code;
    }
    return REF_OK;
}

/* Hilbert index of a 10-bit cell triple: Skilling's transform (J. Skilling,  */
/* "Programming the Hilbert curve", AIP Conf. Proc. 707, 2004), written as   */
/* the paper's loops over the axes: undo excess work from the top bit down,  */
/* Gray encode, then read the transposed coordinates' bits MSB first with    */
/* axis 0 leading each triple.                                               */
uint32_t ref_hilbert_cell(uint32_t cx, uint32_t cy, uint32_t cz) {
    uint32_t X[3] = {cx & 1023u, cy & 1023u, cz & 1023u};
    const uint32_t M = 1u << 9;
    for (uint32_t Q = M; Q > 1u; Q >>= 1) {
        const uint32_t P = Q - 1u;
        for (int i = 0; i < 3; ++i) {
            if (X[i] & Q) {
                X[0] ^= P;                 /* invert */
            } else {
                uint32_t t = (X[0] ^ X[i]) & P;  /* exchange */
                X[0] ^= t;
                X[i] ^= t;
            }
        }
    }
    for (int i = 1; i < 3; ++i) X[i] ^= X[i - 1];
    uint32_t t = 0;
    for (uint32_t Q = M; Q > 1u; Q >>= 1)
        if (X[2] & Q) t ^= Q - 1u;
    for (int i = 0; i < 3; ++i) X[i] ^= t;
    uint32_t key = 0;
    for (int b = 9; b >= 0; --b)
        for (int i = 0; i < 3; ++i) key = (key << 1) | ((X[i] >> b) & 1u);
    return key;
}

void ref_hilbert_cells(const uint32_t *cells, int64_t n, int32_t *codes) {
    for (int64_t i = 0; i < n; ++i)
        codes[i] = (int32_t)ref_hilbert_cell(cells[3 * i], cells[3 * i + 1], cells[3 * i + 2]);
}

int ref_morton_hilbert(const float *xyz, int64_t n, const float lo[3], const float scale[3],
                       int32_t *codes) {
    for (int64_t i = 0; i < n; ++i) {
        uint32_t c[3];
        for (int d = 0; d < 3; ++d) {
            float x = xyz[3 * i + d];
            if (!isfinite(x)) return REF_ERR_DOMAIN;
            float t = (x - lo[d]) * scale[d];
            c[d] = (t <= 0.0f) ? 0u : (t >= 1023.0f ? 1023u : (uint32_t)floorf(t));
        }
        codes[i] = (int32_t)ref_hilbert_cell(c[0], c[1], c[2]);
    }
    return REF_OK;
}

/* ------------------------------------------------------------------------- */
/* A6: sortperm on signed Int32 codes (BVHTraverse.jl:570-573).  Julia's      */
/* sortperm is stable, so ties keep ascending original index.  Stable LSD    */
/* radix sort on the sign-flipped key, 8-bit digits.                         */
/* ------------------------------------------------------------------------- */
void ref_sortperm(const int32_t *codes, int64_t n, int32_t *perm) {
    uint32_t *k0 = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)(n ? n : 1));
    uint32_t *k1 = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)(n ? n : 1));
    int32_t *p1 = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n ? n : 1));
    for (int64_t i = 0; i < n; ++i) {
        k0[i] = (uint32_t)codes[i] ^ 0x80000000u;
        perm[i] = (int32_t)i;
    }
    uint32_t *ks = k0, *kd = k1;
    int32_t *ps = perm, *pd = p1;
    for (int shift = 0; shift < 32; shift += 8) {
        int64_t cnt[257];
        memset(cnt, 0, sizeof(cnt));
        for (int64_t i = 0; i < n; ++i) cnt[((ks[i] >> shift) & 0xFFu) + 1]++;
        if (cnt[((ks[0] >> shift) & 0xFFu) + 1] == n) continue; /* constant digit */
        for (int b = 0; b < 256; ++b) cnt[b + 1] += cnt[b];
        for (int64_t i = 0; i < n; ++i) {
            uint32_t b = (ks[i] >> shift) & 0xFFu;
            int64_t o = cnt[b]++;
            kd[o] = ks[i];
            pd[o] = ps[i];
        }
        uint32_t *tk = ks; ks = kd; kd = tk;
        int32_t *tp = ps; ps = pd; pd = tp;
    }
    if (ps != perm) memcpy(perm, ps, sizeof(int32_t) * (size_t)n);
    free(k0); free(k1); free(p1);
}

/* ------------------------------------------------------------------------- */
/* Julia min/max for Float32 (Base: diff = x-y, pick by signbit(diff)):       */
/* ordinary ordering with -0 < +0; NaN inputs give NaN.                       */
/* ------------------------------------------------------------------------- */
static inline float jl_min(float x, float y) {
    if (isnan(x) || isnan(y)) return x - y;
    if (x < y) return x;
    if (y < x) return y;
    return signbit(x) ? x : y;
}
static inline float jl_max(float x, float y) {
    if (isnan(x) || isnan(y)) return x - y;
    if (x > y) return x;
    if (y > x) return y;
    return signbit(x) ? y : x;
}

/* ------------------------------------------------------------------------- */
/* A7: leafcluster_primitives (BVHTraverse.jl:376-410): leaf l covers sorted  */
/* atoms l*apl .. l*apl+apl-1, box = [min(x - r/2), max(x + r/2)].            */
/* A17: cluster_primitives (BVHTraverse.jl:341-374): the same with +- r.      */
/* ------------------------------------------------------------------------- */
static void leaf_boxes_h(const float *xyz_sorted, int32_t nL, int32_t apl, float h,
                         ref_node *nodes) {
    for (int32_t l = 0; l < nL; ++l) {
        const float *p0 = xyz_sorted + 3 * (int64_t)l * apl;
        float mn[3], mx[3];
        for (int d = 0; d < 3; ++d) { mn[d] = p0[d] - h; mx[d] = p0[d] + h; }
        for (int32_t a = 0; a < apl; ++a) {
            const float *p = xyz_sorted + 3 * ((int64_t)l * apl + a);
            for (int d = 0; d < 3; ++d) {
                mn[d] = jl_min(mn[d], p[d] - h);
                mx[d] = jl_max(mx[d], p[d] + h);
            }
        }
        ref_node *k = &nodes[l];
        for (int d = 0; d < 3; ++d) { k->min[d] = mn[d]; k->max[d] = mx[d]; }
        k->left = 0;
        k->skip = 0;
    }
}

void ref_leaf_boxes(const float *xyz_sorted, int32_t nL, int32_t apl, float r,
                    ref_node *nodes) {
    leaf_boxes_h(xyz_sorted, nL, apl, r / 2.0f /* neighbor_distance/2 in Float32 */, nodes);
}

void ref_cluster_boxes(const float *xyz_sorted, int32_t nL, int32_t apl, float r,
                       ref_node *nodes) {
    leaf_boxes_h(xyz_sorted, nL, apl, r, nodes);
}

/* ------------------------------------------------------------------------- */
/* A8: exptdelta (BVHTraverse.jl:676-696), 1-based leaf i.                    */
/* ------------------------------------------------------------------------- */
static inline int32_t ref_delta(int32_t i, const int32_t *codes_sorted, int32_t nL,
                                int32_t apl) {
    if (i >= nL || i < 1) return INT32_MAX;
    int32_t x = codes_sorted[(int64_t)(i - 1) * apl] ^ codes_sorted[(int64_t)i * apl];
    /* x + (x==0)*(typemin + (i xor (i+1))) - 1, with Int32 wraparound */
    uint32_t t = (uint32_t)x;
    if (x == 0) t += (uint32_t)INT32_MIN + (uint32_t)(i ^ (i + 1));
    t -= 1u;
    return (int32_t)t;
}

int32_t ref_delta_public(int32_t i, const int32_t *codes_sorted, int32_t nL, int32_t apl) {
    return ref_delta(i, codes_sorted, nL, apl);
}

/* ------------------------------------------------------------------------- */
/* A9/A10: bvh_interior! (BVHTraverse.jl:704-846) and                         */
/* bounding_volume_hierarchy! (BVHTraverse.jl:849-860).  Apetrei/ArborX       */
/* bottom-up build: one walker per leaf, the first walker to reach an         */
/* internal node stops, the second merges its sibling's box and climbs.      */
/* ------------------------------------------------------------------------- */
typedef struct {
    ref_node *keys;
    _Atomic int32_t *store;
    const int32_t *codes;
    int32_t nL, apl;
} build_ctx;

#define NODE(k) (c->keys[(k) - 1])

static void ref_bvh_interior(build_ctx *c, int32_t i) {
    const int32_t nL = c->nL, apl = c->apl;
    const int32_t *codes = c->codes;
    int32_t rangel = i, ranger = i;
    int32_t dell = ref_delta(rangel - 1, codes, nL, apl);
    int32_t delr = ref_delta(ranger, codes, nL, apl);
    float bmin[3], bmax[3];
    for (int d = 0; d < 3; ++d) { bmin[d] = NODE(i).min[d]; bmax[d] = NODE(i).max[d]; }

    /* leaf rope (BVHTraverse.jl:717-741) */
    if (i == nL) {
        NODE(i).skip = 0;
    } else {
        int32_t ir = i + 1;
        NODE(i).skip = (delr < ref_delta(ir, codes, nL, apl)) ? ir : ir + nL;
    }

    for (;;) {
        int32_t leftChild;
        if (delr < dell) { /* isLeftChild: the sibling lies to the right */
            leftChild = i;
            int32_t split = ranger;
            int32_t expected = 0;
            atomic_compare_exchange_strong(&c->store[split - 1], &expected, rangel);
            ranger = expected;
            if (ranger == 0) return; /* first to arrive: stop */
            delr = ref_delta(ranger, codes, nL, apl);
            int32_t rightChild = split + 1;
            if (rightChild != ranger) rightChild += nL;
            for (int d = 0; d < 3; ++d) {
                bmax[d] = jl_max(NODE(rightChild).max[d], bmax[d]);
                bmin[d] = jl_min(NODE(rightChild).min[d], bmin[d]);
            }
        } else { /* right child: the sibling lies to the left */
            int32_t split = rangel - 1;
            int32_t expected = 0;
            atomic_compare_exchange_strong(&c->store[split - 1], &expected, ranger);
            rangel = expected;
            if (rangel == 0) return;
            dell = ref_delta(rangel - 1, codes, nL, apl);
            leftChild = split;
            if (leftChild != rangel) leftChild += nL;
            for (int d = 0; d < 3; ++d) {
                bmax[d] = jl_max(NODE(leftChild).max[d], bmax[d]);
                bmin[d] = jl_min(NODE(leftChild).min[d], bmin[d]);
            }
        }
        int32_t q = (delr < dell) ? ranger : rangel;
        int32_t parent = q + nL;
        ref_node *P = &NODE(parent);
        P->left = leftChild;
        for (int d = 0; d < 3; ++d) { P->min[d] = bmin[d]; P->max[d] = bmax[d]; }
        if (ranger == nL) {
            P->skip = 0;
        } else {
            int32_t rr = ranger + 1;
            P->skip = (delr < ref_delta(rr, codes, nL, apl)) ? rr : rr + nL;
        }
        i = parent;
        if (i == 1 + nL) return; /* reached the root */
    }
}

typedef struct {
    build_ctx *c;
    int32_t lo, hi;
} build_task;

static void *build_worker(void *arg) {
    build_task *t = (build_task *)arg;
    for (int32_t i = t->lo; i <= t->hi; ++i) ref_bvh_interior(t->c, i);
    return NULL;
}

/* The build over leaf boxes of half-width h; store_out (nL-1 Int32, may be
 * NULL) receives the CAS slots as the build leaves them (TreeData's store,
 * BVHTraverse.jl:521-525, 542).  With one thread the walkers run for
 * i = 1..nL in order, so every slot holds the first leaf of its node's left
 * child (that child completes first). */
static int build_tree_h(const float *xyz_sorted, const int32_t *codes_sorted, int32_t N,
                        int32_t apl, float h, ref_node *nodes, int override_root, int nthreads,
                        int32_t *store_out) {
    if (apl < 1 || N % apl != 0) return REF_ERR_ARG;
    int32_t nL = N / apl;
    if (nL < 2) return REF_ERR_ARG;
    memset(nodes, 0, sizeof(ref_node) * (size_t)(2 * nL - 1));
    leaf_boxes_h(xyz_sorted, nL, apl, h, nodes);
    _Atomic int32_t *store = (_Atomic int32_t *)calloc((size_t)(nL - 1), sizeof(int32_t));
    build_ctx c = {nodes, store, codes_sorted, nL, apl};
    if (nthreads <= 1) {
        for (int32_t i = 1; i <= nL; ++i) ref_bvh_interior(&c, i);
    } else {
        /* @batch over 1:nL in contiguous chunks (BVHTraverse.jl:851) */
        pthread_t th[256];
        build_task tk[256];
        if (nthreads > 256) nthreads = 256;
        int32_t per = (nL + nthreads - 1) / nthreads;
        int nt = 0;
        for (int t = 0; t < nthreads; ++t) {
            int32_t lo = 1 + t * per, hi = lo + per - 1;
            if (hi > nL) hi = nL;
            if (lo > hi) break;
            tk[nt] = (build_task){&c, lo, hi};
            pthread_create(&th[nt], NULL, build_worker, &tk[nt]);
            ++nt;
        }
        for (int t = 0; t < nt; ++t) pthread_join(th[t], NULL);
    }
    if (store_out)
        for (int32_t k = 0; k < nL - 1; ++k) store_out[k] = atomic_load(&store[k]);
    free((void *)store);
    if (override_root) {
        /* bounding_volume_hierarchy! overwrites the root box (BVHTraverse.jl:856-858) */
        ref_node *root = &nodes[nL];
        for (int d = 0; d < 3; ++d) { root->min[d] = 0.0f; root->max[d] = 1.0f; }
    }
    return REF_OK;
}

int ref_build_tree(const float *xyz_sorted, const int32_t *codes_sorted, int32_t N,
                   int32_t apl, float r, ref_node *nodes, int override_root,
                   int nthreads) {
    return build_tree_h(xyz_sorted, codes_sorted, N, apl, r / 2.0f, nodes, override_root, nthreads,
                        NULL);
}

/* ------------------------------------------------------------------------- */
/* A12-A16: leafneighbor_traverse (BVHTraverse.jl:1236-1323) with the scalar */
/* pair test (onecluster/twocluster_proximitytest!, :1021-1055: s < r*r) or  */
/* the SIMD test (ext/NaiveSIMD.jl:507-608: sqrt(s) < r).                     */
/* ------------------------------------------------------------------------- */
typedef struct {
    ref_pair *v;
    int64_t n, cap;
} pairvec;

static inline void pv_push(pairvec *pv, int32_t i, int32_t j, float d) {
    if (pv->n == pv->cap) {
        pv->cap = pv->cap ? pv->cap * 2 : 1024;
        pv->v = (ref_pair *)realloc(pv->v, sizeof(ref_pair) * (size_t)pv->cap);
    }
    pv->v[pv->n].i = i;
    pv->v[pv->n].j = j;
    pv->v[pv->n].d = d;
    pv->n++;
}

/* s = sum((a - b).^2) with StaticArrays' left fold: ((dx*dx + dy*dy) + dz*dz). */
static inline float ref_sqdist(const float *a, const float *b) {
    float dx = a[0] - b[0];
    float dy = a[1] - b[1];
    float dz = a[2] - b[2];
    float sx = dx * dx, sy = dy * dy, sz = dz * dz;
    float s = sx + sy;
    s = s + sz;
    return s;
}

static inline int ref_accept(float s, float r, float r2, int predicate, float *d) {
    if (predicate == REF_PRED_SCALAR) {
        if (s < r2) { *d = sqrtf(s); return 1; }
        return 0;
    }
    float q = sqrtf(s);
    if (q < r) { *d = q; return 1; }
    return 0;
}

static inline int ref_overlap(const ref_node *a, const ref_node *b) {
    /* all((A.min .< B.max) .& (A.max .> B.min)) (BVHTraverse.jl:1094-1098) */
    int ok = 1;
    for (int d = 0; d < 3; ++d) ok &= (a->min[d] < b->max[d]) & (a->max[d] > b->min[d]);
    return ok;
}

typedef struct {
    const ref_node *nodes;
    const float *xyz;
    const int32_t *idx;
    int32_t nL, apl;
    float r, r2;
    int predicate;
} trav_ctx;

/* 4 x 4 tile of twocluster_proximitytest! / the SIMD tile (ext/NaiveSIMD.jl:
 * 567-608) as four 4-lane vectors (one per query row): the same IEEE
 * operations per pair as ref_sqdist / ref_accept, so the same pairs and
 * distance bits in the same (row-major) order, computed 4 pairs at a time. */
typedef float v4f __attribute__((vector_size(16)));
typedef int32_t v4i __attribute__((vector_size(16)));

static inline void ref_tile4(const trav_ctx *c, int64_t qlo, int64_t tlo, pairvec *out) {
    const float *t = c->xyz + 3 * tlo;
    const v4f tx = {t[0], t[3], t[6], t[9]}, ty = {t[1], t[4], t[7], t[10]},
              tz = {t[2], t[5], t[8], t[11]};
    const v4f r2 = {c->r2, c->r2, c->r2, c->r2}, rr = {c->r, c->r, c->r, c->r};
    for (int a = 0; a < 4; ++a) {
        const float *p = c->xyz + 3 * (qlo + a);
        const v4f dx = p[0] - tx, dy = p[1] - ty, dz = p[2] - tz;
        const v4f sx = dx * dx, sy = dy * dy, sz = dz * dz;
        const v4f s = (sx + sy) + sz;
        v4f d;
        v4i keep;
        if (c->predicate == REF_PRED_SCALAR) {
            keep = s < r2;
            d = s;
        } else {
            d = __builtin_ia32_sqrtps(s);
            keep = d < rr;
        }
        if (!(keep[0] | keep[1] | keep[2] | keep[3])) continue;
        for (int b = 0; b < 4; ++b)
            if (keep[b])
                pv_push(out, c->idx[qlo + a], c->idx[tlo + b],
                        c->predicate == REF_PRED_SCALAR ? sqrtf(d[b]) : d[b]);
    }
}

static void ref_traverse_leaf(const trav_ctx *c, int32_t q, pairvec *out, int64_t *tests) {
    const int32_t apl = c->apl;
    const ref_node *ql = &c->nodes[q - 1];
    const int64_t qlo = (int64_t)(q - 1) * apl;
    float d;
    for (int32_t a = 0; a < apl - 1; ++a)
        for (int32_t b = a + 1; b < apl; ++b) {
            float s = ref_sqdist(c->xyz + 3 * (qlo + a), c->xyz + 3 * (qlo + b));
            if (ref_accept(s, c->r, c->r2, c->predicate, &d))
                pv_push(out, c->idx[qlo + a], c->idx[qlo + b], d);
        }
    int32_t t = ql->skip;
    while (t != 0) {
        const ref_node *tn = &c->nodes[t - 1];
        if (ref_overlap(ql, tn)) {
            if (tn->left == 0) {
                const int64_t tlo = (int64_t)(t - 1) * apl;
                if (tests) (*tests)++;
                if (apl == 4) {
                    ref_tile4(c, qlo, tlo, out);
                    t = tn->skip;
                    continue;
                }
                for (int32_t a = 0; a < apl; ++a)
                    for (int32_t b = 0; b < apl; ++b) {
                        float s = ref_sqdist(c->xyz + 3 * (qlo + a), c->xyz + 3 * (tlo + b));
                        if (ref_accept(s, c->r, c->r2, c->predicate, &d))
                            pv_push(out, c->idx[qlo + a], c->idx[tlo + b], d);
                    }
                t = tn->skip;
            } else {
                t = tn->left;
            }
        } else {
            t = tn->skip;
        }
    }
}

typedef struct {
    const trav_ctx *c;
    int chunk, nthreads;
    pairvec out;
    int64_t tests;
} trav_task;

static void *trav_worker(void *arg) {
    trav_task *t = (trav_task *)arg;
    /* for q in chunk:threads:nL (BVHTraverse.jl:1255-1256) */
    for (int32_t q = t->chunk; q <= t->c->nL; q += t->nthreads)
        ref_traverse_leaf(t->c, q, &t->out, &t->tests);
    return NULL;
}

int64_t ref_traverse(const ref_node *nodes, const float *xyz_sorted,
                     const int32_t *idx_sorted, int32_t N, int32_t apl, float r,
                     int predicate, ref_pair *out, int64_t cap, int nthreads,
                     int64_t *leaf_pair_tests) {
    if (apl < 1 || N % apl != 0 || N / apl < 2) return REF_ERR_ARG;
    trav_ctx c = {nodes, xyz_sorted, idx_sorted, N / apl, apl, r, r * r, predicate};
    if (nthreads < 1) nthreads = 1;
    if (nthreads > 256) nthreads = 256;
    trav_task tk[256];
    pthread_t th[256];
    for (int t = 0; t < nthreads; ++t) {
        tk[t] = (trav_task){&c, t + 1, nthreads, {NULL, 0, 0}, 0};
        if (nthreads == 1) trav_worker(&tk[t]);
        else pthread_create(&th[t], NULL, trav_worker, &tk[t]);
    }
    if (nthreads > 1)
        for (int t = 0; t < nthreads; ++t) pthread_join(th[t], NULL);
    /* reduce(vcat, neighbor_vec) (BVHTraverse.jl:1314) */
    int64_t total = 0, tests = 0;
    for (int t = 0; t < nthreads; ++t) { total += tk[t].out.n; tests += tk[t].tests; }
    if (leaf_pair_tests) *leaf_pair_tests = tests;
    if (out && total <= cap) {
        int64_t o = 0;
        for (int t = 0; t < nthreads; ++t) {
            if (tk[t].out.n) memcpy(out + o, tk[t].out.v, sizeof(ref_pair) * (size_t)tk[t].out.n);
            o += tk[t].out.n;
        }
    }
    for (int t = 0; t < nthreads; ++t) free(tk[t].out.v);
    return total;
}

/* ------------------------------------------------------------------------- */
/* A11 + A16: leafTreeData + leafbuild_traverse_bvh in one call               */
/* (BVHTraverse.jl:544-597, 1423-1428; ext/NaiveSIMD.jl:424-480, 706-710).   */
/* ------------------------------------------------------------------------- */
static double now_s(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

int64_t ref_leafbuild_traverse_bvh(const float *xyz, int32_t N, int32_t apl, float r,
                                   int predicate, int morton_mode, const float lo[3],
                                   const float scale[3], ref_pair *out, int64_t cap,
                                   int nthreads, ref_tree_out *tree, double *timing) {
    if (apl < 1 || N % apl != 0 || N / apl < 2) return REF_ERR_ARG;
    int32_t nL = N / apl;
    double t0 = now_s();
    int32_t *codes = (int32_t *)malloc(sizeof(int32_t) * (size_t)N);
    int rc = (morton_mode == REF_MORTON_MASKED) ? ref_morton_masked(xyz, N, codes)
                                                : ref_morton_interleave(xyz, N, lo, scale, codes);
    if (rc != REF_OK) { free(codes); return rc; }
    int32_t *perm = (int32_t *)malloc(sizeof(int32_t) * (size_t)N);
    ref_sortperm(codes, N, perm);
    float *xs = (float *)malloc(sizeof(float) * 3 * (size_t)N);
    int32_t *cs = (int32_t *)malloc(sizeof(int32_t) * (size_t)N);
    for (int32_t s = 0; s < N; ++s) {
        int32_t o = perm[s];
        cs[s] = codes[o];
        xs[3 * s + 0] = xyz[3 * o + 0];
        xs[3 * s + 1] = xyz[3 * o + 1];
        xs[3 * s + 2] = xyz[3 * o + 2];
    }
    ref_node *nodes = (ref_node *)malloc(sizeof(ref_node) * (size_t)(2 * nL - 1));
    double t1 = now_s();
    ref_build_tree(xs, cs, N, apl, r, nodes, 1, nthreads);
    double t2 = now_s();
    int64_t cnt = ref_traverse(nodes, xs, perm, N, apl, r, predicate, out, cap, nthreads, NULL);
    double t3 = now_s();
    if (tree) {
        if (tree->nodes) memcpy(tree->nodes, nodes, sizeof(ref_node) * (size_t)(2 * nL - 1));
        if (tree->perm) memcpy(tree->perm, perm, sizeof(int32_t) * (size_t)N);
        if (tree->codes_sorted) memcpy(tree->codes_sorted, cs, sizeof(int32_t) * (size_t)N);
    }
    if (timing) { timing[0] = t1 - t0; timing[1] = t2 - t1; timing[2] = t3 - t2; }
    free(codes); free(perm); free(xs); free(cs); free(nodes);
    return cnt;
}

/* ------------------------------------------------------------------------- */
/* A17: the point-query traversals.                                          */
/* REF_QUERY_OWN_LEAF: neighbor_traverse (BVHTraverse.jl:1125-1234) with     */
/*   proximity_test! (:1057-1087) and overlap_test (:1089-1092): query q     */
/*   (sorted, 1-based 1..N-1) starts at its own leaf ceil(q/apl); a node is  */
/*   entered iff all(min .< p .< max); a leaf tests its atoms s with         */
/*   q < s (q <= s and a different atom), s = sum((p - x_s).^2) < fl(r*r).  */
/* REF_QUERY_ROOT: gpuneighbor_traverse! (ext/NaiveKA.jl:352-452): every     */
/*   query starts at the root and keeps subjects whose ORIGINAL index        */
/*   exceeds the query's.                                                    */
/* Pairs are (original query index, original subject index, sqrt(s)); one    */
/* thread gives the order q = 1, 2, ... and walk order within a query.       */
/* ------------------------------------------------------------------------- */
typedef struct {
    const ref_node *nodes;
    const float *xyz;
    const int32_t *idx;
    int32_t N, nL, apl, mode;
    float r2;
} pq_ctx;

static void ref_point_query(const pq_ctx *c, int32_t q, pairvec *out) {
    const float *p = c->xyz + 3 * (int64_t)q;
    int32_t t = (c->mode == REF_QUERY_ROOT) ? c->nL + 1 : q / c->apl + 1;
    int64_t visits = 0;
    while (t != 0) {
        if (++visits > 2 * (int64_t)c->nL) break; /* malformed tree */
        const ref_node *k = &c->nodes[t - 1];
        int inside = 1;
        for (int d = 0; d < 3; ++d) inside &= (k->min[d] < p[d]) & (p[d] < k->max[d]);
        if (!inside) { t = k->skip; continue; }
        if (k->left != 0) { t = k->left; continue; }
        const int64_t lo = (int64_t)(t - 1) * c->apl;
        for (int32_t e = 0; e < c->apl; ++e) {
            const int64_t s = lo + e;
            const int keep = (c->mode == REF_QUERY_ROOT) ? (c->idx[q] < c->idx[s]) : (s > q);
            if (!keep) continue;
            const float d2 = ref_sqdist(p, c->xyz + 3 * s);
            if (d2 < c->r2) pv_push(out, c->idx[q], c->idx[s], sqrtf(d2));
        }
        t = k->skip;
    }
}

typedef struct {
    const pq_ctx *c;
    int chunk, nthreads;
    pairvec out;
} pq_task;

static void *pq_worker(void *arg) {
    pq_task *t = (pq_task *)arg;
    /* for query_index in chunk:threads:N-1 (BVHTraverse.jl:1143); the root
     * form queries every atom (ndrange = atom_count, NaiveKA.jl:540) */
    const int32_t last = (t->c->mode == REF_QUERY_ROOT) ? t->c->N : t->c->N - 1;
    for (int32_t q1 = t->chunk; q1 <= last; q1 += t->nthreads) ref_point_query(t->c, q1 - 1, &t->out);
    return NULL;
}

int64_t ref_neighbor_traverse(const ref_node *nodes, const float *xyz_sorted,
                              const int32_t *idx_sorted, int32_t N, int32_t apl, float r,
                              int mode, ref_pair *out, int64_t cap, int nthreads) {
    if (apl < 1 || N % apl != 0 || N / apl < 2) return REF_ERR_ARG;
    pq_ctx c = {nodes, xyz_sorted, idx_sorted, N, N / apl, apl, mode, r * r};
    if (nthreads < 1) nthreads = 1;
    if (nthreads > 256) nthreads = 256;
    pq_task tk[256];
    pthread_t th[256];
    for (int t = 0; t < nthreads; ++t) {
        tk[t] = (pq_task){&c, t + 1, nthreads, {NULL, 0, 0}};
        if (nthreads == 1) pq_worker(&tk[t]);
        else pthread_create(&th[t], NULL, pq_worker, &tk[t]);
    }
    if (nthreads > 1)
        for (int t = 0; t < nthreads; ++t) pthread_join(th[t], NULL);
    /* reduce(vcat, neighbor_vec) (BVHTraverse.jl:1178) */
    int64_t total = 0;
    for (int t = 0; t < nthreads; ++t) total += tk[t].out.n;
    if (out && total <= cap) {
        int64_t o = 0;
        for (int t = 0; t < nthreads; ++t) {
            if (tk[t].out.n) memcpy(out + o, tk[t].out.v, sizeof(ref_pair) * (size_t)tk[t].out.n);
            o += tk[t].out.n;
        }
    }
    for (int t = 0; t < nthreads; ++t) free(tk[t].out.v);
    return total;
}

/* TreeData (BVHTraverse.jl:500-543) + neighbor_traverse = build_traverse_bvh
 * (:1416-1421): codes, stable sortperm, +- r leaf boxes (cluster_primitives),
 * the build with its store, the root box overwritten; then the walk.  The
 * root form (gpubvh_neighborlist, NaiveKA.jl:470-558) keeps the root's own
 * box (its build kernel does not overwrite it).  store (nL-1) may be NULL. */
int64_t ref_build_traverse_bvh(const float *xyz, int32_t N, int32_t apl, float r, int mode,
                               int morton_mode, const float lo[3], const float scale[3],
                               ref_pair *out, int64_t cap, int nthreads, ref_tree_out *tree,
                               int32_t *store) {
    if (apl < 1 || N % apl != 0 || N / apl < 2) return REF_ERR_ARG;
    int32_t nL = N / apl;
    int32_t *codes = (int32_t *)malloc(sizeof(int32_t) * (size_t)N);
    int rc = (morton_mode == REF_MORTON_MASKED) ? ref_morton_masked(xyz, N, codes)
                                                : ref_morton_interleave(xyz, N, lo, scale, codes);
    if (rc != REF_OK) { free(codes); return rc; }
    int32_t *perm = (int32_t *)malloc(sizeof(int32_t) * (size_t)N);
    ref_sortperm(codes, N, perm);
    float *xs = (float *)malloc(sizeof(float) * 3 * (size_t)N);
    int32_t *cs = (int32_t *)malloc(sizeof(int32_t) * (size_t)N);
    for (int32_t s = 0; s < N; ++s) {
        int32_t o = perm[s];
        cs[s] = codes[o];
        for (int d = 0; d < 3; ++d) xs[3 * s + d] = xyz[3 * o + d];
    }
    ref_node *nodes = (ref_node *)malloc(sizeof(ref_node) * (size_t)(2 * nL - 1));
    build_tree_h(xs, cs, N, apl, r, nodes, mode == REF_QUERY_OWN_LEAF, nthreads, store);
    int64_t cnt = ref_neighbor_traverse(nodes, xs, perm, N, apl, r, mode, out, cap, nthreads);
    if (tree) {
        if (tree->nodes) memcpy(tree->nodes, nodes, sizeof(ref_node) * (size_t)(2 * nL - 1));
        if (tree->perm) memcpy(tree->perm, perm, sizeof(int32_t) * (size_t)N);
        if (tree->codes_sorted) memcpy(tree->codes_sorted, cs, sizeof(int32_t) * (size_t)N);
    }
    free(codes); free(perm); free(xs); free(cs); free(nodes);
    return cnt;
}

/* ------------------------------------------------------------------------- */
/* quantized_positions! (BVHTraverse.jl:290-314): per axis d = x, y, z in     */
/* turn, a stable sort of the atoms by coordinate d (Julia isless: -0 < +0,   */
/* NaN after everything) starting from the order the previous axis left,     */
/* and q[i][d] = 1-based rank of atom i in that sort.                         */
/* ------------------------------------------------------------------------- */
typedef struct {
    uint32_t key;
    int32_t pos; /* place in the order before this sort: makes qsort stable */
    int32_t atom;
} rank_item;

static int rank_cmp(const void *a, const void *b) {
    const rank_item *x = (const rank_item *)a, *y = (const rank_item *)b;
    if (x->key != y->key) return x->key < y->key ? -1 : 1;
    return (x->pos > y->pos) - (x->pos < y->pos);
}

/* total order of isless on Float32 as an unsigned key */
static inline uint32_t isless_key(float v) {
    if (isnan(v)) return 0xFFFFFFFFu;
    uint32_t b;
    memcpy(&b, &v, 4);
    return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}

void ref_quantized_positions(const float *xyz, int32_t N, int32_t *q) {
    rank_item *it = (rank_item *)malloc(sizeof(rank_item) * (size_t)(N > 0 ? N : 1));
    for (int32_t i = 0; i < N; ++i) it[i].atom = i;
    for (int d = 0; d < 3; ++d) {
        for (int32_t k = 0; k < N; ++k) {
            it[k].pos = k;
            it[k].key = isless_key(xyz[3 * (int64_t)it[k].atom + d]);
        }
        qsort(it, (size_t)N, sizeof(rank_item), rank_cmp);
        for (int32_t k = 0; k < N; ++k) q[3 * (int64_t)it[k].atom + d] = k + 1;
    }
    free(it);
}

/* ------------------------------------------------------------------------- */
/* A25: unique_pairs + threshold_pairs (AllToAll.jl:65-88): all i<j in        */
/* original order, d = sqrt(sum((a_i - a_j).^2)), keep d < threshold.         */
/* ------------------------------------------------------------------------- */
int64_t ref_bruteforce(const float *xyz, int32_t N, float r, ref_pair *out, int64_t cap) {
    int64_t n = 0;
    for (int32_t i = 0; i < N - 1; ++i)
        for (int32_t j = i + 1; j < N; ++j) {
            float d = sqrtf(ref_sqdist(xyz + 3 * (int64_t)i, xyz + 3 * (int64_t)j));
            if (d < r) {
                if (out && n < cap) { out[n].i = i; out[n].j = j; out[n].d = d; }
                ++n;
            }
        }
    return n;
}
