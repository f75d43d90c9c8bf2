/*
 * refcpu.h -- declarations of the CPU oracle (test infrastructure only).
 * See refcpu.c / oracle64.c for the reference citations.
 */
#ifndef REFCPU_H
#define REFCPU_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { REF_OK = 0, REF_ERR_ARG = -1, REF_ERR_DOMAIN = -2 };
enum { REF_PRED_SCALAR = 0, REF_PRED_SIMD = 1 };
enum { REF_MORTON_MASKED = 0, REF_MORTON_INTERLEAVE = 1 };
enum { REF_QUERY_OWN_LEAF = 0, REF_QUERY_ROOT = 1 };

/* GridKey{Float32,Int32} bit layout (BVHTraverse.jl:102-109): 32 bytes. */
typedef struct {
    float min[3];
    float max[3];
    int32_t left;
    int32_t skip;
} ref_node;

/* Tuple{Int32,Int32,Float32}: (i, j, d), 12 bytes. */
typedef struct {
    int32_t i;
    int32_t j;
    float d;
} ref_pair;

typedef struct {
    ref_node *nodes;      /* 2nL-1 entries, or NULL */
    int32_t *perm;        /* N entries: sorted slot -> original index, or NULL */
    int32_t *codes_sorted;/* N entries, or NULL */
} ref_tree_out;

float ref_binwidth(void);
int ref_morton_masked(const float *xyz, int64_t n, int32_t *codes);
uint32_t ref_hilbert_cell(uint32_t cx, uint32_t cy, uint32_t cz);
void ref_hilbert_cells(const uint32_t *cells, int64_t n, int32_t *codes);
int ref_morton_hilbert(const float *xyz, int64_t n, const float lo[3], const float scale[3],
                       int32_t *codes);
int ref_morton_interleave(const float *xyz, int64_t n, const float lo[3],
                          const float scale[3], int32_t *codes);
void ref_sortperm(const int32_t *codes, int64_t n, int32_t *perm);
void ref_leaf_boxes(const float *xyz_sorted, int32_t nL, int32_t apl, float r,
                    ref_node *nodes);
int32_t ref_delta_public(int32_t i, const int32_t *codes_sorted, int32_t nL, int32_t apl);
int ref_build_tree(const float *xyz_sorted, const int32_t *codes_sorted, int32_t N,
                   int32_t apl, float r, ref_node *nodes, int override_root,
                   int nthreads);
int64_t ref_traverse(const ref_node *nodes, const float *xyz_sorted,
                     const int32_t *idx_sorted, int32_t N, int32_t apl, float r,
                     int predicate, ref_pair *out, int64_t cap, int nthreads,
                     int64_t *leaf_pair_tests);
int64_t ref_leafbuild_traverse_bvh(const float *xyz, int32_t N, int32_t apl, float r,
                                   int predicate, int morton_mode, const float lo[3],
                                   const float scale[3], ref_pair *out, int64_t cap,
                                   int nthreads, ref_tree_out *tree, double *timing);
int64_t ref_bruteforce(const float *xyz, int32_t N, float r, ref_pair *out, int64_t cap);
/* cluster_primitives: leaf boxes of +- r (BVHTraverse.jl:341-374). */
void ref_cluster_boxes(const float *xyz_sorted, int32_t nL, int32_t apl, float r,
                       ref_node *nodes);
/* neighbor_traverse / gpuneighbor_traverse! over a tree (mode REF_QUERY_*). */
int64_t ref_neighbor_traverse(const ref_node *nodes, const float *xyz_sorted,
                              const int32_t *idx_sorted, int32_t N, int32_t apl, float r,
                              int mode, ref_pair *out, int64_t cap, int nthreads);
/* TreeData + neighbor_traverse (build_traverse_bvh); store: nL-1 CAS slots. */
int64_t ref_build_traverse_bvh(const float *xyz, int32_t N, int32_t apl, float r, int mode,
                               int morton_mode, const float lo[3], const float scale[3],
                               ref_pair *out, int64_t cap, int nthreads, ref_tree_out *tree,
                               int32_t *store);
/* quantized_positions!: q (N x 3) = 1-based rank per axis, cascaded stable sorts. */
void ref_quantized_positions(const float *xyz, int32_t N, int32_t *q);

/* ---- oracle64.c: FP64 textbook physics ---------------------------------- */
typedef struct {
    double eps, sigma;     /* Lennard-Jones */
    double rc_lj;          /* LJ cutoff (0 disables LJ) */
    double k_coul;         /* Coulomb constant */
    double rc_coul;        /* Coulomb cutoff (0 disables Coulomb) */
    int32_t shift_lj;      /* 1: energy-shifted LJ (U(rc)=0) */
    int32_t shift_coul;    /* 1: energy-shifted Coulomb */
} o64_params;

/* Forces, potential energy and virial sum_{i<j} r_ij . F_ij over every pair
 * with r < cutoff (cell list).  F is N x 3.  Returns number of pairs inside
 * max(rc_lj, rc_coul). */
int64_t o64_forces(const double *x, const double *q, int32_t N, const o64_params *p,
                   double *F, double *U, double *W);
/* Same physics over an explicit half pair list (i, j) (0-based, unordered). */
void o64_forces_pairs(const double *x, const double *q, int32_t N, const int32_t *pi,
                      const int32_t *pj, int64_t npairs, const o64_params *p, double *F,
                      double *U, double *W);
/* Velocity-Verlet (kick-drift-force-kick) with optional reflecting walls
 * [lo, hi] applied after the drift.  energy (nsteps+1) x 3 rows of
 * (kinetic, potential, virial) if non-NULL. */
int o64_vv_run(double *x, double *v, const double *q, const double *m, int32_t N,
               const o64_params *p, double dt, int32_t nsteps, int32_t walls, double lo,
               double hi, double *F, double *energy);
/* The same with the velocity rescaling of rescale_velocity!
 * (src/Simulator.jl:119-144) every thermostat_every steps (0 = off): target
 * temperature t0, damping gamma, T from the kinetic energy (k_B = 1). */
int o64_vv_run_thermo(double *x, double *v, const double *q, const double *m, int32_t N,
                      const o64_params *p, double dt, int32_t nsteps, int32_t walls, double lo,
                      double hi, int32_t thermostat_every, double t0, double gamma, double *F,
                      double *energy);

/* ---- threaded FP32 CPU baseline (reference path + textbook forces) ------ */
typedef struct {
    float eps, sigma, rc, k_coul, rc_coul;
    float dt, skin, lo, hi;
    int32_t apl, rebuild_every, walls, predicate, morton_mode, nthreads;
} base_params;
/* Runs nsteps of the reference CPU path on FP32 state (x, v: N x 3).
 * times[0..3] = seconds in (morton+sort, build, traverse, force+integrate);
 * pairs_eval = sum over steps of unique in-cutoff pairs. */
int base_run(float *x, float *v, const float *q, int32_t N, const base_params *p,
             int32_t nsteps, double *times, int64_t *pairs_eval);

#ifdef __cplusplus
}
#endif
#endif
