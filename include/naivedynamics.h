/*
 * naivedynamics.h -- C ABI of libnaivehip.so, the gfx950 HIP implementation of
 * NaiveDynamics.jl's per-step hot path (LBVH neighbour search, Lennard-Jones +
 * Coulomb pair forces, velocity Verlet).
 *
 * Conventions
 *  - Plain pointers and sizes only.  "d_" arguments are device pointers
 *    (allocate them with nd_malloc or any HIP allocator on the context's
 *    device); "h_" arguments are host pointers.
 *  - Every call returns an nd_status; on failure nd_last_error() describes it.
 *    A Julia caller turns a non-zero status into error(), which is how the
 *    reference reports bad input (BVHTraverse.jl:74-80, InexactError at :266).
 *  - Atom and pair indices are 0-based.  Tree node numbers keep the
 *    reference's 1-based GridKey numbering (leaves 1..nL, internal nodes
 *    nL+1..2nL-1, root nL+1, 0 = sentinel); node k is stored in slot k-1.
 *  - Calls on one context are stream-ordered on the context's HIP stream and
 *    are not reentrant; use one context per host thread.  Calls documented as
 *    synchronous return after the work has finished.
 */
#ifndef NAIVEDYNAMICS_H
#define NAIVEDYNAMICS_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    ND_OK = 0,
    ND_ERR_ARG = -1,      /* invalid argument or specification */
    ND_ERR_DOMAIN = -2,   /* coordinate outside the quantisation domain / non-finite */
    ND_ERR_CAPACITY = -3, /* output buffer too small; the required count is returned */
    ND_ERR_HIP = -4,      /* HIP runtime failure */
    ND_ERR_STATE = -5     /* call order violated (e.g. no tree built yet) */
} nd_status;

typedef enum { ND_PRED_SCALAR = 0, ND_PRED_SIMD = 1 } nd_predicate;
typedef enum { ND_MORTON_MASKED = 0, ND_MORTON_INTERLEAVE = 1, ND_MORTON_HILBERT = 2 } nd_morton_mode;
typedef enum { ND_PREC_FP32 = 0, ND_PREC_MIXED = 1, ND_PREC_FP64 = 2 } nd_precision;
typedef enum { ND_QUERY_OWN_LEAF = 0, ND_QUERY_ROOT = 1 } nd_query_mode;

/* GridKey{Float32,Int32} (BVHTraverse.jl:102-109), same 32-byte layout. */
typedef struct {
    float min[3];
    float max[3];
    int32_t left;
    int32_t skip;
} nd_node;

/* Tuple{Int32,Int32,Float32} = (i, j, d) (BVHTraverse.jl:1314), 12 bytes. */
typedef struct {
    int32_t i;
    int32_t j;
    float d;
} nd_pair;

/* SpheresBVHSpecs{Float32,Int32} (BVHTraverse.jl:63-69). */
typedef struct {
    float neighbor_distance;
    int32_t atom_count;
    int32_t leaves_count;
    int32_t branches_count;
    int32_t atomsperleaf;
} nd_bvh_spec;

typedef struct nd_ctx nd_ctx;

/* ---- context, device memory, timing (no reference counterpart: the Julia
 *      shim needs them to stage host arrays, as NaiveKA.jl:507-510 does with
 *      Adapt) -------------------------------------------------------------- */
int nd_ctx_create(int32_t device, nd_ctx **out);
int nd_ctx_destroy(nd_ctx *ctx);
const char *nd_last_error(const nd_ctx *ctx);
int nd_version(void);
int nd_malloc(nd_ctx *ctx, size_t bytes, void **d_ptr);
int nd_free(nd_ctx *ctx, void *d_ptr);
int nd_memcpy_h2d(nd_ctx *ctx, void *d_dst, const void *h_src, size_t bytes);
int nd_memcpy_d2h(nd_ctx *ctx, void *h_dst, const void *d_src, size_t bytes);
int nd_memcpy_d2d(nd_ctx *ctx, void *d_dst, const void *d_src, size_t bytes);
int nd_memset(nd_ctx *ctx, void *d_dst, int32_t byte_value, size_t bytes);
int nd_synchronize(nd_ctx *ctx);
/* The context's HIP stream (hipStream_t) in *stream: every call on the
 * context is ordered on it, so a caller can order its own HIP work or
 * capture a graph around stream-ordered calls. */
int nd_ctx_stream(nd_ctx *ctx, void **stream);
/* HIP events on the context's stream: nd_timer_start records, nd_timer_stop
 * records, synchronises and returns the elapsed milliseconds. */
int nd_timer_start(nd_ctx *ctx);
int nd_timer_stop(nd_ctx *ctx, float *ms);

/* ---- specification ------------------------------------------------------
 * SpheresBVHSpecs(; neighbor_distance, atom_count, floattype=Float32,
 * atomsperleaf) (BVHTraverse.jl:71-94): ND_ERR_ARG unless atomsperleaf
 * divides atom_count and there are at least two leaves. */
int nd_bvh_spec_init(nd_bvh_spec *spec, float neighbor_distance, int64_t atom_count,
                     int64_t atomsperleaf);

/* ---- neighbour-search stages (synchronous) ------------------------------ */

/* mortoncodes! (BVHTraverse.jl:257-288).  ND_MORTON_MASKED reproduces the
 * reference key q_d = floor(x_d / Float32(1/1023)), code = OR_d (q_d & M_d)
 * bit for bit and fails with ND_ERR_DOMAIN where the reference throws
 * InexactError.  ND_MORTON_INTERLEAVE is this library's 30-bit interleaved key
 * over the box [lo, lo + 1024/scale) (lo3/scale3 may be NULL: the bounding
 * box of the input is used); ND_MORTON_HILBERT is the 30-bit Hilbert index of
 * the same 10-bit cells (Skilling's transform), the simulation's default key:
 * its fixed-size leaves are compact runs of space.  d_xyz is N x 3 float
 * (AoS).  The one-call neighbour list takes the first two modes. */
int nd_mortoncodes(nd_ctx *ctx, const float *d_xyz, int32_t n, int32_t mode,
                   const float *h_lo3, const float *h_scale3, int32_t *d_codes);

/* sortperm(codes) + permute! (BVHTraverse.jl:570-573): stable ascending sort
 * of signed Int32 keys; d_perm[s] = original index of sorted slot s. */
int nd_sortperm(nd_ctx *ctx, const int32_t *d_codes, int32_t n, int32_t *d_codes_sorted,
                int32_t *d_perm);

/* leafcluster_primitives + bounding_volume_hierarchy!
 * (BVHTraverse.jl:376-410, 704-860): leaf boxes of +-neighbor_distance/2 around
 * the sorted atoms, Apetrei bottom-up build with ropes, exact box refit.
 * d_nodes has 2*leaves_count-1 entries.  override_root=1 writes the root box
 * as [0,1]^3 as the reference does (BVHTraverse.jl:856-858). */
int nd_bounding_volume_hierarchy(nd_ctx *ctx, const float *d_xyz_sorted,
                                 const int32_t *d_codes_sorted, const nd_bvh_spec *spec,
                                 int32_t override_root, nd_node *d_nodes);

/* leafneighbor_traverse (BVHTraverse.jl:1236-1323) with ND_PRED_SCALAR
 * (s < r*r, :1021-1055) or simdneighbor_traverse (NaiveSIMD.jl:610-704) with
 * ND_PRED_SIMD (sqrt(s) < r).  Pairs are (original index of the query-leaf
 * atom, original index of the other atom, Float32 distance), in the order a
 * single-threaded reference run emits them.  If cap < required, returns
 * ND_ERR_CAPACITY with *count = required and writes nothing. */
int nd_leafneighbor_traverse(nd_ctx *ctx, const nd_node *d_nodes, const float *d_xyz_sorted,
                             const int32_t *d_index_sorted, const nd_bvh_spec *spec,
                             int32_t predicate, nd_pair *d_out, int64_t cap, int64_t *count);

/* ---- one-call neighbour list ---------------------------------------------
 * leafbuild_traverse_bvh (BVHTraverse.jl:1423-1428) / simdbuild_traverse_bvh
 * (NaiveSIMD.jl:706-710) / gpubvh_neighborlist (NaiveKA.jl:470-558): Morton,
 * sort, tree, traversal on device.  Capacity protocol as above. The tree of the
 * last call stays on the device for nd_tree_export. */
int nd_leafbuild_traverse_bvh(nd_ctx *ctx, const float *d_xyz, const nd_bvh_spec *spec,
                              int32_t predicate, int32_t morton_mode, nd_pair *d_out,
                              int64_t cap, int64_t *count);
/* Stream-ordered form of nd_leafbuild_traverse_bvh (no host synchronisation,
 * for hipGraph capture and pipelines): *d_count (device int64) receives the
 * required pair count and the pairs are written only if it is <= cap
 * (otherwise nothing is written, as in the synchronous form); *d_status
 * (device int32) receives status bits, 1 = a coordinate outside the key
 * domain (the synchronous form's ND_ERR_DOMAIN), 4 = a malformed tree.  Same
 * pairs in the same order as the synchronous form; the key sort covers all
 * 32 bits instead of probing which digits vary.  Its work buffers belong to
 * this form alone (no other call resizes them); they grow on the first call
 * for a larger size, so call once at the largest size before capturing a
 * graph, which then stays valid. */
int nd_leafbuild_traverse_bvh_async(nd_ctx *ctx, const float *d_xyz, const nd_bvh_spec *spec,
                                    int32_t predicate, int32_t morton_mode, nd_pair *d_out,
                                    int64_t cap, int64_t *d_count, int32_t *d_status);
/* No reference counterpart: for callers that cannot allocate device memory
 * (code under hipGraph capture, serving loops, harnesses that forbid
 * hipMalloc).  nd_leafbuild_traverse_bvh_async with its work buffers in a
 * static array of the library (32 MiB per GPU, held until the process ends)
 * instead of allocated memory: it calls no allocator.  Sizes whose work
 * buffers exceed the array (about 290,000 atoms at 4 atoms per leaf, more for
 * larger leaves) are rejected with ND_ERR_CAPACITY.  Every call on one GPU
 * uses that GPU's array, so calls that overlap on one GPU (two contexts or
 * streams) race; the array keeps the last call's tree for nd_tree_export. */
int nd_leafbuild_traverse_bvh_noalloc(nd_ctx *ctx, const float *d_xyz, const nd_bvh_spec *spec,
                                      int32_t predicate, int32_t morton_mode, nd_pair *d_out,
                                      int64_t cap, int64_t *d_count, int32_t *d_status);
/* Host-buffer form of the same call (stages h_xyz in and pairs out). */
int nd_leafbuild_traverse_bvh_host(nd_ctx *ctx, const float *h_xyz, const nd_bvh_spec *spec,
                                   int32_t predicate, int32_t morton_mode, nd_pair *h_out,
                                   int64_t cap, int64_t *count);
/* treedata of gpubvh_neighborlist (NaiveKA.jl:557) / leafTreeData
 * (BVHTraverse.jl:596): nodes (2nL-1), sort permutation and sorted codes of
 * the last nd_leafbuild_traverse_bvh call, copied to host (any may be NULL). */
int nd_tree_export(nd_ctx *ctx, nd_node *h_nodes, int32_t *h_perm, int32_t *h_codes_sorted);

/* ---- point-query neighbour list -----------------------------------------
 * neighbor_traverse (BVHTraverse.jl:1125-1234, with proximity_test! :1057-1087
 * and overlap_test :1089-1092) for ND_QUERY_OWN_LEAF: query atom q (sorted
 * order, all but the last) walks from its own leaf; a node is entered iff q
 * lies strictly inside its box on every axis; a leaf keeps its atoms s after
 * q with s_qs < fl(r*r).  ND_QUERY_ROOT: gpuneighbor_traverse!
 * (ext/NaiveKA.jl:352-452), every atom walks from the root and keeps subjects
 * whose original index exceeds its own.  Pairs (original query index,
 * original subject index, sqrt(s_qs)) query-major in walk order (the order of
 * a single-threaded reference run).  Capacity protocol as above. */
int nd_neighbor_traverse(nd_ctx *ctx, const nd_node *d_nodes, const float *d_xyz_sorted,
                         const int32_t *d_index_sorted, const nd_bvh_spec *spec, int32_t query_mode,
                         nd_pair *d_out, int64_t cap, int64_t *count);
/* build_traverse_bvh (BVHTraverse.jl:1416-1421) = TreeData (:500-543:
 * codes, stable sort, cluster_primitives' +- r leaf boxes, the build, the
 * root box overwritten) + neighbor_traverse, for ND_QUERY_OWN_LEAF; the tree
 * and traversal of gpubvh_neighborlist (ext/NaiveKA.jl:470-558; root box
 * kept) for ND_QUERY_ROOT.  The tree stays for nd_tree_export and
 * nd_tree_store_export. */
int nd_build_traverse_bvh(nd_ctx *ctx, const float *d_xyz, const nd_bvh_spec *spec, int32_t query_mode,
                          int32_t morton_mode, nd_pair *d_out, int64_t cap, int64_t *count);
/* store of TreeData / gpubvh_neighborlist's treedata (BVHTraverse.jl:521-542,
 * NaiveKA.jl:557): the nL-1 compare-and-swap slots of the last one-call
 * build, as the build left them.  Slot k (split after leaf k+1, 1-based
 * leaves) holds the range end parked by the child that arrived first: the
 * left child's first leaf or the right child's last leaf (a one-thread
 * reference build always leaves the former; this build's lanes race). */
int nd_tree_store_export(nd_ctx *ctx, int32_t *h_store);
/* quantized_positions! (BVHTraverse.jl:290-314): d_quantized (N x 3 int32)
 * = 1-based rank of each atom in a stable sort by x, then by y over that
 * order, then by z over that order (Julia isless: -0 before +0). */
int nd_quantized_positions(nd_ctx *ctx, const float *d_xyz, int32_t n, int32_t *d_quantized);

/* threshold_pairs(unique_pairs(x), threshold) (AllToAll.jl:65-88): all i<j
 * with sqrt(s) < threshold, in (i, j) lexicographic order.  O(N^2). */
int nd_unique_threshold_pairs(nd_ctx *ctx, const float *d_xyz, int32_t n, float threshold,
                              nd_pair *d_out, int64_t cap, int64_t *count);

/* neighborlist(x, cutoff) of CellListMap, the cell-list search the reference
 * benchmarks its LBVH against (assets/BVHBenchSuite.jl:307; src/MDInput.jl:290):
 * every unordered pair of d_xyz (N x 3 float, AoS) accepted by predicate
 * (ND_PRED_SCALAR: s < cutoff^2, ND_PRED_SIMD: sqrt(s) < cutoff, with s folded
 * as ((dx^2 + dy^2) + dz^2)), once, as (min index, max index, Float32
 * distance) with 0-based original indices.  Uniform cells of side >= cutoff
 * over the bounding box; atoms with a non-finite coordinate pair with none.
 * The pair set and distances equal nd_unique_threshold_pairs' (ND_PRED_SIMD)
 * bit for bit; the order is deterministic (cell order of the earlier atom).
 * Capacity protocol as above. */
int nd_cell_neighborlist(nd_ctx *ctx, const float *d_xyz, int32_t n, float cutoff,
                         int32_t predicate, nd_pair *d_out, int64_t cap, int64_t *count);
/* Host-buffer form of the same call (stages h_xyz in and pairs out). */
int nd_cell_neighborlist_host(nd_ctx *ctx, const float *h_xyz, int32_t n, float cutoff,
                              int32_t predicate, nd_pair *h_out, int64_t cap, int64_t *count);

/* ---- forces over a pair list --------------------------------------------- */
typedef struct {
    float eps, sigma; /* Lennard-Jones */
    float rc;         /* LJ cutoff */
    float k_coul;     /* Coulomb constant */
    float rc_coul;    /* Coulomb cutoff */
    int32_t shift_lj, shift_coul; /* energy shift at the cutoff */
} nd_force_params;

/* force_lennardjones!(force, pairslist, position) (Forces.jl:15-45): zeroes
 * d_force (N x 3 float) and adds the textbook LJ force of every listed pair
 * with r < rc to both atoms (Newton's third law).  *h_energy (may be NULL)
 * receives the pair energy in FP64. */
int nd_force_lennardjones(nd_ctx *ctx, float *d_force, const nd_pair *d_pairs, int64_t npairs,
                          const float *d_xyz, int32_t n, const nd_force_params *p,
                          double *h_energy);
/* force_coulomb!(force, pairslist, charge) (Forces.jl:56-66): textbook cutoff
 * Coulomb k q_i q_j r_ij / r^3; positions are needed for the direction. */
int nd_force_coulomb(nd_ctx *ctx, float *d_force, const nd_pair *d_pairs, int64_t npairs,
                     const float *d_xyz, const float *d_charge, int32_t n,
                     const nd_force_params *p, double *h_energy);
/* sum_forces!(force, f1, f2) (Forces.jl:68-74). */
int nd_sum_forces(nd_ctx *ctx, float *d_force, const float *d_f1, const float *d_f2, int32_t n);

/* ---- simulation (simulate! / simulate_bvh!, Simulator.jl:154-256, 327-379) */
typedef struct {
    int32_t n;            /* atoms (2 .. 2^26) */
    int32_t precision;    /* nd_precision */
    int32_t morton_mode;  /* key for the simulation's LBVH (nd_morton_mode) */
    int32_t atomsperleaf; /* LBVH leaf size (must divide n) */
    int32_t rebuild_every;/* neighbour-list rebuild period in steps (>= 1) */
    int32_t walls;        /* 1: reflecting walls at [lo, hi]^3 (boundary_reflect!) */
    int32_t energy_every; /* record (KE, PE, virial, pairs) every k steps; 0 = never */
    int32_t thermostat_every; /* velocity rescale period (rescale_velocity!); 0 = off */
    int32_t use_graph;    /* 1: replay each rebuild window as a captured hipGraph */
    int32_t max_neighbors;/* initial per-atom list capacity (0 = estimate) */
    int32_t resort_every; /* full re-sort + tree build period in steps (a multiple of
                             rebuild_every; 0 = every rebuild).  Rebuilds in between
                             keep the atom order and tree topology and refit the
                             boxes to the current positions. */
    int32_t list_walk;    /* list walk: 0 = two-phase child-box walk (node frontier, then
                             queued leaves), 1 = rope walk, 2 = one-node child-box walk,
                             3 = uniform cell list over [lo, hi]^3 (no LBVH; the
                             comparison point; needs hi > lo) */
    int32_t pair_batch;   /* step kernel: neighbours per software-pipeline batch, 2, 4 or
                             8 (FP32; MIXED and FP64 use 4); 0 = 4 */
    int32_t walk_front;   /* two-phase walk: tree records fetched per round trip, 8 or
                             16; 0 = 8 */
    int32_t atom_lanes;   /* step kernel: lanes per atom, 1 or 2 (two lanes split an
                             atom's list and add their sums in a fixed order); 0 = 1 */
    double eps, sigma, rc; /* LJ */
    double k_coul, rc_coul;/* Coulomb (rc_coul = 0 disables) */
    double skin;           /* list radius = max(rc, rc_coul) + skin */
    double dt;
    double lo, hi;         /* box for walls and the interleaved key */
    double temperature, vdamp; /* thermostat target and damping gamma */
} nd_sim_params;

typedef struct {
    int64_t steps;            /* steps integrated so far */
    int64_t rebuilds;         /* neighbour-list builds so far */
    int64_t list_entries;     /* directed entries in the current list */
    int32_t max_neighbors;    /* largest per-atom count at the last build */
    int32_t capacity;         /* current per-atom capacity */
    int64_t pairs_in_cutoff;  /* unique pairs with r < cutoff at the last energy step */
    int64_t stale_reruns;     /* nd_sim_run calls redone because an atom had moved half the
                                 skin (by its path length since the list build) before a
                                 scheduled rebuild */
    int64_t forced_rebuilds;  /* list rebuilds outside the rebuild_every schedule in those
                                 reruns (made before the first force the stale list would
                                 have served) */
    int64_t graph_recaptures; /* window graphs captured again because a buffer they use had
                                 moved since capture (a guard: expected 0) */
    int32_t table_spills;     /* table builds' workgroups (summed over builds) whose
                                 records exceeded the LDS record buffer (expanded from
                                 memory instead) */
    int32_t table_records_big; /* 1: table builds hold up to 5,120 records per workgroup in
                                  LDS (chosen when the first checked build spilled), else 3,072 */
} nd_sim_stats;

typedef struct nd_sim nd_sim;

int nd_sim_create(nd_ctx *ctx, const nd_sim_params *p, nd_sim **out);
int nd_sim_destroy(nd_sim *sim);
/* State in original atom order, FP64 host arrays (x, v: N x 3; q, m: N; q or m
 * may be NULL for 0 and 1). */
int nd_sim_set_state(nd_sim *sim, const double *h_x, const double *h_v, const double *h_q,
                     const double *h_m);
int nd_sim_get_state(nd_sim *sim, double *h_x, double *h_v, double *h_f);
/* Advance nsteps velocity-Verlet steps (kick-drift-force-kick), rebuilding
 * the neighbour list every rebuild_every steps.  The list holds the pairs
 * within max(rc, rc_coul) + skin; each atom's path length since the build
 * is tracked, and if an atom reaches skin / 2 before a scheduled rebuild the
 * steps are run again from the call's starting state with a rebuild before
 * every force that list could no longer serve (stats: stale_reruns,
 * forced_rebuilds).  Returns after completion. */
int nd_sim_run(nd_sim *sim, int32_t nsteps);
/* No reference counterpart (the reference has no captured run loop): sizes
 * the neighbour list and computes the current force if that is not done yet
 * (as nd_sim_run(sim, 0) does), then captures and instantiates every rebuild-
 * window hipGraph that nd_sim_run can replay (every buffer parity and window
 * kind), so that later runs spend no time on capture.  Advances no step.
 * Graphs are dropped again when the list capacity grows or the state or
 * replicas are replaced. */
int nd_sim_prepare(nd_sim *sim);
/* Energy rows (step, kinetic, potential, virial, pairs_in_cutoff) recorded
 * since the last call; *rows receives the number written. */
int nd_sim_energies(nd_sim *sim, double *h_rows, int32_t max_rows, int32_t *rows);
int nd_sim_stats_get(nd_sim *sim, nd_sim_stats *out);
/* Average device time per launch of the fused force kernel over the last
 * nd_sim_run, measured with HIP events around each launch when enabled. */
int nd_sim_kernel_timing(nd_sim *sim, int32_t enable, double *avg_force_ms,
                         int64_t *launches);
/* Advances nsteps (2..rebuild_every) steps as one rebuild window: rebuild,
 * force, kick-drift, then nsteps-1 fused force+integrate launches back to
 * back between two HIP events, then the final half kick.  *avg_fused_ms is
 * the event time divided by nsteps-1 (the fused kernel's average duration). */
int nd_sim_time_fused_window(nd_sim *sim, int32_t nsteps, double *avg_fused_ms);

/* ---- batched replicas (no reference counterpart: the reference simulates
 * one system per simulate! call, Simulator.jl:154-256) ----
 * Declares the atoms of the simulation to be nrep independent replicas packed
 * in space: h_rep[i] is the replica of atom i (original order; -1 = an inert
 * padding atom without walls) and h_boxes[6 r .. 6 r + 5] = (lo_x, lo_y, lo_z,
 * hi_x, hi_y, hi_z) is replica r's wall box.  Replica boxes must be separated
 * by more than the list radius so that no pair crosses replicas; with walls
 * on, each atom reflects off its own replica's box instead of [lo, hi]^3.
 * Call after nd_sim_create and before nd_sim_run; nrep = 0 clears it. */
int nd_sim_set_replicas(nd_sim *sim, int32_t nrep, const int32_t *h_rep, const double *h_boxes);

/* Trajectory frame for the poslog of simulate! (Simulator.jl:167, 245):
 * writes the current positions in original atom order to device memory,
 * d_out = n x 3 float (as_double = 0) or double (as_double = 1).
 * Asynchronous on the context's stream (no host copy, no synchronisation),
 * so frames can be collected in a device ring and copied out in blocks. */
int nd_sim_record_positions(nd_sim *sim, void *d_out, int32_t as_double);

/* ---- differentiable simulation: adjoint velocity Verlet (no reference
 * counterpart: NaiveDynamics.jl has no gradients; this differentiates the
 * simulate! step of Simulator.jl:198-225 with textbook forces) ----
 * Forward: unit masses, no walls, shifted-force LJ (eps, sigma, rc) and
 * shifted-force Coulomb (k_coul, rc_coul; 0 disables), velocity Verlet with
 * step dt for nsteps steps in FP32 or FP64 state, pair lists of radius
 * max(rc, rc_coul, rdf_rmax + rdf_h) + skin rebuilt every rebuild_every steps,
 * (x, v) checkpoints every checkpoint_every steps (a multiple of
 * rebuild_every).  Loss: L = sum_b (g_b - t_b)^2 over nbins bins of width
 * rdf_rmax / nbins, g_b = V / (n(n-1)/2 4 pi r_b^2 dr) sum_{i<j} K(r_ij - r_b),
 * K(d) = (1 - (d/h)^2)^3 for |d| < h = rdf_h, V = volume. */
typedef struct {
    int32_t n;               /* atoms (a multiple of 4, >= 8) */
    int32_t precision;       /* ND_PREC_FP32 or ND_PREC_FP64 (state and forward pair maths) */
    int32_t nsteps;
    int32_t checkpoint_every;
    int32_t rebuild_every;
    int32_t nbins;
    double eps, sigma, rc;   /* shifted-force LJ */
    double k_coul, rc_coul;  /* shifted-force Coulomb */
    double skin, dt;
    double rdf_rmax, rdf_h, volume;
} nd_adj_params;

typedef struct nd_adj nd_adj;

int nd_adj_create(nd_ctx *ctx, const nd_adj_params *p, nd_adj **out);
int nd_adj_destroy(nd_adj *adj);
/* Replace eps and sigma (the other parameters are fixed at creation). */
int nd_adj_set_potential(nd_adj *adj, double eps, double sigma);
/* Run the forward model from (x0, v0) (n x 3, original order) with charges q
 * (n, NULL = 0); writes the final radial distribution (nbins) and the loss
 * against h_target (nbins; NULL: no loss, *h_loss = 0). */
int nd_adj_forward(nd_adj *adj, const double *h_x0, const double *h_v0, const double *h_q,
                   const double *h_target, double *h_rdf, double *h_loss);
/* Final (x, v) of the last forward run. */
int nd_adj_final_state(nd_adj *adj, double *h_x, double *h_v);
/* Gradients of the last forward run's loss: dL/dx0, dL/dv0 (n x 3), dL/dq
 * (n) and dL/d(eps, sigma) (2); any output may be NULL.  Recomputes each
 * checkpoint segment; a new forward run is needed before the next call. */
int nd_adj_backward(nd_adj *adj, double *h_dx0, double *h_dv0, double *h_dq, double *h_dparams);

#ifdef __cplusplus
}
#endif
#endif
