// nd_sim.hip -- the per-step hot path of simulate! / simulate_bvh!
// (src/Simulator.jl:154-256, 327-379): LBVH neighbour list every k steps, then
// one fused kernel per step that evaluates Lennard-Jones + Coulomb pair
// forces and advances velocity Verlet (kick-drift-force-kick).
//
// HBM layout (all per-atom arrays in the current LBVH (Morton) order):
//   pos[2]  float4 / double4 (x, y, z, charge), double-buffered: a step reads
//           pos[cur] for every neighbour and writes the drifted atom to pos[alt]
//   vel[2]  float4 / double4 (vx, vy, vz, 1/m), double-buffered for re-sorting
//   frc     float4 / double4 force of the last completed step
//   orig[2] int32 original atom index of each slot
//   nbr     int32 neighbour slots, one block of [cap][64] per 64 consecutive
//           atoms: row k of a block holds the k-th neighbours of the block's
//           atoms (one contiguous 256-byte read per k for a wave), and rows of
//           different k are adjacent rather than a power of two apart
//   ncount  int32 [n]
//   tables  per workgroup of 256 consecutive atoms: the distinct neighbours of
//           its atoms (sorted indices, at most lcap of them) and their count
//           (tcount; -1: too many, the workgroup reads nbr directly)
//   nbr16   the same lists as 16-bit positions in the workgroup's table, two
//           per int32: one block of [ceil(cap/2)][64] per 64 list columns
//   slot_info  uint32 per list column: (list length << 8) | atom within the
//           workgroup; a workgroup built from LDS records orders its columns
//           by decreasing list length, others keep atom order
// Rebuild (every rebuild_every steps): key -> radix sort -> permute state ->
// LBVH (leaf boxes +-r_list/2, nd_lbvh.hip) -> one lane per atom walks the
// tree from the root and appends every slot j != i with |x_i - x_j|^2 <
// r_list^2 to its list (full list: each pair appears in both atoms' lists, so
// the force kernel needs no atomics and is deterministic).
//
// Precision forms: FP32 (float state and pair maths), MIXED (pair maths in
// FP32 on exact fixed-point coordinate differences, per-atom force / position
// / velocity accumulation in FP64), FP64 (everything in double).  Energies and the virial
// are reduced per workgroup in FP64 and summed with Neumaier compensation.
#include <cmath>
#include <cstddef>
#include <type_traits>
#include <vector>

#include "nd_internal.hpp"

namespace {

constexpr int kModeForce = 0;  // F(x_n), store F, energy at n
constexpr int kModeFused = 1;  // F(x_n), finish v_n, [energy], start step: v_{n+1/2}, x_{n+1}
constexpr int kModeFinal = 2;  // F(x_n), finish v_n, [energy], store F

// The first three words are the list walk's counters (k_nlist3 takes them as
// an int32 triple: maximum count, overflow flag, status bits).
struct SimMisc {
    int32_t maxcnt;
    int32_t overflow;
    int32_t status;
    int32_t rowctr;
    int64_t step;
    int64_t partial_step;
    double ke;
    // set by a drift when an atom's path length since the last list build
    // reaches half the skin while the list is still to be used (see path_add)
    int32_t stale;
    // workgroups of a table build whose records did not fit in LDS (the
    // slower global-records path); never reset, read at a checked rebuild
    int32_t rec_spill;
    double pad[2];
};
static_assert(offsetof(SimMisc, maxcnt) == 0 && offsetof(SimMisc, overflow) == 4 &&
                  offsetof(SimMisc, status) == 8,
              "the walk's counter triple");

struct StepParams {
    int32_t n, cap, walls, pad;
    float eps24, eps4, sigma2, rc2, rcc2, rcmax2, kc, elj_shift, ecoul_inv_rc;
    float dtf, lof, hif;
    float fx_inv;  // fixed-point unit in coordinate units (MIXED)
    float path_scale;  // path-length units per coordinate unit (2^14 units = skin / 2)
    double deps24, deps4, dsigma2, drc2, drcc2, drcmax2, dkc, delj_shift, decoul_inv_rc;
    double dt, lo, hi;
    double fx_c, fx_s;  // fixed-point frame: X = round((x - fx_c) * fx_s) (MIXED)
};

// MIXED keeps a fixed-point copy of the positions for its pair kernel:
// int32 (X, Y, Z) with X = round((x - c) s) and the charge's bits in w.
// A neighbour gather then moves 16 bytes (as in FP32), and X_i - X_j is exact
// in integers, so the coordinate difference reaches the FP32 pair maths with
// two roundings (int -> float, times 1/s) on top of the frame's resolution
// 1/s, instead of the error of FP32 absolute coordinates.  A coordinate
// outside the frame sets the domain flag.
__device__ __forceinline__ int32_t to_fix(double x, const StepParams &P, SimMisc *misc);

// Wall boxes of packed replicas: atom slot i belongs to replica
// rep[orig[i]] (-1: no walls) with box boxes[6 r .. 6 r + 5]; rep == nullptr
// means one system with the box [lo, hi]^3.
struct Walls {
    const int32_t *orig;
    const int32_t *rep;
    const double *boxes;
};

// boundary_reflect! (Simulator.jl:81-111): a coordinate past a wall is put
// on the wall and its velocity component reversed.
template <class P4, class PosS>
__device__ __forceinline__ void reflect(P4 &x, P4 &v, const StepParams &P, const Walls &W, int32_t i) {
    PosS lo[3], hi[3];
    if (W.rep) {
        const int32_t r = W.rep[W.orig[i]];
        if (r < 0) return;
        const double *b = W.boxes + 6 * (int64_t)r;
#pragma unroll
        for (int d = 0; d < 3; ++d) { lo[d] = (PosS)b[d]; hi[d] = (PosS)b[3 + d]; }
    } else {
        const PosS l = sizeof(PosS) == 4 ? (PosS)P.lof : (PosS)P.lo;
        const PosS h = sizeof(PosS) == 4 ? (PosS)P.hif : (PosS)P.hi;
#pragma unroll
        for (int d = 0; d < 3; ++d) { lo[d] = l; hi[d] = h; }
    }
    if (x.x < lo[0]) { x.x = lo[0]; v.x = -v.x; }
    if (x.x > hi[0]) { x.x = hi[0]; v.x = -v.x; }
    if (x.y < lo[1]) { x.y = lo[1]; v.y = -v.y; }
    if (x.y > hi[1]) { x.y = hi[1]; v.y = -v.y; }
    if (x.z < lo[2]) { x.z = lo[2]; v.z = -v.z; }
    if (x.z > hi[2]) { x.z = hi[2]; v.z = -v.z; }
}

// List validity.  Each drift adds the atom's displacement |x_new - x_old|
// (after the walls) to its path length since the last list build, kept as a
// 16-bit count of units of (skin / 2) / 2^14, rounded up and saturating; the
// path length bounds how far the atom can be from where it was when the list
// was built.  While every atom has moved less than skin / 2, every pair
// within the cutoff was within cutoff + skin at the build, so it is in the
// list.  A drift whose positions will be used with the current list (check;
// false right before a rebuild) flags misc->stale when the atom reaches
// 2^14 units; the host then runs the steps again, rebuilding where needed.
// first: the first drift after a build (the path starts at 0).
__device__ __forceinline__ void path_add(uint16_t *__restrict__ path, int32_t i, float d, float scale,
                                         int32_t first, int32_t check, SimMisc *misc) {
    const float t = fminf(ceilf(d * scale), 65535.0f);  // NaN -> 65535
    const uint32_t q = first ? 0u : (uint32_t)path[i];
    const uint32_t s = min(q + (uint32_t)t, 65535u);
    path[i] = (uint16_t)s;
    if (check && s >= 16384u) atomicOr(&misc->stale, 1);
}

template <class P4>
__device__ __forceinline__ float displacement(const P4 &a, const P4 &b) {
    const double dx = (double)a.x - (double)b.x, dy = (double)a.y - (double)b.y,
                 dz = (double)a.z - (double)b.z;
    return (float)sqrt(dx * dx + dy * dy + dz * dz);
}
template <>
__device__ __forceinline__ float displacement<float4>(const float4 &a, const float4 &b) {
    const float dx = a.x - b.x, dy = a.y - b.y, dz = a.z - b.z;
    return sqrtf(dx * dx + dy * dy + dz * dz);
}

template <class S> struct V4;
template <> struct V4<float> { using T = float4; };
template <> struct V4<double> { using T = double4; };

__device__ __forceinline__ float rsq_(float x) { return __builtin_amdgcn_rsqf(x); }
__device__ __forceinline__ double rsq_(double x) { return 1.0 / sqrt(x); }

template <class C> struct PP;
template <> struct PP<float> {
    __device__ static float eps24(const StepParams &p) { return p.eps24; }
    __device__ static float eps4(const StepParams &p) { return p.eps4; }
    __device__ static float sigma2(const StepParams &p) { return p.sigma2; }
    __device__ static float rc2(const StepParams &p) { return p.rc2; }
    __device__ static float rcc2(const StepParams &p) { return p.rcc2; }
    __device__ static float rcmax2(const StepParams &p) { return p.rcmax2; }
    __device__ static float kc(const StepParams &p) { return p.kc; }
    __device__ static float eljs(const StepParams &p) { return p.elj_shift; }
    __device__ static float ecs(const StepParams &p) { return p.ecoul_inv_rc; }
};
template <> struct PP<double> {
    __device__ static double eps24(const StepParams &p) { return p.deps24; }
    __device__ static double eps4(const StepParams &p) { return p.deps4; }
    __device__ static double sigma2(const StepParams &p) { return p.dsigma2; }
    __device__ static double rc2(const StepParams &p) { return p.drc2; }
    __device__ static double rcc2(const StepParams &p) { return p.drcc2; }
    __device__ static double rcmax2(const StepParams &p) { return p.drcmax2; }
    __device__ static double kc(const StepParams &p) { return p.dkc; }
    __device__ static double eljs(const StepParams &p) { return p.delj_shift; }
    __device__ static double ecs(const StepParams &p) { return p.decoul_inv_rc; }
};

// Sum of 4 doubles over the workgroup; lane 0 of wave 0 holds the result.
__device__ __forceinline__ void block_sum4(double v[4], double (*sh)[4]) {
#pragma unroll
    for (int c = 0; c < 4; ++c)
        for (int off = 32; off > 0; off >>= 1) v[c] += __shfl_down(v[c], off, 64);
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (lane == 0)
#pragma unroll
        for (int c = 0; c < 4; ++c) sh[w][c] = v[c];
    __syncthreads();
    if (threadIdx.x == 0)
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            double s = 0.0;
            for (int k = 0; k < (int)(blockDim.x >> 6); ++k) s += sh[k][c];
            v[c] = s;
        }
}

// Workgroups are dispatched round-robin over the 8 XCDs of an MI355X; remap
// the workgroup index so that XCD x processes one contiguous slab of atoms.
// Atoms are in Morton order, so each XCD's neighbour gathers then stay in a
// compact region whose positions fit its own L2.
__device__ __forceinline__ int32_t xcd_block(int32_t b, int32_t nblocks) {
    constexpr int32_t kXcd = 8;
    const int32_t per = nblocks / kXcd, rem = nblocks % kXcd;
    const int32_t x = b % kXcd, k = b / kXcd;
    return x * per + (x < rem ? x : rem) + k;
}

// One neighbour's contribution: LJ (r < rc) + Coulomb (r < rc_coul), both
// evaluated without branches and masked, so the loads of the next neighbours
// can be issued before this one finishes.  One reciprocal square root per
// pair gives 1/r, 1/r^2 and 1/r^6; with c12 = 48 eps sigma^12 and
// c6 = 24 eps sigma^6
//   F/r = (1/r^2) [ (1/r^6)(c12/r^6 - c6) + k qi qj / r ]
// (the same force as 24 eps (2 (sigma/r)^12 - (sigma/r)^6) / r^2 + k qi qj / r^3).
// ONE: both cutoffs are equal, so a single compare masks the sum.
template <class PosS, class CalcS, class AccS, bool ENERGY>
struct PairAcc {
    using P4 = typename V4<PosS>::T;
    CalcS rc2, rcc2, c12, c6, e12, e6, kcq, eljs, ecs;
    P4 pi;
    int4 pfx;     // MIXED: fixed-point self position
    float inv;    // MIXED: fixed-point unit
    AccS fx = 0, fy = 0, fz = 0, e = 0, w = 0;
    // MIXED: FP32 force sums over two batches of neighbours (8 pairs), added
    // to the FP64 sums by flush() (one conversion and FP64 add per 8 pairs)
    static constexpr bool kBatchSums = sizeof(AccS) != sizeof(CalcS);
    CalcS bx = 0, by = 0, bz = 0;
    int32_t np = 0;
    __device__ __forceinline__ void flush() {
        if constexpr (kBatchSums) {
            fx += (AccS)bx;
            fy += (AccS)by;
            fz += (AccS)bz;
            bx = by = bz = 0;
        }
    }
    __device__ __forceinline__ void init(const StepParams &P) {
        using Q = PP<CalcS>;
        rc2 = Q::rc2(P);
        rcc2 = Q::rcc2(P);
        const CalcS s2 = Q::sigma2(P), s6 = s2 * s2 * s2;
        c6 = Q::eps24(P) * s6;
        c12 = (CalcS)2 * c6 * s6;
        e6 = Q::eps4(P) * s6;
        e12 = e6 * s6;
        kcq = Q::kc(P) * (CalcS)pi.w;
        eljs = Q::eljs(P);
        ecs = Q::ecs(P);
    }
    // `valid` false: a padding slot (the neighbour is the atom itself), which
    // contributes nothing
    template <bool ONE> __device__ __forceinline__ void add(const P4 &pj, bool valid) {
        add_d<ONE>((CalcS)(pi.x - pj.x), (CalcS)(pi.y - pj.y), (CalcS)(pi.z - pj.z), (CalcS)pj.w,
                   valid);
    }
    // fixed-point neighbour (MIXED): exact integer difference, rounded once to
    // FP32, in fixed-point units.  The unit inv is a power of two, so scaling
    // r^2 by inv^2 and the force factor by inv gives bit for bit the values
    // of scaling each difference, with one multiply fewer per pair.
    template <bool ONE> __device__ __forceinline__ void add(const int4 &pj, bool valid) {
        const CalcS ux = (CalcS)(float)(int32_t)((uint32_t)pfx.x - (uint32_t)pj.x);
        const CalcS uy = (CalcS)(float)(int32_t)((uint32_t)pfx.y - (uint32_t)pj.y);
        const CalcS uz = (CalcS)(float)(int32_t)((uint32_t)pfx.z - (uint32_t)pj.z);
        add_d<ONE, true>(ux, uy, uz, (CalcS)__int_as_float(pj.w), valid);
    }
    // SCALED: dx, dy, dz are in units of inv (MIXED)
    template <bool ONE, bool SCALED = false>
    __device__ __forceinline__ void add_d(const CalcS dx, const CalcS dy, const CalcS dz, const CalcS qj,
                                          const bool valid) {
        CalcS r2 = dx * dx + dy * dy + dz * dz;
        if constexpr (SCALED) r2 *= (CalcS)inv * (CalcS)inv;
        const CalcS ir = rsq_(r2);
        const CalcS ir2 = ir * ir;
        const CalcS ir6 = ir2 * ir2 * ir2;
        const CalcS qq = kcq * qj;
        const bool in_lj = (r2 < rc2) & valid, in_c = ONE ? in_lj : (r2 < rcc2) & valid;
        // both terms are computed for every neighbour and then masked with a
        // select (a conditional expression over unevaluated operands would
        // become a branch per pair)
        const CalcS lj_all = ir6 * (c12 * ir6 - c6);
        CalcS fs;
        if (ONE) {
            const CalcS t = (qq * ir + lj_all) * ir2;
            fs = in_lj ? t : (CalcS)0;
        } else {
            // the outer select also clears padding slots (r = 0: 0 * inf)
            const CalcS lj = in_lj ? lj_all : (CalcS)0;
            const CalcS qm = in_c ? qq : (CalcS)0;
            const CalcS t = (qm * ir + lj) * ir2;
            fs = (in_lj | in_c) ? t : (CalcS)0;
        }
        const CalcS fd = SCALED ? fs * (CalcS)inv : fs;
        if constexpr (kBatchSums) {
            bx += fd * dx;
            by += fd * dy;
            bz += fd * dz;
        } else {
            fx += fd * dx;
            fy += fd * dy;
            fz += fd * dz;
        }
        if (ENERGY) {
            const CalcS elj_all = ir6 * (e12 * ir6 - e6) - eljs;
            const CalcS ec_all = qq * (ir - ecs);
            const CalcS elj = in_lj ? elj_all : (CalcS)0;
            const CalcS ec = in_c ? ec_all : (CalcS)0;
            e += (AccS)(elj + ec);
            w += (AccS)(fs * r2);
            np += (in_lj | in_c) ? 1 : 0;
        }
    }
};

constexpr int kPairBatch = 4;  // neighbours per pipeline batch (FP32 may pick 2 or 8)
// Rows of a 64-atom block of the 16-bit list: two entries per row for up to
// cap entries, plus two rows of sentinel entries (the step kernel's last
// batch may run up to three entries past the wave's longest list).
__host__ __device__ __forceinline__ int32_t rows16(int32_t cap) { return (cap + 1) / 2 + 2; }
// Word 3 of a raw buffer resource on gfx9-family GPUs (32-bit data format;
// untyped loads ignore it otherwise).
constexpr int32_t kBufWord3 = 0x00020000;

// PosS: storage scalar of positions/velocities; CalcS: pair arithmetic;
// AccS: per-atom accumulation.  FP32 = <float,float,float>,
// MIXED = <double,float,double>, FP64 = <double,double,double>.
// LPA (lanes per atom) 1: one lane walks an atom's whole list, 256 threads
// per workgroup.  LPA 2: 512 threads for the same 256 atoms and table; waves
// 2b and 2b + 1 take the 32 + 32 atoms of list block b, and lanes l and l + 32
// of a wave split one atom's entries (entries 4m, 4m + 1 and 4m + 2, 4m + 3,
// i.e. alternate 16-bit list rows), then add their sums in a fixed order:
// twice the waves in flight for the same work (one lane per atom gives 16
// waves per CU at 262,144 atoms).
template <class PosS, class CalcS, class AccS, int MODE, bool ENERGY, int PB = kPairBatch, int LPA = 1>
__global__ void __launch_bounds__(nd::kBlock * LPA)
k_step(const typename V4<PosS>::T *__restrict__ pos_in, typename V4<PosS>::T *__restrict__ pos_out,
       typename V4<PosS>::T *__restrict__ vel, typename V4<PosS>::T *__restrict__ frc,
       const int32_t *__restrict__ nbr, StepParams P, double *__restrict__ partials,
       SimMisc *__restrict__ misc, const int4 *__restrict__ fix_in, int4 *__restrict__ fix_out, Walls W, uint16_t *__restrict__ path, int32_t first, int32_t check,
       const uint32_t *__restrict__ nbr16, const int32_t *__restrict__ tables,
       const int32_t *__restrict__ tcount, int32_t lcap, const uint32_t *__restrict__ slot_info) {
    using P4 = typename V4<PosS>::T;
    // MIXED gathers the fixed-point copy (16 B per neighbour)
    constexpr bool FX = sizeof(PosS) == 8 && sizeof(CalcS) == 4;
    using G = std::conditional_t<FX, int4, P4>;
    constexpr int NT = nd::kBlock * LPA;  // threads per workgroup of 256 atoms
    __shared__ double sh[NT / 64][4];
    extern __shared__ __align__(16) char s_table[];  // lcap neighbour records (G)
    const int32_t blk = xcd_block(blockIdx.x, gridDim.x);
    const int32_t tid = threadIdx.x;
    const int32_t half = LPA == 2 ? (tid >> 5) & 1 : 0;
    // the list column (slot) of this lane, and the atom whose list the table
    // build put there: slot_info = (list length << 8) | atom within the
    // workgroup (the table build orders a workgroup's lists by length, so
    // that each wave's lists are of similar length)
    const int32_t is = LPA == 2 ? blk * nd::kBlock + ((tid >> 7) << 6) + (((tid >> 6) & 1) << 5) + (tid & 31)
                                : blk * nd::kBlock + tid;
    const uint32_t sinfo = slot_info[is];
    const int32_t i = blk * nd::kBlock + (int32_t)(sinfo & 0xFFu);
    const int32_t n = P.n;
    AccS fx = 0, fy = 0, fz = 0;
    double pe = 0.0, vir = 0.0, ke = 0.0, np = 0.0;
    // list length (the wave's block of the list holds rows of 64 int32), and
    // the wave's longest list: rows past it are never needed, so the
    // look-ahead re-reads the last needed row (a cache hit) instead
    const int32_t cnt = i < n ? min((int32_t)(sinfo >> 8), P.cap) : 0;
    // this lane's share of the list: entries e with (e >> 1) & 1 == half
    // (LPA 2; local entry k is e = 4 (k >> 1) + 2 half + (k & 1)), or all
    const int32_t cl = LPA == 1 ? cnt : ((cnt >> 2) << 1) + min(max((cnt & 3) - 2 * half, 0), 2);
    auto entry_of = [&](int32_t k) -> int32_t { return LPA == 1 ? k : ((k >> 1) << 2) + 2 * half + (k & 1); };
    int32_t wmax = cnt;
    for (int off = 32; off > 0; off >>= 1) wmax = max(wmax, __shfl_xor(wmax, off, 64));
    wmax = __builtin_amdgcn_readfirstlane(wmax);
    const int32_t last = max(wmax - 1, 0);
    const __amdgpu_buffer_rsrc_t prs = __builtin_amdgcn_make_buffer_rsrc(
        FX ? (void *)fix_in : (void *)pos_in, 0, n * (int32_t)sizeof(G), kBufWord3);
    // neighbour record j (a sorted index; -1 gives zeros)
    auto gather = [&](int32_t j) -> G {
        const uint32_t off = (uint32_t)j * (uint32_t)sizeof(G);
        G g;
        if constexpr (sizeof(G) == 16) {
            const auto a = __builtin_amdgcn_raw_buffer_load_b128(prs, off, 0, 0);
            __builtin_memcpy(&g, &a, 16);
        } else {
            const auto a = __builtin_amdgcn_raw_buffer_load_b128(prs, off, 0, 0);
            const auto b = __builtin_amdgcn_raw_buffer_load_b128(prs, off + 16u, 0, 0);
            __builtin_memcpy(&g, &a, 16);
            __builtin_memcpy(reinterpret_cast<char *>(&g) + 16, &b, 16);
        }
        return g;
    };
    // this workgroup's neighbourhood table (k_table) staged in LDS; -1: none
    // (more distinct neighbours than lcap), the lists hold sorted indices.
    // Slot U is the sentinel every padding entry names.  FP64: a position so
    // far away that the cutoff masks it, so its table loop needs no per-pair
    // count check (measured faster).  FP32 and MIXED keep the count check (for
    // FP32 dropping it measured slower: the compiler's schedule of the loop
    // waits more; MIXED's fixed-point differences would wrap) and a zero slot.
    const int32_t U = __builtin_amdgcn_readfirstlane(tcount[blk]);
    G *lt = reinterpret_cast<G *>(s_table);
    constexpr bool kSentinel = sizeof(G) == 32;
    // this lane's list: block is >> 6 (rows of 64 columns), column is & 63
    const int32_t wave = __builtin_amdgcn_readfirstlane(is >> 6);
    const uint32_t lane4 = (uint32_t)(is & 63) * 4u;
    // 16-bit table positions, two per int32 row entry; rows up to the wave's
    // last, then two sentinel rows, which only the unchecked (FP64) loop
    // reads: the checked loops re-read the wave's last row for their
    // look-ahead past it
    const int32_t cap2 = rows16(P.cap);
    const __amdgpu_buffer_rsrc_t l16 = __builtin_amdgcn_make_buffer_rsrc(
        (void *)(nbr16 + (int64_t)wave * cap2 * 64), 0, cap2 * 256, kBufWord3);
    const int32_t last2 = kSentinel ? max((wmax + 1) >> 1, 1) + 1 : last >> 1;
    auto entries16 = [&](int32_t r, int32_t *j) {
        static_assert(PB % 2 == 0, "two entries per word");
#pragma unroll
        for (int u = 0; u < PB; u += 2) {
            const int32_t q = entry_of(r + u) >> 1;  // both entries of the row are this lane's
            const uint32_t row = (uint32_t)(q < last2 ? q : last2);
            const uint32_t v = __builtin_amdgcn_raw_buffer_load_b32(l16, lane4, row * 256u, 0);
            j[u] = (int32_t)(v & 0xFFFFu);
            j[u + 1] = (int32_t)(v >> 16);
        }
    };
    // the first two batches of the lane's entries are in flight while the
    // table is staged (waves whose block holds an atom: uniform per wave)
    int32_t pja[PB], pjb[PB];
    if (U >= 0 && (wave << 6) < n) {
        entries16(0, pja);
        entries16(PB, pjb);
    }
    if (U >= 0) {
        // kStage index loads, then kStage gathers, then the LDS writes: two
        // memory latencies per kStage slots instead of two per slot.  The
        // loads are unconditional, so all of them are in flight together: a
        // slot past U reads past the buffer range of the table's used part
        // (zeros, no memory access) and then gathers index -1 (the same).
        const __amdgpu_buffer_rsrc_t trs =
            __builtin_amdgcn_make_buffer_rsrc((void *)(tables + (int64_t)blk * lcap), 0, U * 4, kBufWord3);
        constexpr int kStage = 8;
        for (int32_t t0 = tid; t0 < U; t0 += kStage * NT) {
            int32_t j[kStage];
#pragma unroll
            for (int u = 0; u < kStage; ++u) {
                const int32_t t = t0 + u * NT;
                const int32_t v = (int32_t)__builtin_amdgcn_raw_buffer_load_b32(trs, (uint32_t)t * 4u, 0, 0);
                j[u] = t < U ? v : -1;
            }
            G g[kStage];
#pragma unroll
            for (int u = 0; u < kStage; ++u) g[u] = gather(j[u]);
#pragma unroll
            for (int u = 0; u < kStage; ++u) {
                const int32_t t = t0 + u * NT;
                if (t < U) lt[t] = g[u];
            }
        }
        if (tid == 0) {
            G z = gather(-1);
            if constexpr (kSentinel) {
                z.x = z.y = z.z = (PosS)1e18;
                z.w = 0;
            }
            lt[U] = z;
        }
        __syncthreads();
    }
    if (i < n) {
        PairAcc<PosS, CalcS, AccS, ENERGY> acc;
        acc.pi = pos_in[i];
        if (FX) {
            acc.pfx = fix_in[i];
            acc.inv = P.fx_inv;
        }
        acc.init(P);
        // Software pipeline over batches of PB neighbours: while
        // batch b's pairs are computed, the records of batch b + 1 and the
        // list entries of batch b + 2 are in flight, so one memory latency
        // (not a list load followed by a gather) separates consecutive
        // batches.  List rows go through buffer loads (the row offset is a
        // scalar operand).  Direct lists: slots past the atom's count gather
        // index -1, past the end of the buffer, which the hardware range check
        // answers with zeros and no memory access; table lists: they read
        // table record 0.  The pair maths masks them either way.
        // Two register sets (a, b) take turns, so no batch is copied.  pa,
        // pb: the first two batches of entries already loaded (or null).
        auto pipeline = [&](auto one, auto entries, auto record, auto checked, const int32_t *pa,
                            const int32_t *pb) {
            constexpr bool ONE = decltype(one)::value;
            constexpr bool CHK = decltype(checked)::value;
            constexpr int B = PB;
            int32_t ja[B], jb[B];
            G qa[B], qb[B];
            if (pa) {
#pragma unroll
                for (int u = 0; u < B; ++u) ja[u] = pa[u];
            } else {
                entries(0, ja);
            }
#pragma unroll
            for (int u = 0; u < B; ++u) qa[u] = record(ja[u]);
            if (pb) {
#pragma unroll
                for (int u = 0; u < B; ++u) jb[u] = pb[u];
            } else {
                entries(B, jb);
            }
            for (int32_t k = 0;; k += 2 * B) {
#pragma unroll
                for (int u = 0; u < B; ++u) qb[u] = record(jb[u]);
                entries(k + 2 * B, ja);
#pragma unroll
                for (int u = 0; u < B; ++u) acc.template add<ONE>(qa[u], CHK ? k + u < cl : true);
                if (k + B >= cl) break;
#pragma unroll
                for (int u = 0; u < B; ++u) qa[u] = record(ja[u]);
                entries(k + 3 * B, jb);
#pragma unroll
                for (int u = 0; u < B; ++u) acc.template add<ONE>(qb[u], CHK ? k + B + u < cl : true);
                acc.flush();  // MIXED: FP32 sums over 2 batches, then FP64
                if (k + 2 * B >= cl) break;
            }
            acc.flush();  // after an exit at the first batch (a repeat adds zeros)
        };
        auto run = [&](auto one) {
            if (U >= 0) {
                pipeline(one, entries16, [&](int32_t j) -> G { return lt[j]; },
                         std::integral_constant<bool, !kSentinel>{}, pja, pjb);
            } else {
                const __amdgpu_buffer_rsrc_t lrs = __builtin_amdgcn_make_buffer_rsrc(
                    (void *)(nbr + (int64_t)wave * P.cap * 64), 0, P.cap * 256, kBufWord3);
                auto entries = [&](int32_t r, int32_t *j) {
#pragma unroll
                    for (int u = 0; u < PB; ++u) {
                        const int32_t e = entry_of(r + u);
                        const uint32_t row = (uint32_t)(e < last ? e : last);
                        const int32_t v = (int32_t)__builtin_amdgcn_raw_buffer_load_b32(lrs, lane4, row * 256u, 0);
                        j[u] = e < cnt ? v : -1;
                    }
                };
                pipeline(one, entries, gather, std::true_type{}, nullptr, nullptr);
            }
        };
        // equal cutoffs (uniform over the launch): one mask per pair
        if (acc.rc2 == acc.rcc2) run(std::true_type{});
        else run(std::false_type{});
        fx = acc.fx;
        fy = acc.fy;
        fz = acc.fz;
        if constexpr (LPA == 2) {
            // half 0 adds half 1's sums (a fixed order: deterministic)
            const AccS ox = __shfl_xor(fx, 32, 64), oy = __shfl_xor(fy, 32, 64), oz = __shfl_xor(fz, 32, 64);
            fx = fx + ox;
            fy = fy + oy;
            fz = fz + oz;
        }
        const AccS epair = acc.e, vpair = acc.w;
        const int32_t npair = acc.np;
        const P4 pi = acc.pi;
        // each unordered pair is visited from both atoms: halve the pair sums
        pe = 0.5 * (double)epair;
        vir = 0.5 * (double)vpair;
        np = 0.5 * (double)npair;

        if (half == 0 && (MODE == kModeForce || MODE == kModeFinal)) {
            P4 f;
            f.x = (PosS)fx; f.y = (PosS)fy; f.z = (PosS)fz; f.w = 0;
            frc[i] = f;
        }
        if (half != 0) {
            // the pair sums (pe, vir, np) of this half are kept; one lane per
            // atom integrates
        } else if (MODE != kModeForce) {
            P4 v = vel[i];
            const PosS h = (PosS)0.5 * (sizeof(PosS) == 4 ? (PosS)P.dtf : (PosS)P.dt) * v.w;
            v.x += h * (PosS)fx;
            v.y += h * (PosS)fy;
            v.z += h * (PosS)fz;
            if (ENERGY) ke = 0.5 / (double)v.w * ((double)v.x * v.x + (double)v.y * v.y + (double)v.z * v.z);
            if (MODE == kModeFused) {
                const PosS dt = sizeof(PosS) == 4 ? (PosS)P.dtf : (PosS)P.dt;
                v.x += h * (PosS)fx;
                v.y += h * (PosS)fy;
                v.z += h * (PosS)fz;
                P4 x = pi;
                x.x += dt * v.x;
                x.y += dt * v.y;
                x.z += dt * v.z;
                if (P.walls) reflect<P4, PosS>(x, v, P, W, i);
                pos_out[i] = x;
                path_add(path, i, displacement(x, pi), P.path_scale, first, check, misc);
                if constexpr (FX)
                    fix_out[i] = make_int4(to_fix(x.x, P, misc), to_fix(x.y, P, misc),
                                           to_fix(x.z, P, misc), __float_as_int((float)x.w));
            }
            vel[i] = v;
        } else if (ENERGY) {
            const P4 v = vel[i];
            ke = 0.5 / (double)v.w * ((double)v.x * v.x + (double)v.y * v.y + (double)v.z * v.z);
        }
    }
    if (ENERGY) {
        double acc[4] = {pe, vir, ke, np};
        block_sum4(acc, sh);
        if (tid == 0) {
#pragma unroll
            for (int c = 0; c < 4; ++c) partials[4 * (int64_t)blockIdx.x + c] = acc[c];
        }
    }
    if (blockIdx.x == 0 && tid == 0) {
        if (ENERGY) misc->partial_step = misc->step;
        if (MODE == kModeFused) misc->step += 1;
    }
}

// Kick-drift without a force evaluation (first step of a run, or after the
// thermostat): v += dt/2 F/m; x += dt v; walls.  Optional velocity scale
// beta = sqrt(1 + gamma (T0/T - 1)) from the last recorded kinetic energy.
template <class PosS>
__global__ void __launch_bounds__(nd::kBlock)
k_kick_drift(const typename V4<PosS>::T *__restrict__ pos_in, typename V4<PosS>::T *__restrict__ pos_out,
             typename V4<PosS>::T *__restrict__ vel, const typename V4<PosS>::T *__restrict__ frc,
             StepParams P, SimMisc *__restrict__ misc, int32_t rescale, double t0, double gamma,
             int4 *__restrict__ fix_out, Walls W, uint16_t *__restrict__ path, int32_t first,
             int32_t check) {
    using P4 = typename V4<PosS>::T;
    const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < P.n) {
        P4 v = vel[i];
        if (rescale) {
            // rescale_velocity! (Simulator.jl:119-144) with T = 2 KE / (3 N k_B)
            const double T = 2.0 * misc->ke / (3.0 * (double)P.n);
            const double beta = T > 0.0 ? sqrt(fmax(0.0, 1.0 + gamma * (t0 / T - 1.0))) : 1.0;
            v.x = (PosS)(v.x * beta);
            v.y = (PosS)(v.y * beta);
            v.z = (PosS)(v.z * beta);
        }
        const P4 f = frc[i];
        const PosS dt = sizeof(PosS) == 4 ? (PosS)P.dtf : (PosS)P.dt;
        const PosS h = (PosS)0.5 * dt * v.w;
        v.x += h * f.x;
        v.y += h * f.y;
        v.z += h * f.z;
        const P4 x0 = pos_in[i];
        P4 x = x0;
        x.x += dt * v.x;
        x.y += dt * v.y;
        x.z += dt * v.z;
        if (P.walls) reflect<P4, PosS>(x, v, P, W, i);
        pos_out[i] = x;
        path_add(path, i, displacement(x, x0), P.path_scale, first, check, misc);
        if (fix_out)
            fix_out[i] = make_int4(to_fix((double)x.x, P, misc), to_fix((double)x.y, P, misc),
                                   to_fix((double)x.z, P, misc), __float_as_int((float)x.w));
        vel[i] = v;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) misc->step += 1;
}

__device__ __forceinline__ int32_t to_fix(double x, const StepParams &P, SimMisc *misc) {
    const double t = rint((x - P.fx_c) * P.fx_s);
    if (!(t > -2147483520.0 && t < 2147483520.0)) {
        atomicOr(&misc->status, nd::kStatusDomain);
        return 0;
    }
    return (int32_t)t;
}

// Neumaier-compensated sum of the per-workgroup partials; appends one energy
// row (step, KE, PE, virial, pairs) and keeps KE for the thermostat.  Each
// lane sums a strided share with compensation, then the (sum, error) pairs
// are merged pairwise with an exact TwoSum in a fixed tree order.
__device__ __forceinline__ void two_sum_merge(double &s, double &c, double s2, double c2) {
    const double t = s + s2;
    const double bp = t - s;
    const double err = (s - (t - bp)) + (s2 - bp);
    s = t;
    c = c + c2 + err;
}

// Largest XCD slab the walk order places (blocks of 256 atoms per XCD).
constexpr int kOrderSlab = 1024;

// The next walk's block order for XCD slab x, by one workgroup (the blocks
// xcd_block gives that XCD: slab x starts at x per + min(x, rem) and holds
// per + (x < rem) blocks).  A block's weight is its four waves' work
// estimates from the walk that just ran; blocks are ranked by weight (ties
// by index, so the ranks are a permutation) and rank r goes to slab position
// snake(r): position k is the workgroup the XCD's dispatcher places k-th,
// on CU k mod cus in its (k / cus)-th slot, and odd slots take their ranks in
// reverse, so CU c of a 4-slot slab gets ranks c, 2 cus - 1 - c, 2 cus + c
// and 4 cus - 1 - c: equal rank sums, the heaviest block in the oldest slot.
__device__ __forceinline__ void walk_order_slab(const int32_t *__restrict__ work, int32_t nblocks, int32_t cus,
                                                int32_t *__restrict__ order, int32_t x, int32_t *wt) {
    constexpr int32_t kXcd = 8;
    const int32_t per = nblocks / kXcd, rem = nblocks % kXcd;
    const int32_t base = x * per + (x < rem ? x : rem), S = per + (x < rem ? 1 : 0);
    if (S > kOrderSlab || cus < 1) return;  // uniform: the order stays as it is
    for (int32_t t = threadIdx.x; t < S; t += blockDim.x) {
        const int32_t *q = work + 4 * (int64_t)(base + t);
        wt[t] = q[0] + q[1] + q[2] + q[3];
    }
    __syncthreads();
    for (int32_t t = threadIdx.x; t < S; t += blockDim.x) {
        const int32_t v = wt[t];
        int32_t r = 0;
        for (int32_t j = 0; j < S; ++j) {
            const int32_t u = wt[j];
            r += (u > v) | ((u == v) & (j < t));
        }
        const int32_t slot = r / cus, c = r - slot * cus;
        const int32_t m = min(cus, S - slot * cus);
        const int32_t k = (slot & 1) ? slot * cus + (m - 1 - c) : r;
        order[base + k] = base + t;
    }
}

// Block 0: the energy row.  Blocks 1-8 (launched after a fused walk): the
// next walk's block order, one XCD slab each (walk_order_slab), so the order
// costs no launch of its own.
__global__ void __launch_bounds__(nd::kBlock)
k_energy_finalize(const double *__restrict__ partials, int32_t nblocks, double *__restrict__ rows,
                  int32_t max_rows, SimMisc *__restrict__ misc, const int32_t *__restrict__ work,
                  int32_t wblocks, int32_t cus, int32_t *__restrict__ order) {
    if (blockIdx.x > 0) {
        __shared__ int32_t wt[kOrderSlab];
        walk_order_slab(work, wblocks, cus, order, (int32_t)blockIdx.x - 1, wt);
        return;
    }
    __shared__ double ss[4][nd::kBlock], cc[4][nd::kBlock];
    double s[4] = {0, 0, 0, 0}, c[4] = {0, 0, 0, 0};
    for (int32_t b = threadIdx.x; b < nblocks; b += blockDim.x)
#pragma unroll
        for (int q = 0; q < 4; ++q) two_sum_merge(s[q], c[q], partials[4 * (int64_t)b + q], 0.0);
#pragma unroll
    for (int q = 0; q < 4; ++q) { ss[q][threadIdx.x] = s[q]; cc[q][threadIdx.x] = c[q]; }
    __syncthreads();
    for (int half = nd::kBlock / 2; half > 0; half >>= 1) {
        if ((int)threadIdx.x < half) {
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                double a = ss[q][threadIdx.x], e = cc[q][threadIdx.x];
                two_sum_merge(a, e, ss[q][threadIdx.x + half], cc[q][threadIdx.x + half]);
                ss[q][threadIdx.x] = a;
                cc[q][threadIdx.x] = e;
            }
        }
        __syncthreads();
    }
    if (threadIdx.x != 0) return;
    double tot[4];
    for (int q = 0; q < 4; ++q) tot[q] = ss[q][0] + cc[q][0];
    const int32_t r = misc->rowctr;
    if (r < max_rows) {
        double *row = rows + 5 * (int64_t)r;
        row[0] = (double)misc->partial_step;
        row[1] = tot[2];
        row[2] = tot[0];
        row[3] = tot[1];
        row[4] = tot[3];
    }
    misc->rowctr = r + 1;
    misc->ke = tot[2];
}

// Key of every atom from its (possibly FP64) position rounded to FP32.
template <class PosS>
__global__ void __launch_bounds__(nd::kBlock)
k_sim_keys(const typename V4<PosS>::T *__restrict__ pos, int32_t n, int32_t mode, float lo,
           float scale, int32_t drop, int32_t *__restrict__ codes, int32_t *__restrict__ idx,
           SimMisc *__restrict__ misc, int32_t *__restrict__ slots, int32_t nslots) {
    const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) { misc->maxcnt = 0; }
    if (i < nslots) slots[i] = 0;  // the tree build's CAS slots (nslots < n)
    if (i >= n) return;
    const auto p = pos[i];
    const float x = (float)p.x, y = (float)p.y, z = (float)p.z;
    int32_t code;
    bool ok;
    if (mode == ND_MORTON_MASKED) {
        int32_t qx = 0, qy = 0, qz = 0;
        ok = nd::floor_i32(x / nd::kBinwidth, qx);
        ok &= nd::floor_i32(y / nd::kBinwidth, qy);
        ok &= nd::floor_i32(z / nd::kBinwidth, qz);
        code = (qx & nd::kMagic0) | (qy & nd::kMagic1) | (qz & nd::kMagic2);
    } else {
        ok = __builtin_isfinite(x) && __builtin_isfinite(y) && __builtin_isfinite(z);
        const uint32_t cx = nd::cell10(x, lo, scale), cy = nd::cell10(y, lo, scale),
                       cz = nd::cell10(z, lo, scale);
        code = (int32_t)((mode == ND_MORTON_HILBERT
                              ? nd::hilbert30(cx, cy, cz)
                              : (nd::spread10(cx) | (nd::spread10(cy) << 1) | (nd::spread10(cz) << 2))) >>
                         drop);
    }
    codes[i] = code;
    idx[i] = i;
    if (!ok) atomicOr(&misc->status, nd::kStatusDomain);
}

// Gather the state into the new order; also writes the FP32 copy of the
// positions that the tree build and traversal read, and (MIXED) the
// fixed-point copy that the pair kernel gathers.
// SEG > 0 (a power of two <= 64, the leaf size): also the leaf boxes of the
// new order, the FP32 positions +- h reduced over each leaf's SEG lanes with
// Julia's min / max, as the separate leaf-box kernel would compute them.
template <class PosS, int SEG>
__global__ void __launch_bounds__(nd::kBlock)
k_permute(const int32_t *__restrict__ perm, int32_t n, const typename V4<PosS>::T *__restrict__ pos_in,
          typename V4<PosS>::T *__restrict__ pos_out, const typename V4<PosS>::T *__restrict__ vel_in,
          typename V4<PosS>::T *__restrict__ vel_out, const int32_t *__restrict__ orig_in,
          int32_t *__restrict__ orig_out, float4 *__restrict__ pos32, int4 *__restrict__ fix_out,
          StepParams P, SimMisc *__restrict__ misc, float h, nd_node *__restrict__ nodes,
          float4 *__restrict__ tq, float4 *__restrict__ lc) {
    const int32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    const bool act = s < n;
    const float inf = __int_as_float(0x7F800000);
    float mn0 = inf, mn1 = inf, mn2 = inf, mx0 = -inf, mx1 = -inf, mx2 = -inf;
    float px = 0.0f, py = 0.0f, pz = 0.0f;
    if (act) {
        const int32_t o = perm[s];
        const auto p = pos_in[o];
        pos_out[s] = p;
        vel_out[s] = vel_in[o];
        orig_out[s] = orig_in[o];
        px = (float)p.x; py = (float)p.y; pz = (float)p.z;
        if (pos32) pos32[s] = make_float4(px, py, pz, (float)p.w);
        if (fix_out)
            fix_out[s] = make_int4(to_fix((double)p.x, P, misc), to_fix((double)p.y, P, misc),
                                   to_fix((double)p.z, P, misc), __float_as_int((float)p.w));
        mn0 = px - h; mn1 = py - h; mn2 = pz - h;
        mx0 = px + h; mx1 = py + h; mx2 = pz + h;
    }
    if constexpr (SEG > 0) {
#pragma unroll
        for (int off = 1; off < SEG; off <<= 1) {
            mn0 = jl_min(mn0, __shfl_xor(mn0, off, 64));
            mn1 = jl_min(mn1, __shfl_xor(mn1, off, 64));
            mn2 = jl_min(mn2, __shfl_xor(mn2, off, 64));
            mx0 = jl_max(mx0, __shfl_xor(mx0, off, 64));
            mx1 = jl_max(mx1, __shfl_xor(mx1, off, 64));
            mx2 = jl_max(mx2, __shfl_xor(mx2, off, 64));
        }
        if (act && (s & (SEG - 1)) == 0) {
            float *q = reinterpret_cast<float *>(nodes + s / SEG);
            *reinterpret_cast<float4 *>(q) = make_float4(mn0, mn1, mn2, mx0);
            *reinterpret_cast<float2 *>(q + 4) = make_float2(mx1, mx2);
        }
        if (tq != nullptr) {
            // leaf-local copies for the two-phase walk (as k_leaf_chunk_boxes
            // writes them in a refit)
            const float cx = 0.5f * (mn0 + mx0), cy = 0.5f * (mn1 + mx1), cz = 0.5f * (mn2 + mx2);
            const float qx = px - cx, qy = py - cy, qz = pz - cz;
            float r2 = __builtin_fmaf(qz, qz, __builtin_fmaf(qy, qy, qx * qx));
            if (act) tq[s] = make_float4(-2.0f * qx, -2.0f * qy, -2.0f * qz, r2);
            if (!act) r2 = 0.0f;
#pragma unroll
            for (int off = 1; off < SEG; off <<= 1) r2 = fmaxf(r2, __shfl_xor(r2, off, 64));
            if (act && (s & (SEG - 1)) == 0) lc[s / SEG] = make_float4(cx, cy, cz, r2);
        }
    }
}

// One wave per 64 consecutive atoms (Morton-adjacent, so spatially compact):
// the wave walks the tree from the root as a unit.  The current node is
// wave-uniform (one scalar load per visited node instead of 64 divergent
// pointer chases); every lane tests its own box (atom +- r/2) against it and
// the wave descends where any lane overlaps.  At a leaf each overlapping lane
// appends every slot j != i with |x_i - x_j|^2 < r^2 to its column.  A lane
// gets exactly the list its own walk would give: leaves it does not overlap
// are skipped by its own box test.
// A = leaf size known at compile time (0: runtime `apl`).  With A > 0 the
// first min(A, 8) atoms of a leaf are fetched in the same batch of scalar
// loads as its node (leaves are numbered 1..nL, so a leaf is recognised
// before its node arrives), so a leaf visit costs one memory latency.
template <int A>
__global__ void __launch_bounds__(nd::kBlock)
k_nlist(const float4 *__restrict__ pos, const nd_node *__restrict__ nodes, int32_t n, int32_t nL,
        int32_t apl, float h, float rl2, int32_t cap, int32_t *__restrict__ nbr,
        int32_t *__restrict__ ncount, SimMisc *__restrict__ misc) {
    constexpr int PF = (A > 0 && A < 8) ? A : 8;
    const int32_t la = A > 0 ? A : apl;
    // contiguous Morton slab per XCD: its subtree and positions stay in its L2
    const int32_t i = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    const bool active = i < n;
    // the wave's block of the list: cap rows of 64 int32 (row k = k-th
    // neighbours of the wave's 64 atoms), so rows never sit a power of two apart
    int32_t *lst = nbr + ((int64_t)(i >> 6) * cap) * 64 + (i & 63);
    const float4 p = active ? pos[i] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    const float qmn0 = p.x - h, qmn1 = p.y - h, qmn2 = p.z - h;
    const float qmx0 = p.x + h, qmx1 = p.y + h, qmx2 = p.z + h;
    int32_t cnt = 0;
    auto test = [&](const float4 &q, int32_t j) {
        const float dx = p.x - q.x, dy = p.y - q.y, dz = p.z - q.z;
        const float s = dx * dx + dy * dy + dz * dz;
        if (j != i && s < rl2) {
            if (cnt < cap) lst[cnt * 64] = j;
            ++cnt;
        }
    };
    int32_t t = nL + 1;  // root
    int32_t visits = 0;
    while (t != 0) {
        t = __builtin_amdgcn_readfirstlane(t);
        // each node is visited at most once per walk (DFS order)
        if (++visits > 2 * nL || t < 1 || t > 2 * nL - 1) {
            if ((threadIdx.x & 63) == 0) atomicOr(&misc->status, nd::kStatusTree);
            break;
        }
        const float4 *np4 = reinterpret_cast<const float4 *>(nodes + (t - 1));
        const float4 a = np4[0];
        const float4 b = np4[1];
        const bool leaf = t <= nL;
        const int32_t base = (t - 1) * la;
        float4 q[PF];
        if (A > 0 && leaf) {
#pragma unroll
            for (int m = 0; m < PF; ++m) q[m] = pos[base + m];
        }
        const bool ov = active & (qmn0 < a.w) & (qmx0 > a.x) & (qmn1 < b.x) & (qmx1 > a.y) &
                        (qmn2 < b.y) & (qmx2 > a.z);
        const int32_t left = __float_as_int(b.z), skip = __float_as_int(b.w);
        if (__ballot(ov) == 0ull) {
            t = skip;
        } else if (left != 0) {
            t = left;
        } else {
            if (ov) {
                if (A > 0) {
#pragma unroll
                    for (int m = 0; m < PF; ++m) test(q[m], base + m);
                    for (int m0 = PF; m0 < A; m0 += 8) {
                        float4 r8[8];
#pragma unroll
                        for (int m = 0; m < 8; ++m) r8[m] = pos[base + m0 + m];
#pragma unroll
                        for (int m = 0; m < 8; ++m) test(r8[m], base + m0 + m);
                    }
                } else {
                    for (int32_t m = 0; m < la; ++m) test(pos[base + m], base + m);
                }
            }
            t = skip;
        }
    }
    if (active) ncount[i] = cnt;
    // wave-level max, one atomic per wave
    int32_t m = cnt;
    for (int off = 32; off > 0; off >>= 1) m = max(m, __shfl_down(m, off, 64));
    if ((threadIdx.x & 63) == 0 && m > 0) {
        // the counter only grows: skip the (serialising) atomic when it is
        // already at least m
        if (m > __hip_atomic_load(&misc->maxcnt, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
            atomicMax(&misc->maxcnt, m);
        if (m > cap) atomicOr(&misc->overflow, 1);
    }
}

// Child-box walk: the same wave-cooperative walk over the child-box form of
// the tree (each internal node carries both children's boxes), depth first
// with a wave-uniform stack in LDS.  One fetch tests both children, so only
// internal nodes that some lane overlaps are ever fetched: about half the
// dependent fetches of the rope walk.  The next node's fetch is issued
// before the leaf tests of the current one.
constexpr int kStackDepth = 96;

template <int A>
__global__ void __launch_bounds__(nd::kBlock)
k_nlist2(const float4 *__restrict__ pos, const float4 *__restrict__ wide, int32_t n, int32_t nL,
         int32_t apl, float h, float rl2, int32_t cap, int32_t *__restrict__ nbr,
         int32_t *__restrict__ ncount, SimMisc *__restrict__ misc) {
    __shared__ int32_t stk[nd::kBlock / 64][kStackDepth];
    int32_t *st = stk[threadIdx.x >> 6];
    const int32_t la = A > 0 ? A : apl;
    const int32_t i = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    const bool active = i < n;
    int32_t *lst = nbr + ((int64_t)(i >> 6) * cap) * 64 + (i & 63);
    const float4 p = active ? pos[i] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    const float qmn0 = p.x - h, qmn1 = p.y - h, qmn2 = p.z - h;
    const float qmx0 = p.x + h, qmx1 = p.y + h, qmx2 = p.z + h;
    int32_t cnt = 0;
    auto test = [&](const float4 &q, int32_t j) {
        const float dx = p.x - q.x, dy = p.y - q.y, dz = p.z - q.z;
        const float s = dx * dx + dy * dy + dz * dz;
        if (j != i && s < rl2) {
            if (cnt < cap) lst[cnt * 64] = j;
            ++cnt;
        }
    };
    auto leaf_work = [&](int32_t leaf, bool ov) {
        const int32_t base = (leaf - 1) * la;
        if (A > 0) {
            for (int m0 = 0; m0 < A; m0 += 8) {
                constexpr int kB = A <= 0 ? 1 : (A < 8 ? A : 8);
                float4 q[kB];
#pragma unroll
                for (int m = 0; m < kB; ++m) q[m] = pos[base + m0 + m];
                if (ov) {
#pragma unroll
                    for (int m = 0; m < kB; ++m) test(q[m], base + m0 + m);
                }
            }
        } else if (ov) {
            for (int32_t m = 0; m < la; ++m) test(pos[base + m], base + m);
        }
    };
    int32_t sp = 0;
    // the root nL + 1 is internal (nL >= 2): its record is wide[0]
    float4 a = wide[0], b = wide[1], c = wide[2], d = wide[3];
    for (int32_t iters = 0;; ++iters) {
        const int32_t Lc = __builtin_amdgcn_readfirstlane(__float_as_int(d.x));
        const int32_t Rc = __builtin_amdgcn_readfirstlane(__float_as_int(d.y));
        // every internal node is fetched at most once; children are node numbers
        if (iters >= nL || Lc < 1 || Lc > 2 * nL - 1 || Rc < 1 || Rc > 2 * nL - 1) {
            if ((threadIdx.x & 63) == 0) atomicOr(&misc->status, nd::kStatusTree);
            break;
        }
        const bool ovL = active & (qmn0 < a.w) & (qmx0 > a.x) & (qmn1 < b.x) & (qmx1 > a.y) &
                         (qmn2 < b.y) & (qmx2 > a.z);
        const bool ovR = active & (qmn0 < c.y) & (qmx0 > b.z) & (qmn1 < c.z) & (qmx1 > b.w) &
                         (qmn2 < c.w) & (qmx2 > c.x);
        const bool anyL = __ballot(ovL) != 0ull, anyR = __ballot(ovR) != 0ull;
        const bool leafL = Lc <= nL, leafR = Rc <= nL;
        int32_t next = 0;
        if (anyL && !leafL) next = Lc;
        if (anyR && !leafR) {
            if (next == 0) {
                next = Rc;
            } else {
                if (sp >= kStackDepth) {
                    if ((threadIdx.x & 63) == 0) atomicOr(&misc->status, nd::kStatusTree);
                    break;
                }
                st[sp++] = Rc;
            }
        }
        if (next == 0 && sp > 0) next = __builtin_amdgcn_readfirstlane(st[--sp]);
        if (next != 0) {
            // issue the next record's fetch before this node's leaf tests
            const float4 *nw = wide + 4 * (int64_t)(next - nL - 1);
            a = nw[0]; b = nw[1]; c = nw[2]; d = nw[3];
        }
        if (anyL && leafL) leaf_work(Lc, ovL);
        if (anyR && leafR) leaf_work(Rc, ovR);
        if (next == 0) break;
    }
    if (active) ncount[i] = cnt;
    int32_t m = cnt;
    for (int off = 32; off > 0; off >>= 1) m = max(m, __shfl_down(m, off, 64));
    if ((threadIdx.x & 63) == 0 && m > 0) {
        // the counter only grows: skip the (serialising) atomic when it is
        // already at least m
        if (m > __hip_atomic_load(&misc->maxcnt, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
            atomicMax(&misc->maxcnt, m);
        if (m > cap) atomicOr(&misc->overflow, 1);
    }
}

using nd::wave_lds_sync;

// Two-phase walk over four-wide records for leaves of A atoms (A = 4, 8, 16,
// 32).  Every internal node has a record holding its grandchildren (a child
// that is a leaf stands for itself): up to 4 entries with boxes and node
// numbers.  Phase 1 pops up to kFront records from the wave's LDS stack and
// fetches them together (lane l loads float4 l & 7 of records (l >> 3) + 8 u),
// so one memory latency covers up to kFront records (two tree levels each); every lane
// tests its sphere of radius r_list against each entry's box, overlapped internal entries are pushed,
// and an overlapped leaf is queued for each group of 8 lanes (8 consecutive
// atoms) with an overlapping lane, together with the group's 8 overlap bits.
// Phase 2 drains the 8 group queues side by side: per step each group stages
// the atoms of its next leaf in its own LDS slot (the following leaf's load
// already in flight) and each lane whose bit is set tests them, collecting
// hits in a bit mask that it then appends.  A group of 8 Morton-consecutive
// atoms overlaps about as many leaves as its busiest lane, so the wave does
// far fewer leaf steps than a walk that tests the whole wave against every
// leaf any lane overlaps.  A lane gets exactly the neighbour set of its own
// single-atom walk.
// The stack never overflows for trees of depth <= 63 (any LBVH over 30-bit
// keys, i.e. <= 32 record levels): when fewer than kStackDfs free entries
// remain the frontier shrinks to one record (plain depth-first order), which
// adds at most 3 entries per record level.  Queue entries pack
// (leaf << 8) | bits, so leaves must number below 2^23.
// kFront: records per phase-1 round trip (8, or 16 where the autotuner picks it)
constexpr int kFront = 8, kStack2 = 256, kStackDfs = 128, kGroupQ = 128;
constexpr int32_t kMaxLeaves3 = 1 << 23;

// records per atom in the walk's record buffer: cap rounded up to 8
__host__ __device__ __forceinline__ int32_t rec_stride(int32_t cap) { return (cap + 7) & ~7; }

// REC: instead of list rows, each atom gets one record per leaf it has
// neighbours in, (group g, mask): its neighbours there are the sorted indices
// 32 g + b for the set bits b of the mask (leaves of A <= 32 atoms start at
// multiples of A, so a leaf lies in one 32-atom group).  Records of an atom
// are contiguous (recs[i * rec_stride(cap) + k], k < nrecs[i]; blocks of 64 B
// multiples, so a reader fetches two records per 16-byte load); each holds
// at least one neighbour, so there are never more records than list entries.
// LDS of the walk: per wave a staging buffer (8 group slots of A atoms, or
// F fetched records), a node stack and 8 group queues with dummy slots.
template <int A, int F> struct Walk3Lds {
    static constexpr int kSlot = 4 * A + 4;  // floats per group slot (16 B pad: no bank conflicts)
    static constexpr int kBufF = 8 * kSlot > 32 * F ? 8 * kSlot : 32 * F;  // records: 8 float4 each
    static constexpr int kW = nd::kBlock / 64;
    __align__(16) float buf[kW][kBufF];
    int32_t stk[kW][kStack2];
    int32_t q[kW][8 * kGroupQ + 64];
};

template <int A, bool REC, int F, bool LOCAL = false>
__device__ __forceinline__ void
walk3_body(const float4 *__restrict__ pos, const float4 *__restrict__ wide, int32_t n, int32_t nL, float h,
           float rl2, int32_t cap, int32_t *__restrict__ nbr, int32_t *__restrict__ ncount,
           int32_t *__restrict__ ctr, uint2 *__restrict__ recs, int32_t *__restrict__ nrecs,
           Walk3Lds<A, F> &L, int32_t blk, int32_t *__restrict__ work,
           const float4 *__restrict__ tq = nullptr, const float4 *__restrict__ lc = nullptr) {
    static_assert(A == 4 || A == 8 || A == 16 || A == 32, "leaf size");
    constexpr int NV = (A + 7) / 8;      // atoms a lane stages per leaf (1, 2 or 4)
    constexpr int kSlot = Walk3Lds<A, F>::kSlot;
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int grp = lane >> 3, gl = lane & 7;
    float *buf = L.buf[w];
    float4 *rec = reinterpret_cast<float4 *>(buf);
    int32_t *stk = L.stk[w], *gq = L.q[w];

    const int32_t i = blk * nd::kBlock + threadIdx.x;
    const bool active = i < n;
    // this wave's block of the list as a buffer resource: a row at or past cap
    // lies outside it, and the hardware drops such a store (no branch per hit)
    const __amdgpu_buffer_rsrc_t lrs = __builtin_amdgcn_make_buffer_rsrc(
        (void *)(nbr + (int64_t)__builtin_amdgcn_readfirstlane(i >> 6) * cap * 64), 0, cap * 256, kBufWord3);
    uint32_t voff = (uint32_t)lane * 4u;  // byte offset of row cnt in this lane's column
    int32_t rcnt = 0, nrec = 0;           // REC: list entries and records so far
    uint32_t pend_g = ~0u, pend_b = 0u;    // REC: the record being built (leaves of one group merge)
    uint2 *myrec = recs + (REC ? (int64_t)i * rec_stride(cap) : 0);
    const float4 p = active ? pos[i] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    // A lane overlaps a node when its sphere of radius r_list meets the
    // node's atoms' box (the stored box less the padding h = r_list / 2 on
    // each side), not merely the padded box: at 32 atoms per leaf about a
    // quarter fewer leaves are queued.  The distance to the unpadded box is
    // under-estimated by a slack covering the rounding of the padding and of
    // the differences (8 ulp of the largest magnitude involved) and compared
    // with rl2 plus 8 ulp, so every leaf holding a neighbour still passes;
    // lanes without an atom overlap nothing.
    const float slack = (fmaxf(fmaxf(fabsf(p.x), fabsf(p.y)), fabsf(p.z)) + 4.0f * h) * 0x1p-20f;
    const float hc = h - slack, rlb2 = rl2 * (1.0f + 0x1p-20f);
    bool bad = false;

    int32_t glen = 0;  // lanes 0..7: length of queue `lane`
    int32_t nqu = 0;   // leaves queued since the last drain (bounds every glen)
    // branch-free: lanes with nothing to queue write a private dummy slot
    auto enqueue = [&](int32_t leaf, uint64_t mk) {
        const uint32_t bits = (uint32_t)(mk >> (8 * lane)) & 0xFFu;
        const bool put = (lane < 8) & (bits != 0u);
        gq[put ? lane * kGroupQ + glen : 8 * kGroupQ + lane] = (leaf << 8) | (int32_t)bits;
        glen += put ? 1 : 0;
        ++nqu;
    };

    int32_t sp = 1, fetched = 0;
    int32_t rounds = 0, leaf_steps = 0;  // wave-uniform work counters (phase 1, phase 2)
    stk[0] = nL + 1;  // the root is internal (nL >= 2)
    wave_lds_sync();
    while (sp > 0 || nqu > 0) {
        if (nqu > 0 && (sp == 0 || nqu > kGroupQ - 4 * F)) {
            // ---- phase 2: drain the group queues ----
            const int32_t len = __shfl(glen, grp, 64);
            int32_t steps = len;
            for (int off = 32; off > 0; off >>= 1) steps = max(steps, __shfl_xor(steps, off, 64));
            steps = __builtin_amdgcn_readfirstlane(steps);
            leaf_steps += steps;
            float *slot = buf + grp * kSlot;
            // atom gl + 8 r of the group's leaf k; lanes past the leaf size load atom 0
            // (named registers, not an array: the loads stay in flight across the step)
            // LOCAL: the leaf-local copies (-2 q', |q'|^2) and the leaf's
            // (centre, R^2), loaded with the atoms
            const float4 *src = (LOCAL ? tq : pos) + (gl < A ? gl : 0);
            const float4 zero = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            float4 v0 = zero, v1 = zero, v2 = zero, v3 = zero, cn = zero;
            auto load4 = [&](int32_t e) {
                const float4 *lp = src + ((e >> 8) - 1) * A;
                v0 = lp[0];
                if constexpr (NV > 1) v1 = lp[8];
                if constexpr (NV > 2) { v2 = lp[16]; v3 = lp[24]; }
                if constexpr (LOCAL) cn = lc[(e >> 8) - 1];
            };
            if (len > 0) load4(gq[grp * kGroupQ]);
            for (int32_t k = 0; k < steps; ++k) {
                wave_lds_sync();  // the previous step's reads are done
                if (k < len && gl < A) {
                    float4 *dst = reinterpret_cast<float4 *>(slot) + gl;
                    dst[0] = v0;
                    if constexpr (NV > 1) dst[8] = v1;
                    if constexpr (NV > 2) { dst[16] = v2; dst[24] = v3; }
                }
                wave_lds_sync();
                const float4 cc = cn;  // the staged leaf's centre and R^2 (LOCAL)
                const int32_t e = k < len ? gq[grp * kGroupQ + k] : 0;
                if (k + 1 < len) load4(gq[grp * kGroupQ + k + 1]);
                uint32_t hits = 0u;
                const int32_t base = ((e >> 8) - 1) * A;
                // the list's pair test: |p - q|^2 < rl2, in this order of operations
                auto exact = [&](const float4 &q) -> bool {
                    const float dx = p.x - q.x, dy = p.y - q.y, dz = p.z - q.z;
                    const float s = dx * dx + dy * dy + dz * dz;
                    return s < rl2;
                };
                if ((e >> gl) & 1) {
                    if constexpr (LOCAL) {
                        // |p - q|^2 - rl2 = |a|^2 - rl2 + (-2 q').a + |q'|^2 with a = p - c:
                        // three FMAs per atom against the staged copy.  Its distance
                        // from the exact test's value is below 16 u (|a| + R + r)^2
                        // (u = 2^-24: the roundings of a, q', |q'|^2, the FMAs and
                        // of the exact test itself), which E bounds from above
                        // (3 (x^2 + y^2 + z^2) >= (x + y + z)^2); a wave with any
                        // test within E of the radius redoes the leaf exactly from
                        // the positions, so the list is the exact test's.
                        const float ax = p.x - cc.x, ay = p.y - cc.y, az = p.z - cc.z;
                        const float cp = __builtin_fmaf(az, az, __builtin_fmaf(ay, ay, __builtin_fmaf(ax, ax, -rl2)));
                        const float E = 48.0f * 0x1p-24f * (cp + 2.0f * rl2 + cc.w);
                        uint64_t band = 0;
#pragma unroll
                        for (int m = A - 1; m >= 0; --m) {
                            const float4 q = *reinterpret_cast<const float4 *>(slot + 4 * m);
                            const float t =
                                __builtin_fmaf(ax, q.x, __builtin_fmaf(ay, q.y, __builtin_fmaf(az, q.z, q.w))) + cp;
                            // t's sign bit (a zero t is within E: the exact test decides),
                            // shifted in below the bits of atoms m + 1 .. A - 1 by one
                            // funnel shift, so atom m ends at bit m
                            hits = __builtin_amdgcn_alignbit(hits, __float_as_uint(t), 31);
                            band |= __builtin_amdgcn_ballot_w64(fabsf(t) < E);
                        }
                        if (band != 0) {  // rare: a test within E of the radius
                            hits = 0u;
#pragma unroll 1
                            for (int m = 0; m < A; ++m) hits |= (exact(pos[base + m]) ? 1u : 0u) << m;
                        }
                    } else {
#pragma unroll
                        for (int m = 0; m < A; ++m)
                            hits |= (exact(*reinterpret_cast<const float4 *>(slot + 4 * m)) ? 1u : 0u) << m;
                    }
                    const uint32_t self = (uint32_t)(i - base);
                    if (self < (uint32_t)A) hits &= ~(1u << self);
                    if constexpr (REC) {
                        if (hits) {
                            const uint32_t g = (uint32_t)(base >> 5), b = hits << (base & 31);
                            if (g == pend_g) {
                                pend_b |= b;
                            } else {
                                if (pend_g != ~0u) {
                                    if (nrec < cap) myrec[nrec] = make_uint2(pend_g, pend_b);
                                    ++nrec;
                                }
                                pend_g = g;
                                pend_b = b;
                            }
                            rcnt += __popc(hits);
                        }
                    } else {
                        while (hits) {
                            const int32_t m = __builtin_ctz(hits);
                            hits &= hits - 1u;
                            __builtin_amdgcn_raw_buffer_store_b32(base + m, lrs, voff, 0, 0);
                            voff += 256u;
                        }
                    }
                }
            }
            glen = 0;
            nqu = 0;
            wave_lds_sync();
            continue;
        }
        // ---- phase 1: expand up to F records of the stack top ----
        ++rounds;
        // (each record pushes at most 4 entries: k pops leave room for them)
        const int32_t room = (kStack2 - kStackDfs - sp) / 3;
        int32_t k = sp < F ? sp : F;
        if (k > room) k = room > 1 ? room : 1;
        fetched += k;
        // more records than the tree has internal nodes, or a stack that could overflow
        if (fetched > nL || sp + 3 * k > kStack2) { bad = true; break; }
        sp -= k;
        {
            // lane l loads float4 l & 7 of records (l >> 3) + 8 u, u < F / 8
            float4 v[F / 8];
#pragma unroll
            for (int u = 0; u < F / 8; ++u) {
                const int32_t slot = (lane >> 3) + 8 * u;
                v[u] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
                if (slot < k) v[u] = wide[8 * (int64_t)(stk[sp + slot] - nL - 1) + (lane & 7)];
            }
            wave_lds_sync();
#pragma unroll
            for (int u = 0; u < F / 8; ++u) rec[lane + 64 * u] = v[u];
            wave_lds_sync();
        }
        const uint32_t nn = (uint32_t)(2 * nL - 1);  // valid node numbers are 1 .. 2 nL - 1
        auto entry = [&](float mnx, float mny, float mnz, float mxx, float mxy, float mxz, float idf) {
            const int32_t c = __builtin_amdgcn_readfirstlane(__float_as_int(idf));
            if (c == 0) return;
            if ((uint32_t)(c - 1) >= nn) { bad = true; return; }
            const float ex = fmaxf(fmaxf(mnx - p.x, p.x - mxx), -hc) + hc;
            const float ey = fmaxf(fmaxf(mny - p.y, p.y - mxy), -hc) + hc;
            const float ez = fmaxf(fmaxf(mnz - p.z, p.z - mxz), -hc) + hc;
            const bool ov = active & (ex * ex + ey * ey + ez * ez < rlb2);
            const uint64_t mk = __builtin_amdgcn_ballot_w64(ov);
            if (mk) {
                if (c <= nL) enqueue(c, mk);
                else stk[sp++] = c;
            }
        };
        for (int32_t r = 0; r < k; ++r) {
            const float4 mnx = rec[8 * r], mny = rec[8 * r + 1], mnz = rec[8 * r + 2];
            const float4 mxx = rec[8 * r + 3], mxy = rec[8 * r + 4], mxz = rec[8 * r + 5];
            const float4 ids = rec[8 * r + 6];
            // an internal node has two children, so a record has >= 2 entries
            if (__builtin_amdgcn_readfirstlane(__float_as_int(ids.y)) == 0) { bad = true; break; }
            entry(mnx.x, mny.x, mnz.x, mxx.x, mxy.x, mxz.x, ids.x);
            entry(mnx.y, mny.y, mnz.y, mxx.y, mxy.y, mxz.y, ids.y);
            entry(mnx.z, mny.z, mnz.z, mxx.z, mxy.z, mxz.z, ids.z);
            entry(mnx.w, mny.w, mnz.w, mxx.w, mxy.w, mxz.w, ids.w);
            if (bad) break;
        }
        if (bad) break;
        wave_lds_sync();
    }
    if (bad && lane == 0) atomicOr(&ctr[2], nd::kStatusTree);
    // the wave's work estimate for the next walk's block placement: a CU's
    // walk time follows its waves' leaf steps and round trips (~0.12 and
    // ~0.28 us each when four workgroups share the CU)
    if (work != nullptr && lane == 0) work[4 * blk + w] = 5 * leaf_steps + 12 * rounds;
    if (REC && pend_g != ~0u) {
        if (nrec < cap) myrec[nrec] = make_uint2(pend_g, pend_b);
        ++nrec;
    }
    const int32_t cnt = REC ? rcnt : (int32_t)((voff - (uint32_t)lane * 4u) >> 8);
    if (active) {
        ncount[i] = cnt;
        if constexpr (REC) nrecs[i] = nrec;
    }
    int32_t m = cnt;
    for (int off = 32; off > 0; off >>= 1) m = max(m, __shfl_down(m, off, 64));
    if (lane == 0 && m > 0) {
        // the counter only grows: skip the (serialising) atomic when it is
        // already at least m
        if (m > __hip_atomic_load(&ctr[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) atomicMax(&ctr[0], m);
        if (m > cap) atomicOr(&ctr[1], 1);
    }
}

template <int A, bool REC = false, int F = kFront, bool LOCAL = false>
__global__ void __launch_bounds__(nd::kBlock)
k_nlist3(const float4 *__restrict__ pos, const float4 *__restrict__ wide, int32_t n, int32_t nL,
         float h, float rl2, int32_t cap, int32_t *__restrict__ nbr, int32_t *__restrict__ ncount,
         int32_t *__restrict__ ctr, uint2 *__restrict__ recs = nullptr, int32_t *__restrict__ nrecs = nullptr,
         const float4 *__restrict__ tq = nullptr, const float4 *__restrict__ lc = nullptr) {
    __shared__ Walk3Lds<A, F> L;
    walk3_body<A, REC, F, LOCAL>(pos, wide, n, nL, h, rl2, cap, nbr, ncount, ctr, recs, nrecs, L,
                                 xcd_block(blockIdx.x, gridDim.x), nullptr, tq, lc);
}

// ---- neighbourhood tables ----
// The step kernel stages, per workgroup, the positions of every distinct
// neighbour of its 256 atoms in LDS once and then reads each pair's
// neighbour from there: about 930 positions for 256 atoms at density 0.8
// and r_list = 2.8 (19,400 list entries), so a workgroup gathers each
// neighbour once instead of once per list entry, and a list entry shrinks to
// a 16-bit position in the table.  The table is built from a hash over
// 32-atom groups of sorted indices (key: index >> 5, value: a 32-bit mask of
// the group's atoms in use), a scan of the masks' population counts, and the
// position of sorted index j is its group's base plus the number of set mask
// bits below j & 31.  Table order is hash-slot order; each group's atoms are
// consecutive in it.
constexpr int kHashSlots = 1024;

// The list walks' last kernel (a table build) also copies the stale flag the
// integrators raise to pinned host memory, where the run loop looks for it
// (one thread; the flag only goes from 0 to 1 between rebuilds).
__device__ __forceinline__ void stale_to_host(const SimMisc *__restrict__ misc, volatile int32_t *hstale) {
    if (blockIdx.x == 0 && threadIdx.x == 0 && misc->stale) *hstale = 1;
}

__device__ __forceinline__ uint32_t group_hash(int32_t g) {
    return ((uint32_t)g * 2654435761u) >> (32 - 10);  // kHashSlots = 2^10
}

// A workgroup's table under construction: per hash slot the group (-1:
// empty), the mask of its atoms in use and (after finish()) the table
// position of the group's first atom in use.
struct TableLds {
    int32_t key[kHashSlots];
    uint32_t mask[kHashSlots];
    int32_t base[kHashSlots];
    int32_t part[nd::kBlock / 64];
    int32_t fail;

    __device__ void init() {
        for (int32_t t = threadIdx.x; t < kHashSlots; t += nd::kBlock) { key[t] = -1; mask[t] = 0u; }
        if (threadIdx.x == 0) fail = 0;
    }
    // returns the group's slot (-1: the hash is full)
    __device__ int32_t insert(int32_t g, uint32_t bits) {
        uint32_t h = group_hash(g);
        for (int probe = 0; probe < kHashSlots; ++probe) {
            // a slot's key, once set, never changes: a plain read settles
            // most inserts (the group is there, or another group is), and
            // only an empty slot takes the compare-and-swap
            int32_t cur = __hip_atomic_load(&key[h], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            if (cur == -1) cur = atomicCAS(&key[h], -1, g);
            if (cur == -1 || cur == g) {
                atomicOr(&mask[h], bits);
                return (int32_t)h;
            }
            h = (h + 1u) & (kHashSlots - 1);
        }
        fail = 1;  // more than kHashSlots groups
        return -1;
    }
    // After every insert (and a barrier): group bases by an exclusive scan of
    // the masks' population counts; returns the table size, or -1 when the
    // table does not fit (lcap) or the hash overflowed.  Ends with a barrier.
    __device__ int32_t finish(int32_t lcap) {
        constexpr int kPer = kHashSlots / nd::kBlock;
        const int32_t tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
        int32_t c[kPer], tot = 0;
#pragma unroll
        for (int k = 0; k < kPer; ++k) { c[k] = __popc(mask[kPer * tid + k]); tot += c[k]; }
        int32_t incl = tot;
        for (int off = 1; off < 64; off <<= 1) {
            const int32_t y = __shfl_up(incl, off, 64);
            if (lane >= off) incl += y;
        }
        if (lane == 63) part[w] = incl;
        __syncthreads();
        int32_t wbase = 0, U = 0;
#pragma unroll
        for (int k = 0; k < nd::kBlock / 64; ++k) {
            wbase += k < w ? part[k] : 0;
            U += part[k];
        }
        int32_t b = wbase + incl - tot;
#pragma unroll
        for (int k = 0; k < kPer; ++k) { base[kPer * tid + k] = b; b += c[k]; }
        const bool bad = fail != 0 || U > lcap;
        __syncthreads();
        return bad ? -1 : U;
    }
    // the table: sorted indices of the groups' atoms in use, in slot order.
    // A wave writes the runs of its lanes' non-empty slots two at a time,
    // a half-wave per run (lane b of the half writes the entry of bit b), so
    // the stores of a run are consecutive rather than one lane's serial loop.
    __device__ void write(int32_t *tab) const {
        constexpr int kPer = kHashSlots / nd::kBlock;
        const int lane = threadIdx.x & 63, bit = lane & 31;
#pragma unroll
        for (int k = 0; k < kPer; ++k) {
            const int32_t sl = kPer * threadIdx.x + k;
            const uint32_t m = mask[sl];
            const int32_t b = base[sl], g32 = key[sl] << 5;
            uint64_t todo = __builtin_amdgcn_ballot_w64(m != 0u);
            while (todo) {  // wave-uniform
                const int s0 = (int)__builtin_ctzll(todo);
                todo &= todo - 1u;
                const int s1 = todo ? (int)__builtin_ctzll(todo) : s0;
                const bool two = todo != 0u;
                todo &= todo - 1u;
                const uint32_t m0 = (uint32_t)__builtin_amdgcn_readlane((int)m, s0);
                const uint32_t m1 = two ? (uint32_t)__builtin_amdgcn_readlane((int)m, s1) : 0u;
                const int32_t b0 = __builtin_amdgcn_readlane(b, s0), b1 = __builtin_amdgcn_readlane(b, s1);
                const int32_t g0 = __builtin_amdgcn_readlane(g32, s0), g1 = __builtin_amdgcn_readlane(g32, s1);
                const uint32_t mm = lane < 32 ? m0 : m1;
                if ((mm >> bit) & 1u)
                    tab[(lane < 32 ? b0 : b1) + __popc(mm & ((1u << bit) - 1u))] = (lane < 32 ? g0 : g1) + bit;
            }
        }
    }
    __device__ int32_t slot(int32_t g) const {
        uint32_t h = group_hash(g);
        while (key[h] != g) h = (h + 1u) & (kHashSlots - 1);  // present: inserted
        return (int32_t)h;
    }
    // table position of sorted index j (its group was inserted)
    __device__ uint32_t local(int32_t j) const {
        const int32_t h = slot(j >> 5);
        return (uint32_t)(base[h] + __popc(mask[h] & ((1u << (j & 31)) - 1u)));
    }
};

// Rows are read kRowBatch at a time so that many loads are in flight per wave.
constexpr int kRowBatch = 16;

// Tables from list rows (the rope, one-node and cell walks).
__global__ void __launch_bounds__(nd::kBlock)
k_table(const int32_t *__restrict__ nbr, const int32_t *__restrict__ ncount, int32_t n, int32_t cap,
        int32_t cap2, int32_t lcap, int32_t eshift, uint32_t *__restrict__ nbr16, int32_t *__restrict__ tables,
        int32_t *__restrict__ tcount, const SimMisc *__restrict__ misc, volatile int32_t *hstale,
        uint32_t *__restrict__ slot_info) {
    stale_to_host(misc, hstale);
    __shared__ TableLds T;
    const int32_t blk = xcd_block(blockIdx.x, gridDim.x);
    const int32_t lane = threadIdx.x & 63;
    const int32_t i = blk * nd::kBlock + threadIdx.x;
    T.init();
    __syncthreads();
    const int32_t cnt = i < n ? min(ncount[i], cap) : 0;
    slot_info[i] = ((uint32_t)cnt << 8) | (uint32_t)threadIdx.x;  // lists stay in atom order
    const int32_t *row = nbr + ((int64_t)(i >> 6) * cap) * 64 + lane;
    int32_t wmax = cnt;
    for (int off = 32; off > 0; off >>= 1) wmax = max(wmax, __shfl_xor(wmax, off, 64));
    wmax = __builtin_amdgcn_readfirstlane(wmax);
    // consecutive entries of a list mostly share a group (the walks emit a
    // leaf's atoms together): one hash update per run of a group
    {
        int32_t cg = -1;
        uint32_t bits = 0u;
        for (int32_t r0 = 0; r0 < wmax; r0 += kRowBatch) {
            int32_t jj[kRowBatch];
#pragma unroll
            for (int u = 0; u < kRowBatch; ++u) jj[u] = r0 + u < cnt ? row[(int64_t)(r0 + u) * 64] : -1;
#pragma unroll
            for (int u = 0; u < kRowBatch; ++u) {
                const int32_t j = jj[u];
                if (j < 0) continue;
                const int32_t g = j >> 5;
                if (g != cg) {
                    if (cg >= 0) T.insert(cg, bits);
                    cg = g;
                    bits = 0u;
                }
                bits |= 1u << (j & 31);
            }
        }
        if (cg >= 0) T.insert(cg, bits);
    }
    __syncthreads();
    const int32_t U = T.finish(lcap);
    if (threadIdx.x == 0) tcount[blk] = U;
    if (U < 0) return;  // uniform: the step kernel reads the rows directly
    T.write(tables + (int64_t)blk * lcap);
    // waves past the last atom own no list block
    if (((i >> 6) << 6) >= n) return;
    // 16-bit lists, two entries per int32, one row of 64 per step of the
    // wave, then two rows of sentinels; row 0 is written even for a wave
    // without neighbours (the step kernel's look-ahead reads it).  Entries are
    // table positions << eshift; past an atom's list they name the sentinel
    // slot U.
    uint32_t *out = nbr16 + ((int64_t)(i >> 6) * cap2) * 64 + lane;
    const int32_t nrows = max((wmax + 1) >> 1, 1) + 2;
    const uint32_t sent = (uint32_t)U << eshift;
    for (int32_t q0 = 0; q0 < nrows; q0 += kRowBatch / 2) {
        int32_t jj[kRowBatch];
#pragma unroll
        for (int u = 0; u < kRowBatch; ++u) {
            const int32_t r = 2 * q0 + u;
            jj[u] = r < cnt ? row[(int64_t)r * 64] : -1;
        }
#pragma unroll
        for (int u = 0; u < kRowBatch; u += 2) {
            const int32_t q = q0 + u / 2;
            if (q >= nrows) break;
            const uint32_t a = jj[u] >= 0 ? T.local(jj[u]) << eshift : sent;
            const uint32_t b = jj[u + 1] >= 0 ? T.local(jj[u + 1]) << eshift : sent;
            out[(int64_t)q * 64] = a | (b << 16);
        }
    }
}

// Tables from the two-phase walk's records (k_nlist3<A, true>): a record per
// (atom, leaf with neighbours), ~9 per atom at density 0.8, instead of a list
// entry per neighbour (~76), so the walk writes and this kernel reads ~4x
// fewer bytes.  While the records go into the hash they are copied to LDS
// with their hash slot; once the table is laid out every lane expands its own
// records into its list, reading per record the slot's table base and mask
// in use from LDS and per entry a population count (a lane reading its
// records from memory instead would stall its whole wave at each one).  The
// LDS of a workgroup (36.9 KB) lets four of them share a CU, so the 1,024
// workgroups of 262,144 atoms run in one round on 256 CUs.
// A workgroup whose table does not fit (lcap) gets list rows of sorted
// indices (nbr) instead; one whose records do not fit in LDS (RL records)
// expands them from memory.  RL = 3072 (36.9 KB of LDS: four workgroups per
// CU) unless the first checked build of a simulation saw more than 1/64 of
// its workgroups spill, as on dense blobs (about 13 records per atom), which
// then take RL = 5120 (53.3 KB: three per CU).
constexpr int kRecLds = 3072, kRecLdsBig = 5120;

// LDS of a table build from records: the hash and RL records (hash slot, mask)
template <int RL> struct TableRecLds {
    TableLds T;
    uint2 rec[RL];
    int32_t wsum[nd::kBlock / 64];
};

template <int RL>
__device__ __forceinline__ void
table_rec_body(const uint2 *__restrict__ recs, const int32_t *__restrict__ nrecs,
               const int32_t *__restrict__ ncount, int32_t n, int32_t cap, int32_t cap2, int32_t lcap,
               int32_t eshift, uint32_t *__restrict__ nbr16, int32_t *__restrict__ tables,
               int32_t *__restrict__ tcount, int32_t *__restrict__ nbr, SimMisc *__restrict__ misc,
               volatile int32_t *hstale, uint32_t *__restrict__ slot_info, TableRecLds<RL> &S, int32_t blk) {
    stale_to_host(misc, hstale);
    constexpr int kW = nd::kBlock / 64;
    TableLds &T = S.T;
    uint2 *s_rec = S.rec;        // (hash slot, mask)
    int32_t *s_wsum = S.wsum;
    const int32_t tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int32_t i = blk * nd::kBlock + tid;
    T.init();
    const bool active = i < n;
    const int32_t cnt = active ? min(ncount[i], cap) : 0;
    const int32_t nr = active ? min(nrecs[i], cap) : 0;
    const uint2 *my = recs + (active ? (int64_t)i * rec_stride(cap) : 0);
    // this lane's first LDS record: exclusive scan of the record counts
    int32_t incl = nr;
    for (int off = 1; off < 64; off <<= 1) {
        const int32_t y = __shfl_up(incl, off, 64);
        if (lane >= off) incl += y;
    }
    if (lane == 63) s_wsum[w] = incl;
    __syncthreads();
    int32_t roff = incl - nr, rtot = 0;
#pragma unroll
    for (int k = 0; k < kW; ++k) {
        roff += k < w ? s_wsum[k] : 0;
        rtot += s_wsum[k];
    }
    const bool in_lds = rtot <= RL;  // uniform over the workgroup
    if (!in_lds && tid == 0) atomicAdd(&misc->rec_spill, 1);
    // records are fetched kRecBatch at a time, two per 16-byte load (the
    // atom's block is 64-byte aligned; slots past nr are read and ignored)
    constexpr int kRecBatch = 8;
    auto fetch = [&](int32_t k0, uint2 *r) {
        const uint4 *src = reinterpret_cast<const uint4 *>(my + k0);
#pragma unroll
        for (int u = 0; u < kRecBatch / 2; ++u) {
            const uint4 v = k0 + 2 * u < nr ? src[u] : make_uint4(0u, 0u, 0u, 0u);
            r[2 * u] = make_uint2(v.x, v.y);
            r[2 * u + 1] = make_uint2(v.z, v.w);
        }
    };
    for (int32_t k0 = 0; k0 < nr; k0 += kRecBatch) {
        uint2 r[kRecBatch];
        fetch(k0, r);
#pragma unroll
        for (int u = 0; u < kRecBatch; ++u)
            if (k0 + u < nr) {
                const int32_t h = T.insert((int32_t)r[u].x, r[u].y);
                if (in_lds) s_rec[roff + k0 + u] = make_uint2((uint32_t)h, r[u].y);
            }
    }
    __syncthreads();
    // U >= 0 only if every group found a hash slot and the table fits
    const int32_t U = T.finish(lcap);  // ends with a barrier
    const bool lds_ok = in_lds && U >= 0;
    if (tid == 0) tcount[blk] = U;
    if (U >= 0) T.write(tables + (int64_t)blk * lcap);
    // Slot order.  A wave of the step kernel runs as long as its longest
    // list, so a workgroup expanded from LDS lays its 256 lists out by
    // decreasing length (ties by atom): column t of the workgroup's list
    // blocks holds the list of atom `a`.  A bitonic sort of the keys
    // (0xFFFF - length) << 8 | atom, one per lane: shuffles within a wave,
    // LDS (the hash keys, dead after the table write) across waves.  Other
    // workgroups keep atom order.
    int32_t a = tid, my_cnt = cnt, my_roff = roff, my_nr = nr;
    if (lds_ok) {  // uniform over the workgroup
        uint32_t key = ((uint32_t)(0xFFFF - min(cnt, 0xFFFF)) << 8) | (uint32_t)tid;
        uint32_t *sk = reinterpret_cast<uint32_t *>(T.key);
        int32_t *s_ro = T.key + nd::kBlock, *s_nr = T.key + 2 * nd::kBlock, *s_cn = T.key + 3 * nd::kBlock;
        __syncthreads();  // the table write's reads of the hash keys are done
        s_ro[tid] = roff;
        s_nr[tid] = nr;
        s_cn[tid] = cnt;
#pragma unroll
        for (int k = 2; k <= nd::kBlock; k <<= 1) {
#pragma unroll
            for (int j = k >> 1; j > 0; j >>= 1) {
                uint32_t other;
                if (j >= 64) {
                    __syncthreads();
                    sk[tid] = key;
                    __syncthreads();
                    other = sk[tid ^ j];
                } else {
                    other = (uint32_t)__shfl_xor((int)key, j, 64);
                }
                const bool up = (tid & k) == 0, lower = (tid & j) == 0;
                key = (lower == up) ? min(key, other) : max(key, other);
            }
        }
        __syncthreads();  // s_ro / s_nr / s_cn are written
        a = (int32_t)(key & 0xFFu);
        my_roff = s_ro[a];
        my_nr = s_nr[a];
        my_cnt = s_cn[a];
    }
    slot_info[i] = ((uint32_t)my_cnt << 8) | (uint32_t)a;
    if (((i >> 6) << 6) >= n) return;  // waves past the last atom own no list block
    int32_t wmax = my_cnt;
    for (int off = 32; off > 0; off >>= 1) wmax = max(wmax, __shfl_xor(wmax, off, 64));
    wmax = __builtin_amdgcn_readfirstlane(wmax);
    if (lds_ok) {
        // the entries of atom a in record order (every record has a set bit)
        int32_t rk = my_roff, cb = 0;
        uint32_t bits = 0u, gm = 0u;
        const uint32_t sent = (uint32_t)U << eshift;
        auto pop = [&]() -> uint32_t {  // table entry; the sentinel past the end
            if (bits == 0u) {
                if (rk >= my_roff + my_nr) return sent;
                const uint2 r = s_rec[rk];
                ++rk;
                cb = T.base[r.x];
                gm = T.mask[r.x];
                bits = r.y;
            }
            const int32_t b = __builtin_ctz(bits);
            bits &= bits - 1u;
            return (uint32_t)(cb + __popc(gm & ((1u << b) - 1u))) << eshift;
        };
        // 16-bit lists, two entries per int32, one row of 64 per step of the
        // wave, then two rows of sentinels; row 0 is written even for a wave
        // without neighbours (the step kernel's look-ahead reads it)
        uint32_t *out = nbr16 + ((int64_t)(i >> 6) * cap2) * 64 + lane;
        const int32_t nrows = max((wmax + 1) >> 1, 1) + 2;
        for (int32_t r = 0; r < nrows; ++r) {
            const uint32_t a = pop(), b = pop();
            out[(int64_t)r * 64] = a | (b << 16);
        }
        return;
    }
    // Records in memory (too many for LDS): each lane expands its own, the
    // next batch of kRecBatch records in flight while the current one is
    // used (a lane waiting on a load would stall its whole wave), into table
    // positions (a hash lookup per record) or, without a table, sorted indices.
    const bool table = U >= 0;
    uint2 q[kRecBatch], nq[kRecBatch];
    int32_t qn = 0, nqn = 0, rk = 0, cb = 0, gbase = 0;
    uint32_t bits = 0u, gm = 0u;
    auto prefetch = [&]() {
        nqn = min(kRecBatch, nr - rk);
        if (nqn > 0) {
            fetch(rk, nq);
            rk += nqn;
        }
    };
    prefetch();
    auto pop = [&]() -> int32_t {  // table position / sorted index; -1 when done
        while (bits == 0u) {
            if (qn == 0) {
                if (nqn <= 0) return -1;
#pragma unroll
                for (int u = 0; u < kRecBatch; ++u) q[u] = nq[u];
                qn = nqn;
                prefetch();
            }
            gbase = (int32_t)q[0].x;
            bits = q[0].y;
#pragma unroll
            for (int u = 0; u + 1 < kRecBatch; ++u) q[u] = q[u + 1];
            --qn;
            if (table) {
                const int32_t h = T.slot(gbase);
                cb = T.base[h];
                gm = T.mask[h];
            }
        }
        const int32_t b = __builtin_ctz(bits);
        bits &= bits - 1u;
        return table ? cb + __popc(gm & ((1u << b) - 1u)) : (gbase << 5) + b;
    };
    if (table) {
        uint32_t *out = nbr16 + ((int64_t)(i >> 6) * cap2) * 64 + lane;
        const int32_t nrows = max((wmax + 1) >> 1, 1) + 2;
        auto enc = [&](int32_t a) -> uint32_t { return (uint32_t)(a < 0 ? U : a) << eshift; };
        for (int32_t r = 0; r < nrows; ++r) {
            const int32_t a = pop(), b = pop();
            out[(int64_t)r * 64] = enc(a) | (enc(b) << 16);
        }
    } else {
        int32_t *out = nbr + ((int64_t)(i >> 6) * cap) * 64 + lane;
        for (int32_t r = 0; r < wmax; ++r) out[(int64_t)r * 64] = pop();
    }
}

template <int RL>
__global__ void __launch_bounds__(nd::kBlock)
k_table_rec(const uint2 *__restrict__ recs, const int32_t *__restrict__ nrecs,
            const int32_t *__restrict__ ncount, int32_t n, int32_t cap, int32_t cap2, int32_t lcap,
            int32_t eshift, uint32_t *__restrict__ nbr16, int32_t *__restrict__ tables,
            int32_t *__restrict__ tcount, int32_t *__restrict__ nbr, SimMisc *__restrict__ misc,
            volatile int32_t *hstale, uint32_t *__restrict__ slot_info) {
    __shared__ TableRecLds<RL> S;
    table_rec_body<RL>(recs, nrecs, ncount, n, cap, cap2, lcap, eshift, nbr16, tables, tcount, nbr, misc,
                       hstale, slot_info, S, xcd_block(blockIdx.x, gridDim.x));
}

// The two-phase walk and the table build in one kernel: a workgroup walks
// for its 256 atoms (the same 256 atoms as a step-kernel workgroup), then,
// after a barrier, builds their table in the same LDS.  A workgroup that
// finishes its walk early builds its table while slower ones still walk,
// and one launch replaces two.  Each lane reads back only the records and
// counts it wrote itself, so the barrier (for the LDS) is the only
// synchronisation needed.
//
// Block placement: workgroup b walks atom block order[xcd_block(b)], where
// order is a permutation of each XCD's slab of blocks made (walk_order_slab,
// after the force step that follows a walk) from the previous walk's work
// estimates (the identity at first).
// All workgroups of a launch are resident at once (four per CU), so the
// kernel ends with its busiest CU; the order gives every CU of an XCD a
// similar total.  Any permutation gives the same lists and tables.
template <int A, int F>
__global__ void __launch_bounds__(nd::kBlock, 4)  // four waves per SIMD: all 4 workgroups per CU resident
k_walk_table(const float4 *__restrict__ pos, const float4 *__restrict__ wide, int32_t n, int32_t nL, float h,
             float rl2, int32_t cap, int32_t *__restrict__ nbr, int32_t *__restrict__ ncount,
             int32_t *__restrict__ ctr, uint2 *__restrict__ recs, int32_t *__restrict__ nrecs, int32_t cap2,
             int32_t lcap, int32_t eshift, uint32_t *__restrict__ nbr16, int32_t *__restrict__ tables,
             int32_t *__restrict__ tcount, SimMisc *__restrict__ misc, volatile int32_t *hstale,
             uint32_t *__restrict__ slot_info, const int32_t *__restrict__ order, int32_t *__restrict__ work,
             const float4 *__restrict__ tq, const float4 *__restrict__ lc) {
    union Lds {
        Walk3Lds<A, F> walk;
        TableRecLds<kRecLds> table;
    };
    __shared__ Lds L;
    const int32_t blk = __builtin_amdgcn_readfirstlane(order[xcd_block(blockIdx.x, gridDim.x)]);
    walk3_body<A, true, F, true>(pos, wide, n, nL, h, rl2, cap, nbr, ncount, ctr, recs, nrecs, L.walk, blk, work,
                                 tq, lc);
    __syncthreads();  // every wave's walk is done with the LDS
    table_rec_body<kRecLds>(recs, nrecs, ncount, n, cap, cap2, lcap, eshift, nbr16, tables, tcount, nbr, misc,
                            hstale, slot_info, L.table, blk);
}

// The walk order alone (when the step after a walk writes no energy row;
// otherwise k_energy_finalize's blocks 1-8 compute it).
__global__ void __launch_bounds__(nd::kBlock)
k_walk_order(const int32_t *__restrict__ work, int32_t nblocks, int32_t cus, int32_t *__restrict__ order) {
    __shared__ int32_t wt[kOrderSlab];
    walk_order_slab(work, nblocks, cus, order, blockIdx.x, wt);
}


// Cell-list walk (list_walk 3), the comparison point for the LBVH walks: the
// same list (every other atom with r^2 < r_list^2, as sorted indices) from the
// 27 cells of side >= r_list around the atom's own cell, scanned as 9 runs of
// 3 cells that are contiguous in the cell-sorted copy cpos (w = sorted index,
// cstart = first position of each cell).  One lane per atom; appends go to
// the same [cap][64] blocks as the LBVH walks', with the same counters.
__global__ void __launch_bounds__(nd::kBlock)
k_cell_rows(const float4 *__restrict__ pos, const float4 *__restrict__ cpos,
            const int32_t *__restrict__ cstart, int32_t n, nd::CellGrid cg, float rl2, int32_t cap,
            int32_t *__restrict__ nbr, int32_t *__restrict__ ncount, int32_t *__restrict__ ctr) {
    const int32_t i = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    const int lane = threadIdx.x & 63;
    const bool active = i < n;
    // this wave's block of the list as a buffer resource: a row at or past cap
    // lies outside it, and the hardware drops such a store
    const __amdgpu_buffer_rsrc_t lrs = __builtin_amdgcn_make_buffer_rsrc(
        (void *)(nbr + (int64_t)__builtin_amdgcn_readfirstlane(i >> 6) * cap * 64), 0, cap * 256, kBufWord3);
    uint32_t voff = (uint32_t)lane * 4u;
    if (active) {
        const float4 p = pos[i];
        const int32_t cx = nd::cell_coord(cg, p.x, 0), cy = nd::cell_coord(cg, p.y, 1),
                      cz = nd::cell_coord(cg, p.z, 2);
        const int32_t x0 = cx > 0 ? cx - 1 : 0, x1 = cx + 1 < cg.g[0] ? cx + 1 : cg.g[0] - 1;
        const int32_t y0 = cy > 0 ? cy - 1 : 0, y1 = cy + 1 < cg.g[1] ? cy + 1 : cg.g[1] - 1;
        const int32_t z0 = cz > 0 ? cz - 1 : 0, z1 = cz + 1 < cg.g[2] ? cz + 1 : cg.g[2] - 1;
        for (int32_t z = z0; z <= z1; ++z)
            for (int32_t y = y0; y <= y1; ++y) {
                const int32_t b = (z * cg.g[1] + y) * cg.g[0];
                const int32_t e = cstart[b + x1 + 1];
                for (int32_t t = cstart[b + x0]; t < e; ++t) {
                    const float4 q = cpos[t];
                    const float dx = p.x - q.x, dy = p.y - q.y, dz = p.z - q.z;
                    const float s = dx * dx + dy * dy + dz * dz;
                    const int32_t j = __float_as_int(q.w);
                    if ((s < rl2) & (j != i)) {
                        __builtin_amdgcn_raw_buffer_store_b32(j, lrs, voff, 0, 0);
                        voff += 256u;
                    }
                }
            }
    }
    const int32_t cnt = (int32_t)((voff - (uint32_t)lane * 4u) >> 8);
    if (active) ncount[i] = cnt;
    int32_t m = cnt;
    for (int off = 32; off > 0; off >>= 1) m = max(m, __shfl_down(m, off, 64));
    if (lane == 0 && m > 0) {
        // the counter only grows: skip the (serialising) atomic when it is
        // already at least m
        if (m > __hip_atomic_load(&ctr[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) atomicMax(&ctr[0], m);
        if (m > cap) atomicOr(&ctr[1], 1);
    }
}

// Trajectory frame: slot s's position to row orig[s] of out (n x 3).
template <class PosS, class OutT>
__global__ void __launch_bounds__(nd::kBlock)
k_record(const typename V4<PosS>::T *__restrict__ pos, const int32_t *__restrict__ orig, int32_t n,
         OutT *__restrict__ out) {
    const int32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const auto p = pos[s];
    OutT *o = out + 3 * (int64_t)orig[s];
    o[0] = (OutT)p.x;
    o[1] = (OutT)p.y;
    o[2] = (OutT)p.z;
}

// FP32 copy of FP64 positions for a box refit without a re-sort.
__global__ void __launch_bounds__(nd::kBlock)
k_pos32(const double4 *__restrict__ pos, int32_t n, float4 *__restrict__ out) {
    const int32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const double4 p = pos[s];
    out[s] = make_float4((float)p.x, (float)p.y, (float)p.z, (float)p.w);
}

// Copies a raised stale flag to host memory (the first kernel of every
// rebuild, so a run on the host sees a stale list a window or two later).
__global__ void k_misc_reset_counters(SimMisc *misc) {
    if (threadIdx.x == 0) { misc->overflow = 0; misc->status = 0; misc->maxcnt = 0; }
}

}  // namespace

namespace nd {

void launch_list_walk(hipStream_t st, const nd_node *nodes, float4 *rec, const float4 *pos, int32_t n,
                      int32_t nL, int32_t apl, float rlist, int32_t cap, int32_t *nbr, int32_t *ncount,
                      int32_t *ctr) {
    if (nL >= kMaxLeaves3) throw Error(ND_ERR_ARG, "list walk: too many leaves");
    if (cap < 1 || cap > kMaxListRows) throw Error(ND_ERR_CAPACITY, "list walk: capacity out of range");
    launch_lbvh_wide4(st, nodes, nL, rec);
    const float hl = rlist / 2.0f, rl2 = rlist * rlist;
    const unsigned g = grid_for(n);
    switch (apl) {
        case 4: k_nlist3<4><<<g, kBlock, 0, st>>>(pos, rec, n, nL, hl, rl2, cap, nbr, ncount, ctr); break;
        case 8: k_nlist3<8><<<g, kBlock, 0, st>>>(pos, rec, n, nL, hl, rl2, cap, nbr, ncount, ctr); break;
        case 16: k_nlist3<16><<<g, kBlock, 0, st>>>(pos, rec, n, nL, hl, rl2, cap, nbr, ncount, ctr); break;
        case 32: k_nlist3<32><<<g, kBlock, 0, st>>>(pos, rec, n, nL, hl, rl2, cap, nbr, ncount, ctr); break;
        default: throw Error(ND_ERR_ARG, "list walk: leaf size must be 4, 8, 16 or 32");
    }
    ND_HIP(hipGetLastError());
}

}  // namespace nd

struct nd_sim {
    nd_ctx *ctx = nullptr;
    nd_sim_params p{};
    int32_t n = 0, nL = 0, apl = 0;
    bool wide = false;  // FP64 storage (MIXED, FP64)
    nd::DevBuf pos[2], vel[2], frc, orig[2], pos32, fix[2], codes, perm, nodes, store, kids, nbr,
        ncount, misc, partials, rows, snap_pos, snap_vel, snap_frc, snap_orig, rep, boxes;
    nd::DevBuf wsnap_pos, wsnap_vel, wsnap_orig;  // start of a window in a checked rerun
    nd::DevBuf cpos, cstart;  // cell-list walk: cell-sorted positions, cell starts
    nd::DevBuf path;          // uint16 path length of each atom since the last build
    nd::DevBuf nbr16, tables, tcount;  // neighbourhood tables and 16-bit lists (k_table)
    nd::DevBuf slot_info;              // per list column: (length << 8) | atom in the workgroup
    nd::DevBuf walk_order;             // fused walk: atom block of each slab position
    nd::DevBuf walk_work;              // fused walk: per wave work estimate of the last walk
    nd::DevBuf leaf_q;                 // two-phase walk: per atom (-2 q', |q'|^2) about its leaf's centre
    nd::DevBuf leaf_c;                 // two-phase walk: per leaf (centre, largest |q'|^2)
    int32_t cus_per_xcd = 32;          // compute units per XCD (the walk order's slot width)
    bool order_pending = false;        // a fused walk ran; its work estimates are not yet ordered
    nd::DevBuf recs, nrecs;   // the two-phase walk's (group, mask) records per atom
    int32_t lcap = 0;         // table capacity (records) per workgroup
    int32_t pair_batch = kPairBatch;  // step kernel: neighbours per pipeline batch (FP32)
    int32_t walk_front = kFront;      // two-phase walk: records per phase-1 round trip
    int32_t atom_lanes = 1;           // step kernel: lanes per atom (1 or 2)
    nd::CellGrid cg{};
    int32_t nrep = 0;  // packed replicas (0: one system)
    nd::SortBufs sort;
    int cur = 0;   // position buffer (flips every drift and every re-sort)
    int vcur = 0;  // velocity / original-index buffer (flips only on a re-sort)
    int32_t cap = 0;
    int32_t max_rows = 1 << 16;
    int64_t step = 0, rebuilds = 0, last_build_step = -1;
    int32_t since = 0;  // drifts launched since the last list build
    int64_t stale_reruns = 0, forced_rebuilds = 0;
    int64_t graph_recaptures = 0;  // graphs dropped because a buffer moved (expected 0)
    // stale flag in pinned host memory (written by the table kernel that ends
    // each rebuild) and one event per window in flight, so a run can stop
    // early when a list went stale
    volatile int32_t *hstale = nullptr;
    static constexpr int kInFlight = 3;
    hipEvent_t wev[kInFlight + 1] = {};
    bool have_force = false, list_valid = false, have_state = false;
    // the tree in nodes/store matches the current atom order (a refit may reuse it)
    bool tree_valid = false;
    // table builds hold up to kRecLdsBig records per workgroup in LDS (set
    // when a checked build saw many workgroups spill at kRecLds)
    bool rec_big = false;
    float rlist = 0.0f;
    StepParams sp{};
    // graphs of one rebuild window, per buffer parity (bits 0-1) and kind
    // (bit 2: window that re-sorts)
    hipGraphExec_t gexec[8] = {};
    uint64_t gsig[8] = {};  // buffer set each graph was captured with
    // kernel timing
    bool timing = false;
    std::vector<hipEvent_t> ev_pool;
    size_t ev_used = 0;
    double force_ms = 0.0;
    int64_t force_launches = 0;
    int64_t last_list_entries = 0;
    int32_t last_maxcnt = 0;
    int64_t pairs_last = 0;
};

namespace {

size_t pos_bytes(const nd_sim *s) { return (s->wide ? sizeof(double4) : sizeof(float4)) * (size_t)s->n; }

Walls walls_of(nd_sim *s) {
    return Walls{s->orig[s->vcur].as<int32_t>(), s->nrep > 0 ? s->rep.as<int32_t>() : nullptr,
                 s->nrep > 0 ? s->boxes.as<double>() : nullptr};
}

// fixed-point position buffer b (MIXED only)
int4 *fix_ptr(nd_sim *s, int b) { return s->p.precision == ND_PREC_MIXED ? s->fix[b].as<int4>() : nullptr; }

void drop_graphs(nd_sim *s) {
    for (auto &g : s->gexec)
        if (g) { (void)hipGraphExecDestroy(g); g = nullptr; }
}

// Every device buffer of a simulation (teardown releases them all; the
// buffer-set signature covers them all).
template <class F>
void for_each_buf(nd_sim *s, F &&f) {
    nd::DevBuf *bufs[] = {&s->pos[0], &s->pos[1], &s->vel[0], &s->vel[1], &s->frc, &s->orig[0],
                          &s->orig[1], &s->pos32, &s->fix[0], &s->fix[1], &s->codes, &s->perm, &s->nodes,
                          &s->store, &s->kids, &s->rep, &s->boxes, &s->nbr, &s->ncount, &s->misc,
                          &s->partials, &s->rows, &s->snap_pos, &s->snap_vel, &s->snap_frc, &s->snap_orig,
                          &s->sort.keys_alt, &s->sort.vals_alt, &s->sort.hist, &s->sort.flags, &s->cpos,
                          &s->cstart, &s->path, &s->nbr16, &s->tables, &s->tcount, &s->slot_info, &s->recs, &s->nrecs,
                          &s->wsnap_pos, &s->wsnap_vel, &s->wsnap_orig, &s->walk_order, &s->walk_work,
                          &s->leaf_q, &s->leaf_c};
    for (auto *b : bufs) f(*b);
}

// FNV-1a over the address and size of every buffer a captured launch may
// use: a graph bakes the addresses in, so it may be replayed only while this
// is unchanged.
uint64_t buf_signature(nd_sim *s) {
    // the snapshot buffers are copied by the host, never by a captured launch
    const nd::DevBuf *host_only[] = {&s->snap_pos, &s->snap_vel, &s->snap_frc, &s->snap_orig,
                                     &s->wsnap_pos, &s->wsnap_vel, &s->wsnap_orig};
    uint64_t h = 1469598103934665603ull;
    for_each_buf(s, [&](const nd::DevBuf &b) {
        for (const nd::DevBuf *x : host_only)
            if (x == &b) return;
        const uint64_t w[2] = {(uint64_t)(uintptr_t)b.p, (uint64_t)b.bytes};
        for (uint64_t v : w)
            for (int k = 0; k < 8; ++k) h = (h ^ ((v >> (8 * k)) & 0xFF)) * 1099511628211ull;
    });
    return h;
}

// Host-issued copies of the simulation, each checked against the live
// allocation it names: a copy into a released buffer or past a buffer's end
// is an error here rather than a fault inside the runtime.
void check_span(const nd::DevBuf &b, size_t off, size_t bytes) {
    if (!b.p || off > b.bytes || bytes > b.bytes - off)
        throw nd::Error(ND_ERR_STATE, "simulation: copy outside a live device buffer");
}

void copy_d2d(const nd::DevBuf &dst, const nd::DevBuf &src, size_t bytes, hipStream_t st) {
    if (!bytes) return;
    check_span(dst, 0, bytes);
    check_span(src, 0, bytes);
    ND_HIP(hipMemcpyAsync(dst.p, src.p, bytes, hipMemcpyDeviceToDevice, st));
}

void copy_d2h(void *dst, const nd::DevBuf &src, size_t bytes, hipStream_t st) {
    if (!bytes) return;
    check_span(src, 0, bytes);
    ND_HIP(hipMemcpyAsync(dst, src.p, bytes, hipMemcpyDeviceToHost, st));
}

void copy_h2d(const nd::DevBuf &dst, const void *src, size_t bytes, hipStream_t st, size_t off = 0) {
    if (!bytes) return;
    check_span(dst, off, bytes);
    ND_HIP(hipMemcpyAsync(static_cast<char *>(dst.p) + off, src, bytes, hipMemcpyHostToDevice, st));
}

void alloc_lists(nd_sim *s, int32_t cap) {
    // a wave's block of the list (cap rows of 256 B) is addressed through a
    // buffer resource with a 31-bit byte range
    if (cap > nd::kMaxListRows)
        throw nd::Error(ND_ERR_CAPACITY, "simulation: 2^23 or more neighbours per atom");
    s->cap = cap;
    s->sp.cap = cap;
    // one block of cap rows x 64 slots per 64 atoms (n rounded up to 64)
    const size_t n64 = ((size_t)s->n + 63) & ~(size_t)63;
    s->nbr.reserve(sizeof(int32_t) * (size_t)cap * n64);
    s->nbr16.reserve(sizeof(uint32_t) * (size_t)rows16(cap) * n64);
    s->recs.reserve(sizeof(uint2) * (size_t)rec_stride(cap) * n64);
    drop_graphs(s);
}

hipEvent_t next_event(nd_sim *s) {
    if (s->ev_used == s->ev_pool.size()) {
        hipEvent_t e;
        ND_HIP(hipEventCreate(&e));
        s->ev_pool.push_back(e);
    }
    return s->ev_pool[s->ev_used++];
}

// ---- launch sequences (asynchronous) ----

int32_t resort_period(const nd_sim *s) {
    return s->p.resort_every > 0 ? s->p.resort_every : s->p.rebuild_every;
}

// The two-phase walk (records of four-wide nodes, leaf-local copies) runs
// the simulation's list when asked for and the leaf size and count allow it.
bool two_phase_walk(const nd_sim *s) {
    return s->p.list_walk == 0 && s->nL < kMaxLeaves3 &&
           (s->apl == 4 || s->apl == 8 || s->apl == 16 || s->apl == 32);
}

// A rebuild at step k re-sorts (new keys, order and topology) when the
// period says so or when no tree matches the current order.
bool resort_due(const nd_sim *s, int64_t k) { return !s->tree_valid || k % resort_period(s) == 0; }

void launch_sort_and_build(nd_sim *s);

// 16-bit list entries are table slot numbers (a shift of 4 would make them
// LDS byte offsets of 16-byte records, one instruction fewer per pair, which
// measured slower at config 2 and is not used).
int32_t entry_shift(const nd_sim *) { return 0; }

// Neighbourhood tables and 16-bit lists from the walk's lists.
void launch_tables(nd_sim *s) {
    k_table<<<nd::grid_for(s->n), nd::kBlock, 0, s->ctx->stream>>>(
        s->nbr.as<int32_t>(), s->ncount.as<int32_t>(), s->n, s->cap, rows16(s->cap), s->lcap, entry_shift(s),
        s->nbr16.as<uint32_t>(), s->tables.as<int32_t>(), s->tcount.as<int32_t>(), s->misc.as<SimMisc>(),
        s->hstale, s->slot_info.as<uint32_t>());
    ND_HIP(hipGetLastError());
}

// Neighbour-list rebuild: a full re-sort + tree build, or a box refit of the
// current tree, followed by the list walk.
void launch_rebuild(nd_sim *s, bool full) {
    s->since = 0;  // the next drift starts every path length at 0
    hipStream_t st = s->ctx->stream;
    const int32_t n = s->n;
    const unsigned g = nd::grid_for(n);
    SimMisc *misc = s->misc.as<SimMisc>();
    if (s->p.list_walk == 3) {
        // cell list: the atom order (for the force kernel) still follows the
        // sort key; the cell index is rebuilt from the current positions
        if (full) {
            launch_sort_and_build(s);
            s->tree_valid = true;
        } else if (s->wide) {
            k_pos32<<<g, nd::kBlock, 0, st>>>(s->pos[s->cur].as<double4>(), n, s->pos32.as<float4>());
        }
        const float4 *p4 = s->wide ? s->pos32.as<float4>() : s->pos[s->cur].as<float4>();
        float4 *cpos = s->cpos.as<float4>();
        int32_t *cstart = s->cstart.as<int32_t>();
        nd::launch_cell_index(st, s->sort, reinterpret_cast<const float *>(p4), 4, n, s->cg,
                              s->codes.as<int32_t>(), s->perm.as<int32_t>(), cpos, cstart);
        k_cell_rows<<<g, nd::kBlock, 0, st>>>(p4, cpos, cstart, n, s->cg, s->rlist * s->rlist, s->cap,
                                              s->nbr.as<int32_t>(), s->ncount.as<int32_t>(),
                                              reinterpret_cast<int32_t *>(misc));
        ND_HIP(hipGetLastError());
        launch_tables(s);
        return;
    }
    const bool two_phase = two_phase_walk(s);
    bool records_done = false;  // the four-wide records already hold the refitted boxes
    if (full) {
        launch_sort_and_build(s);
        s->tree_valid = true;
    } else {
        if (s->wide)
            k_pos32<<<g, nd::kBlock, 0, st>>>(s->pos[s->cur].as<double4>(), n, s->pos32.as<float4>());
        const float *p32 = s->wide ? reinterpret_cast<const float *>(s->pos32.p)
                                   : reinterpret_cast<const float *>(s->pos[s->cur].p);
        if (two_phase) {
            nd::launch_lbvh_refit_wide4(st, p32, 4, s->nL, s->apl, s->rlist, s->nodes.as<nd_node>(),
                                        s->store.as<int32_t>(), s->kids.as<float4>(), s->leaf_q.as<float4>(),
                                        s->leaf_c.as<float4>());
            records_done = true;
        } else {
            nd::launch_lbvh_refit(st, p32, 4, s->nL, s->apl, s->rlist, s->nodes.as<nd_node>(),
                                  s->store.as<int32_t>());
        }
    }
    const float4 *p4 = s->wide ? s->pos32.as<float4>() : s->pos[s->cur].as<float4>();
    const float hl = s->rlist / 2.0f, rl2 = s->rlist * s->rlist;
    int32_t *nb = s->nbr.as<int32_t>(), *nc = s->ncount.as<int32_t>();
    int32_t *ctr = reinterpret_cast<int32_t *>(misc);  // maxcnt, overflow, status
    if (s->p.list_walk == 1) {
        const nd_node *nodes = s->nodes.as<nd_node>();
        switch (s->apl) {
            case 4: k_nlist<4><<<g, nd::kBlock, 0, st>>>(p4, nodes, n, s->nL, 4, hl, rl2, s->cap, nb, nc, misc); break;
            case 8: k_nlist<8><<<g, nd::kBlock, 0, st>>>(p4, nodes, n, s->nL, 8, hl, rl2, s->cap, nb, nc, misc); break;
            case 16: k_nlist<16><<<g, nd::kBlock, 0, st>>>(p4, nodes, n, s->nL, 16, hl, rl2, s->cap, nb, nc, misc); break;
            case 32: k_nlist<32><<<g, nd::kBlock, 0, st>>>(p4, nodes, n, s->nL, 32, hl, rl2, s->cap, nb, nc, misc); break;
            default: k_nlist<0><<<g, nd::kBlock, 0, st>>>(p4, nodes, n, s->nL, s->apl, hl, rl2, s->cap, nb, nc, misc);
        }
    } else {
        float4 *kids = s->kids.as<float4>();
        if (two_phase) {
            if (!records_done) nd::launch_lbvh_wide4(st, s->nodes.as<nd_node>(), s->nL, kids);
        } else {
            nd::launch_lbvh_wide(st, s->nodes.as<nd_node>(), s->nL, kids);
        }
        uint2 *rc = s->recs.as<uint2>();
        int32_t *nr = s->nrecs.as<int32_t>();
        if (two_phase) {
            uint32_t *si = s->slot_info.as<uint32_t>();
            uint32_t *n16 = s->nbr16.as<uint32_t>();
            // leaf-local copies of the atoms for the walk's leaf tests (written
            // with the leaf boxes, by the permutation or the refit)
            const float4 *tql = s->leaf_q.as<float4>(), *lcl = s->leaf_c.as<float4>();
            int32_t *tb = s->tables.as<int32_t>(), *tc = s->tcount.as<int32_t>();
            const int32_t c2 = rows16(s->cap), es = entry_shift(s);
            if (!s->rec_big) {
                // walk and table build in one kernel
                const int32_t *ord = s->walk_order.as<int32_t>();
                int32_t *wk = s->walk_work.as<int32_t>();
#define ND_WALKTAB(A, F)                                                                                  \
    k_walk_table<A, F><<<g, nd::kBlock, 0, st>>>(p4, kids, n, s->nL, hl, rl2, s->cap, nb, nc, ctr, rc, nr, c2, \
                                                 s->lcap, es, n16, tb, tc, misc, s->hstale, si, ord, wk, tql, lcl)
                if (s->walk_front == 16) {
                    switch (s->apl) {
                        case 4: ND_WALKTAB(4, 16); break;
                        case 8: ND_WALKTAB(8, 16); break;
                        case 16: ND_WALKTAB(16, 16); break;
                        default: ND_WALKTAB(32, 16); break;
                    }
                } else {
                    switch (s->apl) {
                        case 4: ND_WALKTAB(4, 8); break;
                        case 8: ND_WALKTAB(8, 8); break;
                        case 16: ND_WALKTAB(16, 8); break;
                        default: ND_WALKTAB(32, 8); break;
                    }
                }
#undef ND_WALKTAB
                ND_HIP(hipGetLastError());
                s->order_pending = true;  // the next force step's tail orders the next walk
                return;
            }
            // dense workloads: the walk, then the table build with the larger
            // record buffer (its LDS would lower the walk's occupancy)
#define ND_NLIST3(A, F) k_nlist3<A, true, F><<<g, nd::kBlock, 0, st>>>(p4, kids, n, s->nL, hl, rl2, s->cap, nb, nc, ctr, rc, nr)
            if (s->walk_front == 16) {
                switch (s->apl) {
                    case 4: ND_NLIST3(4, 16); break;
                    case 8: ND_NLIST3(8, 16); break;
                    case 16: ND_NLIST3(16, 16); break;
                    default: ND_NLIST3(32, 16); break;
                }
            } else {
                switch (s->apl) {
                    case 4: ND_NLIST3(4, 8); break;
                    case 8: ND_NLIST3(8, 8); break;
                    case 16: ND_NLIST3(16, 8); break;
                    default: ND_NLIST3(32, 8); break;
                }
            }
#undef ND_NLIST3
            ND_HIP(hipGetLastError());
            k_table_rec<kRecLdsBig><<<g, nd::kBlock, 0, st>>>(rc, nr, nc, n, s->cap, c2, s->lcap, es, n16, tb, tc,
                                                              nb, misc, s->hstale, si);
            ND_HIP(hipGetLastError());
            return;
        }
        switch (s->apl) {
            case 4: k_nlist2<4><<<g, nd::kBlock, 0, st>>>(p4, kids, n, s->nL, 4, hl, rl2, s->cap, nb, nc, misc); break;
            case 8: k_nlist2<8><<<g, nd::kBlock, 0, st>>>(p4, kids, n, s->nL, 8, hl, rl2, s->cap, nb, nc, misc); break;
            case 16: k_nlist2<16><<<g, nd::kBlock, 0, st>>>(p4, kids, n, s->nL, 16, hl, rl2, s->cap, nb, nc, misc); break;
            case 32: k_nlist2<32><<<g, nd::kBlock, 0, st>>>(p4, kids, n, s->nL, 32, hl, rl2, s->cap, nb, nc, misc); break;
            default: k_nlist2<0><<<g, nd::kBlock, 0, st>>>(p4, kids, n, s->nL, s->apl, hl, rl2, s->cap, nb, nc, misc);
        }
    }
    ND_HIP(hipGetLastError());
    launch_tables(s);
}

// New keys, stable sort, state permutation, tree build.
void launch_sort_and_build(nd_sim *s) {
    hipStream_t st = s->ctx->stream;
    const int32_t n = s->n;
    const unsigned g = nd::grid_for(n);
    SimMisc *misc = s->misc.as<SimMisc>();
    const float lo = s->p.lo;
    const float scale = s->p.hi > s->p.lo ? 1024.0f / (s->p.hi - s->p.lo) : 0.0f;
    int32_t *codes = s->codes.as<int32_t>(), *perm = s->perm.as<int32_t>();
    // Key width: with at most 2^20 atoms the top 20 bits of the interleaved or
    // Hilbert key (cells of 1/101 of the box side, about one atom each at the
    // densities here) order the atoms as well as all 30 do for the tree and
    // the walk, and take two radix passes instead of three; the tree is built
    // on the same truncated keys.  The masked key keeps its 30 bits.
    const int32_t key_bits = (s->p.morton_mode != ND_MORTON_MASKED && n <= (1 << 20)) ? 20 : 30;
    if (s->wide)
        k_sim_keys<double><<<g, nd::kBlock, 0, st>>>(s->pos[s->cur].as<double4>(), n, s->p.morton_mode,
                                                     lo, scale, 30 - key_bits, codes, perm, misc,
                                                     s->store.as<int32_t>(), s->nL - 1);
    else
        k_sim_keys<float><<<g, nd::kBlock, 0, st>>>(s->pos[s->cur].as<float4>(), n, s->p.morton_mode,
                                                    lo, scale, 30 - key_bits, codes, perm, misc,
                                                    s->store.as<int32_t>(), s->nL - 1);
    // keys are < 2^key_bits, so the sign-flipped key's higher bits are constant
    nd::radix_sort_keys10(st, s->sort, codes, perm, n, key_bits, &codes, &perm);
    const int a = s->cur ^ 1, va = s->vcur ^ 1;
    // leaf boxes in the permutation for power-of-two leaves up to 64 (the
    // cell walk needs no tree, so none there)
    const bool boxes = s->p.list_walk != 3 && s->apl <= 64 && (s->apl & (s->apl - 1)) == 0;
    // the two-phase walk's leaf-local copies, written with the leaf boxes
    float4 *lq = two_phase_walk(s) ? s->leaf_q.as<float4>() : nullptr;
    float4 *lcn = two_phase_walk(s) ? s->leaf_c.as<float4>() : nullptr;
    const float h = s->rlist / 2.0f;  // as launch_lbvh computes it
    nd_node *nodes = s->nodes.as<nd_node>();
#define ND_PERMUTE(PT, SEG)                                                                              \
    k_permute<PT, SEG><<<g, nd::kBlock, 0, st>>>(                                                        \
        perm, n, s->pos[s->cur].as<V4<PT>::T>(), s->pos[a].as<V4<PT>::T>(), s->vel[s->vcur].as<V4<PT>::T>(), \
        s->vel[va].as<V4<PT>::T>(), s->orig[s->vcur].as<int32_t>(), s->orig[va].as<int32_t>(),           \
        s->wide ? s->pos32.as<float4>() : nullptr, s->wide ? fix_ptr(s, a) : nullptr, s->sp, misc, h, nodes, \
        lq, lcn)
#define ND_PERMUTE_SEG(PT)                                        \
    switch (boxes ? s->apl : 0) {                                \
        case 1: ND_PERMUTE(PT, 1); break;                         \
        case 2: ND_PERMUTE(PT, 2); break;                         \
        case 4: ND_PERMUTE(PT, 4); break;                         \
        case 8: ND_PERMUTE(PT, 8); break;                         \
        case 16: ND_PERMUTE(PT, 16); break;                       \
        case 32: ND_PERMUTE(PT, 32); break;                       \
        case 64: ND_PERMUTE(PT, 64); break;                       \
        default: ND_PERMUTE(PT, 0);                               \
    }
    if (s->wide) {
        ND_PERMUTE_SEG(double)
    } else {
        ND_PERMUTE_SEG(float)
    }
#undef ND_PERMUTE_SEG
#undef ND_PERMUTE
    s->cur = a;
    s->vcur = va;
    if (s->p.list_walk == 3) {  // the cell list needs no tree
        ND_HIP(hipGetLastError());
        return;
    }
    const float *p32 = s->wide ? reinterpret_cast<const float *>(s->pos32.p)
                               : reinterpret_cast<const float *>(s->pos[s->cur].p);
    nd::launch_lbvh(st, p32, 4, codes, s->nL, s->apl, s->rlist, 0, s->nodes.as<nd_node>(),
                    s->store.as<int32_t>(), &misc->status, true, boxes);
    ND_HIP(hipGetLastError());
}

// check: the positions this launch drifts to will be used with the current
// list (the next step is no rebuild step); a fused launch counts as a drift.
template <class PosS, class CalcS, class AccS, int MODE, bool E, int PB, int LPA>
void launch_step_pb(nd_sim *s, bool check) {
    using P4 = typename V4<PosS>::T;
    hipStream_t st = s->ctx->stream;
    const unsigned g = nd::grid_for(s->n);
    const int a = s->cur ^ 1;
    using G = std::conditional_t<sizeof(PosS) == 8 && sizeof(CalcS) == 4, int4, P4>;
    const size_t lds = sizeof(G) * (size_t)(s->lcap + 1);  // the table and its sentinel slot
    if (lds > 65536) {
        // dynamic LDS past 64 KiB must be allowed per kernel (once)
        static bool allowed = false;
        if (!allowed) {
            ND_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(&k_step<PosS, CalcS, AccS, MODE, E, PB, LPA>),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
            allowed = true;
        }
    }
    k_step<PosS, CalcS, AccS, MODE, E, PB, LPA><<<g, nd::kBlock * LPA, lds, st>>>(
        s->pos[s->cur].as<P4>(), s->pos[a].as<P4>(), s->vel[s->vcur].as<P4>(), s->frc.as<P4>(),
        s->nbr.as<int32_t>(), s->sp, s->partials.as<double>(),
        s->misc.as<SimMisc>(), fix_ptr(s, s->cur), fix_ptr(s, a), walls_of(s), s->path.as<uint16_t>(),
        s->since == 0 ? 1 : 0, check ? 1 : 0, s->nbr16.as<uint32_t>(), s->tables.as<int32_t>(),
        s->tcount.as<int32_t>(), s->lcap, s->slot_info.as<uint32_t>());
    if (MODE == kModeFused) s->since++;
}

// Pairs per pipeline batch: FP32 takes the simulation's pair_batch (2, 4 or
// 8, picked by the batched form's autotuner); MIXED and FP64 use 4.
template <class PosS, class CalcS, class AccS, int MODE, bool E>
void launch_step_t(nd_sim *s, bool check) {
    if constexpr (std::is_same_v<CalcS, float> && std::is_same_v<PosS, float>) {
        if (s->atom_lanes == 2) {
            if (s->pair_batch == 2) return launch_step_pb<PosS, CalcS, AccS, MODE, E, 2, 2>(s, check);
            if (s->pair_batch == 8) return launch_step_pb<PosS, CalcS, AccS, MODE, E, 8, 2>(s, check);
            return launch_step_pb<PosS, CalcS, AccS, MODE, E, kPairBatch, 2>(s, check);
        }
        if (s->pair_batch == 2) return launch_step_pb<PosS, CalcS, AccS, MODE, E, 2, 1>(s, check);
        if (s->pair_batch == 8) return launch_step_pb<PosS, CalcS, AccS, MODE, E, 8, 1>(s, check);
    } else {
        if (s->atom_lanes == 2) return launch_step_pb<PosS, CalcS, AccS, MODE, E, kPairBatch, 2>(s, check);
    }
    launch_step_pb<PosS, CalcS, AccS, MODE, E, kPairBatch, 1>(s, check);
}

template <int MODE, bool E>
void launch_step_prec(nd_sim *s, bool check) {
    switch (s->p.precision) {
        case ND_PREC_FP32: launch_step_t<float, float, float, MODE, E>(s, check); break;
        case ND_PREC_MIXED: launch_step_t<double, float, double, MODE, E>(s, check); break;
        default: launch_step_t<double, double, double, MODE, E>(s, check); break;
    }
}

// Whether the positions reached at step k are used with the current list
// (false when step k rebuilds it).
bool uses_list(const nd_sim *s, int64_t k) { return k % s->p.rebuild_every != 0; }

// One force-evaluating kernel for global step k, plus its energy row.
void launch_force_step(nd_sim *s, int mode, int64_t k) {
    const bool e = s->p.energy_every > 0 && (k % s->p.energy_every) == 0;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (s->timing) {
        e0 = next_event(s);
        e1 = next_event(s);
        ND_HIP(hipEventRecord(e0, s->ctx->stream));
    }
    const bool c = uses_list(s, k + 1);  // a fused launch drifts to step k + 1
    if (mode == kModeForce) e ? launch_step_prec<kModeForce, true>(s, c) : launch_step_prec<kModeForce, false>(s, c);
    else if (mode == kModeFused) e ? launch_step_prec<kModeFused, true>(s, c) : launch_step_prec<kModeFused, false>(s, c);
    else e ? launch_step_prec<kModeFinal, true>(s, c) : launch_step_prec<kModeFinal, false>(s, c);
    ND_HIP(hipGetLastError());
    if (s->timing) ND_HIP(hipEventRecord(e1, s->ctx->stream));
    if (mode == kModeFused) s->cur ^= 1;
    // the energy row and, after a fused walk, the next walk's block order
    const bool ord = s->order_pending;
    s->order_pending = false;
    const int32_t nb = (int32_t)nd::grid_for(s->n);
    if (e) {
        k_energy_finalize<<<ord ? 9 : 1, nd::kBlock, 0, s->ctx->stream>>>(
            s->partials.as<double>(), nb, s->rows.as<double>(), s->max_rows, s->misc.as<SimMisc>(),
            s->walk_work.as<int32_t>(), nb, s->cus_per_xcd, s->walk_order.as<int32_t>());
        ND_HIP(hipGetLastError());
    } else if (ord) {
        k_walk_order<<<8, nd::kBlock, 0, s->ctx->stream>>>(s->walk_work.as<int32_t>(), nb, s->cus_per_xcd,
                                                           s->walk_order.as<int32_t>());
        ND_HIP(hipGetLastError());
    }
}

// Kick-drift from step k to k + 1 (after the thermostat's rescale if asked).
void launch_kick_drift(nd_sim *s, bool rescale, int64_t k) {
    hipStream_t st = s->ctx->stream;
    const unsigned g = nd::grid_for(s->n);
    const int a = s->cur ^ 1;
    const int32_t first = s->since == 0 ? 1 : 0, check = uses_list(s, k + 1) ? 1 : 0;
    uint16_t *path = s->path.as<uint16_t>();
    if (s->wide)
        k_kick_drift<double><<<g, nd::kBlock, 0, st>>>(s->pos[s->cur].as<double4>(), s->pos[a].as<double4>(),
                                                       s->vel[s->vcur].as<double4>(), s->frc.as<double4>(), s->sp,
                                                       s->misc.as<SimMisc>(), rescale, s->p.temperature, s->p.vdamp,
                                                       fix_ptr(s, a), walls_of(s), path, first, check);
    else
        k_kick_drift<float><<<g, nd::kBlock, 0, st>>>(s->pos[s->cur].as<float4>(), s->pos[a].as<float4>(),
                                                      s->vel[s->vcur].as<float4>(), s->frc.as<float4>(), s->sp,
                                                      s->misc.as<SimMisc>(), rescale, s->p.temperature, s->p.vdamp,
                                                      nullptr, walls_of(s), path, first, check);
    ND_HIP(hipGetLastError());
    s->cur ^= 1;
    s->since++;
}

bool thermo_step(const nd_sim *s, int64_t k) {
    return s->p.thermostat_every > 0 && k > 0 && (k % s->p.thermostat_every) == 0;
}

// Mid-step state (x_k, v_{k-1/2}) -> (x_{k+1}, v_{k+1/2}) for step index k.
void launch_middle_step(nd_sim *s, int64_t k) {
    if (thermo_step(s, k)) {
        // needs KE(v_k) before the kick: final-mode force, energy row, rescale + kick-drift
        const int32_t ee = s->p.energy_every;
        s->p.energy_every = 1;
        launch_force_step(s, kModeFinal, k);
        s->p.energy_every = ee;
        launch_kick_drift(s, true, k);
    } else {
        launch_force_step(s, kModeFused, k);
    }
}

void read_misc(nd_sim *s, SimMisc &m) {
    copy_d2h(&m, s->misc, sizeof(SimMisc), s->ctx->stream);
    ND_HIP(hipStreamSynchronize(s->ctx->stream));
}

// Eager rebuild with an immediate capacity check (used when the list is stale
// at the start of a run): grows the per-atom capacity until everything fits.
void rebuild_checked(nd_sim *s) {
    for (int attempt = 0; attempt < 8; ++attempt) {
        const int cur0 = s->cur;
        launch_rebuild(s, true);
        SimMisc m;
        read_misc(s, m);
        if (m.status & nd::kStatusDomain)
            throw nd::Error(ND_ERR_DOMAIN, "simulation: coordinate outside the key domain");
        if (m.status & nd::kStatusTree)
            throw nd::Error(ND_ERR_STATE, "simulation: malformed neighbour tree");
        s->last_maxcnt = m.maxcnt;
        if (!s->rec_big && (int64_t)m.rec_spill * 64 > (int64_t)nd::grid_for(s->n)) {
            // later builds hold more records per workgroup in LDS; captured
            // windows used the smaller form
            s->rec_big = true;
            drop_graphs(s);
        }
        if (!m.overflow) {
            s->list_valid = true;
            s->rebuilds++;
            return;
        }
        // The permutation is valid; only the list overflowed.
        (void)cur0;
        alloc_lists(s, m.maxcnt + m.maxcnt / 4 + 8);
        k_misc_reset_counters<<<1, 64, 0, s->ctx->stream>>>(s->misc.as<SimMisc>());
    }
    throw nd::Error(ND_ERR_CAPACITY, "simulation: neighbour list capacity could not be sized");
}

void snapshot(nd_sim *s) {
    hipStream_t st = s->ctx->stream;
    const size_t pb = pos_bytes(s);
    s->snap_pos.reserve(pb);
    s->snap_vel.reserve(pb);
    s->snap_frc.reserve(pb);
    s->snap_orig.reserve(sizeof(int32_t) * (size_t)s->n);
    copy_d2d(s->snap_pos, s->pos[s->cur], pb, st);
    copy_d2d(s->snap_vel, s->vel[s->vcur], pb, st);
    copy_d2d(s->snap_frc, s->frc, pb, st);
    copy_d2d(s->snap_orig, s->orig[s->vcur], sizeof(int32_t) * (size_t)s->n, st);
}

void restore(nd_sim *s) {
    hipStream_t st = s->ctx->stream;
    const size_t pb = pos_bytes(s);
    copy_d2d(s->pos[s->cur], s->snap_pos, pb, st);
    copy_d2d(s->vel[s->vcur], s->snap_vel, pb, st);
    copy_d2d(s->frc, s->snap_frc, pb, st);
    copy_d2d(s->orig[s->vcur], s->snap_orig, sizeof(int32_t) * (size_t)s->n, st);
}

bool graph_eligible(const nd_sim *s) {
    if (!s->p.use_graph || s->timing) return false;
    const int32_t W = s->p.rebuild_every;
    if (s->p.energy_every > 0 && W % s->p.energy_every != 0) return false;
    if (s->p.thermostat_every > 0 && W % s->p.thermostat_every != 0) return false;
    return true;
}

// Window at k0 (k0 % W == 0): rebuild(x_k0) + middle steps k0 .. k0+W-1.
void launch_window(nd_sim *s, int64_t k0, bool full) {
    launch_rebuild(s, full);
    s->rebuilds++;
    s->last_build_step = k0;
    for (int32_t o = 0; o < s->p.rebuild_every; ++o) launch_middle_step(s, k0 + o);
}

// Graph of the window that starts at k0 with position buffer cur, velocity
// buffer vcur and the given kind (re-sort or refit), captured and
// instantiated once; capture executes nothing and leaves the host-side state
// as it found it.
hipGraphExec_t window_graph(nd_sim *s, int64_t k0, int cur, int vcur, bool full) {
    const int key = cur | (vcur << 1) | (full ? 4 : 0);
    const uint64_t sig = buf_signature(s);
    if (s->gexec[key] && s->gsig[key] != sig) {
        // a buffer moved since capture (every reallocation drops the graphs,
        // so this is a guard): never replay stale addresses, capture again
        (void)hipGraphExecDestroy(s->gexec[key]);
        s->gexec[key] = nullptr;
        s->graph_recaptures++;
    }
    if (s->gexec[key]) return s->gexec[key];
    hipStream_t st = s->ctx->stream;
    hipGraph_t graph;
    const int cur0 = s->cur, vcur0 = s->vcur, since0 = s->since;
    const int64_t rb0 = s->rebuilds, lb0 = s->last_build_step;
    const bool tv0 = s->tree_valid, op0 = s->order_pending;
    s->cur = cur;
    s->vcur = vcur;
    s->tree_valid = true;  // a refit window is captured only for a valid tree
    ND_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    try {
        launch_window(s, k0, full);
    } catch (...) {
        (void)hipStreamEndCapture(st, &graph);
        s->cur = cur0;
        s->vcur = vcur0;
        s->since = since0;
        s->rebuilds = rb0;
        s->last_build_step = lb0;
        s->tree_valid = tv0;
        s->order_pending = op0;
        throw;
    }
    s->cur = cur0;
    s->vcur = vcur0;
    s->since = since0;
    s->rebuilds = rb0;
    s->last_build_step = lb0;
    s->tree_valid = tv0;
    s->order_pending = op0;
    ND_HIP(hipStreamEndCapture(st, &graph));
    const hipError_t e = hipGraphInstantiate(&s->gexec[key], graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    ND_HIP(e);
    s->gsig[key] = sig;
    return s->gexec[key];
}

void run_window_graph(nd_sim *s, int64_t k0) {
    const bool full = resort_due(s, k0);
    ND_HIP(hipGraphLaunch(window_graph(s, k0, s->cur, s->vcur, full), s->ctx->stream));
    // Host-side bookkeeping of launch_window: a re-sort swaps the state
    // buffers once and every middle step swaps the positions once more (a
    // fused step, or a final step followed by a kick-drift).
    s->rebuilds++;
    s->last_build_step = k0;
    s->since = s->p.rebuild_every;  // one drift per step of the window
    s->order_pending = false;       // the window's first force step ordered its walk
    if (full) {
        s->cur ^= (1 + s->p.rebuild_every) & 1;
        s->vcur ^= 1;
        s->tree_valid = true;
    } else {
        s->cur ^= s->p.rebuild_every & 1;
    }
}

// Reads and clears the stale flag (synchronises).
bool take_stale(nd_sim *s) {
    SimMisc m;
    read_misc(s, m);
    if (!m.stale) return false;
    m.stale = 0;
    copy_h2d(s->misc, &m.stale, sizeof(int32_t), s->ctx->stream, offsetof(SimMisc, stale));
    ND_HIP(hipStreamSynchronize(s->ctx->stream));
    return true;
}

// State at the start of a window, for a checked rerun: the atoms (positions,
// velocities, original indices in the current buffers) and the host-side
// counters; the stored force is not needed (a window's fused steps compute
// theirs), and the tree and list are rebuilt by the redone window.
struct WindowStart {
    int cur, vcur, since;
    int64_t rebuilds, last_build_step;
    SimMisc misc;
};

WindowStart window_snapshot(nd_sim *s) {
    hipStream_t st = s->ctx->stream;
    const size_t pb = pos_bytes(s);
    s->wsnap_pos.reserve(pb);
    s->wsnap_vel.reserve(pb);
    s->wsnap_orig.reserve(sizeof(int32_t) * (size_t)s->n);
    copy_d2d(s->wsnap_pos, s->pos[s->cur], pb, st);
    copy_d2d(s->wsnap_vel, s->vel[s->vcur], pb, st);
    copy_d2d(s->wsnap_orig, s->orig[s->vcur], sizeof(int32_t) * (size_t)s->n, st);
    WindowStart w{s->cur, s->vcur, s->since, s->rebuilds, s->last_build_step, SimMisc{}};
    read_misc(s, w.misc);  // synchronises
    return w;
}

// Back to the window's start; the next rebuild re-sorts (it also rewrites
// the fixed-point and FP32 copies of the positions from the restored ones).
void window_restore(nd_sim *s, const WindowStart &w) {
    hipStream_t st = s->ctx->stream;
    s->cur = w.cur;
    s->vcur = w.vcur;
    s->since = w.since;
    s->rebuilds = w.rebuilds;
    s->last_build_step = w.last_build_step;
    s->tree_valid = false;
    const size_t pb = pos_bytes(s);
    copy_d2d(s->pos[s->cur], s->wsnap_pos, pb, st);
    copy_d2d(s->vel[s->vcur], s->wsnap_vel, pb, st);
    copy_d2d(s->orig[s->vcur], s->wsnap_orig, sizeof(int32_t) * (size_t)s->n, st);
    SimMisc m = w.misc;
    m.stale = 0;
    copy_h2d(s->misc, &m, sizeof(SimMisc), st);
    ND_HIP(hipStreamSynchronize(st));
}

// Steps before checked_until are checked (a rerun after a stale list): each
// whole rebuild window is still replayed as a graph, then its stale flag is
// read; a window that went stale is restored and redone with eager launches,
// reading the flag after every drift, and a flagged drift makes the next step
// rebuild the list first (a forced rebuild), so no force is evaluated with a
// list that may miss a pair.  Later windows run unchecked, at most kInFlight
// ahead of the device; when the host sees the stale flag of a finished
// window it stops launching and returns false (the caller reruns).
bool run_once(nd_sim *s, int32_t nsteps, int64_t checked_until) {
    const int32_t W = s->p.rebuild_every;
    int64_t k = s->step;
    const bool had_force = s->have_force;
    if (!s->list_valid) {
        // the stored force is in the order before this permutation: recompute it
        rebuild_checked(s);
        s->last_build_step = k;
        s->have_force = false;
    }
    if (!s->have_force) {
        // a recomputed force (after a rerun's re-sort) records no second
        // energy row for step k
        const int32_t ee = s->p.energy_every;
        if (had_force) s->p.energy_every = 0;
        launch_force_step(s, kModeForce, k);
        s->p.energy_every = ee;
        s->have_force = true;
    }
    if (nsteps == 0) return true;
    // step k -> k+1 starts with a kick-drift from the stored force (after the
    // velocity rescale if step k is a thermostat step; the previous run's
    // final step recorded KE(v_k) for it)
    launch_kick_drift(s, thermo_step(s, k), k);
    bool forced = k < checked_until && take_stale(s);
    const int64_t kend = k + nsteps;  // integer step reached at the end
    int64_t j = k + 1;                // next force evaluation index
    const bool use_graph = graph_eligible(s);
    int64_t windows = 0;  // unchecked windows launched
    while (j < kend) {
        const bool checked = j < checked_until;
        if (j % W == 0 || forced) {
            if (use_graph && j % W == 0 && j + W <= kend && !forced) {
                if (!checked) {
                    if (windows >= s->kInFlight) {
                        ND_HIP(hipEventSynchronize(s->wev[(windows - s->kInFlight) % (s->kInFlight + 1)]));
                        if (*s->hstale) {
                            s->step = j;  // where the host stopped (the caller restores)
                            return false;
                        }
                    }
                    run_window_graph(s, j);
                    ND_HIP(hipEventRecord(s->wev[windows % (s->kInFlight + 1)], s->ctx->stream));
                    ++windows;
                    j += W;
                    continue;
                }
                const WindowStart ws = window_snapshot(s);
                run_window_graph(s, j);
                if (!take_stale(s)) {
                    j += W;
                    continue;
                }
                window_restore(s, ws);  // redo this window step by step
            }
            launch_rebuild(s, resort_due(s, j));
            s->rebuilds++;
            if (j % W != 0) s->forced_rebuilds++;
            s->last_build_step = j;
            forced = false;
        }
        launch_middle_step(s, j);
        if (checked) forced = take_stale(s);
        ++j;
    }
    // final integer step: rebuild if due, force + finishing half kick
    if ((kend % W == 0 && s->last_build_step != kend) || forced) {
        launch_rebuild(s, resort_due(s, kend));
        s->rebuilds++;
        if (kend % W != 0) s->forced_rebuilds++;
        s->last_build_step = kend;
    }
    if (thermo_step(s, kend)) {
        const int32_t ee = s->p.energy_every;
        s->p.energy_every = 1;
        launch_force_step(s, kModeFinal, kend);
        s->p.energy_every = ee;
    } else {
        launch_force_step(s, kModeFinal, kend);
    }
    s->step = kend;
    return true;
}

void set_step_params(nd_sim *s) {
    const nd_sim_params &p = s->p;
    StepParams &q = s->sp;
    q = StepParams{};
    q.n = s->n;
    q.cap = s->cap;
    q.walls = p.walls;
    const double eps = p.eps, sig = p.sigma, rc = p.rc, kc = p.k_coul, rcc = p.rc_coul;
    const double rcmax = rc > rcc ? rc : rcc;
    double eljs = 0.0;
    if (rc > 0.0) {
        const double c2 = sig * sig / (rc * rc), c6 = c2 * c2 * c2;
        eljs = 4.0 * eps * c6 * (c6 - 1.0);
    }
    q.deps24 = 24.0 * eps; q.deps4 = 4.0 * eps; q.dsigma2 = sig * sig;
    q.drc2 = rc * rc; q.drcc2 = rcc * rcc; q.drcmax2 = rcmax * rcmax; q.dkc = kc;
    q.delj_shift = 0.0;  // plain (unshifted) LJ energy, as the textbook form
    q.decoul_inv_rc = 0.0;
    (void)eljs;
    q.eps24 = (float)q.deps24; q.eps4 = (float)q.deps4; q.sigma2 = (float)q.dsigma2;
    q.rc2 = (float)q.drc2; q.rcc2 = (float)q.drcc2; q.rcmax2 = (float)q.drcmax2; q.kc = (float)kc;
    q.elj_shift = 0.0f; q.ecoul_inv_rc = 0.0f;
    q.dt = p.dt; q.lo = p.lo; q.hi = p.hi;
    q.dtf = p.dt; q.lof = p.lo; q.hif = p.hi;
    // fixed-point frame of the MIXED pair kernel: centred on the box; with
    // walls it spans the box, without them the box and 1.5 box lengths on
    // each side.  The scale is a power of two, so the unit 1/s is exact in
    // FP32 (an inexact unit would scale every difference by the same
    // relative error and bias the pressure coherently).
    const double span = p.hi > p.lo ? p.hi - p.lo : 1.0;
    const double half = p.walls ? 0.5 * span * 1.001 + 1e-6 : 2.0 * span;
    q.fx_c = 0.5 * (p.lo + p.hi);
    q.fx_s = std::ldexp(1.0, (int)std::floor(std::log2(0.999 * 2147483648.0 / half)));
    q.fx_inv = (float)(1.0 / q.fx_s);
    // 2^14 path-length units make half the skin; with no skin any motion
    // saturates the count (the list is valid only at its own build)
    q.path_scale = p.skin > 0.0 ? (float)(16384.0 / (0.5 * p.skin)) : 3.0e38f;
}

}  // namespace

extern "C" {

int nd_sim_create(nd_ctx *ctx, const nd_sim_params *p, nd_sim **out) {
    ND_API_BEGIN
    ND_REQUIRE(ctx && p && out, "nd_sim_create: null argument");
    *out = nullptr;
    ND_REQUIRE(p->n >= 2, "nd_sim_create: need at least two atoms");
    // the pair kernel addresses positions through a buffer resource whose
    // byte range (32 B per atom in FP64) must fit in 31 bits
    ND_REQUIRE(p->n <= (1 << 26), "nd_sim_create: at most 2^26 atoms");
    ND_REQUIRE(p->atomsperleaf >= 1 && p->n % p->atomsperleaf == 0 && p->n / p->atomsperleaf >= 2,
               "nd_sim_create: atomsperleaf must divide n and give >= 2 leaves");
    ND_REQUIRE(p->precision >= ND_PREC_FP32 && p->precision <= ND_PREC_FP64, "nd_sim_create: precision");
    ND_REQUIRE(p->morton_mode == ND_MORTON_MASKED || p->morton_mode == ND_MORTON_INTERLEAVE ||
                   p->morton_mode == ND_MORTON_HILBERT,
               "nd_sim_create: key mode");
    ND_REQUIRE(p->rebuild_every >= 1, "nd_sim_create: rebuild_every >= 1");
    ND_REQUIRE(p->resort_every >= 0 && p->resort_every % p->rebuild_every == 0,
               "nd_sim_create: resort_every must be a multiple of rebuild_every (or 0)");
    ND_REQUIRE(p->list_walk >= 0 && p->list_walk <= 3, "nd_sim_create: list_walk must be 0, 1, 2 or 3");
    ND_REQUIRE(p->pair_batch == 0 || p->pair_batch == 2 || p->pair_batch == 4 || p->pair_batch == 8,
               "nd_sim_create: pair_batch must be 0, 2, 4 or 8");
    ND_REQUIRE(p->walk_front == 0 || p->walk_front == 8 || p->walk_front == 16,
               "nd_sim_create: walk_front must be 0, 8 or 16");
    ND_REQUIRE(p->atom_lanes >= 0 && p->atom_lanes <= 2, "nd_sim_create: atom_lanes must be 0, 1 or 2");
    ND_REQUIRE(p->list_walk != 3 || p->hi > p->lo, "nd_sim_create: the cell-list walk needs a box (hi > lo)");
    ND_REQUIRE(p->dt > 0.0f && p->rc >= 0.0f && p->rc_coul >= 0.0f && p->skin >= 0.0f,
               "nd_sim_create: dt > 0, cutoffs and skin >= 0");
    ND_REQUIRE(p->rc > 0.0f || p->rc_coul > 0.0f, "nd_sim_create: no interaction enabled");
    ND_REQUIRE(p->hi > p->lo || (p->morton_mode == ND_MORTON_MASKED && !p->walls),
               "nd_sim_create: box hi must exceed lo");
    ND_REQUIRE(p->hi > p->lo || p->precision != ND_PREC_MIXED,
               "nd_sim_create: the mixed form needs a box (hi > lo) for its fixed-point frame");
    ND_HIP(hipSetDevice(ctx->device));
    nd_sim *s = new nd_sim();
    try {
        s->ctx = ctx;
        s->p = *p;
        if (p->pair_batch) s->pair_batch = p->pair_batch;
        if (p->walk_front) s->walk_front = p->walk_front;
        if (p->atom_lanes) s->atom_lanes = p->atom_lanes;
        s->n = p->n;
        s->apl = p->atomsperleaf;
        s->nL = p->n / p->atomsperleaf;
        s->wide = p->precision != ND_PREC_FP32;
        // The list is built from FP32 coordinates; pad its radius by the FP32
        // rounding of coordinates in the box (2^-20 relative) so that it stays
        // a superset of the pairs inside the cutoff in every precision form.
        const double rl = (p->rc > p->rc_coul ? p->rc : p->rc_coul) + p->skin;
        const double span = std::fabs(p->lo) > std::fabs(p->hi) ? std::fabs(p->lo) : std::fabs(p->hi);
        s->rlist = (float)(rl + (span + rl) * 9.5367431640625e-07);
        const size_t pb = pos_bytes(s), n = (size_t)s->n;
        for (int b = 0; b < 2; ++b) {
            s->pos[b].reserve(pb);
            s->vel[b].reserve(pb);
            s->orig[b].reserve(sizeof(int32_t) * n);
        }
        s->frc.reserve(pb);
        if (s->wide) s->pos32.reserve(sizeof(float4) * n);
        if (p->precision == ND_PREC_MIXED)
            for (int b = 0; b < 2; ++b) s->fix[b].reserve(sizeof(int4) * n);
        s->codes.reserve(sizeof(int32_t) * n);
        s->perm.reserve(sizeof(int32_t) * n);
        s->nodes.reserve(sizeof(nd_node) * (size_t)(2 * s->nL - 1));
        s->store.reserve(sizeof(int32_t) * nd::lbvh_scratch_ints(s->nL));
        s->kids.reserve(8 * sizeof(float4) * (size_t)(s->nL - 1));  // four-wide records
        s->ncount.reserve(sizeof(int32_t) * n);
        s->path.reserve(sizeof(uint16_t) * n);
        s->nrecs.reserve(sizeof(int32_t) * n);
        // neighbourhood tables in the step kernel's LDS: records of 16 B
        // (FP32, MIXED) for four workgroups per CU, of 32 B (FP64) for two;
        // each workgroup's share of the CU's LDS holds the table, its sentinel
        // slot and the kernel's static LDS (at most 256 B).  On a 160 KiB CU:
        // 2,543 records (one record more would drop a workgroup per CU).
        {
            int lds_cu = 65536, lds_blk = 65536;
            ND_HIP(hipDeviceGetAttribute(&lds_cu, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor,
                                         ctx->device));
            ND_HIP(hipDeviceGetAttribute(&lds_blk, hipDeviceAttributeMaxSharedMemoryPerBlock, ctx->device));
            const int rec_bytes = p->precision == ND_PREC_FP64 ? 32 : 16;
            const int per_wg = std::min(lds_cu / (rec_bytes == 16 ? 4 : 2), lds_blk) - 256;
            const int fit = per_wg / rec_bytes - 1;  // the sentinel slot
            s->lcap = fit < 2560 ? fit : 2560;
            if (s->lcap < 32) s->lcap = 32;
            // table positions (and the sentinel slot) fit 16-bit entries
            if (s->lcap + 1 > 0xFFFF) throw nd::Error(ND_ERR_STATE, "simulation: table capacity beyond 16-bit entries");
        }
        s->tables.reserve(sizeof(int32_t) * (size_t)nd::grid_for(s->n) * (size_t)s->lcap);
        s->tcount.reserve(sizeof(int32_t) * (size_t)nd::grid_for(s->n));
        s->slot_info.reserve(sizeof(uint32_t) * (size_t)nd::grid_for(s->n) * nd::kBlock);
        {
            // block placement of the fused walk: the identity order, no estimates yet
            const int32_t nb = (int32_t)nd::grid_for(s->n);
            std::vector<int32_t> iota(nb);
            for (int32_t b = 0; b < nb; ++b) iota[b] = b;
            s->walk_order.reserve(sizeof(int32_t) * (size_t)nb);
            s->walk_work.reserve(sizeof(int32_t) * 4 * (size_t)nb);
            s->leaf_q.reserve(sizeof(float4) * (size_t)s->n);
            s->leaf_c.reserve(sizeof(float4) * (size_t)(s->nL > 0 ? s->nL : 1));
            ND_HIP(hipMemcpyAsync(s->walk_order.p, iota.data(), sizeof(int32_t) * (size_t)nb,
                                  hipMemcpyHostToDevice, ctx->stream));
            ND_HIP(hipMemsetAsync(s->walk_work.p, 0, sizeof(int32_t) * 4 * (size_t)nb, ctx->stream));
            ND_HIP(hipStreamSynchronize(ctx->stream));
            int cus = 256;
            ND_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ctx->device));
            s->cus_per_xcd = cus >= 8 ? cus / 8 : 1;
        }
        s->misc.reserve(sizeof(SimMisc));
        s->partials.reserve(sizeof(double) * 4 * (size_t)nd::grid_for(s->n));
        s->rows.reserve(sizeof(double) * 5 * (size_t)s->max_rows);
        const int32_t ntiles = (int32_t)nd::radix_scratch_blocks(s->n);
        s->sort.keys_alt.reserve(sizeof(int32_t) * n);
        s->sort.vals_alt.reserve(sizeof(int32_t) * n);
        s->sort.hist.reserve(sizeof(int32_t) * (1024 * (size_t)ntiles + 1024));  // 10-bit digits
        if (p->list_walk == 3) {
            const double lo3[3] = {p->lo, p->lo, p->lo}, ext = p->hi - p->lo;
            const double ext3[3] = {ext, ext, ext};
            s->cg = nd::make_cell_grid(lo3, ext3, (double)s->rlist, std::min<int64_t>(1 << 22, 2 * (int64_t)n + 64));
            s->cpos.reserve(sizeof(float4) * n);
            s->cstart.reserve(sizeof(int32_t) * ((size_t)s->cg.ncell + 1));
        }
        ND_HIP(hipMemsetAsync(s->misc.p, 0, sizeof(SimMisc), ctx->stream));
        {
            void *h = nullptr;
            ND_HIP(hipHostMalloc(&h, 64, hipHostMallocDefault));
            s->hstale = static_cast<volatile int32_t *>(h);
            *s->hstale = 0;
            for (auto &e : s->wev) ND_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        }
        int32_t cap = p->max_neighbors;
        if (cap <= 0) {
            const double vol = p->hi > p->lo ? (double)(p->hi - p->lo) * (p->hi - p->lo) * (p->hi - p->lo) : 0.0;
            const double rho = vol > 0.0 ? (double)p->n / vol : 1.0;
            const double est = 4.18879 * rho * (double)s->rlist * s->rlist * s->rlist;
            cap = (int32_t)(est * 1.4 + 32.0);
            if (cap > p->n - 1) cap = p->n - 1;
            if (cap < 8) cap = 8;
        }
        alloc_lists(s, cap);
        set_step_params(s);
        ND_HIP(hipStreamSynchronize(ctx->stream));
    } catch (...) {
        nd_sim_destroy(s);
        throw;
    }
    *out = s;
    ND_API_END(ctx)
}

int nd_sim_destroy(nd_sim *s) {
    if (!s) return ND_OK;
    if (s->ctx) {
        (void)hipSetDevice(s->ctx->device);
        (void)hipStreamSynchronize(s->ctx->stream);
    }
    drop_graphs(s);
    for (auto e : s->ev_pool) (void)hipEventDestroy(e);
    for (auto e : s->wev)
        if (e) (void)hipEventDestroy(e);
    if (s->hstale) (void)hipHostFree(const_cast<int32_t *>(s->hstale));
    for_each_buf(s, [](nd::DevBuf &b) { b.release(); });
    delete s;
    return ND_OK;
}

int nd_sim_set_state(nd_sim *s, const double *h_x, const double *h_v, const double *h_q,
                     const double *h_m) {
    nd_ctx *ctx = s ? s->ctx : nullptr;
    ND_API_BEGIN
    ND_REQUIRE(s && h_x && h_v, "nd_sim_set_state: null argument");
    ND_HIP(hipSetDevice(ctx->device));
    const size_t n = (size_t)s->n;
    std::vector<int32_t> orig(n);
    for (size_t i = 0; i < n; ++i) orig[i] = (int32_t)i;
    s->cur = 0;
    s->vcur = 0;
    if (s->wide) {
        std::vector<double4> x(n), v(n);
        for (size_t i = 0; i < n; ++i) {
            const double m = h_m ? h_m[i] : 1.0;
            ND_REQUIRE(m > 0.0, "nd_sim_set_state: masses must be positive");
            x[i] = make_double4(h_x[3 * i], h_x[3 * i + 1], h_x[3 * i + 2], h_q ? h_q[i] : 0.0);
            v[i] = make_double4(h_v[3 * i], h_v[3 * i + 1], h_v[3 * i + 2], 1.0 / m);
        }
        copy_h2d(s->pos[0], x.data(), sizeof(double4) * n, ctx->stream);
        copy_h2d(s->vel[0], v.data(), sizeof(double4) * n, ctx->stream);
        ND_HIP(hipStreamSynchronize(ctx->stream));
    } else {
        std::vector<float4> x(n), v(n);
        for (size_t i = 0; i < n; ++i) {
            const double m = h_m ? h_m[i] : 1.0;
            ND_REQUIRE(m > 0.0, "nd_sim_set_state: masses must be positive");
            x[i] = make_float4((float)h_x[3 * i], (float)h_x[3 * i + 1], (float)h_x[3 * i + 2],
                               h_q ? (float)h_q[i] : 0.0f);
            v[i] = make_float4((float)h_v[3 * i], (float)h_v[3 * i + 1], (float)h_v[3 * i + 2],
                               (float)(1.0 / m));
        }
        copy_h2d(s->pos[0], x.data(), sizeof(float4) * n, ctx->stream);
        copy_h2d(s->vel[0], v.data(), sizeof(float4) * n, ctx->stream);
        ND_HIP(hipStreamSynchronize(ctx->stream));
    }
    copy_h2d(s->orig[0], orig.data(), sizeof(int32_t) * n, ctx->stream);
    ND_HIP(hipMemsetAsync(s->misc.p, 0, sizeof(SimMisc), ctx->stream));
    ND_HIP(hipStreamSynchronize(ctx->stream));
    s->step = 0;
    s->have_force = false;
    s->list_valid = false;
    s->tree_valid = false;
    s->have_state = true;
    s->last_build_step = -1;
    drop_graphs(s);
    ND_API_END(ctx)
}

int nd_sim_get_state(nd_sim *s, double *h_x, double *h_v, double *h_f) {
    nd_ctx *ctx = s ? s->ctx : nullptr;
    ND_API_BEGIN
    ND_REQUIRE(s, "nd_sim_get_state: null simulation");
    ND_REQUIRE(s->have_state, "nd_sim_get_state: no state set");
    ND_HIP(hipSetDevice(ctx->device));
    const size_t n = (size_t)s->n;
    std::vector<int32_t> orig(n);
    copy_d2h(orig.data(), s->orig[s->vcur], sizeof(int32_t) * n, ctx->stream);
    auto unpack = [&](const auto &x, const auto &v, const auto &f) {
        ND_HIP(hipStreamSynchronize(ctx->stream));
        for (size_t k = 0; k < n; ++k) {
            const size_t o = (size_t)orig[k];
            if (h_x) { h_x[3 * o] = x[k].x; h_x[3 * o + 1] = x[k].y; h_x[3 * o + 2] = x[k].z; }
            if (h_v) { h_v[3 * o] = v[k].x; h_v[3 * o + 1] = v[k].y; h_v[3 * o + 2] = v[k].z; }
            if (h_f) {
                if (s->have_force) { h_f[3 * o] = f[k].x; h_f[3 * o + 1] = f[k].y; h_f[3 * o + 2] = f[k].z; }
                else { h_f[3 * o] = h_f[3 * o + 1] = h_f[3 * o + 2] = 0.0; }
            }
        }
    };
    if (s->wide) {
        std::vector<double4> x(n), v(n), f(n);
        copy_d2h(x.data(), s->pos[s->cur], sizeof(double4) * n, ctx->stream);
        copy_d2h(v.data(), s->vel[s->vcur], sizeof(double4) * n, ctx->stream);
        copy_d2h(f.data(), s->frc, sizeof(double4) * n, ctx->stream);
        unpack(x, v, f);
    } else {
        std::vector<float4> x(n), v(n), f(n);
        copy_d2h(x.data(), s->pos[s->cur], sizeof(float4) * n, ctx->stream);
        copy_d2h(v.data(), s->vel[s->vcur], sizeof(float4) * n, ctx->stream);
        copy_d2h(f.data(), s->frc, sizeof(float4) * n, ctx->stream);
        unpack(x, v, f);
    }
    ND_API_END(ctx)
}

int nd_sim_run(nd_sim *s, int32_t nsteps) {
    nd_ctx *ctx = s ? s->ctx : nullptr;
    ND_API_BEGIN
    ND_REQUIRE(s && nsteps >= 0, "nd_sim_run: bad argument");
    ND_REQUIRE(s->have_state, "nd_sim_run: call nd_sim_set_state first");
    ND_HIP(hipSetDevice(ctx->device));
    s->ev_used = 0;
    int64_t checked_until = -1;  // steps before this are run checked
    for (int attempt = 0; attempt < 8; ++attempt) {
        // snapshot for a rerun if the neighbour list overflows or goes stale
        SimMisc m0;
        read_misc(s, m0);
        const int cur0 = s->cur, vcur0 = s->vcur, since0 = s->since;
        const int64_t step0 = s->step, rb0 = s->rebuilds, lb0 = s->last_build_step,
                      fr0 = s->forced_rebuilds;
        const bool hf0 = s->have_force;
        snapshot(s);
        *s->hstale = 0;
        const bool done = run_once(s, nsteps, checked_until);
        const int64_t stop = s->step;
        SimMisc m;
        read_misc(s, m);
        if (m.status & nd::kStatusDomain)
            throw nd::Error(ND_ERR_DOMAIN, "simulation: coordinate outside the key domain");
        if (m.status & nd::kStatusTree)
            throw nd::Error(ND_ERR_STATE, "simulation: malformed neighbour tree");
        s->last_maxcnt = m.maxcnt;
        if (done && !m.overflow && !m.stale) {
            if (s->timing && s->ev_used >= 2) {
                // every force launch recorded an event pair
                for (size_t e = 0; e + 1 < s->ev_used; e += 2) {
                    float ms = 0.0f;
                    ND_HIP(hipEventElapsedTime(&ms, s->ev_pool[e], s->ev_pool[e + 1]));
                    s->force_ms += ms;
                    s->force_launches++;
                }
            }
            return ND_OK;
        }
        // restore the snapshot and run again: with a larger list after an
        // overflow, in checked mode after a stale list (an atom moved half
        // the skin before a scheduled rebuild)
        s->cur = cur0;
        s->vcur = vcur0;
        s->since = since0;
        restore(s);
        if (m.overflow) {
            alloc_lists(s, m.maxcnt + m.maxcnt / 4 + 8);
            set_step_params(s);
            s->list_valid = false;
            s->tree_valid = false;
        } else {
            // later rebuilds of the failed attempt re-sorted the atoms: the
            // rerun builds a list for the restored state first, and checks
            // every window (a system fast enough to outrun the skin once
            // usually does again soon, and an unchecked part would send the
            // run back to its start each time)
            s->list_valid = false;
            s->tree_valid = false;
            checked_until = step0 + nsteps + 1;
            (void)stop;
            s->stale_reruns++;
        }
        m0.overflow = 0;
        m0.maxcnt = 0;
        m0.stale = 0;
        copy_h2d(s->misc, &m0, sizeof(SimMisc), ctx->stream);
        ND_HIP(hipStreamSynchronize(ctx->stream));
        s->step = step0;
        s->rebuilds = rb0;
        s->forced_rebuilds = fr0;
        s->last_build_step = lb0;
        s->have_force = hf0;
        s->ev_used = 0;
    }
    throw nd::Error(ND_ERR_CAPACITY, "nd_sim_run: neighbour list kept overflowing");
    ND_API_END(ctx)
}

int nd_sim_prepare(nd_sim *s) {
    nd_ctx *ctx = s ? s->ctx : nullptr;
    ND_API_BEGIN
    ND_REQUIRE(s, "nd_sim_prepare: null simulation");
    ND_REQUIRE(s->have_state, "nd_sim_prepare: call nd_sim_set_state first");
    ND_HIP(hipSetDevice(ctx->device));
    // size the list (and compute the first force) as a run of zero steps does
    if (!s->list_valid || !s->have_force) {
        const int rc = nd_sim_run(s, 0);
        if (rc != ND_OK) return rc;
    }
    if (!graph_eligible(s)) return ND_OK;
    const int32_t W = s->p.rebuild_every;
    const bool refits = resort_period(s) > W;
    for (int kind = 0; kind < (refits ? 2 : 1); ++kind)
        for (int cur = 0; cur < 2; ++cur)
            for (int vcur = 0; vcur < 2; ++vcur) window_graph(s, W, cur, vcur, kind == 0);
    ND_HIP(hipStreamSynchronize(ctx->stream));
    ND_API_END(ctx)
}

int nd_sim_energies(nd_sim *s, double *h_rows, int32_t max_rows, int32_t *rows) {
    nd_ctx *ctx = s ? s->ctx : nullptr;
    ND_API_BEGIN
    ND_REQUIRE(s && rows, "nd_sim_energies: null argument");
    SimMisc m;
    read_misc(s, m);
    int32_t r = m.rowctr < s->max_rows ? m.rowctr : s->max_rows;
    if (r > max_rows) r = max_rows;
    if (r > 0 && h_rows)
        copy_d2h(h_rows, s->rows, sizeof(double) * 5 * (size_t)r, ctx->stream);
    m.rowctr = 0;
    copy_h2d(s->misc, &m, sizeof(SimMisc), ctx->stream);
    ND_HIP(hipStreamSynchronize(ctx->stream));
    *rows = r;
    if (r > 0 && h_rows) s->pairs_last = (int64_t)h_rows[5 * (size_t)(r - 1) + 4];
    ND_API_END(ctx)
}

int nd_sim_stats_get(nd_sim *s, nd_sim_stats *out) {
    nd_ctx *ctx = s ? s->ctx : nullptr;
    ND_API_BEGIN
    ND_REQUIRE(s && out, "nd_sim_stats_get: null argument");
    std::vector<int32_t> cnt((size_t)s->n);
    int64_t entries = 0;
    if (s->list_valid) {
        copy_d2h(cnt.data(), s->ncount, sizeof(int32_t) * (size_t)s->n, ctx->stream);
        ND_HIP(hipStreamSynchronize(ctx->stream));
        for (int32_t c : cnt) entries += c;
    }
    out->steps = s->step;
    out->rebuilds = s->rebuilds;
    out->list_entries = entries;
    out->max_neighbors = s->last_maxcnt;
    out->capacity = s->cap;
    out->pairs_in_cutoff = s->pairs_last;
    out->stale_reruns = s->stale_reruns;
    out->forced_rebuilds = s->forced_rebuilds;
    out->graph_recaptures = s->graph_recaptures;
    SimMisc m;
    read_misc(s, m);
    out->table_spills = m.rec_spill;
    out->table_records_big = s->rec_big ? 1 : 0;
    ND_API_END(ctx)
}

int nd_sim_time_fused_window(nd_sim *s, int32_t nsteps, double *avg_fused_ms) {
    nd_ctx *ctx = s ? s->ctx : nullptr;
    ND_API_BEGIN
    ND_REQUIRE(s && avg_fused_ms, "nd_sim_time_fused_window: null argument");
    ND_REQUIRE(s->have_state, "nd_sim_time_fused_window: call nd_sim_set_state first");
    ND_REQUIRE(nsteps >= 2 && nsteps <= s->p.rebuild_every,
               "nd_sim_time_fused_window: need 2 <= nsteps <= rebuild_every");
    ND_HIP(hipSetDevice(ctx->device));
    const bool timing = s->timing;
    s->timing = false;
    rebuild_checked(s);  // fresh list for this window (checks capacity)
    s->last_build_step = s->step;
    launch_force_step(s, kModeForce, s->step);
    launch_kick_drift(s, false, s->step);
    ND_HIP(hipEventRecord(ctx->ev0, ctx->stream));
    for (int32_t o = 1; o < nsteps; ++o) launch_step_prec<kModeFused, false>(s, true), s->cur ^= 1;
    ND_HIP(hipGetLastError());
    ND_HIP(hipEventRecord(ctx->ev1, ctx->stream));
    launch_force_step(s, kModeFinal, s->step + nsteps);
    ND_HIP(hipEventSynchronize(ctx->ev1));
    float ms = 0.0f;
    ND_HIP(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    *avg_fused_ms = ms / (double)(nsteps - 1);
    s->step += nsteps;
    s->have_force = true;
    s->timing = timing;
    SimMisc m;
    read_misc(s, m);  // kernels are stream-ordered; this also syncs the final step
    // keep the device step counter in line with the host's
    m.step = s->step;
    copy_h2d(s->misc, &m, sizeof(SimMisc), ctx->stream);
    ND_HIP(hipStreamSynchronize(ctx->stream));
    ND_API_END(ctx)
}

int nd_sim_set_replicas(nd_sim *s, int32_t nrep, const int32_t *h_rep, const double *h_boxes) {
    nd_ctx *ctx = s ? s->ctx : nullptr;
    ND_API_BEGIN
    ND_REQUIRE(s && nrep >= 0, "nd_sim_set_replicas: bad argument");
    ND_HIP(hipSetDevice(ctx->device));
    if (nrep == 0) {
        s->nrep = 0;
        drop_graphs(s);
        return ND_OK;
    }
    ND_REQUIRE(h_rep && h_boxes, "nd_sim_set_replicas: null argument");
    for (int32_t i = 0; i < s->n; ++i)
        ND_REQUIRE(h_rep[i] >= -1 && h_rep[i] < nrep, "nd_sim_set_replicas: replica index out of range");
    for (int32_t r = 0; r < nrep; ++r)
        for (int d = 0; d < 3; ++d)
            ND_REQUIRE(h_boxes[6 * r + 3 + d] > h_boxes[6 * r + d], "nd_sim_set_replicas: empty replica box");
    s->rep.reserve(sizeof(int32_t) * (size_t)s->n);
    s->boxes.reserve(sizeof(double) * 6 * (size_t)nrep);
    copy_h2d(s->rep, h_rep, sizeof(int32_t) * (size_t)s->n, ctx->stream);
    copy_h2d(s->boxes, h_boxes, sizeof(double) * 6 * (size_t)nrep, ctx->stream);
    ND_HIP(hipStreamSynchronize(ctx->stream));
    s->nrep = nrep;
    drop_graphs(s);  // captured launches carry the wall pointers
    ND_API_END(ctx)
}

int nd_sim_record_positions(nd_sim *s, void *d_out, int32_t as_double) {
    nd_ctx *ctx = s ? s->ctx : nullptr;
    ND_API_BEGIN
    ND_REQUIRE(s && d_out, "nd_sim_record_positions: null argument");
    ND_REQUIRE(s->have_state, "nd_sim_record_positions: call nd_sim_set_state first");
    ND_HIP(hipSetDevice(ctx->device));
    const unsigned g = nd::grid_for(s->n);
    hipStream_t st = ctx->stream;
    const int32_t *orig = s->orig[s->vcur].as<int32_t>();
    if (s->wide) {
        if (as_double) k_record<double, double><<<g, nd::kBlock, 0, st>>>(s->pos[s->cur].as<double4>(), orig, s->n, (double *)d_out);
        else k_record<double, float><<<g, nd::kBlock, 0, st>>>(s->pos[s->cur].as<double4>(), orig, s->n, (float *)d_out);
    } else {
        if (as_double) k_record<float, double><<<g, nd::kBlock, 0, st>>>(s->pos[s->cur].as<float4>(), orig, s->n, (double *)d_out);
        else k_record<float, float><<<g, nd::kBlock, 0, st>>>(s->pos[s->cur].as<float4>(), orig, s->n, (float *)d_out);
    }
    ND_HIP(hipGetLastError());
    ND_API_END(ctx)
}

int nd_sim_kernel_timing(nd_sim *s, int32_t enable, double *avg_force_ms, int64_t *launches) {
    nd_ctx *ctx = s ? s->ctx : nullptr;
    ND_API_BEGIN
    ND_REQUIRE(s, "nd_sim_kernel_timing: null simulation");
    if (avg_force_ms) *avg_force_ms = s->force_launches ? s->force_ms / (double)s->force_launches : 0.0;
    if (launches) *launches = s->force_launches;
    if (enable >= 0) {
        s->timing = enable != 0;
        s->force_ms = 0.0;
        s->force_launches = 0;
    }
    ND_API_END(ctx)
}

}  // extern "C"
