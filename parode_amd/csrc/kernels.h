// Device kernels of the LTI scan and their host launchers (gfx950).
//
// Layouts (all row-major, device memory):
//   trajectory X   : [Npad][na]  element-major; column e is the state x_e
//   power table    : P_l = M^(2^l) in FP64, [L][na][na]
//   level tables   : transposed copies PT_l[j][i] = P_l[i][j] (or P_l - I in
//                    delta form), FP32 or FP64, [L][na][na]
//   inputs U       : [rows][m] samples; element e reads rows e*sL + oL + k*rs
//   outputs Y      : [N][nOut] time-major (the reference's xVect layout)
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <map>
#include <utility>
#include <vector>

namespace parode {
namespace k {

// ---------------------------------------------------------------- step maps
// Scalar factors of the RK4 map (RungeKutta4thOrder.cpp:48-57) and of the ABM
// companion blocks (AdamsMoulton.cpp:258-269), evaluated on the host left to
// right; the chunk recurrences' factors of an FP64 ABM map (blocks oldest
// first: x_{k-3} (A^2 only), x_{k-2}, x_{k-1}, x_k).
struct Rk4Coef {
    double h, c2, c3, c4, s1, s2, s3, t0, t1, t2, u0;
};
struct AbmCoef {
    double a0A, a0A2, a1A, a1A2, a2A, a2A2, a3A2, bm1, b0B, b0AB, b1B, b1AB, b2B, b2AB, b3AB;
};
struct AbmFactors {
    float a[4], b[4];
};
Rk4Coef rk4_coefficients(double h);
AbmCoef abm_coefficients(double h);
AbmFactors abm_chunk_factors(double h);

// Z = X * Y.  mirror = true: float products, sequential-k float accumulation,
// no fused multiply-add (the rounding of uBLAS prod() on matrix<float>).
void gemm_seq_f32(hipStream_t s, int r, int kk, int c, const float* X, const float* Y, float* Z);
// FP64 tiled GEMM, Z = X * Y (used for A-products and the power squarings).
void gemm_f64(hipStream_t s, int r, int kk, int c, const double* X, const double* Y, double* Z,
              bool mfma = true);
template <typename T>
void transpose(hipStream_t s, int rows, int cols, const T* in, T* out);
void cast_f32_to_f64(hipStream_t s, long count, const float* in, double* out);
void fill_f64(hipStream_t s, long count, double v, double* out);
void fill_zero_f32(hipStream_t s, long count, float* out);

// RK4 combination (RungeKutta4thOrder.cpp:48-57).  mirror: inputs are the float
// products, combination in double without contraction, one rounding to float;
// M64/Bbar64 receive the float results promoted.  !mirror: FP64 products in,
// FP64 results out.  Bbar32 may be null.
void rk4_combine(hipStream_t s, bool mirror, int n, int m, double h, const void* A, const void* A2,
                 const void* A3, const void* A4, const void* B, const void* AB, const void* A2B,
                 const void* A3B, double* M64, double* Bbar64, float* Bbar32);
// ABM companion (AdamsMoulton.cpp:254-316); M64 is 4n x 4n.
void abm_combine(hipStream_t s, bool mirror, int n, int m, double h, const void* A, const void* A2,
                 const void* B, const void* AB, double* M64, double* Bbar64, float* Bbar32);

// Level tables from the FP64 powers: PT[l] = transpose(P_l - delta*I) as T.
void make_level_table_f32(hipStream_t s, int na, const double* P, bool delta, float* PT);
void make_level_table_f64(hipStream_t s, int na, const double* P, bool delta, double* PT);

// All FP32 level tables [0, nlev) in one launch (PT32 and/or the padded PTm).
void make_level_tables(hipStream_t s, int na, int nap, int nrp, int nlev, int deltaBelow,
                       const double* PT64, float* PT32, float* PTm);
// PT[l + 1] = PT[l]^2 for l < L in one work-group (n_a <= 80), FP64 MFMA.
bool powers_small_supported(int na);
void powers_small(hipStream_t s, int na, int L, double* PT);

// ------------------------------------------------------------------- leaves
// X[e][off + r] = sum_{q<k} Bbar[r][q*m + j] * U[(e*sL + oL + q*rs)*m + j], e < nReal;
// rows off..off+n of X for e in [nReal, nPad) are zero; other rows of X are
// zeroed when zeroOther (the ABM zero fill).  T: X type, TU: U type.
// writeZeros = false writes, of the rows outside [off, off + n), only rows
// [32 floor(off / 32), off) of real elements (all of them for padding
// elements): what a first level that starts at state 32 floor(off / 32) reads.
template <typename T, typename TU, typename TB>
void leaves(hipStream_t s, int n, int m, int k, int na, int off, long nReal, long nPad,
            const TB* Bbar, const TU* U, long sL, long oL, long rs, T* X, bool writeZeros = true);

// Leaves whose Bbar factors as Bf (G (x) I_m): Bbar_q = sum_g G[q][g] Bf_g for
// a kq x kg matrix G (kq <= 5, kg <= 2).  ABM (FP64 map): Bf = [B | AB] and
// G from abm_leaf_factors, so each leaf is 2m products per row instead of 5m,
// in FP64 on the matrix cores.  BfT is Bf transposed, [kg m][n].  Returns
// false (nothing launched) when the shape does not fit the kernel.
struct LeafComb {
    double g[5][2];
};
void abm_leaf_factors(double h, LeafComb& G);
template <typename T, typename TU>
bool leaves_factored(hipStream_t s, int n, int m, int kq, int kg, int na, int off, long nReal,
                     long nPad, const double* BfT, const LeafComb& G, const TU* U, long sL,
                     long oL, long rs, T* X, bool writeZeros = true);

// ------------------------------------------------------------ scan levels
// Up-sweep level with stride d over np pairs: X[R] += P X[L] (R = (k+1)2d-1,
// L = R-d); with delta tables X[R] += X[L] + (P-I) X[L].
// Down-sweep: (X[L], X[R]) <- (X[R], P X[R] + X[L]).
template <typename TX, typename TP>
void level(hipStream_t s, bool down, bool delta, int na, long d, long np, const TP* PT, TX* X);

// The same FP32 level update on the matrix cores (17 <= na <= 1024), reading a
// zero-padded table [mfma_table_rows(na)][32*ceil(na/32)].
bool level_mfma_supported(int na);
int mfma_table_rows(int na);
void make_level_table_padded(hipStream_t s, int na, int nap, int nrp, const double* P, bool delta,
                             float* PTp);
// Structure of the lowest ABM levels (all zero in general): rbeg, the first
// row of the X[L] source / X[R] accumulator (up) or X[L] addend (down) that
// is not zero (rows below are not read); jbeg, the first source state the
// product needs (a multiple of 32, <= rbeg; 0 for the down-sweep); ibase, the
// first row the product forms (a multiple of 32, <= rbeg); rows [clo, ibase)
// are shifts of the source by sh rows, (P x)[i] = x[i + sh], and are copied
// (the down-sweep also copies X[L] <- X[R] there); rows below clo are not
// written; [zlo, zhi) (up only, <= ibase) rows of X[R] written as zero.
struct LevelRows {
    int jbeg = 0, rbeg = 0, ibase = 0, clo = 0, sh = 0, zlo = 0, zhi = 0;
};
// For more than 256 rows (na > 256: level_mfma_splits) the rows are split over
// work-groups, and the down-sweep first copies the pairs' X[R] into S
// (np * na floats).
// tile: 0 by the pair count, 1 the narrow tiles (32 pairs per work-group), 2
// the wide ones.
void level_mfma(hipStream_t s, bool down, bool delta, int na, long d, long np, const float* PT,
                float* X, const LevelRows& lr = LevelRows{}, float* S = nullptr, int tile = 0);
inline bool level_mfma_splits(int na) { return na > 256; }

// The lowest lc levels of the ABM scan (n <= 64) as a sequential recurrence
// per chunk of 2^lc elements on the matrix cores, in delta form with a
// compensated newest block; P64 is the FP64 transposed map (P64[k][i] =
// M[i][k], n_a x n_a).  Pass 1 (pass2 = false, in place of the
// up-sweep's levels 0..lc-1) writes each chunk's aggregate to its last
// element; pass 2 (in place of the down-sweep's levels lc-1..0) starts from
// the state the down-sweep leaves there and writes the last block of every
// real element's state, X[e][3n:4n), for the outputs.  Reads the leaves'
// rows [3n, 4n).
// The same ABM chunk recurrences for an FP64 map, with the companion block
// row factored as x3 + A w1 + A^2 w2 (w1, w2: combinations of the state's four
// blocks with the ABM coefficients of step h): 2n products per row per step.
// A64 and A264 are A and A^2 in FP64, row-major n x n.
bool abm_chunks_factored_supported(int n);   // n <= 256
void abm_chunks_factored(hipStream_t s, bool pass2, int n, int lc, long nReal, long nPad,
                         double h, const double* A64, const double* A264, float* X);

// Chunk recurrences of a dense map P (FP64 transposed, P64[k][i] = P[i][k],
// na x na, na <= 256) over elements at X rows e S + S - 1, e < nElem (nReal
// of them real), chunks of 2^lc elements: pass 1 writes each chunk's aggregate
// to its last element, pass 2 writes every element's state from the value
// there.  RK4 (S = 1, P = M) and the ABM chunk aggregates (S = 2^l, P = M^S).
bool chunks_dense_supported(int na);
void chunks_dense(hipStream_t s, bool pass2, int na, int lc, long S, long nReal, long nElem,
                  const double* P64, float* X);
bool abm_chunks_supported(int n);
void abm_chunks(hipStream_t s, bool pass2, int n, int lc, long nReal, long nPad,
                const double* P64, float* X);

// A level with few pairs as many small K-split tiles (32 x 32 FP32 MFMA when
// TT = float, with PT the padded MFMA table and ldp its row length; 16 x 16
// FP64 MFMA when TT = double, with PT the transposed FP64 power, ldp = na).
// The down-sweep first copies the pairs' X[R] into S (np * na values).
constexpr long kKsMaxPairs = 4096;
// delta: the table holds P - I (FP32 tables of mixed precision's delta levels).
// Down-sweep: gathered = S already holds the pairs' X[R] (written by the level
// above), else it is gathered first; Sn (not S), when given, receives the same
// copy for the next lower level's npn pairs.
template <typename TT, typename TX>
void level_ks(hipStream_t s, bool down, int na, int ldp, long d, long np, const TT* PT, TX* X,
              TX* S, bool delta = false, bool gathered = false, TX* Sn = nullptr, long npn = 0);

// Levels [l0, L) of both sweeps and the root in one work-group (levels >= l64
// from the FP64 table, the others from the FP32 one).
template <typename TX>
void top(hipStream_t s, int na, int l0, int L, int l64, long nReal, const float* PT32,
         const double* PT64, TX* X, const double* seed, double* tot);

// Root of the scan: tot = X[last]; X[last] = seed (seed as TS).
template <typename TX, typename TS>
void root(hipStream_t s, int na, long last, const TS* seed, TX* X, double* tot);
// Carried accumulator for the next batch: seed' = P * seed + tot (FP64).
void carry(hipStream_t s, int na, const double* P, const double* seed, const double* tot,
           double* seedOut);

// ------------------------------------------------------------------ outputs
// Y[e][j] = sum_c C[j][c] X[e][oX*n + c] + sum_c D[j][c] U[(e*sO + oO)*m + c]
// for e < N; with both absent Y[e] = X[e][oX*n ..].  Y stride nOut.
template <typename TX, typename TU>
void outputs(hipStream_t s, int n, int m, int p, int na, int oX, bool zC, bool zD, long N,
             const float* C, const float* D, const TX* X, const TU* U, long sO, long oO,
             float* Y);

// The same outputs in FP64 with one thread per element (X or U in FP64), C^T
// and D^T given in FP64 with rows padded with zeros to pp columns (a multiple
// of 8; pad_cast makes them).  Returns false (nothing launched) when the shape
// does not fit the kernel (states as outputs, or an element's state block too
// large to stage).
void pad_cast(hipStream_t s, int rows, int cols, int pp, const float* in, double* out);
template <typename TX, typename TU>
bool outputs_lanes(hipStream_t s, int n, int m, int p, int pp, int na, int oX, bool zC, bool zD,
                   long N, const double* CT, const double* DT, const TX* X, const TU* U, long sO,
                   long oO, float* Y);

// ---------------------------------------------------------- ABM start-up
// X0 = [x0; x1; x2; x3] by three classic RK4 steps on Uh (9 rows, h/2 grid).
// mirror: float state, double stage combinations (odeint runge_kutta4
// <vector<float>>), float A/B; else FP64 throughout.  Single work-group.
template <typename TU>
void abm_startup(hipStream_t s, bool mirror, int n, int m, double h, const float* A,
                 const float* B, const float* x0, const TU* Uh, double* X0);

// ------------------------------------------------------------ input samples
// out[s][j] = a sin(g) + b cos(g) + c + d t (+ sigma * N(0,1) keyed by
// (seed, s0 + s, j)), t = t0 + (s0 + s) dt, g = w t + phi, coefficient row
// j = (a, b, c, d, w, phi); FP64, S rows of m.  Device-side inputs of served
// requests whose u is a sinusoid family.
void sample_family(hipStream_t s, int m, const double* coef, double t0, double dt, long s0,
                   long S, double sigma, uint64_t seed, double* out);

// Copies a typed vector into FP64.
template <typename T>
void to_f64(hipStream_t s, long count, const T* in, double* out);


// ---------------------------------------------------------- grouped launches
// One launch runs the same kernel for many systems (the served path): each
// system owns a run of the grid's work-groups and an argument record, and
// its work-groups run exactly the single-system body on the system's own
// buffers.  A phase collects the launches of one stage of the scan for all
// systems, keyed by kernel (and pass or sweep direction); the planner packs
// every launch's records and block table into one device buffer and runs the
// launches in key order.
namespace g {
enum Kind {
    kPrep, kGemm, kGemmKs2, kRk4Combine, kAbmCombine, kTranspose, kPowersSmall, kTables,
    kSample, kStartup, kStartOut, kLeavesPlain, kLeavesFac, kLeavesSimt, kOutputsGemm1,
    kOutputsGemm2, kOutputsGemm4,
    kChunksFac16, kChunksFac32, kChunksFac48, kChunksFac64, kChunksFac96, kChunksFac128,
    kChunksTwo160, kChunksTwo192, kChunksTwo224, kChunksTwo256,
    kDense16, kDense32, kDense48, kDense64, kDense96, kDense128, kDense160, kDense192, kDense224,
    kDense256,
    kLevel32, kLevel32W, kLevel32Delta, kLevel32DeltaW, kLevelG, kLevelGDelta, kLevel64, kGather, kTopMfma, kTopSimt, kRoot,
    kOutputsLanes, kOutputsKc8, kOutputsKc16, kOutputsKc24, kOutputsKc32, kOutputsMfma
};
struct Launch {
    int kind = -1;
    bool flag = false;                  // pass 2 of the chunks, or the down-sweep
    size_t argSize = 0;                 // bytes of one argument record
    std::vector<unsigned char> args;    // nsys records
    std::vector<unsigned> start{0};     // nsys + 1: first block of each system
    std::vector<float> cost;            // nsys: the run time of one of a system's blocks
                                        // relative to the others' (0: unknown)
    size_t lds = 0;                     // dynamic LDS bytes (the largest system's)
    int nsys() const { return (int)start.size() - 1; }
    unsigned blocks() const { return start.back(); }
};
// Writes a launch's argument records (args: nsys x argSize bytes) and block
// table (start: nsys + 1) with its systems in decreasing cost (stable), so that
// the longest-running work-groups are dispatched first and the short ones fill
// the compute units around them instead of forming the launch's tail.
void pack_by_cost(const Launch& L, unsigned char* args, unsigned* start);
// Appends from's systems after into's (the same kind, flag and LDS class).
void merge_launch(Launch& into, Launch&& from);
// key: (kind, flag + 2 x LDS size class)
using Phase = std::map<std::pair<int, int>, Launch>;

// Each add_* appends one system's launch of a kernel to a phase (its kind and
// launch shape chosen as the single-system launcher would choose them).
void add_cast(Phase& ph, long count, const float* in, double* out);
void add_fill(Phase& ph, long count, double v, double* out);
// out[r][c] = in[c][r] for c < cols (in: cols x rows), zero up to pp columns
void add_transpose_pad(Phase& ph, int rows, int cols, int pp, const float* in, double* out);
void add_gemm(Phase& ph, int r, int kk, int c, const double* X, const double* Y, double* Z);
// Z = X Y with leading dimensions ldx, ldy, ldz, plus, when add is given,
// add[(i - shift) ldz + j] on the rows i >= shift
void add_gemm_ld(Phase& ph, int r, int kk, int c, const double* X, int ldx, const double* Y,
                 int ldy, double* Z, int ldz, const double* add, int shift);
void add_rk4_combine(Phase& ph, int n, int m, double h, const double* A, const double* A2,
                     const double* A3, const double* A4, const double* B, const double* AB,
                     const double* A2B, const double* A3B, double* MT, double* Bbar);
void add_abm_combine(Phase& ph, int n, int m, double h, const double* A, const double* A2,
                     const double* B, const double* AB, double* MT, double* Bbar);
void add_transpose(Phase& ph, int rows, int cols, const double* in, double* out);
bool add_powers_small(Phase& ph, int na, int L, double* PT);
// (levels [first, nlev) only)
void add_tables(Phase& ph, int na, int nap, int nrp, int nlev, int deltaBelow, const double* PT64,
                float* PT32, float* PTm, int first = 0);
void add_sample(Phase& ph, int m, const double* coef, double t0, double dt, long S, double sigma,
                uint64_t seed, double* out);
void add_startup(Phase& ph, int n, int m, double h, const float* A, const float* B,
                 const float* x0, const double* Uh, double* X0);
void add_start_outputs(Phase& ph, int n, int m, int p, int pp, bool useC, bool useD,
                       const double* CT, const double* DT, const double* X0, const double* U,
                       float* Y);
// whether the leaves of this shape fit the matrix-core leaf kernel's LDS
bool leaves_fit(int n, int m, int kq, int kg, long sL, long rs);
bool add_leaves(Phase& ph, int n, int m, int kq, int kg, int na, int off, long nReal, long nPad,
                const double* Bf, const LeafComb& G, const double* U, long sL, long oL, long rs,
                float* X, bool writeZeros);
bool add_leaves_simt(Phase& ph, int n, int m, int kq, int na, int off, long nReal, long nPad,
                     const double* BbT, const double* U, long sL, long oL, long rs, float* X,
                     bool writeZeros);
// The chunk kernel's kind a system's chunk passes launch (ABM: by n; dense: by
// n_a), or -1 when none covers it.
int chunks_kind(bool abm, int n);
bool add_abm_chunks(Phase& ph, bool pass2, int n, int lc, long nReal, long nPad, double h,
                    const double* A64, const double* A264, float* X);
bool add_dense_chunks(Phase& ph, bool pass2, int na, int lc, long nReal, long nPad,
                      const double* P64, float* X);
// f64: the transposed FP64 power (ldp = na); else the padded FP32 table
void add_level(Phase& ph, bool down, bool f64, bool delta, int na, int ldp, long d, long np,
               const void* PT, float* S, float* X, float* Sn, long npn);
void add_gather(Phase& ph, int na, long d, long np, const float* X, float* S);
void add_top(Phase& ph, int na, int l0, int L, int l64, long nReal, const float* PT32,
             const double* PT64, float* X, const double* seed, double* tot);
void add_root(Phase& ph, int na, long last, const double* seed, float* X, double* tot);
bool add_outputs(Phase& ph, int n, int m, int p, int pp, int na, int oX, bool useC, bool useD,
                 long N, const double* CT, const double* DT, const float* X, const double* U,
                 long sO, long oO, float* Y);
// Launches L with its records at dArgs and its block table at dStart (device).
void run(hipStream_t s, const Launch& L, const void* dArgs, const unsigned* dStart);
}  // namespace g

}  // namespace k
}  // namespace parode
