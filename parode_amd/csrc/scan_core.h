// ScanCore: the device side of one LTI system's parallel-in-time solve.
//
// It owns, for one system shape, the step map (M and Bbar), the FP64 power
// table P_l = M^(2^l) (kept transposed), the per-level tables in the working
// precision and the trajectory X, and it runs the stages of the reference's
// scan driver (Sources/GPUFixedStepIterative.cpp:131-466) as asynchronous
// launches on a caller-supplied stream:
//   build_map    : DistributeA/M construction (RungeKutta4thOrder.cpp:36-70,
//                  AdamsMoulton.cpp:238-316) on the device
//   build_powers : the MM up-sweep products, collapsed to one squaring per level
//   leaves       : BTimesURK4thOrder / BTimesUAdamsMoulton
//   sweep_up     : MVV up-sweep (+ left-leaf copy, implicit in place)
//   sweep_top    : the top of the tree and the root carry with the (I, x0)
//                  seed (:294-337) in one work-group
//   sweep_down   : MVV down-sweep (+ root copy, in place); the MM down-sweep
//                  is dead work and is not performed
//   outputs      : SystemOutput
#pragma once
#include "hip_util.h"

#include <vector>

namespace parode {

class BatchSolve;

enum Method { kRK4 = 0, kABM = 1 };
constexpr int kKsMinStates = 16;   // smallest scan size for the K-split levels
enum Precision { kPlain32 = 0, kMixed = 1, kFp64 = 2 };

struct ScanOptions {
    int precision = kPlain32;
    int top64 = 6;        // mixed: the top levels whose products run in FP64
    int deltaLevels = 0;  // mixed: lowest levels applied as X + (P - I) X
};

class ScanCore {
public:
    // Buffers come from hipMalloc, or from `arena` when given (serving).
    ScanCore(int method, int n, int m, int p, bool zC, bool zD, long nPadMax, ScanOptions opt,
             Arena* arena = nullptr);

    int method() const { return method_; }
    int n() const { return n_; }
    int m() const { return m_; }
    int p() const { return p_; }
    int na() const { return na_; }
    int nOut() const { return (zC_ && zD_) ? n_ : p_; }
    int levels() const { return L_; }
    long nPadMax() const { return nPadMax_; }
    const ScanOptions& options() const { return opt_; }
    bool xIsDouble() const { return opt_.precision == kFp64; }
    size_t footprint() const;
    // p: outputs (C and D, transposed in FP32 and FP64); 0 when not known
    static size_t estimate_footprint(int method, int n, int m, long nPadMax, ScanOptions opt,
                                     int p = 0);

    // With side, the launches that the leaves and the chunk recurrences do not
    // need are forked onto side->stream (fork event recorded on s).
    // join: the level tables are ready (waited before the first tree level);
    // late (optional): the start-up seed and the output tables are ready
    // (waited before the root), else join covers them too
    struct SideStream {
        hipStream_t stream = nullptr;
        hipEvent_t fork = nullptr, join = nullptr, late = nullptr;
    };
    void build_map(hipStream_t s, const float* dA, const float* dB, double h,
                   const SideStream* side = nullptr);
    // withLast: also P_L = M^(2^L), which only carry() between batches reads
    void build_powers(hipStream_t s, bool withLast = true);
    void leaves(hipStream_t s, const void* dU, bool uDouble, long sL, long oL, long rs, long nReal,
                long nPad);
    // The scan of one batch runs as: up-sweep launches for the levels below
    // top_level(), one work-group for the top of the tree (its up levels, the
    // root with the seed, its down levels), down-sweep launches below.
    int top_level(long nReal, long nPad) const;
    // waitTables: an event s waits on before the first launch that reads the
    // level tables (after the first chunk pass, if any)
    void sweep_up(hipStream_t s, long nReal, long nPad, hipEvent_t waitTables = nullptr);
    void sweep_top(hipStream_t s, long nReal, long nPad, const double* dSeed, double* dTot);
    void sweep_down(hipStream_t s, long nReal, long nPad);
    // seedOut = M^nPad * seed + tot (the accumulator carried to the next batch)
    void carry(hipStream_t s, long nPad, const double* dSeed, const double* dTot, double* dOut);
    // C^T, D^T and their padded FP64 copies for the output kernels
    void prepare_outputs(hipStream_t s, const float* dC, const float* dD, bool uDouble);
    // prepared: prepare_outputs ran before on this stream (or one joined to it)
    void outputs(hipStream_t s, const float* dC, const float* dD, const void* dU, bool uDouble,
                 long sO, long oO, long N, float* dY, bool prepared = false);
    // outputs of the ABM start-up states x_0..x_2 (dX0 holds x_0..x_3)
    void startup_outputs(hipStream_t s, const float* dC, const float* dD, const double* dX0,
                         const void* dU, bool uDouble, float* dY, bool prepared = false);
    void abm_startup(hipStream_t s, const float* dA, const float* dB, const float* dx0,
                     const void* dUh, bool uDouble, double h, double* dX0);

    // The whole solve of one system with every array already on the device:
    // step map, powers, (ABM start-up,) leaves, sweeps, top, outputs, for
    // `points` outputs (scan length points, or points - 3 for ABM, at most
    // nPadMax).  Samples on the h/2 grid, (2*points+1) rows of m.  marks, when
    // given, are 7 events recorded before each of the six phases and after the
    // last.  Asynchronous on s.
    // With side, the power table and the ABM start-up run on side.stream
    // (forked from s after the map, joined before the first tree level), in
    // parallel with the leaves and the first chunk pass.
    void solve(hipStream_t s, long points, double h, const float* dA, const float* dB,
               const float* dC, const float* dD, const float* dx0, const void* dU, bool uDouble,
               float* dY, hipEvent_t* marks = nullptr, const SideStream* side = nullptr);

    double* seed() const { return seed_.get(); }
    double* tot() const { return tot_.get(); }
    double* seedNext() const { return seedNext_.get(); }
    const double* powerT(int l) const { return PT64_.get() + (size_t)l * na_ * na_; }
    void* X() const { return xIsDouble() ? (void*)X64_.get() : (void*)X32_.get(); }

    bool structured_level0(long nReal, long nPad) const;
    bool structured_level1(long nReal, long nPad) const;
    // log2 of the chunk length when the lowest levels run as chunk
    // recurrences (0: they run as tree levels)
    // grouped: the rule for a system in a grouped launch (longer chunks)
    int chunk_log(long nReal, long nPad, bool grouped = false) const;

    // Autotuner: times every kernel that can run each level of this shape
    // (K-split tiles, narrow or wide matrix-core tiles, SIMT) on this core's
    // own trajectory buffer (its contents are overwritten) and keeps the
    // fastest per (n_a, precision, level, pairs, sweep) in a process-wide
    // schedule that every later solve of that shape follows.  Returns the
    // number of level launches tuned.  Synchronises the stream.
    int autotune(hipStream_t s, long nReal, long nPad);
    // Level paths of the schedule (kPathAuto: the built-in rule)
    enum LevelPath { kPathAuto = 0, kPathKs = 1, kPathNarrow = 2, kPathWide = 3, kPathSimt = 4 };
    int scheduled_path(bool down, int l, long np) const;
    static void clear_schedule();

private:
    friend class BatchSolve;   // plans grouped launches over the buffers below
    bool level_is_f64(int l) const;
    bool level_is_delta(int l) const;
    void level(hipStream_t s, bool down, int l, long np);
    void level_with(hipStream_t s, bool down, int l, long np, int path);
    // down-sweep levels hi .. lo (descending); consecutive K-split levels pass
    // the copy of their pairs' X[R] down without a gather launch
    void down_levels(hipStream_t s, int hi, int lo, long nReal);
    bool ks_taken(int l, long np, int path) const;
    // the chunk recurrences of the lowest levels (pass 1 or 2)
    void chunks(hipStream_t s, bool pass2, long nReal, long nPad);
    std::vector<int> level_candidates(int l, long np) const;
    void zero_trajectory(hipStream_t s, long nPad);

    int method_, n_, m_, p_, na_;
    bool zC_, zD_;
    long nPadMax_;
    int L_;
    ScanOptions opt_;
    // step map
    DevArray<float> prod32_;      // float products of A (mirror)
    DevArray<double> prod64_;     // FP64 products of A
    DevArray<double> A64_, B64_;
    DevArray<double> Bbar64_, BbT64_;
    // ABM with an FP64 map: [B | AB] transposed and the factors of the leaves
    DevArray<double> BfT64_;
    DevArray<double> CT64_, DT64_;   // C^T and D^T in FP64 for the outputs
    double leafG_[5][2] = {};
    double h_ = 0.0;   // the step of the map built last
    bool leafFac_ = false;
    DevArray<float> Bbar32_, BbT32_;
    DevArray<float> CT_, DT_;        // C^T and D^T for the output stage
    // power table (transposed), L+1 levels, and FP32 level tables
    DevArray<double> PT64_;
    DevArray<float> PT32_;
    DevArray<float> PTm_;            // zero-padded FP32 tables of the MFMA levels
    DevArray<float> Sg_;             // X[R] copy for row-split down-sweep levels (na > 256)
    DevArray<float> Sks32_;          // X[R] copies for K-split down-sweep levels (few
    DevArray<double> Sks64_;         // pairs): two halves of ksHalf_ values, ping-pong
    size_t ksHalf_ = 0;
    struct KsChain {                 // the down-sweep level being launched
        bool gathered = false;       // its half already holds the copy
        int half = 0;
        bool out = false;            // it writes the next level's copy (other half)
        long npOut = 0;
    } ksc_;
    int nap_ = 0, nrp_ = 0;
    // trajectory
    DevArray<float> X32_;
    DevArray<double> X64_;
    DevArray<double> seed_, tot_, seedNext_;
};

}  // namespace parode
