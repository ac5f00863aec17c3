#include "scan_core.h"

#include <algorithm>
#include <array>
#include <mutex>
#include <map>

#include "config.h"
#include "kernels.h"

namespace {
// columns padded to a multiple of 8 (at least 8)
int pad8(int c) { return c > 0 ? (c + 7) / 8 * 8 : 8; }
}  // namespace

namespace parode {

ScanCore::ScanCore(int method, int n, int m, int p, bool zC, bool zD, long nPadMax, ScanOptions opt,
                   Arena* arena)
    : method_(method), n_(n), m_(m), p_(p), na_(method == kABM ? 4 * n : n), zC_(zC), zD_(zD),
      nPadMax_(next_pow2(nPadMax < 1 ? 1 : nPadMax)), L_(ilog2(nPadMax_)), opt_(opt)
{
    if (n <= 0) throw std::invalid_argument("ScanCore: no states");
    const size_t nn = (size_t)n * n, nm = (size_t)n * (m > 0 ? m : 1);
    const size_t na2 = (size_t)na_ * na_;
    const int kq = method == kABM ? 5 : 3;
    if (opt_.precision == kPlain32) {
        prod32_.alloc(4 * nn + 4 * nm, arena);            // A, A2, A3, A4 | B, AB, A2B, A3B (float)
    } else {
        prod64_.alloc(4 * nn + 4 * nm, arena);
        A64_.alloc(nn, arena);
        B64_.alloc(nm, arena);
    }
    Bbar64_.alloc((size_t)n * kq * (m > 0 ? m : 1), arena);
    Bbar32_.alloc((size_t)n * kq * (m > 0 ? m : 1), arena);
    BbT64_.alloc((size_t)n * kq * (m > 0 ? m : 1), arena);
    BbT32_.alloc((size_t)n * kq * (m > 0 ? m : 1), arena);
    if (method == kABM && opt_.precision != kPlain32 && m > 0) BfT64_.alloc(2 * nm, arena);
    CT_.alloc((size_t)n * (p > 0 ? p : 1), arena);
    DT_.alloc((size_t)(m > 0 ? m : 1) * (p > 0 ? p : 1), arena);
    CT64_.alloc((size_t)n * pad8(p), arena);
    DT64_.alloc((size_t)(m > 0 ? m : 1) * pad8(p), arena);
    PT64_.alloc(na2 * (L_ + 1), arena);
    int n32 = 0;
    for (int l = 0; l < L_; ++l)
        if (!level_is_f64(l)) n32 = l + 1;
    if (n32 > 0) PT32_.alloc(na2 * n32, arena);
    if (n32 > 0 && k::level_mfma_supported(na_)) {
        nap_ = (na_ + 31) / 32 * 32;
        nrp_ = k::mfma_table_rows(na_);
        PTm_.alloc((size_t)nap_ * nrp_ * n32, arena);
        // down-sweep copy of X[R] for the row-split MFMA levels
        if (k::level_mfma_splits(na_)) Sg_.alloc((size_t)(nPadMax_ / 2 > 0 ? nPadMax_ / 2 : 1) * na_, arena);
    }
    if (xIsDouble())
        X64_.alloc((size_t)nPadMax_ * na_, arena);
    else
        X32_.alloc((size_t)nPadMax_ * na_, arena);
    // down-sweep copy of X[R] for the K-split levels with few pairs
    if (na_ >= kKsMinStates) {
        ksHalf_ = (size_t)std::min<long>(k::kKsMaxPairs, nPadMax_ / 2 > 0 ? nPadMax_ / 2 : 1) * na_;
        if (xIsDouble())
            Sks64_.alloc(2 * ksHalf_, arena);
        else
            Sks32_.alloc(2 * ksHalf_, arena);
    }
    seed_.alloc(na_, arena);
    tot_.alloc(na_, arena);
    seedNext_.alloc(na_, arena);
}

size_t ScanCore::estimate_footprint(int method, int n, int m, long nPadMax, ScanOptions opt,
                                    int p)
{
    const long nP = next_pow2(nPadMax < 1 ? 1 : nPadMax);
    const int L = ilog2(nP);
    const size_t na = method == kABM ? 4 * n : n;
    const size_t nn = (size_t)n * n, nm = (size_t)n * (m > 0 ? m : 1);
    size_t b = (opt.precision == kPlain32 ? 4 : 8) * (4 * nn + 4 * nm);
    if (opt.precision != kPlain32) b += 8 * (nn + nm);
    b += 24 * (size_t)n * (method == kABM ? 5 : 3) * (m > 0 ? m : 1);
    if (method == kABM && opt.precision != kPlain32 && m > 0) b += 16 * nm;
    b += (size_t)(n + (m > 0 ? m : 1)) * (4 * (p > 0 ? p : 1) + 8 * pad8(p));   // C^T, D^T
    b += 8 * na * na * (L + 1);
    if (opt.precision != kFp64) b += 4 * na * na * L;
    if (opt.precision != kFp64 && na >= 17 && na <= 1024) {
        b += 4 * (size_t)((na + 31) / 32 * 32) * ((na + 255) / 256 * 256) * L;
        if (k::level_mfma_splits((int)na)) b += 4 * (size_t)(nP / 2 > 0 ? nP / 2 : 1) * na;
    }
    b += (opt.precision == kFp64 ? 8 : 4) * (size_t)nP * na;
    if ((int)na >= kKsMinStates)
        b += 2 * (opt.precision == kFp64 ? 8 : 4) *
             (size_t)std::min<long>(k::kKsMaxPairs, nP / 2 > 0 ? nP / 2 : 1) * na;
    b += 24 * na;
    return b + 20 * Arena::kAlign;   // block rounding when carved from an Arena
}

size_t ScanCore::footprint() const
{
    return prod32_.bytes() + prod64_.bytes() + A64_.bytes() + B64_.bytes() + Bbar64_.bytes() +
           Bbar32_.bytes() + BbT64_.bytes() + BbT32_.bytes() + BfT64_.bytes() + CT_.bytes() +
           DT_.bytes() + CT64_.bytes() + DT64_.bytes() +
           PT64_.bytes() + PT32_.bytes() + PTm_.bytes() + Sg_.bytes() + Sks32_.bytes() +
           Sks64_.bytes() + X32_.bytes() + X64_.bytes() + seed_.bytes() +
           tot_.bytes() + seedNext_.bytes();
}

bool ScanCore::level_is_f64(int l) const
{
    if (opt_.precision == kFp64) return true;
    if (opt_.precision == kPlain32) return false;
    return l >= L_ - opt_.top64;
}

// Mixed precision, FP32 levels applied as X + (P - I) X: the lowest
// opt_.deltaLevels, or all of them for wide states (n_a >= 768), where the
// rounding of FP32 sums over n_a near-identity products would otherwise pass
// the 1e-6 bar.
bool ScanCore::level_is_delta(int l) const
{
    constexpr int kWideDeltaStates = 768;
    return opt_.precision == kMixed && !level_is_f64(l) &&
           (l < opt_.deltaLevels || na_ >= kWideDeltaStates);
}

// ------------------------------------------------------------------ step map
void ScanCore::build_map(hipStream_t s, const float* dA, const float* dB, double h,
                         const SideStream* side)
{
    h_ = h;
    const int n = n_, m = m_;
    const size_t nn = (size_t)n * n, nm = (size_t)n * (m > 0 ? m : 1);
    double* MT = PT64_.get();
    if (opt_.precision == kPlain32) {
        // float products with sequential-k accumulation, as the reference's host
        float* A2 = prod32_.get();
        float* A3 = A2 + nn;
        float* A4 = A3 + nn;
        float* AB = A4 + nn;
        float* A2B = AB + nm;
        float* A3B = A2B + nm;
        k::gemm_seq_f32(s, n, n, n, dA, dA, A2);
        if (method_ == kRK4) {
            k::gemm_seq_f32(s, n, n, n, A2, dA, A3);
            k::gemm_seq_f32(s, n, n, n, dA, A3, A4);
            if (m > 0) {
                k::gemm_seq_f32(s, n, n, m, dA, dB, AB);
                k::gemm_seq_f32(s, n, n, m, A2, dB, A2B);
                k::gemm_seq_f32(s, n, n, m, A3, dB, A3B);
            }
            k::rk4_combine(s, true, n, m, h, dA, A2, A3, A4, dB, AB, A2B, A3B, MT, Bbar64_.get(),
                           Bbar32_.get());
        } else {
            if (m > 0) k::gemm_seq_f32(s, n, n, m, dA, dB, AB);
            k::abm_combine(s, true, n, m, h, dA, A2, dB, AB, MT, Bbar64_.get(), Bbar32_.get());
        }
    } else {
        const bool mm = (level_kernel() & 2) == 0;
        double* A = A64_.get();
        double* B = B64_.get();
        double* A2 = prod64_.get();
        double* A3 = A2 + nn;
        double* A4 = A3 + nn;
        double* AB = A4 + nn;
        double* A2B = AB + nm;
        double* A3B = A2B + nm;
        k::cast_f32_to_f64(s, (long)nn, dA, A);
        if (m > 0) k::cast_f32_to_f64(s, (long)n * m, dB, B);
        k::gemm_f64(s, n, n, n, A, A, A2, mm);
        if (method_ == kRK4) {
            k::gemm_f64(s, n, n, n, A2, A, A3, mm);
            k::gemm_f64(s, n, n, n, A, A3, A4, mm);
            if (m > 0) {
                k::gemm_f64(s, n, n, m, A, B, AB, mm);
                k::gemm_f64(s, n, n, m, A2, B, A2B, mm);
                k::gemm_f64(s, n, n, m, A3, B, A3B, mm);
            }
            k::rk4_combine(s, false, n, m, h, A, A2, A3, A4, B, AB, A2B, A3B, MT, Bbar64_.get(),
                           Bbar32_.get());
        } else {
            if (m > 0) k::gemm_f64(s, n, n, m, A, B, AB, mm);
            if (m > 0) {   // leaves as [B | AB] times two combinations of the samples
                k::transpose<double>(s, n, m, B, BfT64_.get());
                k::transpose<double>(s, n, m, AB, BfT64_.get() + (size_t)m * n);
                k::LeafComb G;
                k::abm_leaf_factors(h, G);
                for (int q = 0; q < 5; ++q)
                    for (int g = 0; g < 2; ++g) leafG_[q][g] = G.g[q][g];
                leafFac_ = true;
            }
            // what the factored leaves and the chunk recurrences read is
            // ready: the companion map and B-bar (for the power table, and the
            // unfactored leaves, which large input counts would take) on the
            // side stream
            if (side && m > 0 && m <= 16) {
                PARODE_HIP(hipEventRecord(side->fork, s));
                PARODE_HIP(hipStreamWaitEvent(side->stream, side->fork, 0));
                s = side->stream;
            }
            k::abm_combine(s, false, n, m, h, A, A2, B, AB, MT, Bbar64_.get(), Bbar32_.get());
        }
    }
    if (side && s != side->stream) {
        PARODE_HIP(hipEventRecord(side->fork, s));
        PARODE_HIP(hipStreamWaitEvent(side->stream, side->fork, 0));
    }
    const int kq = method_ == kABM ? 5 : 3;
    if (m > 0) {
        k::transpose<double>(s, n, kq * m, Bbar64_.get(), BbT64_.get());
        k::transpose<float>(s, n, kq * m, Bbar32_.get(), BbT32_.get());
    }
}

void ScanCore::build_powers(hipStream_t s, bool withLast)
{
    const size_t na2 = (size_t)na_ * na_;
    // (P^T)^2 = (P^2)^T: squaring the transposed table keeps it transposed
    const bool mm = (level_kernel() & 2) == 0;
    if (mm && k::powers_small_supported(na_)) {
        k::powers_small(s, na_, L_, PT64_.get());
    } else {
        for (int l = 0; l < (withLast ? L_ : L_ - 1); ++l)
            k::gemm_f64(s, na_, na_, na_, PT64_.get() + l * na2, PT64_.get() + l * na2,
                        PT64_.get() + (l + 1) * na2, mm);
    }
    int n32 = 0, nd = 0;
    for (int l = 0; l < L_; ++l) {
        if (!level_is_f64(l)) n32 = l + 1;
        if (level_is_delta(l)) nd = l + 1;
    }
    k::make_level_tables(s, na_, nap_, nrp_, n32, nd, PT64_.get(), PT32_.get(), PTm_.get());
}

// -------------------------------------------------------------------- leaves
// ABM leaves hold only the last n of the 4n companion rows.  When level 0 of
// the up-sweep runs on the matrix cores as its own launch, it reads just those
// rows (and only the matching columns of P), so the leaves skip the zero rows.
bool ScanCore::structured_level0(long nReal, long nPad) const
{
    return method_ == kABM && !xIsDouble() && L_ > 0 && !level_is_f64(0) && !level_is_delta(0) &&
           (level_kernel() & 1) == 0 && PTm_.get() != nullptr && top_level(nReal, nPad) > 0 &&
           chunk_log(nReal, nPad) == 0;
}

// The lowest levels as chunk recurrences (FP32 levels; RK4 n_a <= 256, ABM
// n <= 64, or n <= 256 with an FP64 map): levels 0..lc-1 over chunks of 2^lc
// elements (ABM: only the
// companion map's dense block row per step).  lc is 7, or less so that a
// launch has at least 1024 chunks (64 work-groups); at least one tree level
// is launched above them.
int ScanCore::chunk_log(long nReal, long nPad, bool grouped) const
{
    constexpr int kChunkLog = 7, kMinChunksLog = 10, kMinGroupedChunksLog = 6;
    if (xIsDouble() || (level_kernel() & 9) != 0) return 0;
    // ABM: n <= 64 with the reference's float map, n <= 128 factored (FP64 map)
    const bool abmOk = opt_.precision == kPlain32 ? k::abm_chunks_supported(n_)
                                                  : k::abm_chunks_factored_supported(n_);
    if (method_ == kABM ? !abmOk : !k::chunks_dense_supported(na_)) return 0;
    // alone, a system needs 2^10 chunks to fill the GPU; in a grouped launch the
    // other systems fill it, so its chunks run up to 2^7 steps (a work-group
    // loads the whole map into registers once for its chunks) below the first
    // tree level that the top or an FP64 table takes
    const int lt = top_level(nReal, nPad);
    int lc = std::min(kChunkLog, ilog2(nPad) - (grouped ? kMinGroupedChunksLog : kMinChunksLog));
    if (grouped) {
        lc = std::min(lc, lt - 1);
        while (lc > 0 && level_is_f64(lc - 1)) --lc;
    }
    if (lc < 2 || lt <= lc) return 0;
    // (the chunk kernels read the FP64 map, so the FP32 tables' form, plain
    // or delta, does not matter below lc)
    for (int l = 0; l < lc; ++l)
        if (level_is_f64(l)) return 0;
    return lc;
}

void ScanCore::chunks(hipStream_t s, bool pass2, long nReal, long nPad)
{
    // an FP64 map: the companion block row factored through A and A^2 (the
    // plain path keeps the reference's float M)
    if (method_ == kABM && opt_.precision != kPlain32)
        k::abm_chunks_factored(s, pass2, n_, chunk_log(nReal, nPad), nReal, nPad, h_, A64_.get(),
                               prod64_.get(), X32_.get());
    else if (method_ == kABM)
        k::abm_chunks(s, pass2, n_, chunk_log(nReal, nPad), nReal, nPad, PT64_.get(), X32_.get());
    else
        k::chunks_dense(s, pass2, na_, chunk_log(nReal, nPad), 1, nReal, nPad, PT64_.get(),
                        X32_.get());
}

void ScanCore::leaves(hipStream_t s, const void* dU, bool uDouble, long sL, long oL, long rs,
                      long nReal, long nPad)
{
    const int kq = method_ == kABM ? 5 : 3;
    const int off = method_ == kABM ? 3 * n_ : 0;
    if (nPad > nPadMax_) throw std::invalid_argument("ScanCore::leaves: batch too long");
    if (leafFac_) {
        const bool wz =
            xIsDouble() || !(structured_level0(nReal, nPad) || chunk_log(nReal, nPad) > 0);
        k::LeafComb G;
        for (int q = 0; q < 5; ++q)
            for (int g = 0; g < 2; ++g) G.g[q][g] = leafG_[q][g];
        bool done;
        if (xIsDouble())
            done = uDouble ? k::leaves_factored<double, double>(s, n_, m_, kq, 2, na_, off, nReal,
                                                                nPad, BfT64_.get(), G,
                                                                (const double*)dU, sL, oL, rs,
                                                                X64_.get(), wz)
                           : k::leaves_factored<double, float>(s, n_, m_, kq, 2, na_, off, nReal,
                                                               nPad, BfT64_.get(), G,
                                                               (const float*)dU, sL, oL, rs,
                                                               X64_.get(), wz);
        else
            done = uDouble ? k::leaves_factored<float, double>(s, n_, m_, kq, 2, na_, off, nReal,
                                                               nPad, BfT64_.get(), G,
                                                               (const double*)dU, sL, oL, rs,
                                                               X32_.get(), wz)
                           : k::leaves_factored<float, float>(s, n_, m_, kq, 2, na_, off, nReal,
                                                              nPad, BfT64_.get(), G,
                                                              (const float*)dU, sL, oL, rs,
                                                              X32_.get(), wz);
        if (done) return;
    }
    if (xIsDouble()) {
        if (uDouble)
            k::leaves<double, double, double>(s, n_, m_, kq, na_, off, nReal, nPad, BbT64_.get(),
                                              (const double*)dU, sL, oL, rs, X64_.get());
        else
            k::leaves<double, float, double>(s, n_, m_, kq, na_, off, nReal, nPad, BbT64_.get(),
                                             (const float*)dU, sL, oL, rs, X64_.get());
    } else if (uDouble) {
        k::leaves<float, double, double>(s, n_, m_, kq, na_, off, nReal, nPad, BbT64_.get(),
                                         (const double*)dU, sL, oL, rs, X32_.get(),
                                         !(structured_level0(nReal, nPad) ||
                                           chunk_log(nReal, nPad) > 0));
    } else {
        k::leaves<float, float, float>(s, n_, m_, kq, na_, off, nReal, nPad, BbT32_.get(),
                                       (const float*)dU, sL, oL, rs, X32_.get(),
                                       !(structured_level0(nReal, nPad) ||
                                         chunk_log(nReal, nPad) > 0));
    }
}

// --------------------------------------------------------------------- sweeps
namespace {
// The autotuned schedule: (n_a, precision, level, pairs, down) -> path
std::mutex g_schedMu;
std::map<std::array<long, 5>, int> g_sched;
}  // namespace

int ScanCore::scheduled_path(bool down, int l, long np) const
{
    std::lock_guard<std::mutex> lk(g_schedMu);
    auto it = g_sched.find({(long)na_, (long)opt_.precision, (long)l, np, down ? 1L : 0L});
    return it == g_sched.end() ? kPathAuto : it->second;
}

void ScanCore::clear_schedule()
{
    std::lock_guard<std::mutex> lk(g_schedMu);
    g_sched.clear();
}

void ScanCore::level(hipStream_t s, bool down, int l, long np)
{
    level_with(s, down, l, np, scheduled_path(down, l, np));
}

// The kernels that can run level l with np pairs: K-split tiles (few pairs),
// matrix-core tiles (FP32 tables), SIMT (FP64 tables).
std::vector<int> ScanCore::level_candidates(int l, long np) const
{
    std::vector<int> c;
    const bool ks = np <= k::kKsMaxPairs && na_ >= kKsMinStates;
    const bool f64 = xIsDouble() || level_is_f64(l);
    const bool mfma = !f64 && (level_kernel() & 1) == 0 && PTm_.get();
    if (ks && (f64 || mfma)) c.push_back(kPathKs);
    if (mfma) {
        c.push_back(kPathNarrow);
        c.push_back(kPathWide);
    } else {
        c.push_back(kPathSimt);
    }
    return c;
}

void ScanCore::level_with(hipStream_t s, bool down, int l, long np, int path)
{
    const size_t na2 = (size_t)na_ * na_;
    const long d = 1L << l;
    const bool ksOk = np <= k::kKsMaxPairs && na_ >= kKsMinStates;
    const bool mfmaOk = !xIsDouble() && !level_is_f64(l) && (level_kernel() & 1) == 0 && PTm_.get();
    int tile = 0;   // matrix-core tiles by the pair count
    if (path == kPathAuto && ksOk) path = kPathKs;
    if (path == kPathNarrow) tile = 1;
    if (path == kPathWide) tile = 2;
    // few pairs: small K-split tiles on the matrix cores, spread over many CUs
    if (path == kPathKs && ksOk) {
        // down: this level's copy of X[R] in one half (gathered first unless
        // the level above wrote it), the next level's, when it is K-split too,
        // into the other
        const size_t in = down ? (size_t)ksc_.half * ksHalf_ : 0;
        const size_t out = (size_t)(1 - ksc_.half) * ksHalf_;
        const bool gathered = down && ksc_.gathered, chain = down && ksc_.out;
        if (xIsDouble()) {
            k::level_ks<double, double>(s, down, na_, na_, d, np, PT64_.get() + l * na2,
                                        X64_.get(), Sks64_.get() + in, false, gathered,
                                        chain ? Sks64_.get() + out : nullptr, ksc_.npOut);
            return;
        }
        if (level_is_f64(l)) {
            k::level_ks<double, float>(s, down, na_, na_, d, np, PT64_.get() + l * na2,
                                       X32_.get(), Sks32_.get() + in, false, gathered,
                                       chain ? Sks32_.get() + out : nullptr, ksc_.npOut);
            return;
        }
        if (mfmaOk) {
            k::level_ks<float, float>(s, down, na_, nap_, d, np,
                                      PTm_.get() + (size_t)l * nap_ * nrp_, X32_.get(),
                                      Sks32_.get() + in, level_is_delta(l), gathered,
                                      chain ? Sks32_.get() + out : nullptr, ksc_.npOut);
            return;
        }
    }
    if (xIsDouble()) {
        k::level<double, double>(s, down, false, na_, d, np, PT64_.get() + l * na2, X64_.get());
    } else if (level_is_f64(l)) {
        k::level<float, double>(s, down, false, na_, d, np, PT64_.get() + l * na2, X32_.get());
    } else if (mfmaOk) {
        k::level_mfma(s, down, level_is_delta(l), na_, d, np,
                      PTm_.get() + (size_t)l * nap_ * nrp_, X32_.get(), k::LevelRows{}, Sg_.get(),
                      tile);
    } else {
        k::level<float, float>(s, down, level_is_delta(l), na_, d, np, PT32_.get() + l * na2,
                               X32_.get());
    }
}

// Pairs of level l that intersect the real elements [0, nReal): the others
// hold only zero leaves (up) or padding states (down) and are skipped.
static long pairs_at(long nReal, int l) { return (nReal + (2L << l) - 1) / (2L << l); }

int ScanCore::top_level(long nReal, long nPad) const
{
    const int L = ilog2(nPad);
    if (level_kernel() & 4) return L;                // fused top switched off
    if ((long)nPad * na_ <= 65536) return 0;         // small: the whole tree in one work-group
    // from 64 states on, the levels with few pairs run as K-split launches
    // over many CUs, faster than streaming each table through one CU
    if (na_ >= 64) return L;
    int l = 0;
    while (l < L && pairs_at(nReal, l) > 32) ++l;
    return l;
}

// Level 1 of the ABM scan on the matrix cores with its inputs nonzero from
// row 2n (level 0 leaves rows below 2n unwritten): rows below 2n of M^2 are
// shifts by 2n, so the product forms only rows from 32 floor(2n / 32) on.
bool ScanCore::structured_level1(long nReal, long nPad) const
{
    const int lt = top_level(nReal, nPad);
    return structured_level0(nReal, nPad) && L_ > 1 && lt > 1 && !level_is_f64(1) &&
           !level_is_delta(1) && (pairs_at(nReal, 1) > k::kKsMaxPairs || na_ < kKsMinStates);
}

// Rows of the structured ABM levels (M: rows below 3n are shifts by n; M^2:
// rows below 2n are shifts by 2n).
static k::LevelRows abm_rows(int n, int level, bool down, bool next)
{
    k::LevelRows r;
    const int b0 = (3 * n) / 32 * 32, b1 = (2 * n) / 32 * 32;
    if (level == 0) {
        r.rbeg = 3 * n;
        r.jbeg = down ? 0 : b0;
        if (down) {            // only the rows the outputs read
            r.ibase = b0;
            r.clo = b0;
        } else if (next) {     // rows from 2n for a structured level 1
            r.ibase = b0;
            r.clo = std::min(2 * n, b0);
            r.sh = n;
            r.zlo = std::min(b1, b0);
            r.zhi = r.clo;
        }
    } else {
        r.rbeg = 2 * n;
        r.jbeg = down ? 0 : b1;
        r.ibase = b1;
        r.clo = 0;
        r.sh = 2 * n;
    }
    return r;
}

void ScanCore::sweep_up(hipStream_t s, long nReal, long nPad, hipEvent_t waitTables)
{
    const int lt = top_level(nReal, nPad);
    int l = 0;
    const int lc = chunk_log(nReal, nPad);
    // the chunk recurrences read only the map; tree levels read the tables
    if (lc) chunks(s, false, nReal, nPad);
    if (waitTables) PARODE_HIP(hipStreamWaitEvent(s, waitTables, 0));
    if (lc) {   // levels below lc: chunk recurrences
        l = lc;
    } else if (lt > 0 && structured_level0(nReal, nPad)) {
        const bool s1 = structured_level1(nReal, nPad);
        k::level_mfma(s, false, false, na_, 1, pairs_at(nReal, 0), PTm_.get(), X32_.get(),
                      abm_rows(n_, 0, false, s1), Sg_.get());
        l = 1;
        if (s1) {
            k::level_mfma(s, false, false, na_, 2, pairs_at(nReal, 1),
                          PTm_.get() + (size_t)nap_ * nrp_, X32_.get(), abm_rows(n_, 1, false, false),
                          Sg_.get());
            l = 2;
        }
    }
    for (; l < lt; ++l) level(s, false, l, pairs_at(nReal, l));
}

// Timings run on zeros (the kernels' speed does not depend on the values).
void ScanCore::zero_trajectory(hipStream_t s, long nPad)
{
    if (xIsDouble())
        k::fill_f64(s, nPad * na_, 0.0, X64_.get());
    else
        k::fill_zero_f32(s, nPad * na_, X32_.get());
}

int ScanCore::autotune(hipStream_t s, long nReal, long nPad)
{
    if (nPad > nPadMax_ || nReal < 1) throw std::invalid_argument("ScanCore::autotune: bad length");
    const int lt = top_level(nReal, nPad);
    const bool s0 = lt > 0 && structured_level0(nReal, nPad);
    const bool s1 = s0 && structured_level1(nReal, nPad);
    const int lc = chunk_log(nReal, nPad);
    const int lo = lc > 0 ? lc : s1 ? 2 : s0 ? 1 : 0;   // the lowest tree level
    hipEvent_t e0 = nullptr, e1 = nullptr;
    PARODE_HIP(hipEventCreate(&e0));
    PARODE_HIP(hipEventCreate(&e1));
    int tuned = 0;
    try {
        for (int l = 0; l < lt; ++l) {
            if ((l == 0 && s0) || (l == 1 && s1) || l < lc) continue;   // fixed launches
            const long np = pairs_at(nReal, l);
            const std::vector<int> cand = level_candidates(l, np);
            if (cand.size() < 2) continue;
            for (int down = 0; down < 2; ++down) {
                int best = kPathAuto;
                float bestMs = 0.0f;
                for (int c : cand) {
                    constexpr int kReps = 3;
                    // a down-sweep K-split level is timed as inside a chain of
                    // them: its copy of X[R] written by the level above (when
                    // there is one), the next level's copy written by it
                    // (when there is one); the copies' contents do not matter here
                    ksc_ = KsChain{};
                    if (down && c == kPathKs) {
                        const long npNext = l - 1 >= lo ? pairs_at(nReal, l - 1) : 0;
                        ksc_.gathered = l + 1 < lt;
                        ksc_.out = npNext > 0 && npNext <= k::kKsMaxPairs;
                        ksc_.npOut = ksc_.out ? npNext : 0;
                    }
                    // a long fill first, so that every timed launch is queued
                    // before the GPU reaches it: the events then time the GPU
                    // alone, as in a solve, not the host's launch rate
                    zero_trajectory(s, nPad);
                    level_with(s, down != 0, l, np, c);   // warm-up
                    PARODE_HIP(hipEventRecord(e0, s));
                    for (int r = 0; r < kReps; ++r) level_with(s, down != 0, l, np, c);
                    PARODE_HIP(hipEventRecord(e1, s));
                    ksc_ = KsChain{};
                    PARODE_HIP(hipEventSynchronize(e1));
                    float ms = 0.0f;
                    PARODE_HIP(hipEventElapsedTime(&ms, e0, e1));
                    if (best == kPathAuto || ms < bestMs) {
                        best = c;
                        bestMs = ms;
                    }
                }
                std::lock_guard<std::mutex> lk(g_schedMu);
                g_sched[{(long)na_, (long)opt_.precision, (long)l, np, (long)down}] = best;
                ++tuned;
            }
        }
    } catch (...) {
        ksc_ = KsChain{};
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        throw;
    }
    PARODE_HIP(hipEventDestroy(e0));
    PARODE_HIP(hipEventDestroy(e1));
    return tuned;
}

void ScanCore::sweep_top(hipStream_t s, long nReal, long nPad, const double* dSeed, double* dTot)
{
    const int L = ilog2(nPad), lt = top_level(nReal, nPad);
    if (lt >= L) {   // no fused top: the root alone
        if (xIsDouble())
            k::root<double, double>(s, na_, nPad - 1, dSeed, X64_.get(), dTot);
        else
            k::root<float, double>(s, na_, nPad - 1, dSeed, X32_.get(), dTot);
        return;
    }
    int l64 = L;
    while (l64 > 0 && level_is_f64(l64 - 1)) --l64;
    if (xIsDouble())
        k::top<double>(s, na_, lt, L, 0, nReal, nullptr, PT64_.get(), X64_.get(), dSeed, dTot);
    else
        k::top<float>(s, na_, lt, L, l64, nReal, PT32_.get(), PT64_.get(), X32_.get(), dSeed,
                      dTot);
}

// Whether level_with(path) runs level l on the K-split kernel.
bool ScanCore::ks_taken(int l, long np, int path) const
{
    if (np > k::kKsMaxPairs || na_ < kKsMinStates) return false;
    if (path != kPathAuto && path != kPathKs) return false;
    const bool mfmaOk = !xIsDouble() && !level_is_f64(l) && (level_kernel() & 1) == 0 && PTm_.get();
    return xIsDouble() || level_is_f64(l) || mfmaOk;
}

void ScanCore::down_levels(hipStream_t s, int hi, int lo, long nReal)
{
    bool haveCopy = false;
    int half = 0;
    for (int l = hi; l >= lo; --l) {
        const long np = pairs_at(nReal, l);
        const int path = scheduled_path(true, l, np);
        const bool ks = ks_taken(l, np, path);
        bool next = false;
        long npNext = 0;
        if (ks && l - 1 >= lo) {
            npNext = pairs_at(nReal, l - 1);
            next = ks_taken(l - 1, npNext, scheduled_path(true, l - 1, npNext));
        }
        ksc_ = KsChain{ks && haveCopy, half, next, next ? npNext : 0};
        level_with(s, true, l, np, path);
        haveCopy = next;
        if (next) half ^= 1;
    }
    ksc_ = KsChain{};
}

void ScanCore::sweep_down(hipStream_t s, long nReal, long nPad)
{
    const int lt = top_level(nReal, nPad);
    if (const int lc = chunk_log(nReal, nPad)) {   // levels below lc: chunk recurrences
        down_levels(s, lt - 1, lc, nReal);
        chunks(s, true, nReal, nPad);
        return;
    }
    const bool s0 = lt > 0 && structured_level0(nReal, nPad);
    const bool s1 = s0 && structured_level1(nReal, nPad);
    down_levels(s, lt - 1, s1 ? 2 : s0 ? 1 : 0, nReal);
    if (s1)
        k::level_mfma(s, true, false, na_, 2, pairs_at(nReal, 1), PTm_.get() + (size_t)nap_ * nrp_,
                      X32_.get(), abm_rows(n_, 1, true, false), Sg_.get());
    // the X[L] leaves of level 0 hold only their last n rows, and the outputs
    // read only the last n rows of each state: the rest is not computed
    if (s0)
        k::level_mfma(s, true, false, na_, 1, pairs_at(nReal, 0), PTm_.get(), X32_.get(),
                      abm_rows(n_, 0, true, false), Sg_.get());
}

void ScanCore::carry(hipStream_t s, long nPad, const double* dSeed, const double* dTot,
                     double* dOut)
{
    k::carry(s, na_, powerT(ilog2(nPad)), dSeed, dTot, dOut);
}

// -------------------------------------------------------------------- outputs
// C^T and D^T (and their FP64 copies padded to pp columns) for the output
// kernels: they depend only on C and D, so a solve makes them on its side
// stream, off the path to the outputs.
void ScanCore::prepare_outputs(hipStream_t s, const float* dC, const float* dD, bool uDouble)
{
    const bool zC = zC_ || !dC, zD = zD_ || !dD;
    if (!zC) k::transpose<float>(s, p_, n_, dC, CT_.get());
    if (!zD && m_ > 0) k::transpose<float>(s, p_, m_, dD, DT_.get());
    if (xIsDouble() || uDouble) {
        const int pp = pad8(p_);
        if (!zC) k::pad_cast(s, n_, p_, pp, CT_.get(), CT64_.get());
        if (!zD && m_ > 0) k::pad_cast(s, m_, p_, pp, DT_.get(), DT64_.get());
    }
}

// prepared: prepare_outputs' tables were made earlier on this stream or on one
// joined to it; else they are made here first.
void ScanCore::outputs(hipStream_t s, const float* dC, const float* dD, const void* dU,
                       bool uDouble, long sO, long oO, long N, float* dY, bool prepared)
{
    const int oX = method_ == kABM ? 3 : 0;
    if (!prepared) prepare_outputs(s, dC, dD, uDouble);
    dC = CT_.get();
    dD = DT_.get();
    // FP64 outputs: one thread per element where the shape fits
    if (xIsDouble() || uDouble) {
        const int pp = pad8(p_);
        bool done;
        if (xIsDouble())
            done = uDouble ? k::outputs_lanes<double, double>(s, n_, m_, p_, pp, na_, oX, zC_, zD_,
                                                               N, CT64_.get(), DT64_.get(),
                                                               X64_.get(), (const double*)dU, sO,
                                                               oO, dY)
                           : k::outputs_lanes<double, float>(s, n_, m_, p_, pp, na_, oX, zC_, zD_,
                                                              N, CT64_.get(), DT64_.get(),
                                                              X64_.get(), (const float*)dU, sO, oO,
                                                              dY);
        else
            done = k::outputs_lanes<float, double>(s, n_, m_, p_, pp, na_, oX, zC_, zD_, N,
                                                   CT64_.get(), DT64_.get(), X32_.get(),
                                                   (const double*)dU, sO, oO, dY);
        if (done) return;
    }
    if (xIsDouble()) {
        if (uDouble)
            k::outputs<double, double>(s, n_, m_, p_, na_, oX, zC_, zD_, N, dC, dD, X64_.get(),
                                       (const double*)dU, sO, oO, dY);
        else
            k::outputs<double, float>(s, n_, m_, p_, na_, oX, zC_, zD_, N, dC, dD, X64_.get(),
                                      (const float*)dU, sO, oO, dY);
    } else if (uDouble) {
        k::outputs<float, double>(s, n_, m_, p_, na_, oX, zC_, zD_, N, dC, dD, X32_.get(),
                                  (const double*)dU, sO, oO, dY);
    } else {
        k::outputs<float, float>(s, n_, m_, p_, na_, oX, zC_, zD_, N, dC, dD, X32_.get(),
                                 (const float*)dU, sO, oO, dY);
    }
}

void ScanCore::startup_outputs(hipStream_t s, const float* dC, const float* dD, const double* dX0,
                               const void* dU, bool uDouble, float* dY, bool prepared)
{
    // x_0, x_1, x_2 are the first three n-blocks of X0; their inputs are
    // rows 0, 2, 4 of the h/2-grid start-up samples (AdamsMoulton.cpp:106-116)
    if (!prepared) {
        if (!zC_) k::transpose<float>(s, p_, n_, dC, CT_.get());
        if (!zD_ && m_ > 0) k::transpose<float>(s, p_, m_, dD, DT_.get());
    }
    if (uDouble)
        k::outputs<double, double>(s, n_, m_, p_, n_, 0, zC_, zD_, 3, CT_.get(), DT_.get(), dX0,
                                   (const double*)dU, 2, 0, dY);
    else
        k::outputs<double, float>(s, n_, m_, p_, n_, 0, zC_, zD_, 3, CT_.get(), DT_.get(), dX0,
                                  (const float*)dU, 2, 0, dY);
}

void ScanCore::abm_startup(hipStream_t s, const float* dA, const float* dB, const float* dx0,
                           const void* dUh, bool uDouble, double h, double* dX0)
{
    const bool mirror = opt_.precision == kPlain32;
    if (uDouble)
        k::abm_startup<double>(s, mirror, n_, m_, h, dA, dB, dx0, (const double*)dUh, dX0);
    else
        k::abm_startup<float>(s, mirror, n_, m_, h, dA, dB, dx0, (const float*)dUh, dX0);
}

// ---------------------------------------------------------------- one solve
void ScanCore::solve(hipStream_t s, long points, double h, const float* dA, const float* dB,
                     const float* dC, const float* dD, const float* dx0, const void* dU,
                     bool uDouble, float* dY, hipEvent_t* marks, const SideStream* side)
{
    const long nScan = method_ == kRK4 ? points : points - 3;
    const long nPad = next_pow2(nScan);
    if (nScan < 1 || nPad > nPadMax_) throw std::invalid_argument("ScanCore::solve: bad length");
    if (side && (!side->stream || !side->fork || !side->join))
        throw std::invalid_argument("ScanCore::solve: incomplete side stream");
    auto tick = [&](int i) {
        if (marks) PARODE_HIP(hipEventRecord(marks[i], s));
    };
    tick(0);
    // the power table and the ABM start-up (and for an FP64 ABM map, the
    // map's last launches): on the side stream when given
    build_map(s, dA, dB, h, side);
    tick(1);
    // the leaves read only what the map's launches on this stream made: they
    // are queued before the side stream's many launches, so this stream does
    // not wait for the host to queue those
    if (method_ == kRK4)
        leaves(s, dU, uDouble, 2, 0, 1, nScan, nPad);
    else
        leaves(s, dU, uDouble, 2, 0, 2, nScan, nPad);
    tick(2);
    hipStream_t ps = side ? side->stream : s;
    build_powers(ps, false);   // one batch: no carry, so no P_L
    // the tree levels wait for the tables only; the root for the seed, the
    // outputs for their tables
    const bool late = side && side->late;
    if (late) PARODE_HIP(hipEventRecord(side->join, ps));
    prepare_outputs(ps, dC, dD, uDouble);
    float* yScan = dY;
    if (method_ == kRK4) {
        k::to_f64<float>(ps, n_, dx0, seed());
    } else {
        abm_startup(ps, dA, dB, dx0, dU, uDouble, h, seed());
        startup_outputs(ps, dC, dD, seed(), dU, uDouble, dY, true);
        yScan = dY + 3L * nOut();
    }
    if (side) PARODE_HIP(hipEventRecord(late ? side->late : side->join, ps));
    sweep_up(s, nScan, nPad, side ? side->join : nullptr);
    tick(3);
    if (late) PARODE_HIP(hipStreamWaitEvent(s, side->late, 0));
    sweep_top(s, nScan, nPad, seed(), tot());
    tick(4);
    sweep_down(s, nScan, nPad);
    tick(5);
    outputs(s, dC, dD, dU, uDouble, 2, method_ == kRK4 ? 0 : 6, nScan, yScan, true);
    tick(6);
}

}  // namespace parode
