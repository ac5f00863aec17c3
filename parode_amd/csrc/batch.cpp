#include "batch.h"

#include <algorithm>
#include <exception>
#include <thread>
#include <cstring>
#include <stdexcept>

#include "config.h"

namespace parode {

namespace {

int pad8(int c) { return c > 0 ? (c + 7) / 8 * 8 : 8; }
long pairs_at(long nReal, int l) { return (nReal + (2L << l) - 1) / (2L << l); }
long ks_half(long nPad, int na) { return std::min<long>(k::kKsMaxPairs, std::max(nPad / 2, 1L)) * na; }
constexpr long kMaxPoints = 1L << 20;   // the first tree level above the chunks has <= 4096 pairs

}  // namespace

bool BatchSolve::eligible(int method, int n, int m, int p, bool zC, bool zD, long points)
{
    if (method != kRK4 && method != kABM) return false;
    if (n < 1 || n > 256 || m < 1 || m > 256 || p < 1 || (zC && zD)) return false;
    if (points < (method == kABM ? 4 : 1) || points > kMaxPoints) return false;
    // with the chunk recurrences off (a kernel switch) level 0 is a tree level
    const long nScan = method == kRK4 ? points : points - 3;
    if ((level_kernel() & 9) != 0 && next_pow2(nScan) / 2 > k::kKsMaxPairs) return false;
    return (level_kernel() & 6) == 0;   // the vector-ALU switches take the per-system path
}

size_t BatchSolve::footprint(int method, int n, int m, int p, long points)
{
    ScanOptions opt;
    opt.precision = kMixed;
    const long nScan = method == kRK4 ? points : points - 3;
    const long nPad = next_pow2(nScan);
    const int na = method == kABM ? 4 * n : n;
    size_t b = ScanCore::estimate_footprint(method, n, m, nPad, opt, p);
    if (na < kKsMinStates) b += Arena::round(2 * ks_half(nPad, na) * sizeof(float));
    return b;
}

BatchSolve::BatchSolve(const std::vector<BatchSystem>& systems, Arena* arena)
{
    if (!arena) throw std::invalid_argument("BatchSolve: needs an arena");
    ScanOptions opt;
    opt.precision = kMixed;
    // stages, in order; the squarings and the tree levels get one phase per level
    enum { kSmp = 0, kIn, kR1, kR2, kR3, kComb, kPost, kPow0 };
    int maxL = 0, upSteps = 0;
    struct Shape {
        long nScan, nPad;
        int L, na, lc, lt;
    };
    std::vector<Shape> sh(systems.size());
    sys_.resize(systems.size());
    for (size_t i = 0; i < systems.size(); ++i) {
        const BatchSystem& q = systems[i];
        if (!eligible(q.method, q.n, q.m, q.p, !q.C, !q.D, q.points) || !(q.h > 0.0) || !q.A ||
            !q.B || !q.x0 || !q.U || !q.Y || (!q.coef && !q.U))
            throw std::invalid_argument("BatchSolve: system outside the grouped path");
        Shape& s = sh[i];
        s.nScan = q.method == kRK4 ? q.points : q.points - 3;
        s.nPad = next_pow2(s.nScan);
        s.L = ilog2(s.nPad);
        s.na = q.method == kABM ? 4 * q.n : q.n;
        sys_[i].core.reset(new ScanCore(q.method, q.n, q.m, q.p, !q.C, !q.D, s.nPad, opt, arena));
        ScanCore& c = *sys_[i].core;
        if (!c.Sks32_.get()) sys_[i].copies.alloc(2 * ks_half(s.nPad, s.na), arena);
        s.lc = c.chunk_log(s.nScan, s.nPad, true);
        s.lt = c.top_level(s.nScan, s.nPad);
        for (int l = s.lc; l < s.lt; ++l)
            if (pairs_at(s.nScan, l) > k::kKsMaxPairs)
                throw std::invalid_argument("BatchSolve: a tree level with too many pairs");
        maxL = std::max(maxL, s.L);
        upSteps = std::max(upSteps, s.lt - s.lc);
    }
    const int kTables = kPow0 + 4 * maxL, kStart = kTables + 1, kLeaves = kStart + 1,
              kChunk1 = kLeaves + 1, kUp0 = kChunk1 + 1, kTop = kUp0 + upSteps,
              kGather = kTop + 1, kDown0 = kGather + 1, kChunk2 = kDown0 + upSteps,
              kOut = kChunk2 + 1;
    // the outputs in three sets, by the stream that runs the system's second
    // chunk pass (kinds in turn), so that each stream's outputs start when its
    // own chunk launches end (sets fold onto fewer streams when fewer run)
    phases_.resize(kOut + kOutSets);
    outPhase_ = kOut;
    auto stage_of = [&](int ph) {
        if (ph <= kIn) return kStageInputs;
        if (ph <= kPost) return kStageMap;
        if (ph < kTables) return kStagePowers;
        if (ph == kTables) return kStageTables;
        if (ph == kStart) return kStageStartup;
        if (ph == kLeaves) return kStageLeaves;
        if (ph == kChunk1) return kStageChunks1;
        if (ph < kTop) return kStageUp;
        if (ph == kTop) return kStageTop;
        if (ph < kChunk2) return kStageDown;
        if (ph == kChunk2) return kStageChunks2;
        return kStageOutputs;
    };
    // one system's launches into the phase set P (systems are planned on a few
    // threads, each into a phase set of its own, merged in system order)
    auto plan = [&](size_t i, std::vector<k::g::Phase>& P) {
        const BatchSystem& q = systems[i];
        const Shape& s = sh[i];
        ScanCore& c = *sys_[i].core;
        const int n = q.n, m = q.m, p = q.p, na = s.na, L = s.L;
        const bool abm = q.method == kABM;
        const size_t nn = (size_t)n * n, nm = (size_t)n * m, na2 = (size_t)na * na;
        double* A = c.A64_.get();
        double* B = c.B64_.get();
        double* A2 = c.prod64_.get();
        double* A3 = A2 + nn;
        double* A4 = A3 + nn;
        double* AB = A4 + nn;
        double* A2B = AB + nm;
        double* A3B = A2B + nm;
        double* PT = c.PT64_.get();
        float* X = c.X32_.get();
        const int pp = pad8(p);
        // inputs, FP64 copies, output tables (RungeKutta4thOrder.cpp:36-70,
        // AdamsMoulton.cpp:238-316 take A and B; SystemOutput.cl:13-60 C and D)
        k::g::Phase& in = P[kIn];
        if (q.coef)
            k::g::add_sample(P[kSmp], m, q.coef, q.t0, 0.5 * q.h, 2 * q.points + 1, q.sigma,
                             q.seed, q.U);
        k::g::add_cast(in, (long)nn, q.A, A);
        k::g::add_cast(in, (long)nm, q.B, B);
        if (abm)
            k::g::add_fill(in, (long)na2, 0.0, PT);   // the companion's shift rows and zeros
        else
            k::g::add_cast(in, n, q.x0, c.seed());
        if (q.C) k::g::add_transpose_pad(in, n, p, pp, q.C, c.CT64_.get());
        if (q.D) k::g::add_transpose_pad(in, m, p, pp, q.D, c.DT64_.get());
        // products of A, in the per-system order
        k::g::add_gemm(P[kR1], n, n, n, A, A, A2);
        k::g::add_gemm(P[kR1], n, n, m, A, B, AB);
        if (!abm) {
            k::g::add_gemm(P[kR2], n, n, n, A2, A, A3);
            k::g::add_gemm(P[kR2], n, n, m, A2, B, A2B);
            k::g::add_gemm(P[kR3], n, n, n, A, A3, A4);
            k::g::add_gemm(P[kR3], n, n, m, A3, B, A3B);
            k::g::add_rk4_combine(P[kComb], n, m, q.h, A, A2, A3, A4, B, AB, A2B, A3B, PT,
                                  c.Bbar64_.get());
            k::g::add_transpose(P[kPost], n, 3 * m, c.Bbar64_.get(), c.BbT64_.get());
        } else {
            k::g::add_abm_combine(P[kComb], n, m, q.h, A, A2, B, AB, PT, c.Bbar64_.get());
            k::g::add_transpose(P[kPost], n, m, B, c.BfT64_.get());
            k::g::add_transpose(P[kPost], n, m, AB, c.BfT64_.get() + nm);
            // B-bar itself for the vector-ALU leaves when the factored ones do not fit
            if (!k::g::leaves_fit(n, m, 5, 2, 2, 2))
                k::g::add_transpose(P[kPost], n, 5 * m, c.Bbar64_.get(), c.BbT64_.get());
        }
        // the FP64 power table P_{l+1} = P_l^2 (transposed), up to P_{L-1}: the
        // top of the tree reads levels below L, and one batch has no carry
        if (!k::g::add_powers_small(P[kPow0], na, L, PT)) {
            for (int l = 0; l + 1 < L; ++l) {
                const double* Q = PT + l * na2;
                double* Q2 = PT + (l + 1) * na2;
                if (!abm) {
                    k::g::add_gemm(P[kPow0 + 4 * l], na, na, na, Q, Q, Q2);
                    continue;
                }
                // ABM: block row i of M^j is the last block row R_{j-3+i} of
                // M^{j-3+i}, and R_{i+1} = R_i M.  Transposed, block column 0 of
                // Q^2 (Q = P_l^T) is Q times Q's block column 0, and block column
                // b is Q_0[:, 3n:4n] times the last n rows of block column b - 1,
                // plus that column shifted down by n rows: 56 n^3 flop instead of
                // the 128 n^3 of the full square.
                k::g::add_gemm_ld(P[kPow0 + 4 * l], na, na, n, Q, na, Q, na, Q2, na,
                                  nullptr, 0);
                for (int b = 1; b < 4; ++b)
                    k::g::add_gemm_ld(P[kPow0 + 4 * l + b], na, n, n, PT + 3 * n, na,
                                      Q2 + 3L * n * na + (b - 1) * n, na, Q2 + b * n, na,
                                      Q2 + (b - 1) * n, n);
            }
        }
        int n32 = 0, nd = 0;
        for (int l = 0; l < L; ++l) {
            if (!c.level_is_f64(l)) n32 = l + 1;
            if (c.level_is_delta(l)) nd = l + 1;
        }
        // (levels below lc are read by no kernel here: the chunk passes take the map)
        k::g::add_tables(P[kTables], na, c.nap_, c.nrp_, n32, nd, PT, c.PT32_.get(),
                         c.PTm_.get(), sh[i].lc);
        // the ABM start-up (AdamsMoulton.cpp:29-104) and its three outputs
        float* yScan = q.Y;
        if (abm) {
            k::g::add_startup(P[kStart], n, m, q.h, q.A, q.B, q.x0, q.U, c.seed());
            k::g::add_start_outputs(P[kOut], n, m, p, pp, q.C != nullptr,
                                    q.D != nullptr, c.CT64_.get(), c.DT64_.get(), c.seed(), q.U,
                                    q.Y);
            yScan = q.Y + 3L * p;
        }
        // leaves (BTimesURK4thOrder.cl / BTimesUAdamsMoulton.cl)
        // (matrix cores: [B | AB] factored for ABM, B-bar for RK4; else the vector ALUs)
        // Only the leaves up to the end of the last chunk holding a real element
        // (of the last level-0 pair without chunks) are read: the chunk passes
        // skip chunks past nScan and the tree levels pairs past it, so the
        // padding's leaves beyond are neither formed nor zeroed.  The tree's
        // aggregates over padding then hold stale values; no real element's
        // prefix reads them (a prefix takes only left-sibling aggregates), and
        // the total (tot) is not used by a grouped solve, which has no carry.
        const long nLeaf = std::min(s.nPad, (s.nScan + (1L << std::max(s.lc, 1)) - 1) >>
                                                std::max(s.lc, 1) << std::max(s.lc, 1));
        bool ok;
        if (abm) {
            k::LeafComb G;
            k::abm_leaf_factors(q.h, G);
            ok = k::g::leaves_fit(n, m, 5, 2, 2, 2)
                     ? k::g::add_leaves(P[kLeaves], n, m, 5, 2, na, 3 * n, s.nScan, nLeaf,
                                        c.BfT64_.get(), G, q.U, 2, 0, 2, X, s.lc == 0)
                     : k::g::add_leaves_simt(P[kLeaves], n, m, 5, na, 3 * n, s.nScan,
                                             nLeaf, c.BbT64_.get(), q.U, 2, 0, 2, X, s.lc == 0);
        } else {
            ok = k::g::leaves_fit(n, m, 3, 0, 2, 1)
                     ? k::g::add_leaves(P[kLeaves], n, m, 3, 0, na, 0, s.nScan, nLeaf,
                                        c.BbT64_.get(), k::LeafComb{}, q.U, 2, 0, 1, X, true)
                     : k::g::add_leaves_simt(P[kLeaves], n, m, 3, na, 0, s.nScan, nLeaf,
                                             c.BbT64_.get(), q.U, 2, 0, 1, X, true);
        }
        if (!ok) throw std::invalid_argument("BatchSolve: leaves do not fit the kernel");
        // chunk recurrences (levels below lc)
        if (s.lc > 0) {
            for (int pass = 0; pass < 2; ++pass) {
                k::g::Phase& ph = P[pass ? kChunk2 : kChunk1];
                ok = abm ? k::g::add_abm_chunks(ph, pass == 1, n, s.lc, s.nScan, s.nPad, q.h, A,
                                                A2, X)
                         : k::g::add_dense_chunks(ph, pass == 1, na, s.lc, s.nScan, s.nPad, PT,
                                                  X);
                if (!ok) throw std::invalid_argument("BatchSolve: chunks do not fit the kernel");
            }
        }
        // tree levels [lc, lt), aligned so that every system's top level runs last
        auto table = [&](int l, bool& f64, int& ldp) -> const void* {
            f64 = c.level_is_f64(l) || !c.PTm_.get();
            ldp = f64 ? na : c.nap_;
            return f64 ? (const void*)(PT + l * na2)
                       : (const void*)(c.PTm_.get() + (size_t)l * c.nap_ * c.nrp_);
        };
        for (int l = s.lc; l < s.lt; ++l) {
            bool f64;
            int ldp;
            const void* T = table(l, f64, ldp);
            k::g::add_level(P[kUp0 + upSteps - (s.lt - l)], false, f64,
                            !f64 && c.level_is_delta(l), na, ldp, 1L << l, pairs_at(s.nScan, l),
                            T, nullptr, X, nullptr, 0);
        }
        // the root with the seed (I, x0) (GPUFixedStepIterative.cpp:294-337)
        if (s.lt < L) {
            int l64 = L;
            while (l64 > 0 && c.level_is_f64(l64 - 1)) --l64;
            k::g::add_top(P[kTop], na, s.lt, L, l64, s.nScan, c.PT32_.get(), PT, X,
                          c.seed(), c.tot());
        } else {
            k::g::add_root(P[kTop], na, s.nPad - 1, c.seed(), X, c.tot());
        }
        // down levels lt-1 .. lc: each writes the next one's copy of X[R]
        float* Sk = c.Sks32_.get() ? c.Sks32_.get() : sys_[i].copies.get();
        const size_t half = c.Sks32_.get() ? c.ksHalf_ : (size_t)ks_half(s.nPad, na);
        if (s.lt - 1 >= s.lc)
            k::g::add_gather(P[kGather], na, 1L << (s.lt - 1), pairs_at(s.nScan, s.lt - 1),
                             X, Sk);
        for (int l = s.lt - 1, j = 0; l >= s.lc; --l, ++j) {
            bool f64;
            int ldp;
            const void* T = table(l, f64, ldp);
            float* S = Sk + (size_t)(j & 1) * half;
            float* Sn = l - 1 >= s.lc ? Sk + (size_t)((j + 1) & 1) * half : nullptr;
            k::g::add_level(P[kDown0 + j], true, f64, !f64 && c.level_is_delta(l), na, ldp,
                            1L << l, pairs_at(s.nScan, l), T, S, X, Sn,
                            Sn ? pairs_at(s.nScan, l - 1) : 0);
        }
        // outputs (SystemOutput.cl:13-60; strides RungeKutta4thOrder.cpp:209,
        // AdamsMoulton.cpp:398)
        const int ck = s.lc > 0 ? k::g::chunks_kind(abm, abm ? n : na) : -1;
        const int set = ck >= 0 ? chunk_lane(ck) : 0;
        if (!k::g::add_outputs(P[kOut + set], n, m, p, pp, na, abm ? 3 : 0, q.C != nullptr,
                               q.D != nullptr, s.nScan, c.CT64_.get(), c.DT64_.get(), X, q.U, 2,
                               abm ? 6 : 0, yScan))
            throw std::invalid_argument("BatchSolve: outputs do not fit the kernel");
    };
    const int parts = (int)std::min<size_t>(8, std::max<size_t>(1, systems.size() / 64));
    if (parts == 1) {
        for (size_t i = 0; i < systems.size(); ++i) plan(i, phases_);
    } else {
        std::vector<std::vector<k::g::Phase>> part(parts, std::vector<k::g::Phase>(phases_.size()));
        std::vector<std::exception_ptr> err(parts);
        std::vector<std::thread> th;
        const size_t per = (systems.size() + parts - 1) / parts;
        for (int t = 0; t < parts; ++t)
            th.emplace_back([&, t] {
                try {
                    for (size_t i = t * per; i < std::min(systems.size(), (t + 1) * per); ++i)
                        plan(i, part[t]);
                } catch (...) {
                    err[t] = std::current_exception();
                }
            });
        for (auto& x : th) x.join();
        for (auto& e : err)
            if (e) std::rethrow_exception(e);
        for (int t = 0; t < parts; ++t)
            for (size_t pi = 0; pi < phases_.size(); ++pi)
                for (auto& kv : part[t][pi]) k::g::merge_launch(phases_[pi][kv.first], std::move(kv.second));
    }

    // pack every launch's records and block table
    auto align = [](size_t v) { return (v + 15) / 16 * 16; };
    size_t off = 0;
    for (size_t pi = 0; pi < phases_.size(); ++pi)
        for (auto& kv : phases_[pi]) {
            const k::g::Launch& Lr = kv.second;
            if (Lr.nsys() <= 0 || Lr.blocks() == 0) continue;
            // streams: the FP64 copies and the map on the main stream; the
            // powers and level tables on the side stream, beside the input
            // samples, the ABM start-up, the leaves and the first chunk pass;
            // the rest after both (roles: 0 copies and map, 5 samples, 1
            // powers and tables, 4 start-up, 2 leaves and chunks, 3 the rest)
            const int pos = (int)pi;
            const int role = pos == kSmp     ? 5
                             : pos <= kPost  ? 0
                             : pos < kStart  ? 1
                             : pos == kStart ? 4
                             : pos <= kChunk1 ? 2
                                              : 3;
            Ref r{&Lr, align(off), 0, stage_of(pos), role, pos};
            r.start = align(r.args + Lr.args.size());
            off = r.start + Lr.start.size() * sizeof(unsigned);
            refs_.push_back(r);
        }
    blob_.assign(align(off), 0);
    // each launch's records with the longest-running work-groups first
    for (const Ref& r : refs_)
        k::g::pack_by_cost(*r.L, blob_.data() + r.args,
                           reinterpret_cast<unsigned*>(blob_.data() + r.start));
}

const char* BatchSolve::stage_name(int stage)
{
    static const char* names[kStages] = {"inputs", "step_map", "powers", "tables", "startup",
                                         "leaves", "chunks_pass1", "up_levels", "top",
                                         "down_levels", "chunks_pass2", "outputs"};
    return stage >= 0 && stage < kStages ? names[stage] : "";
}

// work-groups a phase needs before its launches are spread over fan streams
constexpr long kFanMinBlocks = 4096;

void FanStreams::create()
{
    for (int k = 0; k < kCount; ++k) {
        PARODE_HIP(hipStreamCreateWithFlags(&stream[k], hipStreamNonBlocking));
        PARODE_HIP(hipEventCreateWithFlags(&done[k], hipEventDisableTiming));
    }
    PARODE_HIP(hipEventCreateWithFlags(&start, hipEventDisableTiming));
}

void FanStreams::destroy()
{
    for (int k = 0; k < kCount; ++k) {
        if (done[k]) (void)hipEventDestroy(done[k]);
        if (stream[k]) (void)hipStreamDestroy(stream[k]);
        done[k] = nullptr;
        stream[k] = nullptr;
    }
    if (start) (void)hipEventDestroy(start);
    start = nullptr;
}

void BatchSolve::run(hipStream_t s, void* hostStage, void* devDesc, bool upload,
                     const hipEvent_t* marks, const ScanCore::SideStream* side,
                     const FanStreams* fan) const
{
    if (upload && !blob_.empty()) {
        std::memcpy(hostStage, blob_.data(), blob_.size());
        PARODE_HIP(hipMemcpyAsync(devDesc, hostStage, blob_.size(), hipMemcpyHostToDevice, s));
    }
    const char* base = static_cast<const char*>(devDesc);
    auto launch = [&](hipStream_t st, const Ref& r) {
        k::g::run(st, *r.L, base + r.args, reinterpret_cast<const unsigned*>(base + r.start));
    };
    auto wide = [](int stage) {
        return stage == kStageInputs || stage == kStageLeaves || stage == kStageChunks1 ||
               stage == kStageChunks2 || stage == kStageOutputs;
    };
    // the launches of one role in order; the launches of one wide phase spread
    // over st and f's streams
    auto queue = [&](hipStream_t st, int role, const FanStreams* f) {
        for (size_t i = 0; i < refs_.size();) {
            if (refs_[i].role != role) {
                ++i;
                continue;
            }
            if (refs_[i].stage == kStageChunks2 || refs_[i].stage == kStageOutputs) {
                // the second chunk pass and the outputs: a stream runs its
                // chunk kinds, then the outputs of their systems (set w of
                // the outputs follows the kinds of lane w), joined at the end
                size_t e = i;
                long blocks = 0;
                while (e < refs_.size() && (refs_[e].stage == kStageChunks2 ||
                                            refs_[e].stage == kStageOutputs)) {
                    blocks += refs_[e].L->blocks();
                    ++e;
                }
                const int used = (f && e - i > 1 && blocks >= kFanMinBlocks)
                                     ? FanStreams::kCount : 0;
                if (used > 0) {
                    PARODE_HIP(hipEventRecord(f->start, st));
                    for (int k = 0; k < used; ++k)
                        PARODE_HIP(hipStreamWaitEvent(f->stream[k], f->start, 0));
                }
                for (size_t j = i; j < e; ++j) {
                    const Ref& r = refs_[j];
                    const int lane = r.stage == kStageChunks2 ? chunk_lane(r.L->kind)
                                                              : r.phase - outPhase_;
                    const int w = lane % (used + 1);
                    launch(w == 0 ? st : f->stream[w - 1], r);
                }
                for (int k = 0; k < used; ++k) {
                    PARODE_HIP(hipEventRecord(f->done[k], f->stream[k]));
                    PARODE_HIP(hipStreamWaitEvent(st, f->done[k], 0));
                }
                i = e;
                continue;
            }
            size_t e = i + 1;
            while (e < refs_.size() && refs_[e].phase == refs_[i].phase) ++e;
            const int count = (int)(e - i);
            // spread only phases wide enough to fill the GPU several times over
            // (a small group's fork and joins would cost more than its tails)
            long blocks = 0;
            for (size_t j = i; j < e; ++j) blocks += refs_[j].L->blocks();
            // (the top's launches, one work-group per system, are latency-bound:
            // spread whatever their size)
            const bool top = refs_[i].stage == kStageTop;
            const int used = (f && ((wide(refs_[i].stage) && blocks >= kFanMinBlocks) || top))
                                 ? std::min(count - 1, FanStreams::kCount) : 0;
            if (used > 0) {
                PARODE_HIP(hipEventRecord(f->start, st));
                for (int k = 0; k < used; ++k)
                    PARODE_HIP(hipStreamWaitEvent(f->stream[k], f->start, 0));
            }
            for (size_t j = i; j < e; ++j) {
                const int w = (int)((j - i) % (size_t)(used + 1));
                launch(w == 0 ? st : f->stream[w - 1], refs_[j]);
            }
            for (int k = 0; k < used; ++k) {
                PARODE_HIP(hipEventRecord(f->done[k], f->stream[k]));
                PARODE_HIP(hipStreamWaitEvent(st, f->done[k], 0));
            }
            i = e;
        }
    };
    if (side && !marks) {
        if (!side->stream || !side->fork || !side->join)
            throw std::invalid_argument("BatchSolve::run: incomplete side stream");
        // the power tables start right after the map; the ABM start-up (a few
        // small launches that read the input samples) runs on this stream, which
        // reaches the join first
        queue(s, 0, fan);
        PARODE_HIP(hipEventRecord(side->fork, s));
        PARODE_HIP(hipStreamWaitEvent(side->stream, side->fork, 0));
        queue(side->stream, 1, nullptr);
        PARODE_HIP(hipEventRecord(side->join, side->stream));
        queue(s, 5, fan);
        queue(s, 4, fan);
        queue(s, 2, fan);
        PARODE_HIP(hipStreamWaitEvent(s, side->join, 0));
        queue(s, 3, fan);
        return;
    }
    if (!marks) {
        for (int role : {0, 5, 1, 4, 2, 3}) queue(s, role, fan);
        return;
    }
    int stage = 0;
    PARODE_HIP(hipEventRecord(marks[0], s));
    for (const Ref& r : refs_) {
        while (stage < r.stage) PARODE_HIP(hipEventRecord(marks[++stage], s));
        launch(s, r);
    }
    while (stage < kStages) PARODE_HIP(hipEventRecord(marks[++stage], s));
}

}  // namespace parode

// ------------------------------------------------------------------ C ABI
#include <ParODEEnums.h>
#include <ParODEExt.h>

namespace {

// A device-resident group: its own arena for the scan buffers, the launch
// records on the device and their pinned staging area.
struct Group {
    int device = 0;
    std::unique_ptr<parode::Arena> arena;
    std::unique_ptr<parode::BatchSolve> batch;
    void* desc = nullptr;
    void* stage = nullptr;
    bool uploaded = false;
    hipEvent_t marks[parode::BatchSolve::kStages + 1] = {};
    parode::ScanCore::SideStream side;
    parode::FanStreams fan;
    ~Group()
    {
        batch.reset();
        arena.reset();
        fan.destroy();
        if (side.fork) (void)hipEventDestroy(side.fork);
        if (side.join) (void)hipEventDestroy(side.join);
        if (side.stream) (void)hipStreamDestroy(side.stream);
        for (hipEvent_t e : marks)
            if (e) (void)hipEventDestroy(e);
        if (desc) (void)hipFree(desc);
        if (stage) (void)hipHostFree(stage);
    }
};

}  // namespace

extern "C" {

int ParODE_GroupCreate(int count, const int* solver, const int* nStates, const int* nInputs,
                       const int* nOutputs, const long* points, const double* tStart,
                       const double* tStep, const float* const* dA, const float* const* dB,
                       const float* const* dC, const float* const* dD, const float* const* dx0,
                       const double* const* dCoef, const double* sigma,
                       const unsigned long long* seed, double* const* dU, float* const* dY,
                       void** group)
{
    if (!group) return ParODE_InvalidSystem;
    *group = nullptr;
    if (count < 1 || !solver || !nStates || !nInputs || !nOutputs || !points || !tStep || !dA ||
        !dB || !dx0 || !dU || !dY)
        return ParODE_InvalidSystem;
    std::vector<parode::BatchSystem> sys(count);
    size_t bytes = 0;
    for (int i = 0; i < count; ++i) {
        parode::BatchSystem& q = sys[i];
        q.method = solver[i];
        q.n = nStates[i];
        q.m = nInputs[i];
        q.p = nOutputs[i];
        q.points = points[i];
        q.t0 = tStart ? tStart[i] : 0.0;
        q.h = tStep[i];
        q.A = dA[i];
        q.B = dB[i];
        q.C = dC ? dC[i] : nullptr;
        q.D = dD ? dD[i] : nullptr;
        q.x0 = dx0[i];
        q.coef = dCoef ? dCoef[i] : nullptr;
        q.sigma = sigma ? sigma[i] : 0.0;
        q.seed = seed ? seed[i] : 0;
        q.U = dU[i];
        q.Y = dY[i];
        if (q.method != 0 && q.method != 1) return ParODE_InvalidSolverType;
        if (!parode::BatchSolve::eligible(q.method, q.n, q.m, q.p, !q.C, !q.D, q.points) ||
            !(q.h > 0.0) || !q.A || !q.B || !q.x0 || !q.U || !q.Y)
            return ParODE_InvalidSystem;
        bytes += parode::BatchSolve::footprint(q.method, q.n, q.m, q.p, q.points) +
                 64 * parode::Arena::kAlign;
    }
    try {
        std::unique_ptr<Group> g(new Group);
        PARODE_HIP(hipGetDevice(&g->device));
        g->arena.reset(new parode::Arena(bytes, parode::Arena::kDevice));
        g->batch.reset(new parode::BatchSolve(sys, g->arena.get()));
        const size_t db = g->batch->descriptor_bytes();
        if (db) {
            PARODE_HIP(hipMalloc(&g->desc, db));
            PARODE_HIP(hipHostMalloc(&g->stage, db, hipHostMallocDefault));
        }
        *group = g.release();
        return ParODE_OK;
    } catch (std::exception& e) {
        parode::last_error() = e.what();
        return ParODE_NotEnoughResources;
    }
}

int ParODE_GroupExecute(void* group, void* stream)
{
    Group* g = static_cast<Group*>(group);
    if (!g) return ParODE_InvalidSystem;
    try {
        PARODE_HIP(hipSetDevice(g->device));
        if (!g->side.stream) {
            PARODE_HIP(hipStreamCreate(&g->side.stream));
            PARODE_HIP(hipEventCreateWithFlags(&g->side.fork, hipEventDisableTiming));
            PARODE_HIP(hipEventCreateWithFlags(&g->side.join, hipEventDisableTiming));
        }
        if (!g->fan.start) g->fan.create();
        g->batch->run(static_cast<hipStream_t>(stream), g->stage, g->desc, !g->uploaded, nullptr,
                      &g->side, &g->fan);
        g->uploaded = true;
        return ParODE_OK;
    } catch (std::exception& e) {
        parode::last_error() = e.what();
        return ParODE_NotEnoughResources;
    }
}

int ParODE_GroupExecuteTimed(void* group, void* stream, float* stageMs)
{
    Group* g = static_cast<Group*>(group);
    if (!g || !stageMs) return ParODE_InvalidSystem;
    try {
        PARODE_HIP(hipSetDevice(g->device));
        for (hipEvent_t& e : g->marks)
            if (!e) PARODE_HIP(hipEventCreate(&e));
        g->batch->run(static_cast<hipStream_t>(stream), g->stage, g->desc, !g->uploaded,
                      g->marks);
        g->uploaded = true;
        PARODE_HIP(hipEventSynchronize(g->marks[parode::BatchSolve::kStages]));
        for (int i = 0; i < parode::BatchSolve::kStages; ++i)
            PARODE_HIP(hipEventElapsedTime(&stageMs[i], g->marks[i], g->marks[i + 1]));
        return ParODE_OK;
    } catch (std::exception& e) {
        parode::last_error() = e.what();
        return ParODE_NotEnoughResources;
    }
}

int ParODE_GroupStages(void) { return parode::BatchSolve::kStages; }

const char* ParODE_GroupStageName(int stage) { return parode::BatchSolve::stage_name(stage); }

int ParODE_GroupLaunches(void* group)
{
    Group* g = static_cast<Group*>(group);
    return g ? g->batch->launches() : -1;
}

long ParODE_GroupFootprint(void* group)
{
    Group* g = static_cast<Group*>(group);
    return g ? (long)(g->arena->in_use() + g->batch->descriptor_bytes()) : -1;
}

long ParODE_GroupEstimateBytes(int solver, int nStates, int nInputs, int nOutputs, long points)
{
    if (!parode::BatchSolve::eligible(solver, nStates, nInputs, nOutputs, false, false, points))
        return -1;
    return (long)parode::BatchSolve::footprint(solver, nStates, nInputs, nOutputs, points);
}

int ParODE_GroupDestroy(void* group)
{
    Group* g = static_cast<Group*>(group);
    if (!g) return ParODE_InvalidSystem;
    (void)hipSetDevice(g->device);
    (void)hipDeviceSynchronize();
    delete g;
    return ParODE_OK;
}

}  // extern "C"
