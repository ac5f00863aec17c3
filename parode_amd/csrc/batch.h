// BatchSolve: the mixed-precision scan of many independent systems in grouped
// launches.
//
// A per-system solve (ScanCore::solve) is a chain of 30-60 launches, most of
// them small for a small system.  A batch of hundreds of systems of uneven
// state size n and horizon N instead runs each stage of the scan once for all
// of them: the step maps' products, the FP64 power squarings, the level
// tables, the ABM start-up, the leaves, each chunk-recurrence pass (one
// launch per padded state size), each tree level (the systems' levels aligned
// from the top of their trees), the top of the tree, and the outputs.  Every
// system's work-groups run the same kernel bodies as its own solve would, on
// its own buffers (a ScanCore's, carved from the caller's arena), so its
// results do not depend on which systems share its launches.
//
// Scope: mixed precision (FP64 step map and powers, FP64 top six levels,
// FP32 below), inputs as FP64 samples on the h/2 grid, outputs y = Cx + Du in
// FP32.  eligible() says which shapes the grouped kernels cover; the others
// take the per-system path.
#pragma once
#include <memory>
#include <vector>

#include "hip_util.h"
#include "kernels.h"
#include "scan_core.h"

namespace parode {

// One system of a batch: shape, horizon and its device arrays.
struct BatchSystem {
    int method = 0, n = 0, m = 0, p = 0;
    long points = 0;
    double t0 = 0.0, h = 0.0;
    const float *A = nullptr, *B = nullptr, *C = nullptr, *D = nullptr, *x0 = nullptr;
    const double* coef = nullptr;   // m x 6 sinusoid-family coefficients, or null (U given)
    double sigma = 0.0;
    uint64_t seed = 0;
    double* U = nullptr;            // (2 points + 1) x m samples on the h/2 grid
    float* Y = nullptr;             // points x p outputs
};

// Streams beside the caller's over which the launches of one wide phase (the
// leaves, a chunk pass, the outputs: launches that never read each other's
// results) are spread, so that one launch's last work-groups overlap the next
// launch instead of leaving the GPU part idle.  Default priority.
struct FanStreams {
    static constexpr int kCount = 2;
    hipStream_t stream[kCount] = {};
    hipEvent_t start = nullptr, done[kCount] = {};
    void create();     // throws std::runtime_error on a HIP failure
    void destroy();
};

class BatchSolve {
public:
    static bool eligible(int method, int n, int m, int p, bool zC, bool zD, long points);
    // device bytes one system's scan buffers take from the arena
    static size_t footprint(int method, int n, int m, int p, long points);

    // Plans the grouped solve; every system's scan buffers come from arena
    // (std::bad_alloc when they do not fit: nothing is left allocated).
    BatchSolve(const std::vector<BatchSystem>& systems, Arena* arena);
    BatchSolve(const BatchSolve&) = delete;
    BatchSolve& operator=(const BatchSolve&) = delete;

    // bytes of the launch records (a device buffer and a host staging area)
    size_t descriptor_bytes() const { return blob_.size(); }
    // Copies the records through hostStage (pinned, descriptor_bytes()) to
    // devDesc (device) unless upload is false (they are there from an earlier
    // run), then queues every launch on s.  Both areas must stay untouched
    // until the work on s is done.  marks, when given, are kStages + 1 events
    // recorded before each stage's launches and after the last (every launch
    // then runs on s, one after the other).  Else, with side, the FP64 power
    // tables and the level tables run on side->stream (forked after the step
    // map, joined before the first tree level) beside the input samples, the
    // ABM start-up, the leaves and the first chunk pass; with fan, too, the
    // launches of the inputs, leaves, chunk-pass and output phases (and of the
    // top of the tree) are spread over s and the fan streams (each phase
    // forked from s and joined back to it).
    void run(hipStream_t s, void* hostStage, void* devDesc, bool upload = true,
             const hipEvent_t* marks = nullptr, const ScanCore::SideStream* side = nullptr,
             const FanStreams* fan = nullptr) const;
    int launches() const { return (int)refs_.size(); }

    // Stages of a grouped solve, in launch order (for timing).
    enum Stage {
        kStageInputs, kStageMap, kStagePowers, kStageTables, kStageStartup, kStageLeaves,
        kStageChunks1, kStageUp, kStageTop, kStageDown, kStageChunks2, kStageOutputs, kStages
    };
    static const char* stage_name(int stage);

private:
    struct Planned {
        std::unique_ptr<ScanCore> core;
        DevArray<float> copies;      // K-split copy halves when the core has none
    };
    std::vector<Planned> sys_;
    std::vector<k::g::Phase> phases_;
    std::vector<unsigned char> blob_;
    struct Ref {
        const k::g::Launch* L;
        size_t args, start;
        int stage;
        int role;    // 0 map, 1 side stream, 2 beside the side stream, 3 after the join
        int phase;
    };
    std::vector<Ref> refs_;
    static constexpr int kOutSets = 3;
    int outPhase_ = 0;   // the first of the outputs' kOutSets phases
    // the set (and fan stream) of a chunk kind's second pass
    static int chunk_lane(int kind) { return (kind - k::g::kChunksFac16) % kOutSets; }
};

}  // namespace parode
