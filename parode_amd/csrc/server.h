// Continuous-batching request loop for LTI solves on one GPU.
//
// Requests (one system each: method, shape, horizon, matrices, inputs) are
// submitted from any thread.  One worker thread owns the GPU work.  Whenever
// a lane (a HIP stream) has room, it admits the requests waiting at the head
// of the queue, in arrival order, as one group: with mixed precision, every
// request the grouped kernels cover (BatchSolve::eligible) joins the group
// (up to the free device memory), and the group's scans run as one set of
// grouped launches; any other request runs alone through its own ScanCore.
// A group's work is enqueued on its lane without host synchronisation, so
// groups on different lanes overlap on the device and new arrivals are
// admitted while earlier groups run.  A group's inputs go up in one copy
// (pinned staging -> device), its launch records in another; when its GPU
// work ends, its device memory returns to the arena at once; host outputs
// wait in pinned staging until the caller's wait() copies them out, device
// outputs are written in place.  Inputs are either host samples on the h/2
// grid or the coefficients of the family u_j = a sin(w t + phi) +
// b cos(w t + phi) + c + d t (+ noise), sampled in FP64 on the device.
#pragma once
#include <chrono>
#include <condition_variable>
#include <deque>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "hip_util.h"
#include "batch.h"
#include "scan_core.h"

namespace parode {

struct ServeRequest {
    int solver = 0, n = 0, m = 0, p = 0;
    bool zC = false, zD = false;
    long points = 0;
    double t0 = 0.0, h = 0.0;
    const float *A = nullptr, *B = nullptr, *C = nullptr, *D = nullptr, *x0 = nullptr;
    const double* coef = nullptr;     // m x 6, used when samples is null
    double sigma = 0.0;
    uint64_t seed = 0;
    const double* samples = nullptr;  // host, (2*points+1) x m
    float* Y = nullptr;               // points x nOut: host memory, or device memory when yDevice
    bool yDevice = false;
};

struct ServeStats {
    long submitted = 0, completed = 0, failed = 0, running = 0, queued = 0;
    long peakRunning = 0;
    size_t arenaBytes = 0, arenaPeak = 0;
    long groups = 0, groupedRequests = 0, largestGroup = 0;
};

class Server {
public:
    Server(int lanes, size_t deviceBytes, size_t pinnedBytes, int precision);
    ~Server();
    Server(const Server&) = delete;
    Server& operator=(const Server&) = delete;

    // Returns a ticket (> 0); copies the matrices and coefficients, so the
    // caller keeps only `samples` and `Y` alive until the ticket completes.
    long submit(const ServeRequest& r);
    // The same for count requests, queued together (tickets[i] for rs[i]).
    void submit(const ServeRequest* rs, size_t count, long* tickets);
    // Blocks until the ticket completes; returns its status (0 or an
    // ErrorCode) and the times from submission to completion and to admission.
    int wait(long ticket, double* latencyMs, double* queueMs);
    void drain();
    ServeStats stats();
    static size_t footprint(const ServeRequest& r, int precision);

private:
    struct Job;
    struct Group;
    struct Stage;
    using Clock = std::chrono::steady_clock;
    void loop();
    bool grouped(const Job& j) const;
    // the job of a request (its matrices copied); status InvalidSystem when bad
    std::shared_ptr<Job> make_job(const ServeRequest& r) const;
    // gives j its ticket and queues it (or fails it); mu_ held
    long enqueue(const std::shared_ptr<Job>& j);
    // enqueues the group's work on lanes_[lane]; false: no memory for it now
    bool start(Group& g);
    void finish(Group& g, int status);

    int device_ = 0, precision_ = 0, depth_ = 2;
    size_t groupCap_ = 1024;   // requests per group (halved while memory is short)
    std::vector<hipStream_t> lanes_;
    std::vector<ScanCore::SideStream> sides_;   // per lane: the grouped solves' side stream
    std::vector<FanStreams> fans_;              // per lane: the grouped solves' wide phases
    std::vector<int> load_;
    std::unique_ptr<Arena> arena_, pinned_;
    std::mutex mu_;
    std::condition_variable work_cv_, done_cv_;
    std::deque<std::shared_ptr<Job>> queue_;
    std::unordered_map<long, std::shared_ptr<Job>> jobs_;
    long next_ = 1, running_ = 0, staged_ = 0;
    bool stop_ = false;
    ServeStats st_;
    std::thread worker_;
};

}  // namespace parode
