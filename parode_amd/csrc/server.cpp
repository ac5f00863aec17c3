#include "server.h"

#include <ParODEEnums.h>

#include <algorithm>
#include <exception>
#include <cstring>

#include "batch.h"
#include "config.h"
#include "kernels.h"

namespace parode {

namespace {

size_t up(size_t b) { return Arena::round(b); }

// Byte offsets of one request's uploaded arrays (relative to its place in the
// group's upload region, the same in pinned staging and on the device):
// [A B C D x0 | coef | U when the samples come from the host].
struct Layout {
    size_t A, B, C, D, x0, coef, U, up;
    size_t uBytes, yBytes;
    int nOut;
};

Layout layout(const ServeRequest& r)
{
    Layout L{};
    const size_t n = r.n, m = r.m, p = r.p;
    L.nOut = (r.zC && r.zD) ? r.n : r.p;
    size_t o = 0;
    L.A = o;  o += n * n * 4;
    L.B = o;  o += n * m * 4;
    L.C = o;  o += r.zC ? 0 : p * n * 4;
    L.D = o;  o += r.zD ? 0 : p * m * 4;
    L.x0 = o; o += n * 4;
    o = up(o);
    L.coef = o; o += r.samples ? 0 : m * 6 * 8;
    o = up(o);
    L.uBytes = up((size_t)(2 * r.points + 1) * m * 8);
    L.U = o;  o += r.samples ? L.uBytes : 0;
    L.up = up(o);
    L.yBytes = up((size_t)r.points * L.nOut * 4);
    return L;
}

// Runs fn(0) .. fn(parts - 1), parts - 1 of them on helper threads (inline
// when a thread cannot be started); fn must not throw.
template <typename F>
void run_parts(size_t parts, F&& fn)
{
    std::vector<std::thread> helpers;
    size_t p = 1;
    try {
        for (; p < parts; ++p) helpers.emplace_back(fn, p);
    } catch (...) {
        for (; p < parts; ++p) fn(p);
    }
    fn(0);
    for (auto& t : helpers) t.join();
}

// Throws unless [off, off + bytes) lies inside a block of blockBytes.
void check_range(const char* what, size_t off, size_t bytes, size_t blockBytes)
{
    if (off > blockBytes || bytes > blockBytes - off)
        throw std::logic_error(std::string("Server: ") + what + " outside its block");
}

}  // namespace

// A pinned staging block shared by the requests of one group: it returns to
// the pinned arena when the last of them has handed over its outputs.
struct Server::Stage {
    Arena* arena = nullptr;
    char* p = nullptr;
    size_t bytes = 0;
    ~Stage()
    {
        if (arena) arena->give(p);
    }
};

struct Server::Job {
    long id = 0;
    ServeRequest r;
    std::vector<float> mats;     // A B C D x0 at their Layout offsets
    std::vector<double> coef;
    Layout L{};
    std::shared_ptr<Stage> stage;
    size_t stY = 0;              // its outputs in the stage (host outputs)
    Clock::time_point tSubmit, tAdmit, tDone;
    int status = 1;              // 1 = pending
};

// The unit of GPU work: requests admitted together onto one lane.
struct Server::Group {
    std::vector<std::shared_ptr<Job>> jobs;
    bool grouped = false;
    int lane = -1;
    std::unique_ptr<BatchSolve> batch;   // grouped requests
    std::unique_ptr<ScanCore> core;      // one request of any shape
    char* io = nullptr;                  // device: uploads, samples, host-output staging
    char* desc = nullptr;                // device: the batch's launch records
    std::shared_ptr<Stage> descStage;    // pinned: the launch records on their way up
    hipEvent_t done = nullptr;
    void release(Arena* arena)
    {
        if (done) (void)hipEventDestroy(done);
        done = nullptr;
        batch.reset();
        core.reset();
        descStage.reset();
        arena->give(io);
        arena->give(desc);
        io = desc = nullptr;
    }
};

size_t Server::footprint(const ServeRequest& r, int precision)
{
    const Layout L = layout(r);
    size_t io = L.up + (r.samples ? 0 : L.uBytes) + (r.yDevice ? 0 : L.yBytes);
    if (precision == kMixed &&
        BatchSolve::eligible(r.solver, r.n, r.m, r.p, r.zC, r.zD, r.points))
        return BatchSolve::footprint(r.solver, r.n, r.m, r.p, r.points) + io;
    ScanOptions opt;
    opt.precision = precision;
    const long nScan = r.solver == kRK4 ? r.points : r.points - 3;
    return ScanCore::estimate_footprint(r.solver, r.n, r.m, nScan, opt, r.p) + io;
}

Server::Server(int lanes, size_t deviceBytes, size_t pinnedBytes, int precision)
    : precision_(precision)
{
    if (lanes < 1 || lanes > 16) throw std::invalid_argument("Server: 1..16 lanes");
    PARODE_HIP(hipGetDevice(&device_));
    arena_.reset(new Arena(deviceBytes, Arena::kDevice));
    pinned_.reset(new Arena(pinnedBytes, Arena::kPinnedHost));
    lanes_.resize(lanes);
    load_.assign(lanes, 0);
    sides_.resize(lanes);
    fans_.resize(lanes);
    // a lane's streams are made one after another: the runtime hands the
    // device's hardware queues to new streams in turn, so a lane's main, side
    // and fan streams run on different queues
    for (int l = 0; l < lanes; ++l) {
        PARODE_HIP(hipStreamCreateWithFlags(&lanes_[l], hipStreamNonBlocking));
        ScanCore::SideStream& sd = sides_[l];
        PARODE_HIP(hipStreamCreateWithFlags(&sd.stream, hipStreamNonBlocking));
        PARODE_HIP(hipEventCreateWithFlags(&sd.fork, hipEventDisableTiming));
        PARODE_HIP(hipEventCreateWithFlags(&sd.join, hipEventDisableTiming));
        fans_[l].create();
    }
    st_.arenaBytes = arena_->capacity();
    worker_ = std::thread([this] { loop(); });
}

Server::~Server()
{
    {
        std::lock_guard<std::mutex> g(mu_);
        stop_ = true;
    }
    work_cv_.notify_all();
    if (worker_.joinable()) worker_.join();
    for (auto& s : lanes_) (void)hipStreamDestroy(s);
    for (auto& sd : sides_) {
        if (sd.fork) (void)hipEventDestroy(sd.fork);
        if (sd.join) (void)hipEventDestroy(sd.join);
        if (sd.stream) (void)hipStreamDestroy(sd.stream);
    }
    for (auto& f : fans_) f.destroy();
}

std::shared_ptr<Server::Job> Server::make_job(const ServeRequest& r0) const
{
    auto j = std::make_shared<Job>();
    j->r = r0;
    const ServeRequest& r = j->r;
    const bool bad = (r.solver != kRK4 && r.solver != kABM) || r.n <= 0 || r.m < 0 ||
                     r.points < (r.solver == kABM ? 4 : 1) || !r.A || !r.x0 || !r.Y ||
                     (r.m > 0 && (!r.B || (!r.coef && !r.samples))) || (!r.zC && !r.C) ||
                     (!r.zD && !r.D) || (!(r.zC && r.zD) && r.p <= 0) || !(r.h > 0.0);
    j->L = layout(r);
    const size_t n = r.n, m = r.m, p = r.p;
    if (bad) {
        j->status = ParODE_InvalidSystem;
    } else {
        j->mats.assign(j->L.coef / 4, 0.0f);
        std::memcpy(&j->mats[j->L.A / 4], r.A, n * n * 4);
        if (m) std::memcpy(&j->mats[j->L.B / 4], r.B, n * m * 4);
        if (!r.zC) std::memcpy(&j->mats[j->L.C / 4], r.C, p * n * 4);
        if (!r.zD && m) std::memcpy(&j->mats[j->L.D / 4], r.D, p * m * 4);
        std::memcpy(&j->mats[j->L.x0 / 4], r.x0, n * 4);
        if (m && !r.samples) j->coef.assign(r.coef, r.coef + 6 * m);
    }
    j->r.A = j->r.B = j->r.C = j->r.D = j->r.x0 = nullptr;
    j->r.coef = nullptr;
    j->tSubmit = Clock::now();
    return j;
}

long Server::enqueue(const std::shared_ptr<Job>& j)
{
    j->id = next_++;
    jobs_[j->id] = j;
    ++st_.submitted;
    const size_t stage = j->L.up + (j->r.yDevice ? 0 : j->L.yBytes);
    if (j->status == 1 &&
        (footprint(j->r, precision_) > arena_->capacity() || stage > pinned_->capacity()))
        j->status = ParODE_NotEnoughResources;
    if (j->status != 1) {
        j->tAdmit = j->tDone = j->tSubmit;
        ++st_.failed;
        return j->id;
    }
    queue_.push_back(j);
    return j->id;
}

long Server::submit(const ServeRequest& r)
{
    auto j = make_job(r);
    std::lock_guard<std::mutex> g(mu_);
    const long id = enqueue(j);
    work_cv_.notify_one();
    return id;
}

void Server::submit(const ServeRequest* rs, size_t count, long* tickets)
{
    std::vector<std::shared_ptr<Job>> js(count);
    // a large batch's copies of the matrices split among a few threads
    const size_t parts = std::min<size_t>(8, count / 64 + 1);
    std::vector<std::exception_ptr> err(parts);
    auto copy = [&](size_t p) {
        try {
            for (size_t i = p; i < count; i += parts) js[i] = make_job(rs[i]);
        } catch (...) {
            err[p] = std::current_exception();
        }
    };
    run_parts(parts, copy);
    for (auto& e : err)
        if (e) std::rethrow_exception(e);
    std::lock_guard<std::mutex> g(mu_);
    for (size_t i = 0; i < count; ++i) tickets[i] = enqueue(js[i]);
    work_cv_.notify_one();
}

bool Server::grouped(const Job& j) const
{
    const ServeRequest& r = j.r;
    return precision_ == kMixed &&
           BatchSolve::eligible(r.solver, r.n, r.m, r.p, r.zC, r.zD, r.points);
}

bool Server::start(Group& g)
{
    hipStream_t s = lanes_[g.lane];
    // device io block: [uploads][samples made on the device][host outputs];
    // pinned stage: [uploads][host outputs]; the launch records: a pinned block of their own
    size_t upTot = 0, uTot = 0, yTot = 0;
    std::vector<size_t> upOff(g.jobs.size()), uOff(g.jobs.size()), yOff(g.jobs.size());
    for (size_t i = 0; i < g.jobs.size(); ++i) {
        const Job& j = *g.jobs[i];
        upOff[i] = upTot;
        upTot += j.L.up;
        uOff[i] = uTot;
        uTot += j.r.samples || j.r.m == 0 ? 0 : j.L.uBytes;
        yOff[i] = yTot;
        yTot += j.r.yDevice ? 0 : j.L.yBytes;
    }
    const size_t ioBytes = upTot + uTot + yTot;
    g.io = static_cast<char*>(arena_->take(ioBytes));
    if (!g.io) return false;
    char* dU = g.io + upTot;
    char* dYs = dU + uTot;
    auto fdev = [&](size_t i, size_t off) {
        return reinterpret_cast<const float*>(g.io + upOff[i] + off);
    };
    auto u_of = [&](size_t i) {
        const Job& j = *g.jobs[i];
        return reinterpret_cast<double*>(j.r.samples ? g.io + upOff[i] + j.L.U : dU + uOff[i]);
    };
    auto y_of = [&](size_t i) {
        const Job& j = *g.jobs[i];
        return j.r.yDevice ? j.r.Y : reinterpret_cast<float*>(dYs + yOff[i]);
    };
    // pinned stage: [uploads][host outputs] (the launch records get a block of
    // their own once planned); the uploads go up while the launches are planned
    auto stage = std::make_shared<Stage>();
    stage->bytes = upTot + yTot;
    if (stage->bytes) {
        stage->p = static_cast<char*>(pinned_->take(stage->bytes));
        if (!stage->p) {
            g.release(arena_.get());
            return false;
        }
        stage->arena = pinned_.get();
    }
    // every host copy and every transfer stays inside its own block (checked
    // for all requests before any copy); a large group's copies into the
    // pinned stage are split among a few threads
    for (size_t i = 0; i < g.jobs.size(); ++i) {
        Job& j = *g.jobs[i];
        const ServeRequest& r = j.r;
        check_range("staged matrices", upOff[i], j.mats.size() * 4, upTot);
        check_range("staged coefficients", upOff[i] + j.L.coef, j.coef.size() * 8, upTot);
        if (r.samples && r.m)
            check_range("staged samples", upOff[i] + j.L.U, (size_t)(2 * r.points + 1) * r.m * 8,
                        upTot);
        if (!r.yDevice) {
            check_range("device outputs", upTot + uTot + yOff[i], j.L.yBytes, ioBytes);
            check_range("staged outputs", upTot + yOff[i], j.L.yBytes, upTot + yTot);
        }
        j.stage = stage;
        j.stY = upTot + yOff[i];
    }
    const size_t count = g.jobs.size(), parts = std::min<size_t>(8, count / 64 + 1);
    auto copy = [&](size_t p) {
        for (size_t i = p; i < count; i += parts) {
            const Job& j = *g.jobs[i];
            const ServeRequest& r = j.r;
            std::memcpy(stage->p + upOff[i], j.mats.data(), j.mats.size() * 4);
            if (!j.coef.empty())
                std::memcpy(stage->p + upOff[i] + j.L.coef, j.coef.data(), j.coef.size() * 8);
            if (r.samples && r.m)
                std::memcpy(stage->p + upOff[i] + j.L.U, r.samples,
                            (size_t)(2 * r.points + 1) * r.m * 8);
        }
    };
    run_parts(parts, copy);
    if (upTot) PARODE_HIP(hipMemcpyAsync(g.io, stage->p, upTot, hipMemcpyHostToDevice, s));
    size_t descBytes = 0;
    try {
        if (g.grouped) {
            std::vector<BatchSystem> sys(g.jobs.size());
            for (size_t i = 0; i < g.jobs.size(); ++i) {
                const Job& j = *g.jobs[i];
                BatchSystem& q = sys[i];
                q.method = j.r.solver;
                q.n = j.r.n;
                q.m = j.r.m;
                q.p = j.r.p;
                q.points = j.r.points;
                q.t0 = j.r.t0;
                q.h = j.r.h;
                q.A = fdev(i, j.L.A);
                q.B = fdev(i, j.L.B);
                q.C = j.r.zC ? nullptr : fdev(i, j.L.C);
                q.D = j.r.zD ? nullptr : fdev(i, j.L.D);
                q.x0 = fdev(i, j.L.x0);
                q.coef = j.r.samples ? nullptr
                                     : reinterpret_cast<const double*>(g.io + upOff[i] + j.L.coef);
                q.sigma = j.r.sigma;
                q.seed = j.r.seed;
                q.U = u_of(i);
                q.Y = y_of(i);
            }
            g.batch.reset(new BatchSolve(sys, arena_.get()));
            descBytes = g.batch->descriptor_bytes();
            if (descBytes) {
                g.desc = static_cast<char*>(arena_->take(descBytes));
                if (!g.desc) throw std::bad_alloc();
            }
        } else {
            const Job& j = *g.jobs[0];
            ScanOptions opt;
            opt.precision = precision_;
            const long nScan = j.r.solver == kRK4 ? j.r.points : j.r.points - 3;
            g.core.reset(new ScanCore(j.r.solver, j.r.n, j.r.m, j.r.p, j.r.zC, j.r.zD, nScan, opt,
                                      arena_.get()));
        }
        if (descBytes) {
            g.descStage = std::make_shared<Stage>();
            g.descStage->bytes = Arena::round(descBytes);
            g.descStage->p = static_cast<char*>(pinned_->take(g.descStage->bytes));
            if (!g.descStage->p) throw std::bad_alloc();
            g.descStage->arena = pinned_.get();
            check_range("launch records", 0, descBytes, g.descStage->bytes);
        }
    } catch (std::bad_alloc&) {
        // the upload reads the stage and writes io: it ends before they return
        (void)hipStreamSynchronize(s);
        for (auto& j : g.jobs) j->stage.reset();
        g.release(arena_.get());
        return false;
    }
    if (g.grouped) {
        g.batch->run(s, g.descStage ? g.descStage->p : nullptr, g.desc, true, nullptr,
                     &sides_[g.lane], &fans_[g.lane]);
    } else {
        const Job& j = *g.jobs[0];
        const ServeRequest& r = j.r;
        double* U = u_of(0);
        if (!r.samples && r.m)
            k::sample_family(s, r.m, reinterpret_cast<const double*>(g.io + upOff[0] + j.L.coef), r.t0,
                             0.5 * r.h, 0, 2 * r.points + 1, r.sigma, r.seed, U);
        g.core->solve(s, r.points, r.h, fdev(0, j.L.A), r.m ? fdev(0, j.L.B) : nullptr,
                      r.zC ? nullptr : fdev(0, j.L.C), r.zD ? nullptr : fdev(0, j.L.D),
                      fdev(0, j.L.x0), U, true, y_of(0));
    }
    if (yTot) PARODE_HIP(hipMemcpyAsync(stage->p + upTot, dYs, yTot, hipMemcpyDeviceToHost, s));
    PARODE_HIP(hipEventCreateWithFlags(&g.done, hipEventDisableTiming));
    PARODE_HIP(hipEventRecord(g.done, s));
    const auto now = Clock::now();
    for (auto& j : g.jobs) j->tAdmit = now;
    return true;
}

void Server::finish(Group& g, int status)
{
    g.release(arena_.get());
    std::lock_guard<std::mutex> lk(mu_);
    if (g.lane >= 0) {
        --load_[g.lane];
        running_ -= (long)g.jobs.size();
    }
    const auto now = Clock::now();
    for (auto& jp : g.jobs) {
        Job& j = *jp;
        if (status != ParODE_OK || j.r.yDevice) j.stage.reset();   // nothing to copy out
        if (j.stage) ++staged_;
        j.tDone = now;
        j.status = status;
        if (status == ParODE_OK)
            ++st_.completed;
        else
            ++st_.failed;
    }
    done_cv_.notify_all();
}

void Server::loop()
{
    if (hipSetDevice(device_) != hipSuccess) return;
    std::vector<std::unique_ptr<Group>> running;
    for (;;) {
        bool progressed = false;
        // retire the finished groups
        for (size_t i = 0; i < running.size();) {
            const hipError_t e = hipEventQuery(running[i]->done);
            if (e == hipErrorNotReady) {
                ++i;
                continue;
            }
            finish(*running[i], e == hipSuccess ? ParODE_OK : ParODE_NotEnoughResources);
            running.erase(running.begin() + i);
            progressed = true;
        }
        // admit the head of the queue while a lane has room
        for (;;) {
            std::unique_ptr<Group> g(new Group);
            {
                std::lock_guard<std::mutex> lk(mu_);
                if (queue_.empty()) break;
                int lane = -1;
                for (int l = 0; l < (int)lanes_.size(); ++l)
                    if (load_[l] < depth_ && (lane < 0 || load_[l] < load_[lane])) lane = l;
                if (lane < 0) break;
                g->lane = lane;
                g->grouped = grouped(*queue_.front());
                // a group: the eligible requests at the head, in order, while
                // the free device memory holds them
                size_t bytes = 0;
                const size_t room = arena_->largest_free();
                while (!queue_.empty() && g->jobs.size() < (g->grouped ? groupCap_ : 1)) {
                    const Job& j = *queue_.front();
                    if (grouped(j) != g->grouped) break;
                    const size_t b = footprint(j.r, precision_) + 64 * Arena::kAlign;
                    if (!g->jobs.empty() && bytes + b > room) break;
                    bytes += b;
                    g->jobs.push_back(queue_.front());
                    queue_.pop_front();
                }
                ++load_[lane];
                running_ += (long)g->jobs.size();
            }
            bool ok = false;
            int err = ParODE_OK;
            try {
                ok = start(*g);
            } catch (std::exception& ex) {
                last_error() = ex.what();
                err = ParODE_NotEnoughResources;
                (void)hipStreamSynchronize(lanes_[g->lane]);   // nothing of it may still run
                g->release(arena_.get());
            }
            if (ok) {
                std::lock_guard<std::mutex> lk(mu_);
                st_.peakRunning = std::max(st_.peakRunning, running_);
                st_.arenaPeak = std::max(st_.arenaPeak, arena_->in_use());
                ++st_.groups;
                if (g->grouped) st_.groupedRequests += (long)g->jobs.size();
                st_.largestGroup = std::max(st_.largestGroup, (long)g->jobs.size());
                if (g->grouped && groupCap_ < 1024) groupCap_ *= 2;
                running.push_back(std::move(g));
                progressed = true;
                continue;
            }
            bool requeued = false;
            {
                std::lock_guard<std::mutex> lk(mu_);
                --load_[g->lane];
                running_ -= (long)g->jobs.size();
                // no memory now: back to the head of the queue (a smaller group
                // next time) while anything else holds memory; else a single
                // request can never fit
                if (err == ParODE_OK &&
                    (g->jobs.size() > 1 || !running.empty() || staged_ > 0)) {
                    for (auto it = g->jobs.rbegin(); it != g->jobs.rend(); ++it)
                        queue_.push_front(*it);
                    if (g->jobs.size() > 1) groupCap_ = std::max<size_t>(1, g->jobs.size() / 2);
                    requeued = true;
                }
            }
            if (requeued) {
                if (g->jobs.size() > 1) continue;   // retry at once with the smaller cap
                break;
            }
            g->lane = -1;
            finish(*g, err == ParODE_OK ? ParODE_NotEnoughResources : err);
            progressed = true;
        }
        std::unique_lock<std::mutex> lk(mu_);
        st_.running = running_;
        st_.queued = (long)queue_.size();
        if (stop_ && queue_.empty() && running.empty()) break;
        if (progressed) continue;
        if (running.empty())
            work_cv_.wait_for(lk, std::chrono::milliseconds(2),
                              [&] { return stop_ || !queue_.empty(); });
        else
            work_cv_.wait_for(lk, std::chrono::microseconds(20));
    }
}

int Server::wait(long ticket, double* latencyMs, double* queueMs)
{
    std::unique_lock<std::mutex> lk(mu_);
    auto it = jobs_.find(ticket);
    if (it == jobs_.end()) return ParODE_InvalidSystem;
    std::shared_ptr<Job> j = it->second;
    done_cv_.wait(lk, [&] { return j->status != 1; });
    using ms = std::chrono::duration<double, std::milli>;
    if (latencyMs) *latencyMs = ms(j->tDone - j->tSubmit).count();
    if (queueMs) *queueMs = ms(j->tAdmit - j->tSubmit).count();
    jobs_.erase(ticket);
    lk.unlock();
    // the outputs leave pinned staging on the caller's thread
    if (j->stage) {
        if (j->status == ParODE_OK)
            std::memcpy(j->r.Y, j->stage->p + j->stY, (size_t)j->r.points * j->L.nOut * 4);
        j->stage.reset();
        {
            std::lock_guard<std::mutex> g(mu_);
            --staged_;
        }
        work_cv_.notify_all();   // staging space may have come back
    }
    return j->status;
}

void Server::drain()
{
    std::unique_lock<std::mutex> lk(mu_);
    done_cv_.wait(lk, [&] { return queue_.empty() && st_.completed + st_.failed == st_.submitted; });
}

ServeStats Server::stats()
{
    std::lock_guard<std::mutex> g(mu_);
    return st_;
}

}  // namespace parode

// ------------------------------------------------------------------ C ABI
#include <ParODEExt.h>

extern "C" {

int ParODE_ServerCreate(int lanes, long deviceBytes, long pinnedBytes, int precision,
                        void** server)
{
    if (!server) return ParODE_InvalidSystem;
    *server = nullptr;
    if (lanes < 1 || lanes > 16 || deviceBytes <= 0 || pinnedBytes <= 0 ||
        precision < parode::kPlain32 || precision > parode::kFp64)
        return ParODE_InvalidSystem;
    try {
        *server = new parode::Server(lanes, (size_t)deviceBytes, (size_t)pinnedBytes, precision);
        return ParODE_OK;
    } catch (std::exception& e) {
        parode::last_error() = e.what();
        return ParODE_NotEnoughResources;
    }
}

int ParODE_ServerSubmit(void* server, int solver, int nStates, int nInputs, int nOutputs,
                        int zeroC, int zeroD, long points, double tStart, double tStep,
                        const float* A, const float* B, const float* C, const float* D,
                        const float* x0, const double* coef, double sigma,
                        unsigned long long seed, const double* samples, float* Y, long* ticket)
{
    if (!server || !ticket) return ParODE_InvalidSystem;
    parode::ServeRequest r;
    r.solver = solver;
    r.n = nStates;
    r.m = nInputs;
    r.p = nOutputs;
    r.zC = zeroC != 0;
    r.zD = zeroD != 0;
    r.points = points;
    r.t0 = tStart;
    r.h = tStep;
    r.A = A; r.B = B; r.C = C; r.D = D; r.x0 = x0;
    r.coef = coef;
    r.sigma = sigma;
    r.seed = seed;
    r.samples = samples;
    r.Y = Y;
    try {
        *ticket = static_cast<parode::Server*>(server)->submit(r);
        return ParODE_OK;
    } catch (std::exception& e) {
        parode::last_error() = e.what();
        return ParODE_NotEnoughResources;
    }
}

int ParODE_ServerSubmitDevice(void* server, int solver, int nStates, int nInputs, int nOutputs,
                              int zeroC, int zeroD, long points, double tStart, double tStep,
                              const float* A, const float* B, const float* C, const float* D,
                              const float* x0, const double* coef, double sigma,
                              unsigned long long seed, const double* samples, float* dY,
                              long* ticket)
{
    if (!server || !ticket) return ParODE_InvalidSystem;
    parode::ServeRequest r;
    r.solver = solver;
    r.n = nStates;
    r.m = nInputs;
    r.p = nOutputs;
    r.zC = zeroC != 0;
    r.zD = zeroD != 0;
    r.points = points;
    r.t0 = tStart;
    r.h = tStep;
    r.A = A; r.B = B; r.C = C; r.D = D; r.x0 = x0;
    r.coef = coef;
    r.sigma = sigma;
    r.seed = seed;
    r.samples = samples;
    r.Y = dY;
    r.yDevice = true;
    try {
        *ticket = static_cast<parode::Server*>(server)->submit(r);
        return ParODE_OK;
    } catch (std::exception& e) {
        parode::last_error() = e.what();
        return ParODE_NotEnoughResources;
    }
}

int ParODE_ServerSubmitDeviceBatch(void* server, int count, const int* solver,
                                   const int* nStates, const int* nInputs, const int* nOutputs,
                                   const int* zeroC, const int* zeroD, const long* points,
                                   const double* tStart, const double* tStep,
                                   const float* const* A, const float* const* B,
                                   const float* const* C, const float* const* D,
                                   const float* const* x0, const double* const* coef,
                                   const double* sigma, const unsigned long long* seed,
                                   const double* const* samples, float* const* dY, long* tickets)
{
    if (!server || count < 0 || (count > 0 && (!solver || !nStates || !nInputs || !nOutputs ||
                                               !zeroC || !zeroD || !points || !tStart ||
                                               !tStep || !A || !B || !C || !D || !x0 || !dY ||
                                               !tickets)))
        return ParODE_InvalidSystem;
    std::vector<parode::ServeRequest> rs((size_t)count);
    for (int i = 0; i < count; ++i) {
        parode::ServeRequest& r = rs[(size_t)i];
        r.solver = solver[i];
        r.n = nStates[i];
        r.m = nInputs[i];
        r.p = nOutputs[i];
        r.zC = zeroC[i] != 0;
        r.zD = zeroD[i] != 0;
        r.points = points[i];
        r.t0 = tStart[i];
        r.h = tStep[i];
        r.A = A[i]; r.B = B[i]; r.C = C[i]; r.D = D[i]; r.x0 = x0[i];
        r.coef = coef ? coef[i] : nullptr;
        r.sigma = sigma ? sigma[i] : 0.0;
        r.seed = seed ? seed[i] : 0;
        r.samples = samples ? samples[i] : nullptr;
        r.Y = dY[i];
        r.yDevice = true;
    }
    try {
        static_cast<parode::Server*>(server)->submit(rs.data(), rs.size(), tickets);
        return ParODE_OK;
    } catch (std::exception& e) {
        parode::last_error() = e.what();
        return ParODE_NotEnoughResources;
    }
}

int ParODE_ServerWait(void* server, long ticket, double* timesMs)
{
    if (!server) return ParODE_InvalidSystem;
    double lat = 0, q = 0;
    int st = static_cast<parode::Server*>(server)->wait(ticket, &lat, &q);
    if (timesMs) {
        timesMs[0] = lat;
        timesMs[1] = q;
    }
    return st;
}

int ParODE_ServerWaitBatch(void* server, int count, const long* tickets, int* status,
                           double* timesMs)
{
    if (!server || count < 0 || (count > 0 && (!tickets || !status))) return ParODE_InvalidSystem;
    parode::Server* sv = static_cast<parode::Server*>(server);
    for (int i = 0; i < count; ++i) {
        double lat = 0, q = 0;
        status[i] = sv->wait(tickets[i], &lat, &q);
        if (timesMs) {
            timesMs[2 * i] = lat;
            timesMs[2 * i + 1] = q;
        }
    }
    return ParODE_OK;
}

int ParODE_ServerDrain(void* server)
{
    if (!server) return ParODE_InvalidSystem;
    static_cast<parode::Server*>(server)->drain();
    return ParODE_OK;
}

int ParODE_ServerStats(void* server, long* stats)
{
    if (!server || !stats) return ParODE_InvalidSystem;
    parode::ServeStats s = static_cast<parode::Server*>(server)->stats();
    const long v[8] = {s.submitted, s.completed, s.failed, s.running, s.queued, s.peakRunning,
                       (long)s.arenaBytes, (long)s.arenaPeak};
    for (int i = 0; i < 8; ++i) stats[i] = v[i];
    return ParODE_OK;
}

int ParODE_ServerGroupStats(void* server, long* stats)
{
    if (!server || !stats) return ParODE_InvalidSystem;
    parode::ServeStats s = static_cast<parode::Server*>(server)->stats();
    stats[0] = s.groups;
    stats[1] = s.groupedRequests;
    stats[2] = s.largestGroup;
    return ParODE_OK;
}

long ParODE_ServerFootprint(int solver, int nStates, int nInputs, int nOutputs, long points,
                            int precision)
{
    if ((solver != 0 && solver != 1) || nStates <= 0 || nInputs < 0 ||
        points < (solver == 1 ? 4 : 1))
        return -1;
    parode::ServeRequest r;
    r.solver = solver;
    r.n = nStates;
    r.m = nInputs;
    r.p = nOutputs > 0 ? nOutputs : nStates;
    r.points = points;
    return (long)parode::Server::footprint(r, precision);
}

int ParODE_ServerDestroy(void* server)
{
    if (!server) return ParODE_InvalidSystem;
    delete static_cast<parode::Server*>(server);
    return ParODE_OK;
}

}  // extern "C"
