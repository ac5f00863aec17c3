#include "server.h"

#include <ParODEEnums.h>

#include <algorithm>
#include <cstring>

#include "config.h"
#include "kernels.h"

namespace parode {

namespace {

size_t up(size_t b) { return Arena::round(b); }

// Byte offsets of one request's arrays in its device block (io) and pinned
// staging block.  io: [A B C D x0 | coef | U | Y]; staging holds the same
// prefix (and U when the samples come from the host), then Y.
struct Layout {
    size_t A, B, C, D, x0, coef, U, Y, ioBytes;
    size_t stY, stBytes, h2d;
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
    L.coef = o; o += m * 6 * 8;
    o = up(o);
    L.U = o;  o += (size_t)(2 * r.points + 1) * m * 8;
    o = up(o);
    L.Y = o;  o += (size_t)r.points * L.nOut * 4;
    L.ioBytes = up(o);
    L.h2d = r.samples ? L.Y : L.U;
    L.stY = L.h2d;
    L.stBytes = up(L.stY + (size_t)r.points * L.nOut * 4);
    return L;
}

// Throws unless [off, off + bytes) lies inside a block of blockBytes.
void check_range(const char* what, size_t off, size_t bytes, size_t blockBytes)
{
    if (off > blockBytes || bytes > blockBytes - off)
        throw std::logic_error(std::string("Server: ") + what + " outside its block");
}

}  // namespace

struct Server::Job {
    long id = 0;
    ServeRequest r;
    std::vector<float> mats;
    std::vector<double> coef;
    Layout L{};
    std::unique_ptr<ScanCore> core;
    char* io = nullptr;
    char* stage = nullptr;
    int lane = -1;
    hipEvent_t done = nullptr;
    Clock::time_point tSubmit, tAdmit, tDone;
    int status = 1;   // 1 = pending
};

size_t Server::footprint(const ServeRequest& r, int precision)
{
    ScanOptions opt;
    opt.precision = precision;
    const long nScan = r.solver == kRK4 ? r.points : r.points - 3;
    return ScanCore::estimate_footprint(r.solver, r.n, r.m, nScan, opt, r.p) + layout(r).ioBytes;
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
    for (auto& s : lanes_) PARODE_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    st_.arenaBytes = arena_->capacity();
    workers_ = lanes < 4 ? lanes : 4;
    for (int w = 0; w < workers_; ++w) threads_.emplace_back([this, w] { loop(w); });
}

Server::~Server()
{
    {
        std::lock_guard<std::mutex> g(mu_);
        stop_ = true;
    }
    work_cv_.notify_all();
    for (auto& t : threads_)
        if (t.joinable()) t.join();
    for (auto& s : lanes_) (void)hipStreamDestroy(s);
}

long Server::submit(const ServeRequest& r0)
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
    if (!bad) {
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
    std::lock_guard<std::mutex> g(mu_);
    j->id = next_++;
    jobs_[j->id] = j;
    ++st_.submitted;
    if (bad) {
        j->status = ParODE_InvalidSystem;
        j->tAdmit = j->tDone = j->tSubmit;
        ++st_.failed;
        return j->id;
    }
    if (footprint(r, precision_) > arena_->capacity() || j->L.stBytes > pinned_->capacity()) {
        j->status = ParODE_NotEnoughResources;
        j->tAdmit = j->tDone = j->tSubmit;
        ++st_.failed;
        return j->id;
    }
    queue_.push_back(j);
    work_cv_.notify_one();
    return j->id;
}

bool Server::admit(Job& j, int lane)
{
    const ServeRequest& r = j.r;
    const Layout& L = j.L;
    ScanOptions opt;
    opt.precision = precision_;
    const long nScan = r.solver == kRK4 ? r.points : r.points - 3;
    try {
        j.core.reset(new ScanCore(r.solver, r.n, r.m, r.p, r.zC, r.zD, nScan, opt, arena_.get()));
    } catch (std::bad_alloc&) {
        j.core.reset();
        return false;
    }
    j.io = static_cast<char*>(arena_->take(L.ioBytes));
    j.stage = j.io ? static_cast<char*>(pinned_->take(L.stBytes)) : nullptr;
    if (!j.io || !j.stage) {
        arena_->give(j.io);
        j.io = nullptr;
        j.core.reset();
        return false;
    }
    hipStream_t s = lanes_[lane];
    // every host copy and every transfer stays inside its own block
    const size_t uBytes = (size_t)(2 * r.points + 1) * r.m * 8;
    const size_t yBytes = (size_t)r.points * L.nOut * 4;
    check_range("staged matrices", 0, j.mats.size() * 4, L.stBytes);
    check_range("staged coefficients", L.coef, j.coef.size() * 8, L.stBytes);
    if (r.samples && r.m) check_range("staged samples", L.U, uBytes, L.stBytes);
    check_range("upload", 0, L.h2d, std::min(L.stBytes, L.ioBytes));
    check_range("device outputs", L.Y, yBytes, L.ioBytes);
    check_range("staged outputs", L.stY, yBytes, L.stBytes);
    std::memcpy(j.stage, j.mats.data(), j.mats.size() * 4);
    if (!j.coef.empty()) std::memcpy(j.stage + L.coef, j.coef.data(), j.coef.size() * 8);
    if (r.samples && r.m) std::memcpy(j.stage + L.U, r.samples, uBytes);
    PARODE_HIP(hipMemcpyAsync(j.io, j.stage, L.h2d, hipMemcpyHostToDevice, s));
    if (!r.samples && r.m)
        k::sample_family(s, r.m, reinterpret_cast<const double*>(j.io + L.coef), r.t0, 0.5 * r.h,
                         0, 2 * r.points + 1, r.sigma, r.seed,
                         reinterpret_cast<double*>(j.io + L.U));
    auto f = [&](size_t off) { return reinterpret_cast<const float*>(j.io + off); };
    float* dY = reinterpret_cast<float*>(j.io + L.Y);
    j.core->solve(s, r.points, r.h, f(L.A), r.m ? f(L.B) : nullptr, r.zC ? nullptr : f(L.C),
                  r.zD ? nullptr : f(L.D), f(L.x0), j.io + L.U, true, dY);
    PARODE_HIP(hipMemcpyAsync(j.stage + L.stY, dY, yBytes, hipMemcpyDeviceToHost, s));
    PARODE_HIP(hipEventCreateWithFlags(&j.done, hipEventDisableTiming));
    PARODE_HIP(hipEventRecord(j.done, s));
    j.lane = lane;
    j.tAdmit = Clock::now();
    return true;
}

void Server::finish(Job& j, int status)
{
    if (j.done) (void)hipEventDestroy(j.done);
    j.done = nullptr;
    j.core.reset();
    arena_->give(j.io);
    j.io = nullptr;
    if (status != ParODE_OK) {   // nothing to copy out
        pinned_->give(j.stage);
        j.stage = nullptr;
    }
    std::lock_guard<std::mutex> g(mu_);
    if (j.lane >= 0) {
        --load_[j.lane];
        --running_;
    }
    if (j.stage) ++staged_;
    j.tDone = Clock::now();
    j.status = status;
    if (status == ParODE_OK)
        ++st_.completed;
    else
        ++st_.failed;
    done_cv_.notify_all();
}

void Server::loop(int worker)
{
    if (hipSetDevice(device_) != hipSuccess) return;
    std::vector<std::shared_ptr<Job>> running;   // this worker's requests on the GPU
    for (;;) {
        bool progressed = false;
        // retire the finished requests
        for (size_t i = 0; i < running.size();) {
            Job& j = *running[i];
            const hipError_t e = hipEventQuery(j.done);
            if (e == hipErrorNotReady) {
                ++i;
                continue;
            }
            finish(j, e == hipSuccess ? ParODE_OK : ParODE_NotEnoughResources);
            running.erase(running.begin() + i);
            progressed = true;
        }
        // admit in arrival order while one of this worker's lanes has room
        for (;;) {
            std::shared_ptr<Job> j;
            int lane = -1;
            {
                std::lock_guard<std::mutex> g(mu_);
                if (queue_.empty()) break;
                for (int l = worker; l < (int)lanes_.size(); l += workers_)
                    if (load_[l] < depth_ && (lane < 0 || load_[l] < load_[lane])) lane = l;
                if (lane < 0) break;
                j = queue_.front();
                queue_.pop_front();
                ++load_[lane];
                ++active_;
            }
            bool ok = false;
            int err = ParODE_OK;
            try {
                ok = admit(*j, lane);
            } catch (std::exception& ex) {
                last_error() = ex.what();
                err = ParODE_NotEnoughResources;
                (void)hipStreamSynchronize(lanes_[lane]);   // nothing of it may still run
            }
            if (!ok) {
                bool requeued = false;
                {
                    std::lock_guard<std::mutex> g(mu_);
                    --load_[lane];
                    --active_;
                    // no memory now: back to the head of the queue while
                    // anything else holds memory (running, being admitted, or
                    // outputs not yet collected); else it can never fit
                    if (err == ParODE_OK &&
                        (running_ > 0 || active_ > 0 || staged_ > 0 || !running.empty())) {
                        queue_.push_front(j);
                        requeued = true;
                    }
                }
                if (requeued) break;
                finish(*j, err == ParODE_OK ? ParODE_NotEnoughResources : err);
                progressed = true;
                continue;
            }
            {
                std::lock_guard<std::mutex> g(mu_);
                --active_;
                ++running_;
                st_.peakRunning = std::max(st_.peakRunning, running_);
                st_.arenaPeak = std::max(st_.arenaPeak, arena_->in_use());
            }
            running.push_back(j);
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
            std::memcpy(j->r.Y, j->stage + j->L.stY, (size_t)j->r.points * j->L.nOut * 4);
        pinned_->give(j->stage);
        j->stage = nullptr;
        {
            std::lock_guard<std::mutex> g(mu_);
            --staged_;
        }
        work_cv_.notify_all();   // staging space came back
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
