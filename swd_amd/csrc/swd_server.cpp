// Continuous-batching gradient server.
//
// Requests are forward + gradient solves (the per-iteration gradient of
// SWD.m:171-201) of any model size, nt and shot count.  A scheduler thread
// owns one FD stream on which every admitted request's shots advance in
// shared time-blocked launches: a request joins the running launch sequence
// at the next launch after admission (its jobs carry the global step and
// launch index at which they joined) and leaves it when its phase ends, so
// requests of different sizes and arrival times overlap on the device
// without waiting for each other.  The per-request stages between the FD
// phases (Radon, picking, adjoint sources; image reduction and read-back)
// run on side streams, ordered after the request's last FD launch by events.
//
// Memory: every request's buffers come from one device arena (the HBM
// budget), results from one pinned host arena.  A request is admitted, in
// arrival order, when a slot is free and its buffers fit; it leaves all of
// its memory resident on the device until it finishes.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "swd_engine.h"

namespace swd {

// ---------------------------------------------------------------------------
// Arena
// ---------------------------------------------------------------------------
struct Arena::Impl {
    mutable std::mutex m;
    std::map<size_t, size_t> free_;  // offset -> size
    size_t used = 0;
};

static constexpr size_t kArenaAlign = 256;
static size_t arena_round(size_t n) { return (n + kArenaAlign - 1) / kArenaAlign * kArenaAlign; }

Arena::~Arena()
{
    if (base_) {
        if (host_) (void)hipHostFree(base_);
        else (void)hipFree(base_);
    }
    delete impl_;
}

int Arena::init_device(size_t bytes)
{
    bytes = arena_round(std::max<size_t>(bytes, kArenaAlign));
    void* p = nullptr;
    if (hipMalloc(&p, bytes) != hipSuccess) {
        (void)hipGetLastError();
        set_error("device arena of %zu bytes: hipMalloc failed", bytes);
        return kErrNoMem;
    }
    base_ = (char*)p;
    cap_ = bytes;
    host_ = false;
    impl_ = new Impl;
    impl_->free_[0] = bytes;
    return kOk;
}

int Arena::init_host(size_t bytes)
{
    bytes = arena_round(std::max<size_t>(bytes, kArenaAlign));
    void* p = nullptr;
    if (hipHostMalloc(&p, bytes, hipHostMallocDefault) != hipSuccess) {
        (void)hipGetLastError();
        set_error("pinned host arena of %zu bytes: hipHostMalloc failed", bytes);
        return kErrNoMem;
    }
    base_ = (char*)p;
    cap_ = bytes;
    host_ = true;
    impl_ = new Impl;
    impl_->free_[0] = bytes;
    return kOk;
}

void* Arena::take(size_t n)
{
    if (!impl_) return nullptr;
    n = arena_round(std::max<size_t>(n, 1));
    std::lock_guard<std::mutex> lk(impl_->m);
    for (auto it = impl_->free_.begin(); it != impl_->free_.end(); ++it)
        if (it->second >= n) {
            const size_t off = it->first, sz = it->second;
            impl_->free_.erase(it);
            if (sz > n) impl_->free_[off + n] = sz - n;
            impl_->used += n;
            return base_ + off;
        }
    return nullptr;
}

void Arena::give(void* p, size_t n)
{
    if (!impl_ || !p) return;
    n = arena_round(std::max<size_t>(n, 1));
    size_t off = (size_t)((char*)p - base_);
    std::lock_guard<std::mutex> lk(impl_->m);
    impl_->used -= n;
    auto nx = impl_->free_.lower_bound(off);
    if (nx != impl_->free_.end() && off + n == nx->first) {  // merge with the following block
        n += nx->second;
        nx = impl_->free_.erase(nx);
    }
    if (nx != impl_->free_.begin()) {  // merge with the preceding block
        auto pv = std::prev(nx);
        if (pv->first + pv->second == off) {
            pv->second += n;
            return;
        }
    }
    impl_->free_[off] = n;
}

size_t Arena::used() const
{
    if (!impl_) return 0;
    std::lock_guard<std::mutex> lk(impl_->m);
    return impl_->used;
}

size_t Arena::largest_free() const
{
    if (!impl_) return 0;
    std::lock_guard<std::mutex> lk(impl_->m);
    size_t b = 0;
    for (const auto& f : impl_->free_) b = std::max(b, f.second);
    return b;
}

static thread_local Arena* t_arena = nullptr;
Arena* current_arena() { return t_arena; }
ArenaScope::ArenaScope(Arena* a) : prev(t_arena) { t_arena = a; }
ArenaScope::~ArenaScope() { t_arena = prev; }

// ---------------------------------------------------------------------------
// Server
// ---------------------------------------------------------------------------
namespace {

using Clock = std::chrono::steady_clock;

enum Phase { kQueued = 0, kFwd, kDisp, kAdj, kFinish, kDone, kFailed };

struct Request {
    long long id = 0;
    swd_problem p{};
    std::vector<int> sx, sz, isx;
    std::vector<double> vp, vs, wav, cr0r, cr0l;
    std::unique_ptr<Engine> eng;
    std::atomic<int> phase{kQueued};  // read by waiting callers
    // current FD phase: launches it needs and has been given, whether its jobs
    // are in the phase's launch table, waves per launch, cell-steps per launch
    int total = 0, done = 0;
    bool sel = false;
    long long waves = 0;
    double work_per_launch = 0, work_after = 0;
    double deadline = 0;  // selection order: arrival + slack x the request's work at the nominal rate
    std::atomic<bool> cancel{false};  // set by swd_server_cancel; the scheduler retires it
    std::vector<FdJob<float>> jobs;
    hipEvent_t ev_fd = nullptr, ev = nullptr;  // last FD launch of a phase; end of a side-stream stage
    hipStream_t side = nullptr;
    double* hout = nullptr;  // pinned: dk_vs (nz*nx), res (2*nshots), status (2*nshots ints)
    size_t hbytes = 0;
    int rc = kOk;
    std::string err;
    double t_submit = 0, t_admit = 0, t_done = 0;
    std::vector<double> dk, res_r, res_l;
    double remaining() const { return (total - done) * work_per_launch + work_after; }
};

constexpr int kMaxJobs = 1 << 14;       // shots in one FD launch table
constexpr int kMaxBlocks = 1 << 20;     // workgroups of one launch
constexpr int kStage = 4;               // pinned staging slots per FD phase
constexpr int kMaxLaunchesPerEpoch = 32;
// nominal FD rate (cell-steps per ms) that turns a request's work into its
// deadline: only the order of deadlines matters, so this just scales the slack
constexpr double kNominalCellStepsPerMs = 2.5e8;

}  // namespace

struct Server {
    Arena dev, host;
    hipStream_t fd[2] = {nullptr, nullptr};  // forward and adjoint launch sequences
    std::vector<hipStream_t> side;
    int next_side = 0;
    int nt_blk[2] = {4, 4};
    long long slots = 0;  // resident waves of the GPU for the time-blocked kernels
    std::atomic<long long> slots_override{0};  // set by the caller (0: slots)
    int max_active = 16;
    Clock::time_point t0 = Clock::now();

    std::mutex m;
    std::condition_variable cv_work, cv_done, cv_admit;
    std::deque<Request*> queue;   // submitted, not yet admitted
    std::vector<Request*> ready;  // admitted by the admission thread, not yet seen by the scheduler
    std::vector<Request*> active; // admitted, not finished (admission slots); entered with ready
    Request* admitting = nullptr; // taken off the queue by the admission thread, being set up
    double slack = 2.0;           // deadline = arrival + slack x work / nominal rate
    long long fail_rebuild_at = -1;  // test hook: this launch-table rebuild fails (SWD_SERVER_FAIL_REBUILD)
    std::map<long long, std::unique_ptr<Request>> reqs;
    long long next_id = 1;
    bool stop = false;
    bool mem_blocked = false;  // the head of the queue waits for memory
    std::thread th, th_admit;

    // scheduler-thread state
    int gstep[2] = {0, 0}, glaunch[2] = {0, 0};
    std::vector<Request*> members[2];
    bool reselect = false;
    std::vector<Request*> table_of[2];  // members in the current launch table
    DevBuf djobs[2], dbmap[2];
    int nblocks[2] = {0, 0};
    bool b1_tab[2] = {false, false};  // the launch table has a job with its own b1 plane
    char* stage[2][kStage] = {};
    hipEvent_t stage_ev[2][kStage] = {};
    bool stage_used[2][kStage] = {};
    int stage_next[2] = {0, 0};
    hipEvent_t epoch_ev[2][2] = {};
    bool epoch_used[2][2] = {};
    int epoch[2] = {0, 0};

    // statistics
    long long n_done = 0, n_failed = 0, peak_active = 0;
    std::atomic<long long> n_launch{0}, n_rebuild{0};
    size_t peak_bytes = 0;

    double now_ms() const { return std::chrono::duration<double, std::milli>(Clock::now() - t0).count(); }

    int init(size_t hbm, int maxa, int ntf, int nta);
    void shutdown();
    void loop();
    void admission_loop();
    void fail(Request* r, int rc, const char* what);
    void progress();
    void select(double now);
    int rebuild(int ph);
    int end_phase(Request* r, int ph);
    void retire(Request* r);
    void fail_all(int rc, const char* what);
    void drop_cancelled();
};

int Server::init(size_t hbm, int maxa, int ntf, int nta)
{
    max_active = std::max(1, maxa);
    nt_blk[0] = ntf;
    nt_blk[1] = nta;
    size_t fr = 0, tot = 0;
    SWD_HIP(hipMemGetInfo(&fr, &tot));
    const size_t reserve = (size_t)2 << 30;
    const size_t avail = fr > reserve + ((size_t)64 << 20) ? fr - reserve : fr / 2;
    if (hbm == 0 || hbm > avail) hbm = avail;
    int rc = dev.init_device(hbm);
    // one allocation of (nearly) all free memory can be refused: step down
    while (rc == kErrNoMem && hbm > ((size_t)2 << 30)) {
        hbm = hbm / 8 * 7;
        rc = dev.init_device(hbm);
    }
    if (rc || (rc = host.init_host((size_t)256 << 20))) return rc;
    int dev_id = 0, cus = 0;
    SWD_HIP(hipGetDevice(&dev_id));
    SWD_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev_id));
    const int nb = std::min(tb_blocks_per_cu<float>(false, ntf), tb_blocks_per_cu<float>(true, nta));
    slots = (long long)std::max(cus, 1) * std::max(nb, 1) * kMarchWaves;
    for (auto& s : fd) SWD_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    side.resize(4);
    for (auto& s : side) SWD_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    for (int ph = 0; ph < 2; ++ph) {
        if ((rc = djobs[ph].alloc(kMaxJobs * sizeof(FdJob<float>))) ||
            (rc = dbmap[ph].alloc(kMaxBlocks * sizeof(BlockMap))))
            return rc;
        for (int k = 0; k < kStage; ++k) {
            void* h = nullptr;
            SWD_HIP(hipHostMalloc(&h, kMaxJobs * sizeof(FdJob<float>) + kMaxBlocks * sizeof(BlockMap),
                                  hipHostMallocDefault));
            stage[ph][k] = (char*)h;
            SWD_HIP(hipEventCreateWithFlags(&stage_ev[ph][k], hipEventDisableTiming));
        }
        for (auto& e : epoch_ev[ph]) SWD_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    }
    if (const char* f = getenv("SWD_SERVER_FAIL_REBUILD")) fail_rebuild_at = atoll(f);
    th = std::thread([this] { loop(); });
    th_admit = std::thread([this] { admission_loop(); });
    return kOk;
}

void Server::shutdown()
{
    {
        std::lock_guard<std::mutex> lk(m);
        stop = true;
        for (Request* r : queue) {
            r->rc = kErrArg;
            r->err = "server shut down before the request was admitted";
            r->phase = kFailed;
        }
        queue.clear();
    }
    cv_admit.notify_all();
    cv_work.notify_all();
    if (th_admit.joinable()) th_admit.join();
    if (th.joinable()) th.join();
    cv_done.notify_all();
    for (auto s : fd)
        if (s) (void)hipStreamSynchronize(s);
    for (auto s : side) (void)hipStreamSynchronize(s);
    for (auto& kv : reqs) {
        Request* r = kv.second.get();
        if (r->eng) { r->eng->st = nullptr; r->eng.reset(); }
        if (r->ev) (void)hipEventDestroy(r->ev);
        if (r->ev_fd) (void)hipEventDestroy(r->ev_fd);
        r->ev = r->ev_fd = nullptr;
    }
    for (int ph = 0; ph < 2; ++ph) {
        for (int k = 0; k < kStage; ++k) {
            if (stage[ph][k]) (void)hipHostFree(stage[ph][k]);
            if (stage_ev[ph][k]) (void)hipEventDestroy(stage_ev[ph][k]);
        }
        for (auto& e : epoch_ev[ph])
            if (e) (void)hipEventDestroy(e);
    }
    for (auto s : side) (void)hipStreamDestroy(s);
    for (auto s : fd)
        if (s) (void)hipStreamDestroy(s);
}

void Server::fail(Request* r, int rc, const char* what)
{
    r->rc = rc;
    r->err = std::string(what) + ": " + last_error();
    if (r->eng) {
        if (r->side) (void)hipStreamSynchronize(r->side);
        r->eng->st = nullptr;
        r->eng.reset();
    }
    if (r->hout) { host.give(r->hout, r->hbytes); r->hout = nullptr; }
    r->t_done = now_ms();
    {
        std::lock_guard<std::mutex> lk(m);
        active.erase(std::remove(active.begin(), active.end(), r), active.end());
        ++n_failed;
        mem_blocked = false;
        r->phase = kFailed;
    }
    cv_done.notify_all();
    cv_admit.notify_all();
}

// Admission thread: in arrival order, while a slot is free, take the head
// request off the queue, build its engine in the device arena and upload its
// model, wavelet and observed picks; a request whose buffers do not fit goes
// back to the head of the queue until a running request returns memory (or
// fails when nothing else holds any).  A request enters `active` (and
// `ready`, for the scheduler) only once it is set up, so shutdown, failures of
// the launch sequence and cancellation never see it half-built: until then it
// is `admitting`, which only this thread touches.
void Server::admission_loop()
{
    // fail the request being admitted; it is not in `active`, and once failed a
    // waiter may free it, so it is not touched afterwards
    auto give_up = [&](Request* r, int rc, const char* what) {
        {
            std::lock_guard<std::mutex> lk(m);
            admitting = nullptr;
        }
        fail(r, rc, what);
    };
    for (;;) {
        Request* r = nullptr;
        {
            std::unique_lock<std::mutex> lk(m);
            cv_admit.wait(lk, [&] {
                return stop || (!queue.empty() && (int)active.size() < max_active && !mem_blocked);
            });
            if (stop) return;
            r = queue.front();
            queue.pop_front();
            admitting = r;
        }
        r->p.sx = r->sx.data();
        r->p.sz = r->sz.data();
        r->p.is_index = r->isx.data();
        r->side = side[next_side];
        std::unique_ptr<Engine> e(new Engine);
        int rc;
        {
            ArenaScope scope(&dev);
            rc = e->create(&r->p, r->side);
        }
        if (rc == kErrNoMem) {
            e->st = nullptr;
            e.reset();
            std::unique_lock<std::mutex> lk(m);
            if (!active.empty() && !stop && !r->cancel) {
                queue.push_front(r);  // back at the head, until memory returns
                admitting = nullptr;
                mem_blocked = true;   // cleared when a request returns its memory
                continue;
            }
            const bool quit = stop, cancelled = r->cancel;
            lk.unlock();
            if (quit) {
                set_error("server shut down");
                give_up(r, kErrArg, "server shut down before the request was admitted");
            } else if (cancelled) {
                set_error("cancelled by the caller");
                give_up(r, kErrBusy, "request cancelled");
            } else {
                give_up(r, rc, "request does not fit in the server's HBM budget");
            }
            continue;
        }
        next_side = (next_side + 1) % (int)side.size();
        r->eng = std::move(e);
        if (rc) { give_up(r, rc, "engine create"); continue; }
        Engine& g = *r->eng;
        const size_t S = r->p.nshots;
        r->hbytes = g.splane * sizeof(double) + 2 * S * sizeof(double) + 2 * S * sizeof(int);
        r->hout = (double*)host.take(r->hbytes);
        if (!r->hout) { set_error("pinned result arena exhausted"); give_up(r, kErrNoMem, "admit"); continue; }
        if ((rc = g.set_model(r->vp.data(), r->vs.data())) || (rc = g.set_wavelet(r->wav.data())) ||
            (rc = g.set_observed_picks(r->cr0r.data(), r->cr0l.data()))) {
            give_up(r, rc, "admit");
            continue;
        }
        if (hipEventCreateWithFlags(&r->ev_fd, hipEventDisableTiming) != hipSuccess ||
            hipEventCreateWithFlags(&r->ev, hipEventDisableTiming) != hipSuccess) {
            (void)hipGetLastError();
            set_error("hipEventCreate failed");
            give_up(r, kErrHip, "admit");
            continue;
        }
        r->jobs.clear();
        g.make_jobs(r->jobs, false, true);
        const double work = (double)S * g.nzbc * g.nxbc * (2.0 * r->p.nt - 1);
        bool quit = false, cancelled = false;
        {
            std::lock_guard<std::mutex> lk(m);
            quit = stop;
            cancelled = r->cancel;
            if (!quit && !cancelled) {
                r->deadline = r->t_submit + slack * work / kNominalCellStepsPerMs;
                active.push_back(r);
                peak_active = std::max<long long>(peak_active, (long long)active.size());
                peak_bytes = std::max(peak_bytes, dev.used());
                r->t_admit = now_ms();
                ready.push_back(r);
                admitting = nullptr;
            }
        }
        if (quit) {
            set_error("server shut down");
            give_up(r, kErrArg, "server shut down before the request was admitted");
            continue;
        }
        if (cancelled) {
            set_error("cancelled by the caller");
            give_up(r, kErrBusy, "request cancelled");
            continue;
        }
        cv_work.notify_all();
    }
}

// Start a request's FD phase: its launch count, its per-launch work and its
// wave count for the selection.
static void begin_phase(Request* r, int ph, int NT)
{
    Engine& g = *r->eng;
    const int steps = ph == 0 ? r->p.nt : r->p.nt - 1;
    r->total = div_up(steps, NT);
    r->done = 0;
    r->sel = false;
    const double cells = (double)r->p.nshots * g.nzbc * g.nxbc;
    r->work_per_launch = cells * NT;
    r->work_after = ph == 0 ? cells * (r->p.nt - 1) : 0.0;
    r->waves = (long long)r->p.nshots * div_up(g.nxbc - 2, kMarchLanes - 2 * NT) * div_up(g.nzbc - 2, 32);
}

// Move requests whose side-stream stage has finished to their next phase.
void Server::progress()
{
    std::vector<Request*> act;
    {
        std::lock_guard<std::mutex> lk(m);
        act = active;
        for (Request* r : ready) {
            if (r->phase == kFailed) continue;
            r->phase = kFwd;
            begin_phase(r, 0, nt_blk[0]);
            members[0].push_back(r);
            reselect = true;
        }
        ready.clear();
    }
    for (Request* r : act) {
        if (r->phase != kDisp && r->phase != kFinish) continue;
        const hipError_t q = hipEventQuery(r->ev);
        if (q == hipErrorNotReady) continue;
        if (q != hipSuccess) {
            set_error("stage event: %s", hipGetErrorString(q));
            fail(r, kErrHip, "side stream");
            continue;
        }
        if (r->phase == kDisp) {
            r->jobs.clear();
            r->eng->make_jobs(r->jobs, true, true);
            r->phase = kAdj;
            begin_phase(r, 1, nt_blk[1]);
            members[1].push_back(r);
            reselect = true;
        } else {
            retire(r);
        }
    }
}

void Server::retire(Request* r)
{
    Engine& g = *r->eng;
    const size_t S = r->p.nshots;
    const double* res = r->hout + g.splane;
    const int* stat = (const int*)(res + 2 * S);
    r->dk.assign(r->hout, r->hout + g.splane);
    r->res_r.assign(res, res + S);
    r->res_l.assign(res + S, res + 2 * S);
    r->rc = kOk;
    for (size_t c = 0; c < 2 * S; ++c)
        if (g.disp.cols[c].enabled && stat[c]) {
            r->rc = kErrPick;
            char b[160];
            snprintf(b, sizeof b, "pick of shot %zu (%s side) kept fewer than two columns", c % S,
                     c < S ? "right" : "left");
            r->err = b;
            break;
        }
    host.give(r->hout, r->hbytes);
    r->hout = nullptr;
    g.st = nullptr;  // its side-stream work has finished (event above)
    r->eng.reset();  // buffers return to the arena
    (void)hipEventDestroy(r->ev);
    (void)hipEventDestroy(r->ev_fd);
    r->ev = r->ev_fd = nullptr;
    r->t_done = now_ms();
    {
        std::lock_guard<std::mutex> lk(m);
        if (r->rc == kOk) ++n_done;
        else ++n_failed;
        active.erase(std::remove(active.begin(), active.end(), r), active.end());
        mem_blocked = false;
        r->phase = r->rc == kOk ? kDone : kFailed;
    }
    cv_done.notify_all();
    cv_admit.notify_all();
}

// Choose the requests that the next launches advance: earliest deadline first
// (arrival time + slack x the request's work at a nominal rate) while their
// shots fit in the GPU's resident waves (every further request would only
// lengthen the launches).  A new small request comes before a
// large one that arrived at about the same time, as with least remaining
// work first, but a large request's deadline does not move, so it overtakes
// the requests that keep arriving after it: waits stay bounded by the slack.
void Server::select(double now)
{
    (void)now;
    reselect = false;
    std::vector<Request*> c;
    for (int ph = 0; ph < 2; ++ph)
        for (Request* r : members[ph]) c.push_back(r);
    std::stable_sort(c.begin(), c.end(), [](const Request* a, const Request* b) { return a->deadline < b->deadline; });
    long long waves = 0;
    const long long cap = slots_override.load() > 0 ? slots_override.load() : slots;
    for (Request* r : c) {
        // the earliest deadline always runs; the others join only where their
        // waves still fit, so no launch holds more than the GPU does at once
        // (a request that would overflow it waits instead of slowing every
        // selected request down)
        r->sel = waves == 0 || waves + r->waves <= cap;
        if (r->sel) waves += r->waves;
    }
    for (int ph = 0; ph < 2; ++ph) {
        std::vector<Request*> t;
        for (Request* r : members[ph])
            if (r->sel) t.push_back(r);
        if (t != table_of[ph]) {
            table_of[ph] = t;
            const int rc = rebuild(ph);
            if (rc) fail_all(rc, "FD launch table");
        }
    }
}

// Upload the launch table of one FD phase.  Each selected job's origin is
// set from the launches it has been given, so that the kernels see its own
// step (step - start = done*NT) and ping-pong parity (launch - launch0 =
// done) however long it was paused; a job's state is cleared before its
// first launch.  The table goes up through a pinned staging slot in stream
// order, so launches already queued keep the table they were queued with.
int Server::rebuild(int ph)
{
    if (++n_rebuild == fail_rebuild_at) {
        set_error("launch-table rebuild %lld failed (SWD_SERVER_FAIL_REBUILD)", fail_rebuild_at);
        return kErrArg;
    }
    if (members[ph].empty()) gstep[ph] = glaunch[ph] = 0;
    if (gstep[ph] > (1 << 29)) { gstep[ph] = 0; glaunch[ph] = 0; }  // only differences matter
    const int NT = nt_blk[ph];
    std::vector<FdJob<float>> jobs;
    for (Request* r : table_of[ph]) {
        if (r->done == 0) {
            Engine& g = *r->eng;
            SWD_HIP(hipMemsetAsync(g.state.p, 0, g.state.bytes, fd[ph]));
            if (ph == 0) SWD_HIP(hipMemsetAsync(g.traces.p, 0, g.traces.bytes, fd[ph]));
            else {
                SWD_HIP(hipMemsetAsync(g.imgs.p, 0, g.imgs.bytes, fd[ph]));
                SWD_HIP(hipMemsetAsync(g.imgp.p, 0, g.imgp.bytes, fd[ph]));
            }
        }
        for (auto& J : r->jobs) {
            J.start = gstep[ph] - r->done * NT;
            J.launch0 = glaunch[ph] - r->done;
        }
        jobs.insert(jobs.end(), r->jobs.begin(), r->jobs.end());
    }
    nblocks[ph] = 0;
    if (jobs.empty()) return kOk;
    if ((int)jobs.size() > kMaxJobs) { set_error("more than %d shots in one FD launch", kMaxJobs); return kErrArg; }
    const int rows = auto_band_rows(jobs, ph == 1, NT);
    MarchPlan plan;
    int rc = build_march_plan(jobs, rows, plan, kMarchLanes - 2 * NT, NT, ph == 1);
    if (rc) return rc;
    if (plan.nblocks > kMaxBlocks) { set_error("FD launch of %d workgroups exceeds %d", plan.nblocks, kMaxBlocks); return kErrArg; }
    const int k = stage_next[ph];
    stage_next[ph] = (k + 1) % kStage;
    if (stage_used[ph][k]) SWD_HIP(hipEventSynchronize(stage_ev[ph][k]));
    char* h = stage[ph][k];
    const size_t jb = jobs.size() * sizeof(FdJob<float>), bb = plan.bmap.size() * sizeof(BlockMap);
    memcpy(h, jobs.data(), jb);
    memcpy(h + kMaxJobs * sizeof(FdJob<float>), plan.bmap.data(), bb);
    SWD_HIP(hipMemcpyAsync(djobs[ph].p, h, jb, hipMemcpyHostToDevice, fd[ph]));
    SWD_HIP(hipMemcpyAsync(dbmap[ph].p, h + kMaxJobs * sizeof(FdJob<float>), bb, hipMemcpyHostToDevice, fd[ph]));
    SWD_HIP(hipEventRecord(stage_ev[ph][k], fd[ph]));
    stage_used[ph][k] = true;
    nblocks[ph] = plan.nblocks;
    b1_tab[ph] = jobs_b1(jobs);
    return kOk;
}

// A request's FD phase has been fully launched: order its next stage on its
// side stream after the last launch.
int Server::end_phase(Request* r, int ph)
{
    Engine& g = *r->eng;
    SWD_HIP(hipEventRecord(r->ev_fd, fd[ph]));
    SWD_HIP(hipStreamWaitEvent(r->side, r->ev_fd, 0));
    int rc;
    r->sel = false;
    if (ph == 0) {
        if ((rc = g.dispersion()) || (rc = g.adjoint_source())) return rc;
        SWD_HIP(hipEventRecord(r->ev, r->side));
        r->phase = kDisp;
    } else {
        if ((rc = g.assemble())) return rc;
        const size_t S = r->p.nshots;
        double* res = r->hout + g.splane;
        SWD_HIP(hipMemcpyAsync(r->hout, g.outs.p, g.splane * sizeof(double), hipMemcpyDeviceToHost, r->side));
        SWD_HIP(hipMemcpyAsync(res, g.disp.res.p, 2 * S * sizeof(double), hipMemcpyDeviceToHost, r->side));
        SWD_HIP(hipMemcpyAsync(res + 2 * S, g.disp.status.p, 2 * S * sizeof(int), hipMemcpyDeviceToHost, r->side));
        SWD_HIP(hipEventRecord(r->ev, r->side));
        r->phase = kFinish;
    }
    return kOk;
}

void Server::fail_all(int rc, const char* what)
{
    for (auto s : fd) (void)hipStreamSynchronize(s);
    std::vector<Request*> act;
    {
        std::lock_guard<std::mutex> lk(m);
        act = active;   // every request of `ready` is in `active` too
        ready.clear();
    }
    for (int ph = 0; ph < 2; ++ph) {
        members[ph].clear();
        table_of[ph].clear();
        nblocks[ph] = 0;
    }
    for (Request* r : act) fail(r, rc, what);
}

// Retire cancelled requests: they leave the launch tables, the launches
// already queued (which may still read their buffers) drain, and then their
// memory goes back to the arenas.  The other requests continue from the same
// state: their job origins are recomputed from the launches they were given.
void Server::drop_cancelled()
{
    std::vector<Request*> cut;
    {
        std::lock_guard<std::mutex> lk(m);
        for (Request* r : active)
            if (r->cancel && r->phase != kFailed && r->phase != kDone) cut.push_back(r);
        if (cut.empty()) return;
        ready.erase(std::remove_if(ready.begin(), ready.end(), [](Request* r) { return r->cancel.load(); }),
                    ready.end());
    }
    for (int ph = 0; ph < 2; ++ph) {
        auto gone = [&](Request* r) { return std::find(cut.begin(), cut.end(), r) != cut.end(); };
        members[ph].erase(std::remove_if(members[ph].begin(), members[ph].end(), gone), members[ph].end());
        if (std::any_of(table_of[ph].begin(), table_of[ph].end(), gone)) {
            table_of[ph].erase(std::remove_if(table_of[ph].begin(), table_of[ph].end(), gone), table_of[ph].end());
            const int rc = rebuild(ph);
            if (rc) { fail_all(rc, "FD launch table"); return; }
        }
    }
    for (auto st : fd) (void)hipStreamSynchronize(st);
    for (Request* r : cut) {
        set_error("cancelled by the caller");
        fail(r, kErrBusy, "request cancelled");
    }
    reselect = true;
}

void Server::loop()
{
    for (;;) {
        {
            std::unique_lock<std::mutex> lk(m);
            if (active.empty() && ready.empty()) {
                if (stop && queue.empty()) break;
                cv_work.wait_for(lk, std::chrono::milliseconds(20));
                continue;
            }
        }
        drop_cancelled();
        progress();
        const double now = now_ms();
        if (reselect) select(now);
        for (int ph = 0; ph < 2; ++ph)
            if (gstep[ph] > (1 << 29)) {  // re-origin the table before the counters grow large
                const int rc = rebuild(ph);
                if (rc) fail_all(rc, "FD launch table");
            }
        if (table_of[0].empty() && table_of[1].empty()) {
            std::unique_lock<std::mutex> lk(m);  // side-stream stages or admissions pending
            cv_work.wait_for(lk, std::chrono::microseconds(100));
            continue;
        }
        // launches until the nearest phase end, at most kMaxLaunchesPerEpoch
        int E = kMaxLaunchesPerEpoch;
        for (int ph = 0; ph < 2; ++ph)
            for (Request* r : table_of[ph]) E = std::min(E, r->total - r->done);
        E = std::max(E, 1);
        int rc = kOk;
        for (int k = 0; k < E && rc == kOk; ++k)
            for (int ph = 0; ph < 2 && rc == kOk; ++ph) {
                if (table_of[ph].empty()) continue;
                rc = ph == 0 ? launch_fwd_tb<float>(fd[0], djobs[0].as<FdJob<float>>(), dbmap[0].as<BlockMap>(),
                                                    nblocks[0], gstep[0], glaunch[0], nt_blk[0], b1_tab[0])
                             : launch_adj_tb<float>(fd[1], djobs[1].as<FdJob<float>>(), dbmap[1].as<BlockMap>(),
                                                    nblocks[1], gstep[1], glaunch[1], nt_blk[1], b1_tab[1]);
                gstep[ph] += nt_blk[ph];
                ++glaunch[ph];
                ++n_launch;
            }
        if (rc != kOk) { fail_all(rc, "FD launch"); continue; }
        for (int ph = 0; ph < 2; ++ph) {
            if (table_of[ph].empty()) continue;
            std::vector<Request*> keep;
            for (Request* r : table_of[ph]) r->done += E;
            for (Request* r : members[ph]) {
                if (!r->sel || r->done < r->total) { keep.push_back(r); continue; }
                const int erc = end_phase(r, ph);
                if (erc) fail(r, erc, ph == 0 ? "dispersion stage" : "assembly stage");
                reselect = true;
            }
            members[ph].swap(keep);
            // keep at most two epochs of launches queued ahead of the device
            hipError_t he = hipEventRecord(epoch_ev[ph][epoch[ph]], fd[ph]);
            epoch_used[ph][epoch[ph]] = true;
            epoch[ph] ^= 1;
            if (he == hipSuccess && epoch_used[ph][epoch[ph]]) he = hipEventSynchronize(epoch_ev[ph][epoch[ph]]);
            if (he != hipSuccess) {
                set_error("FD stream: %s", hipGetErrorString(he));
                fail_all(kErrHip, "FD stream");
                break;
            }
        }
    }
}

}  // namespace swd

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
using namespace swd;

struct swd_server {
    Server s;
};

extern "C" {

int swd_server_create(long long hbm_bytes, int max_active, int nt_block_fwd, int nt_block_adj, swd_server** out)
{
    if (!out) { set_error("null output handle"); return kErrArg; }
    *out = nullptr;
    if (hbm_bytes < 0 || max_active < 1 || nt_block_fwd < 2 || nt_block_fwd > 6 || nt_block_adj < 2 ||
        nt_block_adj > 6) {
        set_error("server needs hbm_bytes >= 0, max_active >= 1 and time blocks of 2..6 steps");
        return kErrArg;
    }
    swd_server* h = new (std::nothrow) swd_server;
    if (!h) { set_error("out of host memory"); return kErrNoMem; }
    const int rc = h->s.init((size_t)hbm_bytes, max_active, nt_block_fwd, nt_block_adj);
    if (rc) {
        h->s.shutdown();
        delete h;
        return rc;
    }
    *out = h;
    return kOk;
}

int swd_server_destroy(swd_server* h)
{
    if (!h) return kErrArg;
    h->s.shutdown();
    delete h;
    return kOk;
}

int swd_server_submit(swd_server* h, const swd_problem* p, const double* vp, const double* vs, const double* wavelet,
                      const double* cr0_r, const double* cr0_l, long long* ticket)
{
    if (!h || !p || !vp || !vs || !wavelet || !cr0_r || !cr0_l || !ticket) { set_error("null argument"); return kErrArg; }
    if (p->nshots < 1 || p->nz < 1 || p->nx < 1 || p->nt < 2 || !p->sx || !p->sz) {
        set_error("bad problem");
        return kErrArg;
    }
    int npair = 0;
    int rc = swd_problem_npair(p, &npair);
    if (rc) return rc;
    std::unique_ptr<Request> r(new (std::nothrow) Request);
    if (!r) { set_error("out of host memory"); return kErrNoMem; }
    r->p = *p;
    const size_t S = p->nshots, n = (size_t)p->nz * p->nx, np = (size_t)npair * S;
    r->sx.assign(p->sx, p->sx + S);
    r->sz.assign(p->sz, p->sz + S);
    if (p->is_index) r->isx.assign(p->is_index, p->is_index + S);
    else for (size_t s = 0; s < S; ++s) r->isx.push_back((int)s + 1);
    r->vp.assign(vp, vp + n);
    r->vs.assign(vs, vs + n);
    r->wav.assign(wavelet, wavelet + p->nt);
    r->cr0r.assign(cr0_r, cr0_r + np);
    r->cr0l.assign(cr0_l, cr0_l + np);
    Server& s = h->s;
    {
        std::lock_guard<std::mutex> lk(s.m);
        if (s.stop) { set_error("server is shutting down"); return kErrArg; }
        r->id = s.next_id++;
        r->t_submit = s.now_ms();
        *ticket = r->id;
        s.queue.push_back(r.get());
        s.reqs[r->id] = std::move(r);
    }
    s.cv_admit.notify_all();
    return kOk;
}

int swd_server_wait(swd_server* h, long long ticket, double timeout_ms, double* dk_vs, double* res_r, double* res_l,
                    double* times_ms)
{
    if (!h) return kErrArg;
    Server& s = h->s;
    std::unique_lock<std::mutex> lk(s.m);
    auto it = s.reqs.find(ticket);
    if (it == s.reqs.end()) { set_error("unknown ticket %lld", ticket); return kErrArg; }
    Request* r = it->second.get();
    auto finished = [&] { return r->phase == kDone || r->phase == kFailed; };
    if (timeout_ms < 0) s.cv_done.wait(lk, finished);
    else if (!s.cv_done.wait_for(lk, std::chrono::duration<double, std::milli>(timeout_ms), finished)) {
        set_error("request %lld still running", ticket);
        return kErrBusy;
    }
    if (times_ms) {
        times_ms[0] = r->t_admit > 0 ? r->t_admit - r->t_submit : r->t_done - r->t_submit;
        times_ms[1] = r->t_admit > 0 ? r->t_done - r->t_admit : 0;
        times_ms[2] = r->t_done - r->t_submit;
        times_ms[3] = r->t_submit;
    }
    const int rc = r->rc;
    if (rc == kOk) {
        if (dk_vs) std::copy(r->dk.begin(), r->dk.end(), dk_vs);
        if (res_r) std::copy(r->res_r.begin(), r->res_r.end(), res_r);
        if (res_l) std::copy(r->res_l.begin(), r->res_l.end(), res_l);
    } else {
        set_error("%s", r->err.c_str());
    }
    s.reqs.erase(it);
    return rc;
}

int swd_server_cancel(swd_server* h, long long ticket)
{
    if (!h) return kErrArg;
    Server& s = h->s;
    {
        std::lock_guard<std::mutex> lk(s.m);
        auto it = s.reqs.find(ticket);
        if (it == s.reqs.end()) { set_error("unknown ticket %lld", ticket); return kErrArg; }
        Request* r = it->second.get();
        if (r->phase == kDone || r->phase == kFailed) { set_error("request %lld has finished", ticket); return kErrArg; }
        auto q = std::find(s.queue.begin(), s.queue.end(), r);
        if (q != s.queue.end()) {   // not admitted yet: it ends now
            s.queue.erase(q);
            r->rc = kErrBusy;
            r->err = "request cancelled before admission";
            r->t_done = s.now_ms();
            ++s.n_failed;
            r->phase = kFailed;
        } else {
            r->cancel = true;       // admitting or admitted: the server threads retire it
        }
    }
    s.cv_done.notify_all();
    s.cv_admit.notify_all();
    s.cv_work.notify_all();
    return kOk;
}

int swd_server_set_deadline_slack(swd_server* h, double slack)
{
    if (!h || !(slack > 0)) { set_error("deadline slack must be > 0"); return kErrArg; }
    Server& s = h->s;
    std::lock_guard<std::mutex> lk(s.m);
    s.slack = slack;
    return kOk;
}

int swd_server_set_wave_slots(swd_server* h, long long waves)
{
    if (!h || waves < 0) { set_error("bad wave slots"); return kErrArg; }
    Server& s = h->s;
    std::lock_guard<std::mutex> lk(s.m);
    s.slots_override = waves;
    return kOk;
}

int swd_server_stats(swd_server* h, long long* vals, int n)
{
    if (!h || !vals) return kErrArg;
    Server& s = h->s;
    std::lock_guard<std::mutex> lk(s.m);
    const long long v[10] = {(long long)s.queue.size(), (long long)s.active.size(), s.n_done, s.n_failed,
                             (long long)s.dev.used(), (long long)s.dev.capacity(), s.n_launch.load(), s.peak_active,
                             (long long)s.peak_bytes, s.n_rebuild.load()};
    for (int i = 0; i < n && i < 10; ++i) vals[i] = v[i];
    return kOk;
}

}  // extern "C"
