// Serving loop: continuous batching of SCF-style polarization solves of many
// systems on one GPU.
//
// A request is one system plus an SCF sequence of external fields, each the
// field of a set of point charges (a QM region) at the polarizable sites.
// Requests are served in two lanes by system size, each with its own batch
// engine, stream and worker thread: a lane's worker admits requests while its
// batch slots and the (shared) HBM budget allow, keeps every admitted system
// resident (geometry, lists, static field, last dipoles), and runs ticks of the
// batched mixed solver (batch.h) over the lane's admitted systems at once, so
// that small systems are not paced by the matvecs of large ones.  A preparation
// thread per lane brings a newly admitted system onto the device (upload,
// static field, on the system's own stream) while the workers keep ticking; the
// lane's worker adds it to its batch when it is ready.  A solve that converges is retired at the next
// tick boundary; the request's next solve starts right away from the converged
// dipoles (warm start), and a request whose last solve finished releases its
// slot and memory to the next waiting request.  Column states are read back one
// tick late, so the GPU always has the next tick queued.
//
// Warm steps of an SCF sequence change only the external field, and by the
// same amount for the D and P sets; the correction T^-1 (ext_new - ext_old) is
// therefore solved once, in an extra FP32-refined column, added to both sets,
// and each set is then checked (and if needed refined) against its own FP64
// residual.
//
// Each solve's arithmetic depends only on its own system (batch.h), so a
// request gives the same dipoles and energies alone and in any mix.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <set>
#include <string>
#include <thread>
#include <vector>

#include "batch_engine.h"
#include "ommp_internal.h"

namespace ommp {

#define SCHECK(x) HIP_CHECK(x)

DeviceView device_view(System &s);
size_t device_footprint_estimate(const System &s);

namespace {

double now_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

}  // namespace

struct Request {
    int id = 0;
    System *sys = nullptr;
    int nsteps = 0, nq = 0;
    std::vector<double> charges;  // [nsteps][nq][4]
    double *charges_pin = nullptr;  // pinned copy, uploaded in stream order at admission
    double t_arrival = 0.0;
    // progress (worker thread only)
    int seg = -1;
    int step = 0;
    bool shared = false;   // the shared-correction column of a warm step is running
    int inner_shared = 0;  // its inner iterations (reported with the step)
    long valid_from = 0;
    double t_ready = 0.0;
    size_t footprint = 0;  // bytes held against the budget (the estimate until added)
    double *epin = nullptr;  // pinned [nsteps] energies
    hipEvent_t ev = nullptr;
    // results
    std::vector<double> epin_copy;
    std::vector<double> latency_ms, t_done;
    std::vector<int> outer, inner;
    int status = 0;
    bool done = false;
    int lane = 0;
};

// Systems below kSmallLaneSites polarizable sites are solved in their own
// lane: a batch engine, stream and worker thread of their own, so that their
// ticks (a few small kernels) do not wait for the matvecs of large systems and
// run beside them on the GPU.  Both streams keep the default priority.
constexpr int kSmallLaneSites = 4096;
constexpr int kLanes = 2;

class Server {
    // tolerance of the shared correction relative to the solve's tolerance:
    // its error adds to each set's residual before the FP64 check
    static constexpr double kSharedTolFrac = 0.25;

    struct Lane {
        hipStream_t stream = nullptr;
        BatchEngine *engine = nullptr;
        std::thread worker, prep;
        std::vector<Request *> active;     // this lane's worker
        std::vector<Request *> finishing;  // this lane's worker
        // under mu_
        std::deque<Request *> to_prepare;  // admitted, for the lane's preparation thread
        std::vector<std::pair<Request *, DeviceView>> prepared;
        int preparing = 0;  // admitted, not yet in the batch
        int admitted = 0;   // admitted and not finished (preparing included)
        int pending = 0;    // waiting in pending_
    };

  public:
    Server(double budget, int max_active) : budget_(budget), max_active_(std::max(1, max_active)) {
        hipDeviceProp_t prop;
        if (hipGetDevice(&dev_) == hipSuccess && hipGetDeviceProperties(&prop, dev_) == hipSuccess)
            num_cu_ = prop.multiProcessorCount;
        for (int l = 0; l < kLanes; l++) SCHECK(hipStreamCreateWithFlags(&lanes_[l].stream, hipStreamNonBlocking));
        for (int l = 0; l < kLanes; l++) lanes_[l].worker = std::thread([this, l] { run(l); });
        for (int l = 0; l < kLanes; l++) lanes_[l].prep = std::thread([this, l] { prepare_loop(l); });
    }

    ~Server() {
        {
            std::lock_guard<std::mutex> g(mu_);
            stop_ = true;
            gen_++;
        }
        cv_.notify_all();
        for (Lane &L : lanes_) L.worker.join();
        {
            std::lock_guard<std::mutex> g(mu_);
            stop_prep_ = true;
        }
        cv_.notify_all();
        for (Lane &L : lanes_) L.prep.join();
        if (failed_) fatal(error_);  // a worker thread failed: exit from the caller's thread
        for (Lane &L : lanes_) delete L.engine;
        for (auto &kv : reqs_) delete kv.second;
        for (Lane &L : lanes_) (void)hipStreamDestroy(L.stream);
    }

    int submit(System *s, int nsteps, int nq, const double *charges) {
        if (nsteps < 1) fatal("A request needs at least one solve");
        if (nq < 0 || (long)nq * nsteps > kMaxCharges)
            fatal("Too many point charges in a request (at most " + std::to_string(kMaxCharges) + " in all steps)");
        Request *r = new Request();
        r->sys = s;
        r->nsteps = nsteps;
        r->nq = nq;
        r->lane = s->npol < kSmallLaneSites ? 0 : 1;
        r->charges.assign(charges, charges + 4L * nq * nsteps);
        r->latency_ms.assign(nsteps, 0.0);
        r->t_done.assign(nsteps, 0.0);
        r->outer.assign(nsteps, 0);
        r->inner.assign(nsteps, 0);
        {
            std::unique_lock<std::mutex> g(mu_);
            check_failed(g);
            if (in_service_.count(s)) fatal("A system can be in only one pending or active request");
            in_service_.insert(s);
            r->id = next_id_++;
            r->t_arrival = now_s();
            reqs_[r->id] = r;
            pending_.push_back(r);
            lanes_[r->lane].pending++;
            gen_++;
        }
        cv_.notify_all();
        return r->id;
    }

    void wait(int id) {
        std::unique_lock<std::mutex> lk(mu_);
        cv_.wait(lk, [&] {
            if (failed_) return true;
            if (id >= 0) {
                auto it = reqs_.find(id);
                return it == reqs_.end() || it->second->done;
            }
            for (auto &kv : reqs_)
                if (!kv.second->done) return false;
            return true;
        });
        check_failed(lk);
    }

    int result(int id, double *energies, double *latency_ms, int *iters) {
        std::unique_lock<std::mutex> g(mu_);
        check_failed(g);
        auto it = reqs_.find(id);
        if (it == reqs_.end() || !it->second->done) return -1;
        Request &r = *it->second;
        for (int k = 0; k < r.nsteps; k++) {
            if (energies) energies[k] = r.epin_copy[k];
            if (latency_ms) latency_ms[k] = r.latency_ms[k];
            if (iters) {
                iters[2 * k] = r.outer[k];
                iters[2 * k + 1] = r.inner[k];
            }
        }
        return r.status;
    }

    void stats(double *out) {
        std::lock_guard<std::mutex> g(mu_);
        out[0] = (double)ticks_;
        out[1] = (double)solves_.load();
        out[2] = ticks_ ? (double)col_ticks_ / (double)ticks_ : 0.0;  // mean columns in flight per tick
        out[3] = (double)max_active_seen_;
        out[4] = (double)peak_bytes_;
        out[5] = (double)deferred_;
    }

  private:
    // A failure on a worker or the preparation thread (a fatal error there is
    // thrown as FatalError) is recorded, all threads stop, and every waiting
    // caller wakes up to report it.
    void fail(const std::string &msg) {
        {
            std::lock_guard<std::mutex> g(mu_);
            if (!failed_) error_ = msg;
            failed_ = true;
            gen_++;
        }
        cv_.notify_all();
    }

    // Called with mu_ held by the caller's thread: exits the process with the
    // worker's message after a failure.
    void check_failed(std::unique_lock<std::mutex> &lk) {
        if (!failed_) return;
        const std::string msg = error_;
        lk.unlock();
        fatal(msg);
    }

    // Requests leave the queue in arrival order while their lane has a free
    // slot (admitted + in preparation < max_active) and the budget allows; a
    // request whose lane is full is passed over, one that does not fit the
    // budget stops admission for both lanes until memory is released (with
    // nothing admitted anywhere it is admitted anyway).  The estimated
    // footprint is held from admission on.
    void admit_locked(std::unique_lock<std::mutex> &lk) {
        for (auto it = pending_.begin(); it != pending_.end();) {
            Request *r = *it;
            Lane &L = lanes_[r->lane];
            if (L.admitted >= max_active_) {
                ++it;
                continue;
            }
            const int nsets = r->sys->n_ipd;
            const size_t fp = device_footprint_estimate(*r->sys) +
                              BatchEngine::segment_bytes(r->sys->npol, 1, nsets, false, num_cu_, nsets > 1);
            if (used_ + fp > budget_ && total_admitted_ > 0) {
                deferred_++;
                return;  // wait for memory to be released
            }
            it = pending_.erase(it);
            L.pending--;
            if (r->sys->npol == 0) {
                // nothing to solve: every step has zero energy
                r->epin_copy.assign(r->nsteps, 0.0);
                lk.unlock();
                finish_now(r);
                lk.lock();
                it = pending_.begin();  // the queue may have changed meanwhile
                continue;
            }
            r->footprint = fp;
            used_ += fp;
            L.preparing++;
            L.admitted++;
            total_admitted_++;
            if (total_admitted_ > max_active_seen_) max_active_seen_ = total_admitted_;
            L.to_prepare.push_back(r);
            cv_.notify_all();
        }
    }

    // Preparation thread of lane l: device state and static field of each
    // admitted system (device_view synchronizes the system's own stream), so
    // that a small system is not prepared behind a large one.
    void prepare_loop(int l) {
        FatalToException guard;
        try {
            prepare_body(l);
        } catch (const FatalError &e) {
            fail(e.msg);
        }
    }

    void prepare_body(int l) {
        SCHECK(hipSetDevice(dev_));  // the creating thread's device
        Lane &L = lanes_[l];
        std::unique_lock<std::mutex> lk(mu_);
        for (;;) {
            cv_.wait(lk, [&] { return stop_prep_ || failed_ || !L.to_prepare.empty(); });
            if (failed_ || L.to_prepare.empty()) break;
            Request *r = L.to_prepare.front();
            L.to_prepare.pop_front();
            lk.unlock();
            DeviceView v = device_view(*r->sys);
            lk.lock();
            L.prepared.push_back({r, v});
            gen_++;
            cv_.notify_all();
        }
    }

    // Worker of lane l: prepared systems join its batch.  One that does not fit
    // (no free segment or column range) goes back to the front of the queue.
    void add_prepared_locked(int l, std::unique_lock<std::mutex> &lk) {
        Lane &L = lanes_[l];
        if (L.prepared.empty()) return;
        std::vector<std::pair<Request *, DeviceView>> ready;
        ready.swap(L.prepared);
        lk.unlock();
        std::vector<Request *> back;
        for (auto &rv : ready)
            if (!add_to_batch(l, rv.first, rv.second)) back.push_back(rv.first);
        lk.lock();
        L.preparing -= (int)ready.size();
        for (auto it = back.rbegin(); it != back.rend(); ++it) {
            L.admitted--;
            total_admitted_--;
            if (total_admitted_ == 0) fatal("Cannot allocate device memory for a request of the serving loop");
            used_ -= (*it)->footprint;
            (*it)->footprint = 0;
            pending_.push_front(*it);
            L.pending++;
        }
    }

    bool add_to_batch(int l, Request *r, const DeviceView &v) {
        Lane &L = lanes_[l];
        System &s = *r->sys;
        {
            std::lock_guard<std::mutex> g(mu_);
            if (amoeba_ < 0) amoeba_ = s.amoeba ? 1 : 0;
            if ((s.amoeba ? 1 : 0) != amoeba_) fatal("All systems of a server must use the same force field");
        }
        if (!L.engine) L.engine = new BatchEngine(L.stream, s.amoeba, false, max_active_, 3 * max_active_);
        const int seg = L.engine->add(v, 1, s.n_ipd > 1 ? 1 : 0);
        if (seg < 0) return false;
        SegInfo &S = L.engine->segs[seg];
        r->seg = seg;
        {
            std::lock_guard<std::mutex> g(mu_);
            const size_t held = std::max(r->footprint, v.bytes + S.bytes);
            used_ += held - r->footprint;
            r->footprint = held;
            if (used_ > peak_bytes_) peak_bytes_ = used_;
        }
        SCHECK(hipHostMalloc((void **)&r->epin, sizeof(double) * r->nsteps, hipHostMallocDefault));
        SCHECK(hipEventCreateWithFlags(&r->ev, hipEventDisableTiming));
        // pinned source: the copy stays in stream order behind earlier users of
        // this (pooled) memory
        SCHECK(hipHostMalloc((void **)&r->charges_pin, sizeof(double) * std::max<size_t>(r->charges.size(), 1),
                             hipHostMallocDefault));
        std::memcpy(r->charges_pin, r->charges.data(), sizeof(double) * r->charges.size());
        SCHECK(hipMemcpyAsync(S.charges, r->charges_pin, sizeof(double) * r->charges.size(),
                              hipMemcpyHostToDevice, L.stream));
        const long n = 3L * s.npol;
        SCHECK(hipMemsetAsync(S.dev.x, 0, sizeof(double) * S.ncols * n, L.stream));
        start_step(L, r, true);
        r->t_ready = r->t_arrival;
        L.active.push_back(r);
        return true;
    }

    // First step: x = 0, every set solved.  Later steps (x = the last
    // solve's dipoles, same geometry): with two sets, the shared correction
    // column first; with one set, r = b - T x follows from the last FP64
    // residual and the change of b without another FP64 matvec.
    void start_step(Lane &L, Request *r, bool first) {
        BatchEngine &E = *L.engine;
        SegInfo &S = E.segs[r->seg];
        System &s = *r->sys;
        const long n = 3L * s.npol;
        const double4 *q = S.charges + (long)r->step * r->nq;
        r->inner_shared = 0;
        if (first) {
            launch_charges_field(L.stream, s.npol, S.view.ps64, r->nq, q, S.ext);
            E.set_rhs(r->seg);
            E.start(r->seg, s.tol, s.max_iter, kResXZero);
            r->shared = false;
        } else if (S.nsets > 1) {
            double *bs = S.dev.b + (long)S.nsets * n;
            launch_charges_field(L.stream, s.npol, S.view.ps64, r->nq, q, bs);
            launch_batch_shared_rhs(L.stream, n, S.nsets, S.view.e_m2d, S.ext, S.dev.b, bs);
            SCHECK(hipMemsetAsync(S.dev.x + (long)S.nsets * n, 0, sizeof(double) * n, L.stream));
            E.start_cols(r->seg, S.nsets, 1, kSharedTolFrac * s.tol, s.max_iter, kResXZero, 1);
            r->shared = true;
        } else {
            launch_charges_field(L.stream, s.npol, S.view.ps64, r->nq, q, S.ext);
            E.shift_rhs(r->seg);
            E.start(r->seg, s.tol, s.max_iter, kResKnown);
            r->shared = false;
        }
        r->valid_from = E.next_readback();
    }

    // The shared correction is done: add it to every set and check each set
    // against its FP64 residual.
    void feed_shared(Lane &L, Request *r, const BCol *snap) {
        BatchEngine &E = *L.engine;
        SegInfo &S = E.segs[r->seg];
        System &s = *r->sys;
        const long n = 3L * s.npol;
        r->inner_shared = snap[S.col0 + S.nsets].s[0].inner_total;
        launch_batch_feed(L.stream, n, S.nsets, S.dev.x, S.dev.x + (long)S.nsets * n);
        E.start_cols(r->seg, 0, S.nsets, s.tol, s.max_iter, kResMatvec, 0);
        r->shared = false;
        r->valid_from = E.next_readback();
    }

    // Solve r->step converged (seen in a snapshot): energy, next step or finish.
    void complete_step(Lane &L, Request *r, const BCol *snap, double t) {
        BatchEngine &E = *L.engine;
        SegInfo &S = E.segs[r->seg];
        int outer = 0, inner = 0;
        for (int c = 0; c < S.nsets; c++) {
            const BColSub &cs = snap[S.col0 + c].s[0];
            if (cs.status != 0) r->status = 2;
            outer = std::max(outer, cs.outer + 1);
            inner = std::max(inner, cs.inner_total);
        }
        inner += r->inner_shared;
        const int k = r->step;
        r->outer[k] = outer;
        r->inner[k] = inner;
        r->t_done[k] = t;
        r->latency_ms[k] = (t - r->t_ready) * 1e3;
        E.energies(r->seg);
        SCHECK(hipMemcpyAsync(r->epin + k, S.eout, sizeof(double), hipMemcpyDeviceToHost, L.stream));
        solves_++;
        r->step++;
        if (r->step < r->nsteps) {
            start_step(L, r, false);
            r->t_ready = t;
            return;
        }
        // last solve: dipoles back into the system, slot and memory released
        const long n = 3L * r->sys->npol;
        SCHECK(hipMemcpyAsync(S.view.ipd, S.dev.x, sizeof(double) * S.nsets * n, hipMemcpyDeviceToDevice,
                              L.stream));
        SCHECK(hipEventRecord(r->ev, L.stream));
        E.remove(r->seg);
        r->seg = -1;
        {
            std::lock_guard<std::mutex> g(mu_);
            used_ -= r->footprint;
            L.admitted--;
            total_admitted_--;
            gen_++;
        }
        cv_.notify_all();
        L.finishing.push_back(r);
    }

    void finish_now(Request *r) {
        std::lock_guard<std::mutex> g(mu_);
        r->done = true;
        in_service_.erase(r->sys);
        gen_++;
        cv_.notify_all();
    }

    void reap_finished(Lane &L) {
        for (size_t i = 0; i < L.finishing.size();) {
            Request *r = L.finishing[i];
            if (hipEventQuery(r->ev) != hipSuccess) {
                i++;
                continue;
            }
            r->epin_copy.assign(r->epin, r->epin + r->nsteps);
            (void)hipHostFree(r->epin);
            r->epin = nullptr;
            (void)hipHostFree(r->charges_pin);
            r->charges_pin = nullptr;
            (void)hipEventDestroy(r->ev);
            r->ev = nullptr;
            System &s = *r->sys;
            s.ipd_done = true;
            s.ipd_use_guess = true;
            s.host_ipd_stale = true;
            L.finishing.erase(L.finishing.begin() + i);
            finish_now(r);
        }
    }

    void run(int l) {
        FatalToException guard;
        try {
            run_body(l);
        } catch (const FatalError &e) {
            fail(e.msg);
        }
    }

    void run_body(int l) {
        SCHECK(hipSetDevice(dev_));  // the creating thread's device
        Lane &L = lanes_[l];
        long prev = -1;
        std::unique_lock<std::mutex> lk(mu_);
        for (;;) {
            if (failed_) break;
            admit_locked(lk);
            add_prepared_locked(l, lk);
            if (L.active.empty()) {
                lk.unlock();
                if (!L.finishing.empty()) SCHECK(hipStreamSynchronize(L.stream));
                reap_finished(L);
                lk.lock();
                prev = -1;
                if (stop_ && L.pending == 0 && L.finishing.empty() && L.preparing == 0 && L.prepared.empty()) break;
                // nothing to tick: wait until something changes (a request
                // arrives or is prepared, memory is released, stop); a request
                // of this lane that is being prepared wakes us when it is ready
                const long seen = gen_;
                cv_.wait(lk, [&] { return failed_ || !L.prepared.empty() || gen_ != seen; });
                continue;
            }
            lk.unlock();
            L.engine->tick(refine_eta(), refine_frac());
            const long seq = L.engine->enqueue_readback();
            int live = 0;
            if (prev >= 0) {
                const BCol *snap = L.engine->wait_readback(prev);
                const double t = now_s();
                std::vector<Request *> still;
                for (Request *r : L.active) {
                    if (prev < r->valid_from) {
                        still.push_back(r);
                        continue;
                    }
                    const SegInfo &S = L.engine->segs[r->seg];
                    const int c0 = r->shared ? S.nsets : 0, c1 = r->shared ? S.nsets + 1 : S.nsets;
                    bool all = true;
                    for (int c = c0; c < c1 && all; c++) all = snap[S.col0 + c].s[0].phase == kPhaseDone;
                    if (!all) {
                        live += c1 - c0;
                        still.push_back(r);
                        continue;
                    }
                    if (r->shared) {
                        feed_shared(L, r, snap);
                        still.push_back(r);
                        continue;
                    }
                    complete_step(L, r, snap, t);
                    if (r->seg >= 0) still.push_back(r);
                }
                L.active.swap(still);
            }
            prev = seq;
            reap_finished(L);
            lk.lock();
            ticks_++;
            col_ticks_ += live;
        }
    }

    double budget_;
    int max_active_;
    int num_cu_ = 256;
    int dev_ = 0;
    Lane lanes_[kLanes];
    std::mutex mu_;
    std::condition_variable cv_;
    long gen_ = 0;  // bumped (under mu_) on every change a waiting worker may act on
    bool stop_ = false;
    bool failed_ = false;
    std::string error_;
    int next_id_ = 0;
    int amoeba_ = -1;
    std::map<int, Request *> reqs_;
    std::deque<Request *> pending_;  // arrival order, both lanes
    std::set<System *> in_service_;
    size_t used_ = 0;
    int total_admitted_ = 0;
    bool stop_prep_ = false;  // under mu_
    size_t peak_bytes_ = 0;
    long ticks_ = 0, col_ticks_ = 0, deferred_ = 0;  // under mu_
    std::atomic<long> solves_{0};
    int max_active_seen_ = 0;
};

// C-ABI glue (capi.cpp)
void *server_create(double budget, int max_active) { return new Server(budget, max_active); }
void server_destroy(void *p) { delete static_cast<Server *>(p); }
int server_submit(void *p, System *s, int nsteps, int nq, const double *charges) {
    return static_cast<Server *>(p)->submit(s, nsteps, nq, charges);
}
void server_wait(void *p, int id) { static_cast<Server *>(p)->wait(id); }
int server_result(void *p, int id, double *e, double *lat, int *iters) {
    return static_cast<Server *>(p)->result(id, e, lat, iters);
}
void server_stats(void *p, double *out) { static_cast<Server *>(p)->stats(out); }

}  // namespace ommp
