// sbe_server.cpp -- continuous-batching solve server on one GPU.
//
// Requests (a _tb.dat path, a Settings copy and a precision) flow through
//   submit -> host preparation (thread pool: parse, setup, initial state)
//          -> admission (FIFO; the solve's HBM footprint must fit the budget
//             and a concurrency slot must be free)
//          -> device: upload and launch on the request's own HIP stream,
//             so admitted solves run concurrently and their workgroups share
//             the CUs
//          -> completion (event polled), result kept until released.
// A request that arrives while others are on the GPU is admitted as soon as
// a slot and memory free up; nothing waits for a whole batch to drain.
// Each solve's arithmetic is independent of what else is running, so its
// result is the same alone and in the mix.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "sbe_host.h"
#include "sbe_server.h"

namespace hhg {

namespace {
double now_s() {
    using namespace std::chrono;
    return duration<double>(steady_clock::now().time_since_epoch()).count();
}
}  // namespace

struct ServerRequest {
    long long id = 0;
    std::string path_tb;
    SolveParams params;
    int precision = 0;
    int state = REQ_QUEUED;
    std::string err;
    std::unique_ptr<HostSolve> host;
    DeviceSolve* dev = nullptr;
    size_t bytes = 0;
    double t_submit = 0, t_prepared = 0, t_admit = 0, t_done = 0, device_ms = 0;
    std::vector<double> J;
    long long ksteps = 0;
};

// Device work of a solve in flops of the generic per-k formulation
// 8 (8 nb^3 + 4 ns nb^2 + 48 nb^2) per k-point and RK4 step.
static double work_estimate(const HostSolve& hs) {
    const double nb = hs.d.nb;
    return (double)hs.d.N * hs.d.nt * 8.0 * (8 * nb * nb * nb + 4 * hs.d.ns * nb * nb + 48 * nb * nb);
}

struct Server::Impl {
    ServerConfig cfg;
    std::mutex mu;
    std::condition_variable cv_prep, cv_sched, cv_done;
    std::map<long long, std::shared_ptr<ServerRequest>> reqs;
    std::deque<std::shared_ptr<ServerRequest>> to_prepare, prepared;
    std::vector<std::shared_ptr<ServerRequest>> inflight;
    long long next_id = 1;
    size_t bytes_in_use = 0;
    int preparing = 0;  // requests a worker has taken and not yet handed on
    bool stop = false;
    std::vector<std::thread> workers;
    std::thread scheduler;
    ServerStats stats;

    void prep_loop() {
        for (;;) {
            std::shared_ptr<ServerRequest> r;
            {
                std::unique_lock<std::mutex> lk(mu);
                cv_prep.wait(lk, [&] { return stop || !to_prepare.empty(); });
                if (stop && to_prepare.empty()) return;
                r = to_prepare.front();
                to_prepare.pop_front();
                r->state = REQ_PREPARING;
                ++preparing;
            }
            std::unique_ptr<HostSolve> hs;
            std::string err;
            size_t bytes = 0;
            try {
                TBModel m = read_tb(r->path_tb);
                hs = std::make_unique<HostSolve>(build_solve(m, r->params, false));  // initial state on the device
                bytes = dev_bytes(*hs);
                if (bytes > cfg.hbm_budget) throw std::runtime_error("solve footprint exceeds the server's HBM budget");
            } catch (const std::exception& e) {
                err = e.what();
            }
            {
                std::lock_guard<std::mutex> lk(mu);
                --preparing;
                r->t_prepared = now_s();
                if (!err.empty()) {
                    r->state = REQ_FAILED;
                    r->err = err;
                    r->t_done = r->t_prepared;
                    stats.failed++;
                    cv_done.notify_all();
                } else {
                    r->host = std::move(hs);
                    r->bytes = bytes;
                    r->state = REQ_PREPARED;
                    prepared.push_back(r);
                }
                cv_sched.notify_all();  // new work, or the last preparation of a shutdown ended
            }
        }
    }

    void sched_loop() {
        for (;;) {
            // admit in FIFO order while memory and slots allow
            std::vector<std::shared_ptr<ServerRequest>> admit;
            {
                std::unique_lock<std::mutex> lk(mu);
                // the server is drained when nothing is queued, being prepared,
                // prepared or on the device
                auto drained = [&] {
                    return stop && prepared.empty() && inflight.empty() && to_prepare.empty() && preparing == 0;
                };
                if (inflight.empty()) cv_sched.wait(lk, [&] { return drained() || !prepared.empty(); });
                if (drained()) return;
                size_t used = bytes_in_use;
                int slots = cfg.max_inflight - (int)inflight.size();
                if (cfg.policy == 1 && prepared.size() > 1) {
                    // least estimated work first, with aging
                    const double t = now_s();
                    auto score = [&](const std::shared_ptr<ServerRequest>& r) {
                        return work_estimate(*r->host) / (1.0 + (t - r->t_prepared) / cfg.aging_s);
                    };
                    std::stable_sort(prepared.begin(), prepared.end(),
                                     [&](const auto& a, const auto& b) { return score(a) < score(b); });
                }
                for (auto it = prepared.begin(); it != prepared.end() && slots > 0;) {
                    if (used + (*it)->bytes > cfg.hbm_budget) {
                        if (cfg.policy == 0) break;  // FIFO: wait for memory in order
                        ++it;
                        continue;
                    }
                    auto r = *it;
                    it = prepared.erase(it);
                    used += r->bytes;
                    --slots;
                    r->state = REQ_RUNNING;
                    r->t_admit = now_s();
                    admit.push_back(r);
                }
                bytes_in_use = used;
                stats.peak_inflight = std::max<long long>(stats.peak_inflight, (long long)(inflight.size() + admit.size()));
                stats.peak_bytes = std::max<double>(stats.peak_bytes, (double)used);
            }
            for (auto& r : admit) {
                try {
                    r->dev = dev_create(*r->host);
                    dev_set_shared(r->dev, true);  // runs beside other requests: batch-member launch rules
                    dev_launch(r->dev, r->precision);
                } catch (const std::exception& e) {
                    std::lock_guard<std::mutex> lk(mu);
                    dev_destroy(r->dev);
                    r->dev = nullptr;
                    r->state = REQ_FAILED;
                    r->err = e.what();
                    r->t_done = now_s();
                    bytes_in_use -= r->bytes;
                    stats.failed++;
                    cv_done.notify_all();
                    continue;
                }
                std::lock_guard<std::mutex> lk(mu);
                inflight.push_back(r);
            }
            // completions
            std::vector<std::shared_ptr<ServerRequest>> done;
            {
                std::lock_guard<std::mutex> lk(mu);
                for (auto it = inflight.begin(); it != inflight.end();) {
                    bool fin = false;
                    try {
                        fin = dev_query((*it)->dev);
                    } catch (const std::exception& e) {
                        (*it)->err = e.what();
                        fin = true;
                    }
                    if (fin) {
                        done.push_back(*it);
                        it = inflight.erase(it);
                    } else {
                        ++it;
                    }
                }
            }
            for (auto& r : done) {
                try {
                    if (!r->err.empty()) throw std::runtime_error(r->err);
                    r->J.resize((size_t)6 * r->host->d.nt);
                    r->device_ms = dev_complete(r->dev, r->J.data());
                } catch (const std::exception& e) {
                    r->err = e.what();
                }
                dev_destroy(r->dev);
                r->dev = nullptr;
                std::lock_guard<std::mutex> lk(mu);
                r->t_done = now_s();
                bytes_in_use -= r->bytes;
                if (r->err.empty()) {
                    r->state = REQ_DONE;
                    r->ksteps = (long long)r->host->d.N * r->host->d.nt;
                    stats.completed++;
                    stats.ksteps += (double)r->ksteps;
                } else {
                    r->state = REQ_FAILED;
                    stats.failed++;
                }
                r->host->rho0.clear();
                r->host->rho0.shrink_to_fit();
                r->host->steptab.clear();
                r->host->steptab.shrink_to_fit();
                r->host->psteps.clear();
                r->host->psteps.shrink_to_fit();
                cv_done.notify_all();
            }
            if (!inflight.empty() && done.empty()) std::this_thread::sleep_for(std::chrono::microseconds(200));
        }
    }
};

Server::Server(const ServerConfig& cfg) : impl_(new Impl()) {
    impl_->cfg = cfg;
    if (impl_->cfg.max_inflight < 1) impl_->cfg.max_inflight = 1;
    if (impl_->cfg.host_threads < 1) impl_->cfg.host_threads = 1;
    for (int i = 0; i < impl_->cfg.host_threads; ++i) impl_->workers.emplace_back([this] { impl_->prep_loop(); });
    impl_->scheduler = std::thread([this] { impl_->sched_loop(); });
}

Server::~Server() {
    {
        std::lock_guard<std::mutex> lk(impl_->mu);
        impl_->stop = true;
    }
    impl_->cv_prep.notify_all();
    impl_->cv_sched.notify_all();
    for (auto& t : impl_->workers) t.join();
    impl_->cv_sched.notify_all();
    impl_->scheduler.join();
}

long long Server::submit(const std::string& path_tb, const SolveParams& p, int precision) {
    auto r = std::make_shared<ServerRequest>();
    r->path_tb = path_tb;
    r->params = p;
    r->precision = precision;
    r->t_submit = now_s();
    std::lock_guard<std::mutex> lk(impl_->mu);
    if (impl_->stop) throw std::runtime_error("server is shutting down");
    r->id = impl_->next_id++;
    impl_->reqs[r->id] = r;
    impl_->to_prepare.push_back(r);
    impl_->stats.submitted++;
    impl_->cv_prep.notify_one();
    return r->id;
}

void Server::set_policy(int policy, double aging_s) {
    if (policy != 0 && policy != 1) throw std::runtime_error("admission policy must be 0 (arrival order) or 1 (least work first)");
    if (!(aging_s > 0)) throw std::runtime_error("aging time must be positive");
    std::lock_guard<std::mutex> lk(impl_->mu);
    impl_->cfg.policy = policy;
    impl_->cfg.aging_s = aging_s;
}

static std::shared_ptr<ServerRequest> find(Server::Impl* im, long long id) {
    auto it = im->reqs.find(id);
    if (it == im->reqs.end()) throw std::runtime_error("unknown request id " + std::to_string(id));
    return it->second;
}

int Server::wait(long long id, double timeout_s) {
    std::unique_lock<std::mutex> lk(impl_->mu);
    auto r = find(impl_.get(), id);
    auto pred = [&] { return r->state == REQ_DONE || r->state == REQ_FAILED; };
    if (timeout_s < 0)
        impl_->cv_done.wait(lk, pred);
    else if (!impl_->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), pred))
        return r->state;
    return r->state;
}

int Server::state(long long id) {
    std::lock_guard<std::mutex> lk(impl_->mu);
    return find(impl_.get(), id)->state;
}

std::string Server::error(long long id) {
    std::lock_guard<std::mutex> lk(impl_->mu);
    return find(impl_.get(), id)->err;
}

int Server::nt(long long id) {
    std::lock_guard<std::mutex> lk(impl_->mu);
    auto r = find(impl_.get(), id);
    return r->host ? r->host->d.nt : (int)(r->params.tmax / r->params.dt);
}

void Server::result(long long id, double* t, double* J) {
    std::lock_guard<std::mutex> lk(impl_->mu);
    auto r = find(impl_.get(), id);
    if (r->state != REQ_DONE) throw std::runtime_error("request not done: " + (r->err.empty() ? std::string("pending") : r->err));
    const int nt = r->host->d.nt;
    if (t) std::copy(r->host->t.begin(), r->host->t.end(), t);
    if (J) std::copy(r->J.begin(), r->J.begin() + 6 * (size_t)nt, J);
}

void Server::timing(long long id, double* out) {
    std::lock_guard<std::mutex> lk(impl_->mu);
    auto r = find(impl_.get(), id);
    out[0] = r->t_submit;
    out[1] = r->t_prepared;
    out[2] = r->t_admit;
    out[3] = r->t_done;
    out[4] = r->device_ms;
    out[5] = (double)r->ksteps;
    out[6] = (double)r->bytes;
}

void Server::release(long long id) {
    std::lock_guard<std::mutex> lk(impl_->mu);
    auto r = find(impl_.get(), id);
    if (r->state != REQ_DONE && r->state != REQ_FAILED) throw std::runtime_error("request still in progress");
    impl_->reqs.erase(id);
}

ServerStats Server::stats() {
    std::lock_guard<std::mutex> lk(impl_->mu);
    ServerStats s = impl_->stats;
    s.bytes_in_use = (double)impl_->bytes_in_use;
    s.inflight = (long long)impl_->inflight.size();
    s.queued = (long long)(impl_->to_prepare.size() + impl_->prepared.size());
    return s;
}

}  // namespace hhg
