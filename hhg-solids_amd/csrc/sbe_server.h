// sbe_server.h -- continuous-batching solve server (see sbe_server.cpp).
#pragma once
#include <cstddef>
#include <memory>
#include <string>

#include "sbe_host.h"

namespace hhg {

enum { REQ_QUEUED = 0, REQ_PREPARING = 1, REQ_PREPARED = 2, REQ_RUNNING = 3, REQ_DONE = 4, REQ_FAILED = 5 };

struct ServerConfig {
    size_t hbm_budget = (size_t)64 << 30;  // bytes of device memory the server may hold
    int max_inflight = 8;                  // solves on the device at once (one HIP stream each)
    int host_threads = 4;                  // host preparation workers
    // admission order among prepared requests: 0 = arrival (FIFO), 1 = least
    // estimated device work first, the estimate divided by 1 + wait / aging_s
    // so that a large request is not passed over indefinitely
    int policy = 1;
    double aging_s = 2.0;
};

struct ServerStats {
    long long submitted = 0, completed = 0, failed = 0, peak_inflight = 0, inflight = 0, queued = 0;
    double ksteps = 0, bytes_in_use = 0, peak_bytes = 0;
};

class Server {
  public:
    explicit Server(const ServerConfig& cfg);
    ~Server();  // finishes everything submitted, then stops
    long long submit(const std::string& path_tb, const SolveParams& p, int precision);
    int wait(long long id, double timeout_s);  // returns the request state
    int state(long long id);
    std::string error(long long id);
    int nt(long long id);
    void result(long long id, double* t, double* J /* [nt][6] */);
    // submit, prepared, admitted, done (steady-clock s), device ms, k-steps, bytes
    void timing(long long id, double* out7);
    void release(long long id);
    ServerStats stats();
    // admission order (ServerConfig::policy, aging_s); applies from the next admission
    void set_policy(int policy, double aging_s);
    struct Impl;

  private:
    std::unique_ptr<Impl> impl_;
};

}  // namespace hhg
