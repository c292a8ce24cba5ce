// Serving loop for many steady solves on one GPU (host side); see jm_server.cpp.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "jm_ctx.hpp"
#include "jm_mesh.hpp"

namespace jm {

class Server {
 public:
  enum Status { kQueued = 0, kRunning = 1, kDone = 2, kFailed = 3 };

  struct Request {
    int id = 0;
    const Mesh *mesh = nullptr;  // owned by the caller until the solve is released
    std::string joe_in;
    int precision = 0;
    size_t footprint = 0;
    std::vector<double> init;  // optional initial state (rho, rhou[3], rhoE), reference numbering
    int status = kQueued;
    int fp64_retries = 0;     // 1 once a mixed-precision solve that went non-finite was queued again in FP64
    int steps_abandoned = 0;  // steps of that abandoned mixed-precision attempt
    double t_submit = 0, t_start = 0, t_done = 0;
    std::atomic<int> steps{0};  // written by a stepping thread, read by progress()
    StepInfo last;
    double resid0[5] = {0, 0, 0, 0, 0};
    std::vector<double> rho, rhou, rhoE;
    std::string error;
  };

  // batched: every active solve steps in lockstep on one stream, their
  // linear solves in one launch (one workgroup per solve); otherwise each
  // solve steps on its own stream with the graph-captured device step
  // fp64_retry: a mixed-precision solve whose residual turns non-finite is
  // queued again once, at the front, in FP64 from its initial state
  Server(size_t mem_budget, int max_active, bool batched = false, bool fp64_retry = true);
  ~Server();
  Server(const Server &) = delete;
  Server &operator=(const Server &) = delete;

  int submit(const Mesh *mesh, const std::string &joe_in, int precision, const double *rho, const double *rhou,
             const double *rhoE);
  int status(int id);
  // pseudo-time steps the solve has taken so far
  int progress(int id);
  void wait(int id);
  std::shared_ptr<Request> finished(int id);
  void release(int id);

 private:
  struct Active {
    std::shared_ptr<Request> req;
    std::unique_ptr<Ctx> ctx;
    hipStream_t stream = nullptr;
    bool inflight = false;
    bool dead = false;
    bool captured = false;
  };
  void admit_loop();
  void step_loop();
  void batch_loop(int group);
  void retire(Active &a, bool failed, const std::string &err);
  // the fp64_retry path for a solve that has stopped on step s: true when the
  // solve was queued again (a's context and slot are then released)
  bool retry_in_fp64(Active &a, const StepInfo &s);
  void release_ctx(Active &a);

  size_t budget_;
  int max_active_;
  bool batched_ = false;
  bool fp64_retry_ = true;
  // batched mode: ngroups_ lockstep groups (JOEMI_SERVER_GROUPS, 1 to
  // kMaxGroups), each with its own stepping thread, stream, solver jobs
  // (pinned) and batched-BiCGStab buffers, so one group's kernels fill the
  // gaps of the others (its host polls and synchronisations)
  static constexpr int kMaxGroups = 4;
  int ngroups_ = 2;
  hipStream_t bs_[kMaxGroups] = {};
  KrylovJob *jobs_h_[kMaxGroups] = {};
  KrylovBatchBuf kbuf64_[kMaxGroups], kbuf32_[kMaxGroups];
  std::vector<std::thread> group_threads_;
  std::vector<Active> groups_active_[kMaxGroups];  // each group's stepping thread only
  int device_ = 0;
  std::mutex mu_;
  // held while a context is built (synchronous uploads) and while a solve's
  // first step is captured into a graph: a capture must not overlap another
  // thread's synchronous HIP calls
  std::mutex hip_mu_;
  std::condition_variable cv_;
  bool stop_ = false;
  int next_id_ = 1;
  std::map<int, std::shared_ptr<Request>> reqs_;
  std::deque<int> queue_;
  std::vector<Active> ready_;   // built by the admission thread, picked up by the stepping thread
  std::vector<Active> active_;  // stepping thread only
  int slots_used_ = 0;
  size_t used_ = 0;
  std::thread admit_thread_, step_thread_;
};

}  // namespace jm
