// Serving loop for many steady solves on one GPU (host side); see jm_server.cpp.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
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
    double t_submit = 0, t_start = 0, t_done = 0;
    int steps = 0;
    StepInfo last;
    double resid0[5] = {0, 0, 0, 0, 0};
    std::vector<double> rho, rhou, rhoE;
    std::string error;
  };

  Server(size_t mem_budget, int max_active);
  ~Server();
  Server(const Server &) = delete;
  Server &operator=(const Server &) = delete;

  int submit(const Mesh *mesh, const std::string &joe_in, int precision, const double *rho, const double *rhou,
             const double *rhoE);
  int status(int id);
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
  void retire(Active &a, bool failed, const std::string &err);

  size_t budget_;
  int max_active_;
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
