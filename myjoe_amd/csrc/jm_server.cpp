// Serving loop for many steady solves on one GPU: requests are admitted as
// long as their device footprint fits the memory budget and an active slot
// is free; every active solve keeps its state and Jacobian resident, runs on
// its own HIP stream and advances by graph-captured pseudo-time steps, so the
// steps of different solves overlap on the device. A solve retires when the
// reference's stop rule (doneSolver, plus RESID_ENERG_REL) says so.
//
// Two host threads: an admission thread builds contexts (host numbering,
// uploads, initial state), a stepping thread launches and collects steps.
// Each solve's kernels touch only its own buffers, so its result is the same
// bits alone or in any mix.
//
// Batched mode: the stepping thread advances every active solve by one step
// per round, on the group's stream: each kernel of the step before the
// linear solve launched once for all solves of a kernel configuration
// (StepBatch::pre), the block-Jacobi BiCGStab of all of them (each kernel
// launched once for all solves, nine per iteration whatever their number,
// with a host poll between chunks of iterations), then the update and
// refresh, again one launch per kernel (StepBatch::post), and one
// synchronisation. The tables are rebuilt only when a solve joins or
// leaves. A solve's result is the same bits as its ordinary step's.
//
// A mixed-precision solve (FP32 Jacobian and Krylov vectors) can go
// non-finite where the FP64 one does not: a linear solve that ends at its
// iteration limit keeps its last iterate (as the reference's BiCGStab does),
// and in FP32 that iterate has been far off often enough to push a transonic
// solve to negative pressure. Such a solve is queued again once, at the
// front, in FP64 from its initial state; its report says so.
#include "jm_server.hpp"
#include "jm_param.hpp"

#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <stdexcept>

namespace jm {

namespace {
double now_s() {
  using namespace std::chrono;
  return duration<double>(steady_clock::now().time_since_epoch()).count();
}
}  // namespace

Server::Server(size_t mem_budget, int max_active, bool batched, bool fp64_retry)
    : budget_(mem_budget), max_active_(max_active), batched_(batched), fp64_retry_(fp64_retry) {
  if (max_active < 1) throw std::invalid_argument("max_active must be at least 1");
  if (hipGetDevice(&device_) != hipSuccess) throw std::runtime_error("no HIP device available");
  if (batched_) {
    if (const char *e = std::getenv("JOEMI_SERVER_GROUPS")) ngroups_ = std::atoi(e);
    ngroups_ = std::max(1, std::min(std::min(ngroups_, kMaxGroups), max_active_));
    bool ok = true;
    for (int g = 0; g < ngroups_ && ok; g++)
      ok = hipStreamCreateWithFlags(&bs_[g], hipStreamNonBlocking) == hipSuccess &&
           hipHostMalloc((void **)&jobs_h_[g], sizeof(KrylovJob) * max_active_) == hipSuccess;
    if (!ok) {
      for (int g = 0; g < ngroups_; g++) {
        if (jobs_h_[g]) (void)hipHostFree(jobs_h_[g]);
        if (bs_[g]) (void)hipStreamDestroy(bs_[g]);
      }
      throw std::runtime_error("HIP resources for batched serving");
    }
  }
  admit_thread_ = std::thread([this] { admit_loop(); });
  if (batched_)
    for (int g = 0; g < ngroups_; g++) group_threads_.emplace_back([this, g] { batch_loop(g); });
  else
    step_thread_ = std::thread([this] { step_loop(); });
}

Server::~Server() {
  {
    std::lock_guard<std::mutex> g(mu_);
    stop_ = true;
  }
  cv_.notify_all();
  admit_thread_.join();
  if (step_thread_.joinable()) step_thread_.join();
  for (auto &t : group_threads_) t.join();
  for (auto &a : active_) {
    a.ctx.reset();
    if (a.stream) (void)hipStreamDestroy(a.stream);
  }
  for (auto &a : ready_) {
    a.ctx.reset();
    if (a.stream) (void)hipStreamDestroy(a.stream);
  }
  for (auto &a : groups_active_)
    for (auto &b : a) {
      b.ctx.reset();
      if (b.stream) (void)hipStreamDestroy(b.stream);
    }
  for (int g = 0; g < ngroups_; g++) {
    if (jobs_h_[g]) (void)hipHostFree(jobs_h_[g]);
    if (bs_[g]) (void)hipStreamDestroy(bs_[g]);
  }
}

int Server::submit(const Mesh *mesh, const std::string &joe_in, int precision, const double *rho,
                   const double *rhou, const double *rhoE) {
  if (!mesh) throw std::invalid_argument("null mesh");
  SolverParams p;
  std::vector<ZoneBc> zbc;
  parse_params(joe_in, *mesh, p, zbc);  // reject bad input at submission
  if (batched_ && p.sa) throw Unsupported("batched serving does not carry the Spalart-Allmaras scalar");
  if (batched_ && p.ilu_sub > 0) throw Unsupported("batched serving runs the block-Jacobi preconditioner");
  if (batched_ && precision != 0 && !p.krylov_recurrence)
    throw Unsupported("batched serving takes the FP32 BiCGStab residual by the recurrence "
                      "(GPU_KRYLOV_RESIDUAL = RECURRENCE)");
  if (batched_ && p.mu_ref <= 0.0 && p.flux_kernel != kResidEntry)
    throw Unsupported("batched serving runs the per-entry flux kernel (GPU_FLUX_KERNEL = ENTRY)");
  auto r = std::make_shared<Request>();
  r->mesh = mesh;
  r->joe_in = joe_in;
  r->precision = precision;
  r->footprint = Ctx::footprint(*mesh, precision);
  if (r->footprint > budget_) throw std::invalid_argument("solve does not fit the server's memory budget");
  if (rho) {
    const size_t n = mesh->ncv;
    r->init.resize(5 * n);
    std::memcpy(r->init.data(), rho, n * sizeof(double));
    std::memcpy(r->init.data() + n, rhou, 3 * n * sizeof(double));
    std::memcpy(r->init.data() + 4 * n, rhoE, n * sizeof(double));
  }
  r->t_submit = now_s();
  std::lock_guard<std::mutex> g(mu_);
  r->id = next_id_++;
  reqs_[r->id] = r;
  queue_.push_back(r->id);
  cv_.notify_all();
  return r->id;
}

int Server::status(int id) {
  std::lock_guard<std::mutex> g(mu_);
  auto it = reqs_.find(id);
  if (it == reqs_.end()) throw std::invalid_argument("unknown ticket");
  return it->second->status;
}

int Server::progress(int id) {
  std::shared_ptr<Request> r;
  {
    std::lock_guard<std::mutex> g(mu_);
    auto it = reqs_.find(id);
    if (it == reqs_.end()) throw std::invalid_argument("unknown ticket");
    r = it->second;
  }
  return r->steps.load();
}

void Server::wait(int id) {
  std::unique_lock<std::mutex> g(mu_);
  auto it = reqs_.find(id);
  if (it == reqs_.end()) throw std::invalid_argument("unknown ticket");
  auto r = it->second;
  cv_.wait(g, [&] { return r->status >= kDone; });
}

std::shared_ptr<Server::Request> Server::finished(int id) {
  std::lock_guard<std::mutex> g(mu_);
  auto it = reqs_.find(id);
  if (it == reqs_.end()) throw std::invalid_argument("unknown ticket");
  if (it->second->status < kDone) throw std::logic_error("solve not finished");
  return it->second;
}

void Server::release(int id) {
  std::lock_guard<std::mutex> g(mu_);
  auto it = reqs_.find(id);
  if (it == reqs_.end()) throw std::invalid_argument("unknown ticket");
  if (it->second->status < kDone) throw std::logic_error("solve not finished");
  reqs_.erase(it);
}

// admission: FIFO; the head waits until its footprint fits and a slot is free
void Server::admit_loop() {
  (void)hipSetDevice(device_);
  for (;;) {
    std::shared_ptr<Request> r;
    {
      std::unique_lock<std::mutex> g(mu_);
      cv_.wait(g, [&] {
        if (stop_) return true;
        if (queue_.empty()) return false;
        auto h = reqs_[queue_.front()];
        return slots_used_ < max_active_ && used_ + h->footprint <= budget_;
      });
      if (stop_) return;
      r = reqs_[queue_.front()];
      queue_.pop_front();
      slots_used_++;
      used_ += r->footprint;
      r->status = kRunning;
    }
    Active a;
    a.req = r;
    try {
      std::lock_guard<std::mutex> hg(hip_mu_);
      if (hipStreamCreateWithFlags(&a.stream, hipStreamNonBlocking) != hipSuccess)
        throw std::runtime_error("HIP stream creation failed");
      a.ctx.reset(new Ctx(*r->mesh, r->joe_in, r->precision, a.stream));
      if (r->init.empty()) {
        a.ctx->init_uniform();
      } else {
        const size_t n = r->mesh->ncv;
        a.ctx->set_state(r->init.data(), r->init.data() + n, r->init.data() + 4 * n);
        if (r->precision == 0 || !fp64_retry_) {  // a mixed solve may start again from it
          r->init.clear();
          r->init.shrink_to_fit();
        }
      }
      if (r->t_start == 0) r->t_start = now_s();  // a retry's latency counts from the first admission
      std::lock_guard<std::mutex> g(mu_);
      ready_.push_back(std::move(a));
    } catch (std::exception &e) {
      a.ctx.reset();
      if (a.stream) (void)hipStreamDestroy(a.stream);
      std::lock_guard<std::mutex> g(mu_);
      r->error = e.what();
      r->status = kFailed;
      r->t_done = now_s();
      slots_used_--;
      used_ -= r->footprint;
      cv_.notify_all();
    }
  }
}

void Server::retire(Active &a, bool failed, const std::string &err) {
  auto r = a.req;
  if (!failed) {
    try {
      const size_t n = r->mesh->ncv;
      r->rho.resize(n);
      r->rhou.resize(3 * n);
      r->rhoE.resize(n);
      a.ctx->get_state(r->rho.data(), r->rhou.data(), r->rhoE.data());
    } catch (std::exception &e) {
      failed = true;
      r->error = e.what();
    }
  } else {
    r->error = err;
  }
  release_ctx(a);
  std::lock_guard<std::mutex> g(mu_);
  r->t_done = now_s();
  r->status = failed ? kFailed : kDone;
  r->init.clear();
  r->init.shrink_to_fit();
  slots_used_--;
  used_ -= r->footprint;
  cv_.notify_all();
}

void Server::release_ctx(Active &a) {
  std::lock_guard<std::mutex> hg(hip_mu_);
  a.ctx.reset();
  if (a.stream) (void)hipStreamDestroy(a.stream);
  a.stream = nullptr;
}

bool Server::retry_in_fp64(Active &a, const StepInfo &s) {
  auto r = a.req;
  if (!fp64_retry_ || r->precision == 0 || r->fp64_retries > 0) return false;
  bool finite = true;
  for (int k = 0; k < 5; k++) finite = finite && std::isfinite(s.resid[k]);
  if (finite) return false;
  const size_t fp = Ctx::footprint(*r->mesh, 0);
  if (fp > budget_) return false;  // the FP64 solve could never be admitted
  release_ctx(a);
  std::lock_guard<std::mutex> g(mu_);
  slots_used_--;
  used_ -= r->footprint;
  r->precision = 0;
  r->footprint = fp;
  r->fp64_retries = 1;
  r->steps_abandoned = s.step;
  r->steps.store(0);
  for (int k = 0; k < 5; k++) r->resid0[k] = 0.0;
  r->status = kQueued;
  queue_.push_front(r->id);
  cv_.notify_all();
  return true;
}

void Server::step_loop() {
  (void)hipSetDevice(device_);
  for (;;) {
    {
      std::unique_lock<std::mutex> g(mu_);
      if (stop_) return;
      for (auto &a : ready_) active_.push_back(std::move(a));
      ready_.clear();
      if (active_.empty()) {
        cv_.wait_for(g, std::chrono::milliseconds(2), [&] { return stop_ || !ready_.empty(); });
        continue;
      }
    }
    bool progress = false;
    for (auto &a : active_) {
      if (a.inflight) continue;
      try {
        if (!a.captured) {
          std::lock_guard<std::mutex> hg(hip_mu_);
          a.ctx->device_step_async();
          a.captured = true;
        } else {
          a.ctx->device_step_async();
        }
        a.inflight = true;
        progress = true;
      } catch (std::exception &e) {
        a.dead = true;
        retire(a, true, e.what());
      }
    }
    for (auto &a : active_) {
      if (a.dead || !a.inflight) continue;
      try {
        if (!a.ctx->device_step_ready()) continue;
        StepInfo s = a.ctx->device_step_result();
        a.inflight = false;
        progress = true;
        auto &r = *a.req;
        r.steps.store(s.step);
        r.last = s;
        if (s.step == 1)
          for (int k = 0; k < 5; k++) r.resid0[k] = s.resid[k];
        if (a.ctx->done_solver(s)) {
          a.dead = true;
          if (!retry_in_fp64(a, s)) retire(a, false, "");
        }
      } catch (std::exception &e) {
        a.dead = true;
        retire(a, true, e.what());
      }
    }
    active_.erase(std::remove_if(active_.begin(), active_.end(), [](const Active &a) { return a.dead; }),
                  active_.end());
    if (!progress) std::this_thread::sleep_for(std::chrono::microseconds(20));
  }
}

void Server::batch_loop(int group) {
  (void)hipSetDevice(device_);
  hipStream_t bs = bs_[group];
  KrylovJob *jobs_h = jobs_h_[group];
  KrylovBatchBuf &kb64 = kbuf64_[group], &kb32 = kbuf32_[group];
  std::vector<Active> &active_ = groups_active_[group];
  std::vector<std::unique_ptr<StepBatch>> steps;  // one per kernel configuration present
  int nsteps = 0;
  const int cap = (max_active_ + ngroups_ - 1) / ngroups_;
  std::vector<int> members;  // request ids the tables were built for
  auto fail = [&](Active &a, const char *what) {
    a.dead = true;
    retire(a, true, what);
  };
  for (;;) {
    {
      std::unique_lock<std::mutex> g(mu_);
      if (stop_) break;
      while (!ready_.empty() && (int)active_.size() < cap) {  // each group takes up to its share
        active_.push_back(std::move(ready_.front()));
        ready_.erase(ready_.begin());
      }
      if (active_.empty()) {
        cv_.wait_for(g, std::chrono::milliseconds(2), [&] { return stop_ || !ready_.empty(); });
        continue;
      }
    }
    for (auto &a : active_) {
      if (a.captured) continue;
      try {
        std::lock_guard<std::mutex> hg(hip_mu_);
        a.ctx->batch_prepare();
        a.captured = true;
      } catch (std::exception &e) {
        fail(a, e.what());
      }
    }
    active_.erase(std::remove_if(active_.begin(), active_.end(), [](const Active &a) { return a.dead; }),
                  active_.end());
    if (active_.empty()) continue;
    std::vector<int> now;
    for (auto &a : active_) now.push_back(a.req->id);
    try {
      if (now != members) {
        // the step batches (one per kernel configuration) and the BiCGStab
        // batches (FP64 solves, then FP32 ones) of the new members
        members.clear();
        std::vector<uint32_t> keys;
        for (auto &a : active_)
          if (std::find(keys.begin(), keys.end(), a.ctx->batch_key()) == keys.end())
            keys.push_back(a.ctx->batch_key());
        std::lock_guard<std::mutex> hg(hip_mu_);
        while (steps.size() < keys.size()) steps.emplace_back(new StepBatch);
        nsteps = (int)keys.size();
        for (int k = 0; k < nsteps; k++) {
          std::vector<Ctx *> cs;
          for (auto &a : active_)
            if (a.ctx->batch_key() == keys[k]) cs.push_back(a.ctx.get());
          steps[k]->set(cs.data(), (int)cs.size(), bs);
        }
        int n64 = 0, n32 = 0;
        for (auto &a : active_)
          if (a.ctx->precision() == 0) a.ctx->batch_job(jobs_h[n64++]);
        for (auto &a : active_)
          if (a.ctx->precision() != 0) a.ctx->batch_job(jobs_h[n64 + n32++]);
        setup_krylov_batch(jobs_h, n64, bs, kb64);
        setup_krylov_batch(jobs_h + n64, n32, bs, kb32);
        members = now;
      }
      for (int k = 0; k < nsteps; k++) steps[k]->pre(bs);
      run_krylov_batch(0, bs, kb64);
      run_krylov_batch(1, bs, kb32);
      for (int k = 0; k < nsteps; k++) steps[k]->post(bs);
      if (hipStreamSynchronize(bs) != hipSuccess) throw std::runtime_error("HIP error in a batched step");
    } catch (std::exception &e) {
      members.clear();
      for (auto &a : active_) fail(a, e.what());
      active_.clear();
      continue;
    }
    for (auto &a : active_) {
      try {
        StepInfo s = a.ctx->batch_result();
        auto &r = *a.req;
        r.steps.store(s.step);
        r.last = s;
        if (s.step == 1)
          for (int k = 0; k < 5; k++) r.resid0[k] = s.resid[k];
        if (a.ctx->done_solver(s)) {
          a.dead = true;
          if (!retry_in_fp64(a, s)) retire(a, false, "");
        }
      } catch (std::exception &e) {
        fail(a, e.what());
      }
    }
    const size_t before = active_.size();
    active_.erase(std::remove_if(active_.begin(), active_.end(), [](const Active &a) { return a.dead; }),
                  active_.end());
    // a solve queued again in FP64 comes back under the same id with a new
    // context: the tables must be rebuilt even if the id list matches again
    if (active_.size() != before) members.clear();
  }
  (void)hipStreamSynchronize(bs);
}

}  // namespace jm
