// Many independent instances driven concurrently: the batched-sweep entry
// points of the C ABI.
//
// pf_batch_minimize_fused steps all instances through their minimize
// iterations together: every phase between two waits is recorded per
// instance (pf_launch.hpp) and replayed as one batched launch per kernel, so
// a stage of the iteration is one launch over all instances of a grid size
// instead of one launch each; every instance still computes exactly what it
// computes alone.
//
// pf_batch_step and pf_batch_minimize run each instance on its own stream:
// each instance is a pf_handle with its own HIP stream, so the kernels of
// different instances run concurrently on the GPU while one host thread per
// worker waits on an instance's host round trips (residual tests, Anderson
// coefficients, minimize multipliers).  Workers take instances from a shared
// counter in the caller's order (largest first balances the tail).  Every
// instance computes exactly what it computes alone: nothing is shared but
// the device.
#include <algorithm>
#include <atomic>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "../../include/pfcore.h"
#include "pf_error.hpp"
#include "pf_handle.hpp"
#include "pf_launch.hpp"


namespace {

struct FirstError {
  std::atomic<bool> set{false};
  int code = PF_OK;
  std::string msg;
  void record(int rc, int instance) {
    bool expected = false;
    if (set.compare_exchange_strong(expected, true)) {
      char buf[1024];
      pf_last_error(buf, sizeof buf);
      code = rc;
      msg = "instance " + std::to_string(instance) + ": " + buf;
    }
  }
};

template <typename F>
int run_all(pf_batch* b, F&& work) {
  const int n = (int)b->h.size();
  std::atomic<int> next{0};
  FirstError err;
  auto worker = [&]() {
    for (;;) {
      const int i = next.fetch_add(1);
      if (i >= n || err.set.load()) return;
      const int rc = work(i);
      if (rc != PF_OK) err.record(rc, i);
    }
  };
  const int t = std::max(1, std::min(b->threads, n));
  std::vector<std::thread> pool;
  for (int k = 1; k < t; ++k) pool.emplace_back(worker);
  worker();
  for (auto& th : pool) th.join();
  if (err.set.load()) return pf::fail(err.code, err.msg);
  return PF_OK;
}

// The fused runner's device resources: a stream of default priority, the
// replayer's table staging, and the gather buffer of every instance's device
// scalars (one copy back per wait).
struct RecScope {   // the calling thread records into r while in scope
  explicit RecScope(pf::Recorder* r) {
    r->clear();
    pf::tl_rec = r;
  }
  ~RecScope() { pf::tl_rec = nullptr; }
};

#define PF_FHIP(call)                                                                  \
  do {                                                                                 \
    hipError_t e_ = (call);                                                            \
    if (e_ != hipSuccess)                                                              \
      return pf::fail(PF_E_HIP, std::string(#call) + ": " + hipGetErrorString(e_));  \
  } while (0)

struct Fused {
  pf_batch* b = nullptr;
  int n = 0, ns = 0;
  std::vector<pf::EngineBase*> e;
  std::vector<pf::Recorder> rec;
  hipStream_t stream = nullptr;
  pf::BatchReplayer rep;
  double* gdev = nullptr;
  double* ghost = nullptr;
  ~Fused() {
    if (stream) (void)hipStreamSynchronize(stream);
    if (gdev) (void)hipFree(gdev);
    if (ghost) (void)hipHostFree(ghost);
    if (stream) (void)hipStreamDestroy(stream);
  }
  int init(pf_batch* batch) {
    b = batch;
    n = (int)b->h.size();
    e.resize(n);
    rec.resize(n);
    for (int i = 0; i < n; ++i) e[i] = b->h[i]->eng.get();
    ns = e[0]->mb_nscal();
    PF_FHIP(hipStreamCreate(&stream));
    PF_FHIP(hipMalloc((void**)&gdev, (size_t)n * ns * sizeof(double)));
    PF_FHIP(hipHostMalloc((void**)&ghost, (size_t)n * ns * sizeof(double), hipHostMallocDefault));
    rep.use_graphs = b->graphs != 0;
    b->fused_launches = b->fused_recorded = 0;
    return PF_OK;
  }
  void gather(int i) { e[i]->mb_gather(gdev + (size_t)i * ns); }
  // replay the recorders of `set` as batched launches, copy every instance's
  // scalars back and mirror them on the host
  int replay(const std::vector<int>& set) {
    std::vector<const pf::Recorder*> rs;
    for (int i : set) {
      rs.push_back(&rec[i]);
      b->fused_recorded += (long long)rec[i].recs.size();
    }
    const unsigned long long before = rep.launches;
    const int r = rep.run(stream, rs);
    if (r == -2) {   // phases of different shape: one by one, in order
      for (const pf::Recorder* q : rs) {
        q->replay(stream);
        b->fused_launches += (long long)q->recs.size();
      }
    } else if (r) {
      return pf::fail(PF_E_HIP, "fused batch: staging of the launch tables failed");
    } else {
      b->fused_launches += (long long)(rep.launches - before);
    }
    PF_FHIP(hipGetLastError());
    PF_FHIP(hipMemcpyAsync(ghost, gdev, (size_t)n * ns * sizeof(double), hipMemcpyDeviceToHost, stream));
    PF_FHIP(hipStreamSynchronize(stream));
    for (int i : set) e[i]->mb_scatter(ghost + (size_t)i * ns);
    return PF_OK;
  }
  // the accelerated solves that still run take two iterations more per wait
  int solve_loop(const std::vector<int>& act) {
    int rc;
    for (;;) {
      std::vector<int> more;
      for (int i : act) {
        RecScope sc(&rec[i]);
        int m = 0;
        if ((rc = e[i]->mb_solve_more(&m))) return rc;
        if (m) {
          gather(i);
          more.push_back(i);
        }
      }
      if (more.empty()) return PF_OK;
      if ((rc = replay(more))) return rc;
    }
  }
  void finish() {
    b->fused_graphs_built = (long long)rep.graphs_built;
    b->fused_graph_launches = (long long)rep.graph_launches;
  }
};

int fused_minimize(pf_batch* b, const pf_minimize_opts* opts, pf_minimize_stats* stats) {
  const int n = (int)b->h.size();
  int rc;
  for (int i = 0; i < n; ++i) {
    pf::EngineBase* e = b->h[i]->eng.get();
    if ((rc = pf_minimize_begin(b->h[i], &opts[i]))) return rc;
    if ((rc = e->mb_check())) {
      char buf[512];
      pf_last_error(buf, sizeof buf);
      return pf::fail(rc, "pf_batch_minimize_fused: instance " + std::to_string(i) + ": " + buf);
    }
    if ((rc = e->sync())) return rc;
  }
  Fused f;
  if ((rc = f.init(b))) return rc;
  std::vector<pf_minimize_stats> last(n);
  std::vector<int> iters(n, 0);
  for (;;) {
    std::vector<int> act;
    for (int i = 0; i < n; ++i)
      if (iters[i] < opts[i].max_iter && !f.e[i]->mb_converged()) act.push_back(i);
    if (act.empty()) break;
    // phase a: energies, <h>, the start of the solve with a common number of
    // iterations (those past an instance's convergence do nothing)
    int spec = 1;
    for (int i : act) spec = std::max(spec, f.e[i]->mb_spec());
    for (int i : act) {
      RecScope sc(&f.rec[i]);
      if ((rc = f.e[i]->mb_phase_a(spec))) return rc;
      f.gather(i);
    }
    if ((rc = f.replay(act)) || (rc = f.solve_loop(act))) return rc;
    // phase b: the solve's result through the final scalars
    for (int i : act) {
      RecScope sc(&f.rec[i]);
      if ((rc = f.e[i]->mb_phase_b())) return rc;
      f.gather(i);
    }
    if ((rc = f.replay(act))) return rc;
    for (int i : act) {
      if ((rc = f.e[i]->mb_phase_c(&last[i]))) return rc;
      ++iters[i];
    }
  }
  f.finish();
  if (stats)
    for (int i = 0; i < n; ++i) stats[i] = last[i];
  return PF_OK;
}

// nsteps time steps of every instance together: each step's phases are
// recorded per instance and replayed as batched launches; the shift of a
// step's zero mode is recorded with the next step's start.
int fused_step(pf_batch* b, int nsteps) {
  const int n = (int)b->h.size();
  int rc;
  for (int i = 0; i < n; ++i) {
    pf::EngineBase* e = b->h[i]->eng.get();
    if ((rc = e->sb_check())) {
      char buf[512];
      pf_last_error(buf, sizeof buf);
      return pf::fail(rc, "pf_batch_step_fused: instance " + std::to_string(i) + ": " + buf);
    }
    if ((rc = e->sync())) return rc;
  }
  Fused f;
  if ((rc = f.init(b))) return rc;
  std::vector<int> all(n);
  for (int i = 0; i < n; ++i) all[i] = i;
  for (int k = 0; k <= nsteps; ++k) {
    const bool last = k == nsteps;
    int spec = 1;
    for (int i : all) spec = std::max(spec, f.e[i]->mb_spec());
    for (int i : all) {
      RecScope sc(&f.rec[i]);
      if (k > 0 && (rc = f.e[i]->sb_phase_c())) return rc;   // the previous step's shift and bookkeeping
      if (!last && (rc = f.e[i]->sb_phase_a(spec))) return rc;
      f.gather(i);
    }
    if ((rc = f.replay(all))) return rc;
    if (last) break;
    if ((rc = f.solve_loop(all))) return rc;
    int zspec = 1;
    for (int i : all) zspec = std::max(zspec, f.e[i]->sb_zspec());
    for (int i : all) {
      RecScope sc(&f.rec[i]);
      if ((rc = f.e[i]->sb_phase_b(zspec))) return rc;
      f.gather(i);
    }
    if ((rc = f.replay(all))) return rc;
    for (;;) {   // the zero-mode searches that still run take four steps more per wait
      std::vector<int> more;
      for (int i : all) {
        RecScope sc(&f.rec[i]);
        int m = 0;
        if ((rc = f.e[i]->sb_zm_more(&m))) return rc;
        if (m) {
          f.gather(i);
          more.push_back(i);
        }
      }
      if (more.empty()) break;
      if ((rc = f.replay(more))) return rc;
    }
  }
  f.finish();
  return PF_OK;
}

}  // namespace

extern "C" {

int pf_batch_create(int n, pf_handle* const* handles, int threads, pf_batch** out) {
  if (n < 1 || !handles || !out || threads < 1) return pf::fail(PF_E_ARG, "pf_batch_create: bad argument");
  for (int i = 0; i < n; ++i)
    if (!handles[i]) return pf::fail(PF_E_ARG, "pf_batch_create: null handle");
  pf_batch* b = new pf_batch();
  b->h.assign(handles, handles + n);
  b->threads = threads;
  *out = b;
  return PF_OK;
}

int pf_batch_step(pf_batch* b, int nsteps) {
  if (!b || nsteps < 0) return pf::fail(PF_E_ARG, "pf_batch_step: bad argument");
  return run_all(b, [&](int i) { return pf_step(b->h[i], nsteps); });
}

int pf_batch_minimize(pf_batch* b, const pf_minimize_opts* opts, pf_minimize_stats* stats) {
  if (!b || !opts) return pf::fail(PF_E_ARG, "pf_batch_minimize: bad argument");
  return run_all(b, [&](int i) {
    pf_minimize_stats st;
    std::memset(&st, 0, sizeof st);
    const int rc = pf_minimize(b->h[i], &opts[i], &st);
    if (stats) stats[i] = st;
    return rc;
  });
}

int pf_batch_step_fused(pf_batch* b, int nsteps) {
  if (!b || nsteps < 0) return pf::fail(PF_E_ARG, "pf_batch_step_fused: bad argument");
  return nsteps ? fused_step(b, nsteps) : PF_OK;
}

int pf_batch_minimize_fused(pf_batch* b, const pf_minimize_opts* opts, pf_minimize_stats* stats) {
  if (!b || !opts) return pf::fail(PF_E_ARG, "pf_batch_minimize_fused: bad argument");
  return fused_minimize(b, opts, stats);
}

int pf_batch_fused_launches(const pf_batch* b, long long* batched, long long* recorded, long long* graphs_built,
                            long long* graph_launches) {
  if (!b) return pf::fail(PF_E_ARG, "pf_batch_fused_launches: null batch");
  if (batched) *batched = b->fused_launches;
  if (recorded) *recorded = b->fused_recorded;
  if (graphs_built) *graphs_built = b->fused_graphs_built;
  if (graph_launches) *graph_launches = b->fused_graph_launches;
  return PF_OK;
}

int pf_batch_set_graphs(pf_batch* b, int enable) {
  if (!b || (enable != 0 && enable != 1)) return pf::fail(PF_E_ARG, "pf_batch_set_graphs: bad argument");
  b->graphs = enable;
  return PF_OK;
}

int pf_batch_destroy(pf_batch* b) {
  delete b;
  return PF_OK;
}

}  // extern "C"
