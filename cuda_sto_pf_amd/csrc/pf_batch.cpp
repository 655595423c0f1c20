// Many independent instances driven concurrently: the batched-sweep entry
// points of the C ABI.
//
// Each instance is a pf_handle with its own HIP stream, so the kernels of
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

struct pf_batch {
  std::vector<pf_handle*> h;
  int threads = 1;
};

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

int pf_batch_destroy(pf_batch* b) {
  delete b;
  return PF_OK;
}

}  // extern "C"
