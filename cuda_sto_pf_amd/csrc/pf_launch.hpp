// Kernel launches that can be recorded and replayed as batched launches.
//
// A kernel that takes part in batching is described by a descriptor K:
//   using Args = ...;                       its argument block (trivially copyable)
//   static constexpr int BLOCK;             threads per workgroup
//   __device__ static void body(const Args&, B3 blockIdx, B3 gridDim);
//   static void direct(hipStream_t, B3 grid, const Args&);   the single launch
// The body is the kernel's code with its block index and grid size passed in,
// so the same code runs as the single kernel and, for many instances, inside
// batched_kernel<K>: one launch whose workgroups are the concatenation of every
// instance's grid, each workgroup reading its instance's Args from a table in
// device memory.
//
// pf::launch<K>(stream, grid, args) launches at once, or appends the launch to
// the calling thread's Recorder when one is set (tl_rec).  A batch runner
// records one phase of many instances (each with its own recorder), then
// replays them stage by stage: the k-th launches of all instances are one
// batched launch per distinct kernel (instances of different grid sizes use
// different transform instantiations, so a stage is one launch per size).
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <cstring>
#include <vector>

namespace pf {

// block index / grid size as passed to a kernel body
struct B3 {
  unsigned x, y, z;
};
#define PF_BIDX (::pf::B3{blockIdx.x, blockIdx.y, blockIdx.z})
#define PF_GDIM (::pf::B3{gridDim.x, gridDim.y, gridDim.z})

inline B3 b3(dim3 d) { return B3{d.x, d.y, d.z}; }

// batched table: int start[n+1] (prefix sums of the instance grids' block
// counts), B3 grid[n], Args args[n], each part 16-byte aligned
constexpr int BATCH_MAX = 64;   // instances per batched launch (one wave's ballot)
__host__ __device__ constexpr size_t bt_align(size_t o, size_t a) { return (o + a - 1) / a * a; }
__host__ __device__ constexpr size_t bt_grid_off(int n) { return bt_align((size_t)(n + 1) * sizeof(int), 16); }
template <class A>
__host__ __device__ constexpr size_t bt_args_off(int n) {
  return bt_align(bt_grid_off(n) + (size_t)n * sizeof(B3), alignof(A) > 16 ? alignof(A) : 16);
}
template <class A>
constexpr size_t bt_size(int n) {
  return bt_align(bt_args_off<A>(n) + (size_t)n * sizeof(A), 256);
}

template <class K>
__global__ void __launch_bounds__(K::BLOCK) batched_kernel(const char* __restrict__ tab, int n) {
  // the instance of this workgroup: the last start[i] <= blockIdx.x, found by
  // one ballot over the (at most 64) prefix sums
  const int* start = (const int*)tab;
  const int b = (int)blockIdx.x;
  const int lane = (int)(threadIdx.x & 63);
  const bool le = lane < n && start[lane] <= b;
  const int inst = __popcll(__ballot(le)) - 1;
  const B3 g = ((const B3*)(tab + bt_grid_off(n)))[inst];
  const unsigned l = (unsigned)(b - start[inst]);
  const B3 bid{l % g.x, (l / g.x) % g.y, l / (g.x * g.y)};
  // The table is restrict: stores through the output pointers the arguments
  // hold never change it, as they never change kernel arguments, so the
  // compiler may reuse argument loads as it does in the single kernel (where
  // it otherwise evaluates, and rounds, some expressions differently).
  const auto* args = (const typename K::Args*)(tab + bt_args_off<typename K::Args>(n));
  K::body(args[inst], bid, g);
}

struct LaunchRec {
  void (*direct)(hipStream_t, B3, const void*);
  void (*batched)(hipStream_t, const char*, int, unsigned);
  size_t (*table_size)(int);
  size_t (*args_off)(int);
  B3 grid;
  unsigned asize;
  size_t aoff;   // the Args bytes in the recorder's arena
};

template <class K>
void launch_direct_erased(hipStream_t s, B3 g, const void* a) {
  K::direct(s, g, *(const typename K::Args*)a);
}
template <class K>
void launch_batched_erased(hipStream_t s, const char* tab, int n, unsigned blocks) {
  hipLaunchKernelGGL(batched_kernel<K>, dim3(blocks), dim3(K::BLOCK), 0, s, tab, n);
}
template <class K>
size_t table_size_erased(int n) {
  return bt_size<typename K::Args>(n);
}
template <class K>
size_t args_off_erased(int n) {
  return bt_args_off<typename K::Args>(n);
}

struct Recorder {
  std::vector<LaunchRec> recs;
  std::vector<unsigned char> arena;
  void clear() {
    recs.clear();
    arena.clear();
  }
  template <class K>
  void push(B3 g, const typename K::Args& a) {
    const size_t off = bt_align(arena.size(), 16);
    arena.resize(off + sizeof(a));
    std::memcpy(arena.data() + off, &a, sizeof(a));
    recs.push_back(LaunchRec{&launch_direct_erased<K>, &launch_batched_erased<K>, &table_size_erased<K>,
                             &args_off_erased<K>, g, (unsigned)sizeof(a), off});
  }
  // replay the recorded launches one by one on s (the unbatched order)
  void replay(hipStream_t s) const {
    for (const LaunchRec& r : recs) r.direct(s, r.grid, arena.data() + r.aoff);
  }
};

// the recorder of the calling thread; null: launches go out at once
inline thread_local Recorder* tl_rec = nullptr;

template <class K>
inline void launch(hipStream_t s, B3 g, const typename K::Args& a) {
  if (tl_rec)
    tl_rec->push<K>(g, a);
  else
    K::direct(s, g, a);
}
template <class K>
inline void launch(hipStream_t s, dim3 g, const typename K::Args& a) {
  launch<K>(s, b3(g), a);
}

// Replays the recorders of several instances stage by stage as batched
// launches on one stream.  All recorders must hold the same number of
// launches; stage k of every instance is grouped by kernel.  Tables go
// through one pinned staging buffer and one host-to-device copy per call.
//
// With graphs on, the launches of a call are captured into a HIP graph the
// first time their plan (kernels, instance counts, block counts, table
// offsets) occurs, and later calls with the same plan copy their tables and
// launch that graph: the tables sit at the same device addresses, so the
// captured launches read the new ones.
struct BatchReplayer {
  char* host = nullptr;   // pinned
  char* dev = nullptr;
  size_t cap = 0;
  hipEvent_t copied = nullptr;   // the last staging copy (the staging is reused after it)
  bool pending = false;
  unsigned long long launches = 0, stages = 0, graph_launches = 0, graphs_built = 0;
  bool use_graphs = false;
  struct Cached {
    std::vector<uint64_t> plan;
    hipGraphExec_t exec;
  };
  std::vector<Cached> graphs;   // at most GRAPH_CACHE plans
  static constexpr size_t GRAPH_CACHE = 64;
  void drop_graphs() {
    for (auto& g : graphs) (void)hipGraphExecDestroy(g.exec);
    graphs.clear();
  }
  ~BatchReplayer() {
    if (pending) (void)hipEventSynchronize(copied);
    drop_graphs();
    if (copied) (void)hipEventDestroy(copied);
    if (host) (void)hipHostFree(host);
    if (dev) (void)hipFree(dev);
  }
  int reserve(size_t n) {
    if (n <= cap) return 0;
    size_t c = cap ? cap : (1u << 20);
    while (c < n) c *= 2;
    if (pending) (void)hipEventSynchronize(copied);
    pending = false;
    drop_graphs();   // captured launches point into the old tables
    if (host) (void)hipHostFree(host);
    if (dev) (void)hipFree(dev);
    host = dev = nullptr;
    cap = 0;
    if (hipHostMalloc((void**)&host, c, hipHostMallocDefault) != hipSuccess) return -1;
    if (hipMalloc((void**)&dev, c) != hipSuccess) return -1;
    cap = c;
    return 0;
  }
  // returns 0, or -1 (staging allocation failed), or -2 (recorders differ in length)
  int run(hipStream_t s, const std::vector<const Recorder*>& rs) {
    if (rs.empty()) return 0;
    const size_t L = rs[0]->recs.size();
    for (const Recorder* r : rs)
      if (r->recs.size() != L) return -2;
    struct Group {
      const LaunchRec* first;
      std::vector<int> inst;
      size_t off;
      unsigned blocks;
    };
    std::vector<std::vector<Group>> plan(L);
    size_t total = 0;
    for (size_t k = 0; k < L; ++k) {
      auto& gs = plan[k];
      for (int i = 0; i < (int)rs.size(); ++i) {
        const LaunchRec& r = rs[i]->recs[k];
        Group* g = nullptr;
        for (auto& q : gs)
          if (q.first->batched == r.batched && q.inst.size() < (size_t)BATCH_MAX) {
            g = &q;
            break;
          }
        if (!g) {
          gs.push_back(Group{&r, {}, 0, 0});
          g = &gs.back();
        }
        g->inst.push_back(i);
      }
      for (auto& g : gs) {
        g.off = total;
        total += g.first->table_size((int)g.inst.size());
      }
    }
    if (!copied && hipEventCreateWithFlags(&copied, hipEventDisableTiming) != hipSuccess) return -1;
    if (pending && hipEventSynchronize(copied) != hipSuccess) return -1;
    pending = false;
    if (reserve(total)) return -1;
    for (size_t k = 0; k < L; ++k)
      for (auto& g : plan[k]) {
        const int n = (int)g.inst.size();
        char* t = host + g.off;
        int* start = (int*)t;
        B3* grid = (B3*)(t + bt_grid_off(n));
        char* args = t + g.first->args_off(n);
        unsigned acc = 0;
        for (int j = 0; j < n; ++j) {
          const Recorder* rr = rs[g.inst[j]];
          const LaunchRec& r = rr->recs[k];
          start[j] = (int)acc;
          grid[j] = r.grid;
          acc += r.grid.x * r.grid.y * r.grid.z;
          std::memcpy(args + (size_t)j * r.asize, rr->arena.data() + r.aoff, r.asize);
        }
        start[n] = (int)acc;
        g.blocks = acc;
      }
    if (hipMemcpyAsync(dev, host, total, hipMemcpyHostToDevice, s) != hipSuccess) return -1;
    if (hipEventRecord(copied, s) != hipSuccess) return -1;
    pending = true;
    auto issue = [&]() {
      for (size_t k = 0; k < L; ++k)
        for (auto& g : plan[k])
          if (g.blocks) g.first->batched(s, dev + g.off, (int)g.inst.size(), g.blocks);
    };
    size_t nl = 0;
    for (size_t k = 0; k < L; ++k) nl += plan[k].size();
    launches += nl;
    stages += L;
    if (!use_graphs) {
      issue();
      return 0;
    }
    std::vector<uint64_t> key;
    key.reserve(4 * nl + 1);
    key.push_back(total);
    for (size_t k = 0; k < L; ++k)
      for (auto& g : plan[k]) {
        key.push_back((uint64_t)(uintptr_t)g.first->batched);
        key.push_back(g.inst.size());
        key.push_back(g.blocks);
        key.push_back(g.off);
      }
    for (auto& c : graphs)
      if (c.plan == key) {
        ++graph_launches;
        return hipGraphLaunch(c.exec, s) == hipSuccess ? 0 : -1;
      }
    if (graphs.size() >= GRAPH_CACHE) {
      (void)hipStreamSynchronize(s);
      drop_graphs();
    }
    hipGraph_t graph = nullptr;
    if (hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal) != hipSuccess) return -1;
    issue();
    if (hipStreamEndCapture(s, &graph) != hipSuccess) return -1;
    hipGraphExec_t exec = nullptr;
    const hipError_t ie = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (ie != hipSuccess) return -1;
    graphs.push_back(Cached{std::move(key), exec});
    ++graphs_built;
    ++graph_launches;
    return hipGraphLaunch(exec, s) == hipSuccess ? 0 : -1;
  }
};

}  // namespace pf
