// Learned warm start: network state, GPU training (forward, backward, Adam)
// and the pf_nn_* C ABI.  Kernels are in pf_nn.hpp.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/pfcore.h"
#include "pf_error.hpp"
#include "pf_nn.hpp"

namespace pf {
namespace nn {

#define NN_HIP(expr)                                                                              \
  do {                                                                                            \
    hipError_t e_ = (expr);                                                                       \
    if (e_ != hipSuccess) return fail(PF_E_HIP, std::string("HIP: ") + hipGetErrorString(e_) + \
                                                    " at " #expr);                                \
  } while (0)

static_assert(P_TOTAL == PF_NN_PARAMS, "parameter count of the C ABI");

// Activations of one forward/backward pass over a batch of grids.
// Static scratch of pf_nn_forward_noalloc: [3][NOALLOC_MAX_VOXELS][CH] BF16
// (packed input, then the two activation planes), 24 MiB held by each GPU the
// library runs on until the process ends.  Host code uses the per-device
// address from hipGetSymbolAddress, never the array's name.
constexpr long NOALLOC_MAX_VOXELS = 64L * 64 * 64;
static __device__ __align__(256) uint16_t scratch[3 * NOALLOC_MAX_VOXELS * CH];

constexpr int MAX_DEVICES = 64;
static std::mutex scratch_mu;
static uint16_t* scratch_ptr[MAX_DEVICES] = {};

// The device address of `scratch` on `device` (which must be current),
// looked up once per device.
static int scratch_address(int device, uint16_t** out) {
  if (device < 0 || device >= MAX_DEVICES) return fail(PF_E_ARG, "pf_nn: device index out of range");
  std::lock_guard<std::mutex> lk(scratch_mu);
  if (!scratch_ptr[device]) {
    void* p = nullptr;
    NN_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) return fail(PF_E_HIP, "pf_nn: no device address for the static scratch");
    scratch_ptr[device] = static_cast<uint16_t*>(p);
  }
  *out = scratch_ptr[device];
  return PF_OK;
}

struct Buffers {
  Geom g{};
  uint16_t* act[LAYERS] = {};   // act[0] input features, act[L] output of layer L-1
  uint16_t* grad[2] = {};
  float* y3 = nullptr;
  float* t3 = nullptr;
  float* partials = nullptr;    // wgrad per-wave partials
  float* lossp = nullptr;
  long nwaves = 0;
  int lblocks = 0;
  std::vector<void*> mem;
  ~Buffers() {
    for (void* p : mem) (void)hipFree(p);
  }
  template <typename U>
  int alloc(U** p, size_t n) {
    void* q = nullptr;
    NN_HIP(hipMalloc(&q, n * sizeof(U) + 256));
    mem.push_back(q);
    *p = static_cast<U*>(q);
    return PF_OK;
  }
  int init(int B, int Nx, int Ny, int Nz) {
    g = Geom{B, Nx, Ny, Nz, (long)Nx * Ny * Nz};
    const size_t n16 = (size_t)B * g.S * CH;
    int rc;
    for (int l = 0; l < LAYERS; ++l)
      if ((rc = alloc(&act[l], n16))) return rc;
    if ((rc = alloc(&grad[0], n16)) || (rc = alloc(&grad[1], n16))) return rc;
    if ((rc = alloc(&y3, (size_t)B * OUT_CH * g.S)) || (rc = alloc(&t3, (size_t)B * OUT_CH * g.S))) return rc;
    const long ntiles = (long)B * Nx * Ny * ((Nz + 15) / 16);
    nwaves = std::max<long>(4, std::min<long>(1024, (ntiles + 3) / 4 * 4));
    lblocks = 512;
    if ((rc = alloc(&partials, (size_t)nwaves * P_PER_LAYER)) || (rc = alloc(&lossp, lblocks))) return rc;
    return PF_OK;
  }
};

struct Net {
  int Nx = 0, Ny = 0, Nz = 0, max_batch = 0, capacity = 0, count = 0, pending = -1;
  long S = 0;
  double s_phi = 1.0, s_tau = 0.1, s_out = 1.0;
  int pool = 1;              // 2: the network runs on the grid averaged over 2x2x2 voxels
  float* master = nullptr;   // [P_TOTAL]
  float* m = nullptr;
  float* v = nullptr;
  float* grad = nullptr;
  uint16_t* Wb = nullptr;    // [LAYERS][WK_PER_LAYER] forward order (MFMA B operands)
  uint16_t* Wt = nullptr;    // transposed, tap-flipped
  uint16_t* feat = nullptr;  // dataset [capacity][S][16]
  float* tgt = nullptr;      // dataset [capacity][3][S]
  long adam_t = 0;
  int device = 0;            // the GPU the network lives on
  hipStream_t stream = nullptr;
  Buffers train;
  std::vector<void*> mem;
  ~Net() {
    if (stream) (void)hipStreamSynchronize(stream);
    for (void* p : mem) (void)hipFree(p);
    if (stream) (void)hipStreamDestroy(stream);
  }
  template <typename U>
  int alloc(U** p, size_t n) {
    void* q = nullptr;
    NN_HIP(hipMalloc(&q, n * sizeof(U) + 256));
    mem.push_back(q);
    *p = static_cast<U*>(q);
    return PF_OK;
  }
  int pack() {
    const int n = LAYERS * WK_PER_LAYER;
    hipLaunchKernelGGL(pack_weights_kernel, dim3((n + 255) / 256), dim3(256), 0, stream, master, Wb, Wt);
    NN_HIP(hipGetLastError());
    return PF_OK;
  }
};

// forward: act[0] -> act[1..4] -> y3
static int forward(Net* net, Buffers& bf, hipStream_t s) {
  const Geom& g = bf.g;
  for (int L = 0; L < LAYERS; ++L) {
    const uint16_t* W = net->Wb + (size_t)L * WK_PER_LAYER;
    const float* bias = net->master + (size_t)L * P_PER_LAYER + W_PER_LAYER;
    if (L < LAYERS - 1)
      conv3_launch<EPI_ACT>(s, g, bf.act[L], W, bias, bf.act[L + 1], nullptr, nullptr);
    else
      conv3_launch<EPI_OUT3>(s, g, bf.act[L], W, bias, nullptr, bf.y3, nullptr);
  }
  NN_HIP(hipGetLastError());
  return PF_OK;
}

// loss and backward into grad_out[P_TOTAL]; returns the mean-square loss
static int backward(Net* net, Buffers& bf, hipStream_t s, float inv_t, float* grad_out, float* loss) {
  const Geom& g = bf.g;
  const float inv_count = 1.0f / (float)((double)g.B * g.S * OUT_CH);
  hipLaunchKernelGGL(loss_kernel, dim3(bf.lblocks), dim3(256), 0, s, bf.y3, bf.t3, g, inv_t, inv_count, bf.grad[0],
                     bf.lossp);
  int cur = 0;
  for (int L = LAYERS - 1; L >= 0; --L) {
    hipLaunchKernelGGL(conv3_wgrad_kernel, dim3(bf.nwaves / 4), dim3(256), 0, s, g, bf.act[L], bf.grad[cur],
                       bf.partials);
    hipLaunchKernelGGL(wgrad_reduce_kernel, dim3((P_PER_LAYER + 255) / 256), dim3(256), 0, s, bf.partials, bf.nwaves,
                       grad_out + (size_t)L * P_PER_LAYER);
    if (L > 0) {
      conv3_launch<EPI_GRAD>(s, g, bf.grad[cur], net->Wt + (size_t)L * WK_PER_LAYER, nullptr, bf.grad[1 - cur],
                             nullptr, bf.act[L]);
      cur = 1 - cur;
    }
  }
  NN_HIP(hipGetLastError());
  if (loss) {
    std::vector<float> lp(bf.lblocks);
    NN_HIP(hipMemcpyAsync(lp.data(), bf.lossp, bf.lblocks * sizeof(float), hipMemcpyDeviceToHost, s));
    NN_HIP(hipStreamSynchronize(s));
    double acc = 0.0;
    for (float x : lp) acc += x;
    *loss = (float)(acc * inv_count);
  }
  return PF_OK;
}

static int gather(Net* net, const int* idx, int batch) {
  if (batch < 1 || batch > net->max_batch) return fail(PF_E_ARG, "pf_nn: batch outside [1, max_batch]");
  Buffers& bf = net->train;
  bf.g.B = batch;
  for (int i = 0; i < batch; ++i) {
    const int j = idx[i];
    if (j < 0 || j >= net->count) return fail(PF_E_ARG, "pf_nn: snapshot index out of range");
    NN_HIP(hipMemcpyAsync(bf.act[0] + (size_t)i * net->S * CH, net->feat + (size_t)j * net->S * CH,
                          net->S * CH * sizeof(uint16_t), hipMemcpyDeviceToDevice, net->stream));
    NN_HIP(hipMemcpyAsync(bf.t3 + (size_t)i * OUT_CH * net->S, net->tgt + (size_t)j * OUT_CH * net->S,
                          OUT_CH * net->S * sizeof(float), hipMemcpyDeviceToDevice, net->stream));
  }
  return PF_OK;
}

int infer(Net* net, hipStream_t s, int Nx, int Ny, int Nz, const uint16_t* X0, uint16_t* a, uint16_t* b, float* Y3) {
  if (s != net->stream) NN_HIP(hipStreamSynchronize(net->stream));   // weights are packed on the network's stream
  const Geom g{1, Nx, Ny, Nz, (long)Nx * Ny * Nz};
  const uint16_t* in = X0;
  uint16_t* bufs[2] = {a, b};
  for (int L = 0; L < LAYERS; ++L) {
    const uint16_t* W = net->Wb + (size_t)L * WK_PER_LAYER;
    const float* bias = net->master + (size_t)L * P_PER_LAYER + W_PER_LAYER;
    if (L < LAYERS - 1) {
      conv3_launch<EPI_ACT>(s, g, in, W, bias, bufs[L & 1], nullptr, nullptr);
      in = bufs[L & 1];
    } else {
      conv3_launch<EPI_OUT3>(s, g, in, W, bias, nullptr, Y3, nullptr);
    }
  }
  NN_HIP(hipGetLastError());
  return PF_OK;
}

int pool(const Net* net) { return net->pool; }

void scales(const Net* net, double* s_phi, double* s_tau, double* s_out) {
  *s_phi = net->s_phi;
  *s_tau = net->s_tau;
  *s_out = net->s_out;
}

int dataset_slot(Net* net, int Nx, int Ny, int Nz, uint16_t** feat, float** target) {
  if (Nx != net->Nx || Ny != net->Ny || Nz != net->Nz)
    return fail(PF_E_ARG, "pf_nn_collect: handle grid differs from the training grid");
  if (net->count >= net->capacity) return fail(PF_E_ARG, "pf_nn_collect: dataset is full");
  net->pending = net->count;
  *feat = net->feat + (size_t)net->count * net->S * CH;
  *target = net->tgt + (size_t)net->count * OUT_CH * net->S;
  return PF_OK;
}

int dataset_commit(Net* net) {
  if (net->pending != net->count) return fail(PF_E_ARG, "pf_nn: no pending snapshot");
  ++net->count;
  net->pending = -1;
  return net->count;
}

}  // namespace nn
}  // namespace pf

using pf::fail;
using namespace pf::nn;

struct pf_nn {
  Net net;
};

extern "C" {

int pf_nn_create(int Nx, int Ny, int Nz, int max_batch, int capacity, pf_nn** out) {
  if (!out || Nx < 1 || Ny < 1 || Nz < 1 || max_batch < 1 || capacity < 0)
    return fail(PF_E_ARG, "pf_nn_create: bad argument");
  pf_nn* h = new pf_nn();
  Net& n = h->net;
  n.Nx = Nx;
  n.Ny = Ny;
  n.Nz = Nz;
  n.S = (long)Nx * Ny * Nz;
  n.max_batch = max_batch;
  n.capacity = capacity;
  int rc;
  auto bad = [&](int code) {
    delete h;
    return code;
  };
  if (hipGetDevice(&n.device) != hipSuccess) return bad(fail(PF_E_HIP, "pf_nn_create: device"));
  if (hipStreamCreateWithFlags(&n.stream, hipStreamNonBlocking) != hipSuccess)
    return bad(fail(PF_E_HIP, "pf_nn_create: stream"));
  if ((rc = n.alloc(&n.master, P_TOTAL)) || (rc = n.alloc(&n.m, P_TOTAL)) || (rc = n.alloc(&n.v, P_TOTAL)) ||
      (rc = n.alloc(&n.grad, P_TOTAL)) || (rc = n.alloc(&n.Wb, (size_t)LAYERS * WK_PER_LAYER)) ||
      (rc = n.alloc(&n.Wt, (size_t)LAYERS * WK_PER_LAYER)))
    return bad(rc);
  if (capacity > 0 && ((rc = n.alloc(&n.feat, (size_t)capacity * n.S * CH)) ||
                       (rc = n.alloc(&n.tgt, (size_t)capacity * OUT_CH * n.S))))
    return bad(rc);
  if ((rc = n.train.init(max_batch, Nx, Ny, Nz))) return bad(rc);
  std::vector<float> zero(P_TOTAL, 0.0f);
  if (hipMemcpy(n.master, zero.data(), P_TOTAL * sizeof(float), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(n.m, zero.data(), P_TOTAL * sizeof(float), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(n.v, zero.data(), P_TOTAL * sizeof(float), hipMemcpyHostToDevice) != hipSuccess)
    return bad(fail(PF_E_HIP, "pf_nn_create: init copy"));
  if ((rc = n.pack())) return bad(rc);
  *out = h;
  return PF_OK;
}

int pf_nn_destroy(pf_nn* net) {
  delete net;
  return PF_OK;
}

int pf_nn_set_scales(pf_nn* net, double s_phi, double s_tau, double s_out) {
  if (!net || !(s_out != 0.0)) return fail(PF_E_ARG, "pf_nn_set_scales: bad argument");
  net->net.s_phi = s_phi;
  net->net.s_tau = s_tau;
  net->net.s_out = s_out;
  return PF_OK;
}

int pf_nn_set_pool(pf_nn* net, int factor) {
  if (!net || (factor != 1 && factor != 2)) return fail(PF_E_ARG, "pf_nn_set_pool: the factor is 1 or 2");
  net->net.pool = factor;
  return PF_OK;
}

int pf_nn_get_pool(pf_nn* net, int* factor) {
  if (!net || !factor) return fail(PF_E_ARG, "pf_nn_get_pool: null");
  *factor = net->net.pool;
  return PF_OK;
}

int pf_nn_get_scales(pf_nn* net, double* s_phi, double* s_tau, double* s_out) {
  if (!net || !s_phi || !s_tau || !s_out) return fail(PF_E_ARG, "pf_nn_get_scales: null");
  scales(&net->net, s_phi, s_tau, s_out);
  return PF_OK;
}

int pf_nn_set_params(pf_nn* net, const float* params) {
  if (!net || !params) return fail(PF_E_ARG, "pf_nn_set_params: null");
  Net& n = net->net;
  std::vector<float> zero(P_TOTAL, 0.0f);
  NN_HIP(hipMemcpyAsync(n.master, params, P_TOTAL * sizeof(float), hipMemcpyHostToDevice, n.stream));
  NN_HIP(hipMemcpyAsync(n.m, zero.data(), P_TOTAL * sizeof(float), hipMemcpyHostToDevice, n.stream));
  NN_HIP(hipMemcpyAsync(n.v, zero.data(), P_TOTAL * sizeof(float), hipMemcpyHostToDevice, n.stream));
  n.adam_t = 0;
  int rc = n.pack();
  if (rc) return rc;
  NN_HIP(hipStreamSynchronize(n.stream));
  return PF_OK;
}

int pf_nn_get_params(pf_nn* net, float* params) {
  if (!net || !params) return fail(PF_E_ARG, "pf_nn_get_params: null");
  NN_HIP(hipMemcpyAsync(params, net->net.master, P_TOTAL * sizeof(float), hipMemcpyDeviceToHost, net->net.stream));
  NN_HIP(hipStreamSynchronize(net->net.stream));
  return PF_OK;
}

int pf_nn_dataset_size(pf_nn* net) { return net ? net->net.count : fail(PF_E_ARG, "pf_nn_dataset_size: null"); }

int pf_nn_eval_loss(pf_nn* net, const int* idx, int batch, float* loss) {
  if (!net || !idx || !loss) return fail(PF_E_ARG, "pf_nn_eval_loss: null");
  Net& n = net->net;
  int rc;
  if ((rc = gather(&n, idx, batch)) || (rc = forward(&n, n.train, n.stream)) ||
      (rc = backward(&n, n.train, n.stream, (float)(1.0 / n.s_out), n.grad, loss)))
    return rc;
  return PF_OK;
}

int pf_nn_train_step(pf_nn* net, const int* idx, int batch, float lr, float* loss) {
  if (!net || !idx) return fail(PF_E_ARG, "pf_nn_train_step: null");
  Net& n = net->net;
  int rc;
  float l = 0.f;
  if ((rc = gather(&n, idx, batch)) || (rc = forward(&n, n.train, n.stream)) ||
      (rc = backward(&n, n.train, n.stream, (float)(1.0 / n.s_out), n.grad, &l)))
    return rc;
  ++n.adam_t;
  const float b1 = 0.9f, b2 = 0.999f;
  const float c1 = 1.0f - std::pow(b1, (float)n.adam_t), c2 = 1.0f - std::pow(b2, (float)n.adam_t);
  hipLaunchKernelGGL(adam_kernel, dim3((P_TOTAL + 255) / 256), dim3(256), 0, n.stream, n.master, n.m, n.v, n.grad,
                     P_TOTAL, lr, b1, b2, 1e-8f, c1, c2);
  if ((rc = n.pack())) return rc;
  NN_HIP(hipStreamSynchronize(n.stream));
  if (loss) *loss = l;
  return PF_OK;
}

static int host_pass(pf_nn* net, int batch, int Nx, int Ny, int Nz, const float* in7, const float* target,
                     float* out3, float* loss, float* grad) {
  if (!net || !in7 || batch < 1 || Nx < 1 || Ny < 1 || Nz < 1) return fail(PF_E_ARG, "pf_nn: bad argument");
  Net& n = net->net;
  Buffers bf;
  int rc;
  if ((rc = bf.init(batch, Nx, Ny, Nz))) return rc;
  const long nv = (long)batch * bf.g.S;
  float* d_in = nullptr;
  if ((rc = bf.alloc(&d_in, (size_t)nv * IN_CH))) return rc;
  NN_HIP(hipMemcpyAsync(d_in, in7, (size_t)nv * IN_CH * sizeof(float), hipMemcpyHostToDevice, n.stream));
  hipLaunchKernelGGL(pack_input_kernel, dim3(512), dim3(256), 0, n.stream, (const float*)d_in, nv, bf.act[0]);
  if ((rc = forward(&n, bf, n.stream))) return rc;
  if (out3)
    NN_HIP(hipMemcpyAsync(out3, bf.y3, (size_t)nv * OUT_CH * sizeof(float), hipMemcpyDeviceToHost, n.stream));
  if (grad) {
    NN_HIP(hipMemcpyAsync(bf.t3, target, (size_t)nv * OUT_CH * sizeof(float), hipMemcpyHostToDevice, n.stream));
    float* d_grad = nullptr;
    if ((rc = bf.alloc(&d_grad, P_TOTAL))) return rc;
    if ((rc = backward(&n, bf, n.stream, 1.0f, d_grad, loss))) return rc;
    NN_HIP(hipMemcpyAsync(grad, d_grad, P_TOTAL * sizeof(float), hipMemcpyDeviceToHost, n.stream));
  }
  NN_HIP(hipStreamSynchronize(n.stream));
  return PF_OK;
}

int pf_nn_forward(pf_nn* net, int batch, int Nx, int Ny, int Nz, const float* in7, float* out3) {
  if (!out3) return fail(PF_E_ARG, "pf_nn_forward: null output");
  return host_pass(net, batch, Nx, Ny, Nz, in7, nullptr, out3, nullptr, nullptr);
}

int pf_nn_gradients(pf_nn* net, int batch, int Nx, int Ny, int Nz, const float* in7, const float* target,
                    float* loss, float* grad) {
  if (!target || !grad) return fail(PF_E_ARG, "pf_nn_gradients: null");
  return host_pass(net, batch, Nx, Ny, Nz, in7, target, nullptr, loss, grad);
}

int pf_nn_time_inference(pf_nn* net, int Nx, int Ny, int Nz, int reps, double* us) {
  if (!net || !us || Nx < 1 || Ny < 1 || Nz < 1 || reps < 1) return fail(PF_E_ARG, "pf_nn_time_inference: bad argument");
  Net& n = net->net;
  Buffers bf;
  int rc;
  const size_t S = (size_t)Nx * Ny * Nz;
  uint16_t *x0, *a, *b;
  float* y3;
  if ((rc = bf.alloc(&x0, S * CH)) || (rc = bf.alloc(&a, S * CH)) || (rc = bf.alloc(&b, S * CH)) ||
      (rc = bf.alloc(&y3, S * OUT_CH)))
    return rc;
  NN_HIP(hipMemsetAsync(x0, 0, S * CH * sizeof(uint16_t), n.stream));
  if ((rc = infer(&n, n.stream, Nx, Ny, Nz, x0, a, b, y3))) return rc;   // warm-up
  hipEvent_t e0, e1;
  NN_HIP(hipEventCreate(&e0));
  NN_HIP(hipEventCreate(&e1));
  NN_HIP(hipEventRecord(e0, n.stream));
  for (int r = 0; r < reps; ++r)
    if ((rc = infer(&n, n.stream, Nx, Ny, Nz, x0, a, b, y3))) return rc;
  NN_HIP(hipEventRecord(e1, n.stream));
  NN_HIP(hipEventSynchronize(e1));
  float ms = 0.f;
  NN_HIP(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  *us = 1000.0 * ms / reps;
  return PF_OK;
}

// Inference that allocates nothing (no counterpart in the reference; for
// callers that cannot allocate device memory).  The packed input and the two
// ping-pong activation planes live in one file-scope device array of
// NOALLOC_MAX_VOXELS voxels; every call on a GPU shares that GPU's copy.
int pf_nn_forward_noalloc(pf_nn* net, int Nx, int Ny, int Nz, const float* in7_dev, float* out3_dev) {
  if (!net || !in7_dev || !out3_dev || Nx < 1 || Ny < 1 || Nz < 1)
    return fail(PF_E_ARG, "pf_nn_forward_noalloc: bad argument");
  const long S = (long)Nx * Ny * Nz;
  if (S > pf::nn::NOALLOC_MAX_VOXELS)
    return fail(PF_E_ARG, "pf_nn_forward_noalloc: " + std::to_string(S) + " voxels exceed the " +
                              std::to_string(pf::nn::NOALLOC_MAX_VOXELS) + " (64^3) of its static scratch");
  Net& n = net->net;
  int prev = 0;
  NN_HIP(hipGetDevice(&prev));
  if (prev != n.device) NN_HIP(hipSetDevice(n.device));
  uint16_t* scr = nullptr;
  int rc = pf::nn::scratch_address(n.device, &scr);
  if (!rc) {
    uint16_t* X0 = scr;
    uint16_t* a = scr + (size_t)pf::nn::NOALLOC_MAX_VOXELS * CH;
    uint16_t* b = scr + 2 * (size_t)pf::nn::NOALLOC_MAX_VOXELS * CH;
    hipLaunchKernelGGL(pack_input_kernel, dim3(512), dim3(256), 0, n.stream, in7_dev, S, X0);
    rc = infer(&n, n.stream, Nx, Ny, Nz, X0, a, b, out3_dev);
    if (!rc && hipStreamSynchronize(n.stream) != hipSuccess) rc = fail(PF_E_HIP, "pf_nn_forward_noalloc: sync");
  }
  if (prev != n.device) (void)hipSetDevice(prev);
  return rc;
}

}  // extern "C"

pf::nn::Net* pf_nn_impl(pf_nn* net) { return net ? &net->net : nullptr; }
