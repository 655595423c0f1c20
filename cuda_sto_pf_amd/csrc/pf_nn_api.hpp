// Interface between the learned-warm-start network (pf_nn.hip) and the
// time-step engine / C ABI.  BF16 helpers are inline so every translation
// unit can pack features.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../include/pfcore.h"

namespace pf {
namespace nn {

struct Net;

__device__ __forceinline__ uint16_t f2bf(float f) {
  uint32_t u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);   // round to nearest even (finite inputs)
  return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }
__device__ __forceinline__ uint32_t pk(uint16_t lo, uint16_t hi) { return (uint32_t)lo | ((uint32_t)hi << 16); }

// Inference of one grid on stream s: X0 [S][16] BF16 features -> Y3 [3][S]
// FP32 (network output units); a, b are [S][16] BF16 scratch.
int infer(Net* net, hipStream_t s, int Nx, int Ny, int Nz, const uint16_t* X0, uint16_t* a, uint16_t* b, float* Y3);
void scales(const Net* net, double* s_phi, double* s_tau, double* s_out);
// 1: the network sees the grid itself; 2: the grid averaged over 2x2x2 voxels
// (its output is read back at each of the 8 voxels)
int pool(const Net* net);
// Next dataset slot for a grid of Nx*Ny*Nz (error if full or the grid differs);
// commit() makes it part of the dataset once the engine's stream has finished.
int dataset_slot(Net* net, int Nx, int Ny, int Nz, uint16_t** feat, float** target);
int dataset_commit(Net* net);

}  // namespace nn
}  // namespace pf

// The network behind an ABI handle (pf_nn.hip).
pf::nn::Net* pf_nn_impl(pf_nn* net);
