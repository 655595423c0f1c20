// Engine interface shared by the C ABI and the two precision instantiations.
#pragma once
#include <memory>

#include "../../include/pfcore.h"
#include "pf_checkpoint.hpp"
#include "pf_nn_api.hpp"
#include "pf_physics.hpp"

namespace pf {

PhysParams make_phys(const pf_params& p);
ckpt::Provenance provenance_from_abi(const pf_checkpoint_provenance* q);
bool elastic_state_required(const pf_params* p);

struct EngineBase {
  virtual ~EngineBase() = default;
  virtual int init(const pf_params& p) = 0;
  virtual int set_fields(const double* phi, const double* xB, const double* dYdt, double xB_cap) = 0;
  virtual int step() = 0;
  virtual int get_field(int which, double* out) = 0;
  virtual int keep_stress(int enable) = 0;
  virtual int scalars(pf_scalars* out) = 0;
  virtual long long elastic_state_size() = 0;
  virtual int get_elastic_state(double* out) = 0;
  virtual int set_elastic_state(const double* in) = 0;
  virtual int fft_forward(const double* in, double* out_half) = 0;
  virtual int fft_inverse(const double* in_half, double* out) = 0;
  virtual int sync() = 0;
  virtual int time_kernel(int id, int reps, double* us, double* bytes) = 0;
  virtual int probe_kernels(unsigned mask) = 0;
  virtual int probe_read(int id, long long* launches, long long* worked, double* us, double* bytes) = 0;
  virtual int minimize_begin(const pf_minimize_opts& o) = 0;
  virtual int minimize_iterate(pf_minimize_stats* st) = 0;
  virtual int set_warm_start(const double* u_or_null) = 0;
  virtual int set_warm_start_callback(pf_warm_start_fn fn, void* user, pf_handle* h) = 0;
  virtual int attach_nn(nn::Net* net) = 0;
  virtual int nn_collect(nn::Net* net) = 0;
  virtual int eta_update(const pf_eta_params& q, const double* eta, const double* phi, const double* xa,
                         const double* trs, double* out) = 0;
  virtual int checkpoint_write(const char* path, const pf_checkpoint_provenance* prov) = 0;
  virtual int checkpoint_read(const char* path, const pf_checkpoint_provenance* prov) = 0;
  // fused batch runner (pf_batch_minimize_fused): the phases of a phi-only
  // minimize iteration, recorded for many handles and replayed as batched
  // launches; the runner mirrors the device scalars between phases
  virtual int mb_check() = 0;                  // PF_OK when the handle can run recorded iterations
  virtual int mb_spec() = 0;                   // solver iterations the next solve enqueues first
  virtual int mb_phase_a(int spec) = 0;        // energies, <h>, the solve's start and `spec` iterations
  virtual int mb_solve_more(int* more) = 0;    // two more iterations while the solve runs (*more = 1)
  virtual int mb_phase_b() = 0;                // the solve's result through the final scalars
  virtual int mb_phase_c(pf_minimize_stats* st) = 0;   // stopping rule and statistics (host)
  virtual int mb_converged() = 0;
  virtual int mb_nscal() = 0;                  // doubles of device scalars
  virtual void mb_gather(double* dst) = 0;     // enqueue device scalars -> dst (device)
  virtual void mb_scatter(const double* src) = 0;   // host mirror <- src
  // the phases of a time step (pf_batch_step_fused); the solver's further
  // iterations go through mb_solve_more
  virtual int sb_check() = 0;
  virtual int sb_phase_a(int spec) = 0;        // the solve's start and `spec` iterations
  virtual int sb_zspec() = 0;                  // zero-mode steps the next search enqueues first
  virtual int sb_phase_b(int zspec) = 0;       // the solve's result through the opening of the search
  virtual int sb_zm_more(int* more) = 0;       // four more search steps while it runs (*more = 1)
  virtual int sb_phase_c() = 0;                // the search's result, the shift, the step's bookkeeping
};

std::unique_ptr<EngineBase> make_engine_f64();
std::unique_ptr<EngineBase> make_engine_f32();

}  // namespace pf
