// C ABI of libpfcore: handles, parameters and the physical constants hoisted
// from the per-voxel thermodynamics.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <string>

#include "../../include/pfcore.h"
#include "pf_engine_api.hpp"
#include "pf_error.hpp"
#include "pf_handle.hpp"

namespace pf {

std::string& last_error() {
  static thread_local std::string msg;
  return msg;
}

namespace {
const double kR = 8.31446261815324;   // thermo_utils.h:260

// thermo_utils.h:316-351 (SGTE lattice stabilities, Gierlotka 2010)
double GHSER_Pb(double T) {
  if (T < 600.61)
    return -7650.085 + 101.700244 * T - 24.5242231 * T * std::log(T) - 0.00365895 * T * T -
           2.4395e-7 * std::pow(T, 3);
  return -10531.095 + 154.243182 * T - 32.4913959 * T * std::log(T) + 0.00154613 * T * T +
         8.05448e25 * std::pow(T, -9);
}
double GHSER_Ag(double T) {
  if (T < 1234.93)
    return -7209.512 + 118.202013 * T - 23.8463314 * T * std::log(T) - 0.001790585 * T * T -
           3.98587e-7 * std::pow(T, 3) - 12011.0 / T;
  return -15095.252 + 190.266404 * T - 33.472 * T * std::log(T) + 1.411773e29 * std::pow(T, -9);
}
double GHSER_Te(double T) {
  if (T < 722.66)
    return -10544.679 + 183.372894 * T - 35.6687 * T * std::log(T) + 0.01583435 * T * T -
           5.240417e-6 * std::pow(T, 3) + 155015.0 / T;
  return 9160.595 - 129.265373 * T + 13.004 * T * std::log(T) - 0.0362361 * T * T +
         5.006367e-6 * std::pow(T, 3) - 1.28681e30 * std::pow(T, -9);
}
double G_PbTe(double T) { return (-76063.2138 + 9.67716633 * T) + GHSER_Pb(T) + GHSER_Te(T); }
double G_Ag2Te(double T) {
  double base = -10128.93 - 12.645115 * T;
  double g = base + (2.0 / 3.0) * GHSER_Ag(T) + (1.0 / 3.0) * GHSER_Te(T);
  return 3.0 * g;
}
// thermo_utils.h:453-481
double solve_x_eq(double T) {
  double L = 41212.9 - 18.05 * T, RT = kR * T;
  double x = std::exp(-L / RT);
  if (x < 1e-9) x = 1e-9;
  if (x > 0.99) x = 0.5;
  for (int i = 0; i < 20; ++i) {
    double f = RT * std::log(x) + L * (1.0 - x) * (1.0 - x);
    double df = RT / x - 2.0 * L * (1.0 - x);
    double delta = f / df;
    x -= delta;
    if (x < 1e-10) x = 1e-10;
    if (x > 0.999) x = 0.999;
    if (std::fabs(delta) < 1e-8) break;
  }
  return x;
}
}  // namespace

PhysParams make_phys(const pf_params& p) {
  PhysParams P;
  std::memset(&P, 0, sizeof(P));
  double T = p.temperature_C + 273.15;
  if (T < 1.0) T = 1.0;
  P.T = T;
  P.RT = kR * T;
  P.Lparam = 41212.9 - 18.05 * T;
  P.G_A0 = G_PbTe(T);
  P.G_B0 = G_Ag2Te(T);
  double sc = p.mu_reference_scale;
  if (std::fabs(sc) < 1e-30) sc = (sc >= 0.0) ? 1e-30 : -1e-30;
  P.mu_scale = sc;
  P.convex = p.thermo_convex_extrapolation_enabled ? 1 : 0;
  {
    // constants of the convex continuation at x = 0.09 (thermo_utils.h:400-430)
    const double xc = 0.09;
    P.muA_c = P.G_A0 + P.RT * std::log(1.0 - xc) + P.Lparam * xc * xc;
    P.dmuA_c = -kR * T / (1.0 - xc) + 2.0 * P.Lparam * xc;
    P.muB_c = P.G_B0 + P.RT * std::log(xc) + P.Lparam * (1.0 - xc) * (1.0 - xc);
    P.dmuB_c = kR * T / xc - 2.0 * P.Lparam * (1.0 - xc);
  }
  // mu0_compound (pfparams_refresh_thermo, main_cuda.cu:15604-15614)
  {
    const double xeq = solve_x_eq(T);
    const double muA = mu_A(P, xeq), muB = mu_B(P, xeq);
    double vs = p.v_A + p.v_B;
    if (std::fabs(vs) < 1e-30) vs = (vs >= 0.0) ? 1e-30 : -1e-30;
    P.mu0 = (p.v_A * muA + p.v_B * muB) / vs;
  }
  P.v_A = p.v_A;
  P.v_B = p.v_B;
  P.Vm_c = p.Vm_compound;
  P.Vm_a0 = p.Vm_alpha_0;
  P.dVm = p.dVm_alpha_dxB;
  P.W = p.W;
  P.kappa = p.kappa_phi;
  P.L_phi = p.L_phi;
  P.D_a = p.D_alpha;
  P.D_c = p.D_compound;
  P.eps_iso = p.eps_iso_over_vB;
  P.shift = p.elastic_shift_dimless;
  P.Y_clip = p.Y_clip;
  P.xB_eps = p.xB_eps;
  P.dt = p.dt;
  P.elastic = p.elastic_enabled ? 1 : 0;
  for (int i = 0; i < 21; ++i) {
    P.C[i] = p.S[i];
    P.dC[i] = p.S_p[i];
  }
  for (int i = 0; i < 6; ++i) {
    P.eps00[i] = p.eps00[i];
    P.E0[i] = p.E0[i];
  }
  // zero normal-shear coupling and off-diagonal shear entries in both tensors
  P.ortho = 1;
  for (int i = 0; i < 6; ++i)
    for (int j = std::max(i + 1, 3); j < 6; ++j)
      if (P.C[vidx(i, j)] != 0.0 || P.dC[vidx(i, j)] != 0.0) P.ortho = 0;
  return P;
}

}  // namespace pf


extern "C" {

int pf_version(void) { return PFCORE_ABI_VERSION; }

int pf_last_error(char* buf, size_t n) {
  if (!buf || n == 0) return PF_E_ARG;
  std::strncpy(buf, pf::last_error().c_str(), n - 1);
  buf[n - 1] = '\0';
  return PF_OK;
}

static int validate(const pf_params* p) {
  if (p->Nx < 2 || p->Ny < 2 || p->Nz < 2) return pf::fail(PF_E_ARG, "grid dimensions must be >= 2");
  if (!(p->dt > 0.0) || !(p->dx > 0.0) || !(p->dy > 0.0) || !(p->dz > 0.0))
    return pf::fail(PF_E_ARG, "dt and spacings must be positive");
  if (!(p->W >= 0.0) || !(p->kappa_phi >= 0.0) || !(p->L_phi >= 0.0) || !(p->D_alpha >= 0.0) ||
      !(p->D_compound >= 0.0))
    return pf::fail(PF_E_PARAMS, "physical parameters W, kappa_phi, L_phi, D_alpha, D_compound must be set");
  const bool accel = p->elastic_warm_start_enabled || p->elastic_residual_control_enabled;
  if (p->elastic_enabled && accel && !(p->elastic_warm_start_enabled && p->elastic_residual_control_enabled))
    return pf::fail(PF_E_PARAMS,
                    "accelerated elastic solver requires elastic_warm_start_enabled=1 and "
                    "elastic_residual_control_enabled=1");
  if (p->elastic_iter_max < 1) return pf::fail(PF_E_PARAMS, "elastic_iter_max must be >= 1");
  // accelerated elastic iteration/residual contract (main_cuda.cu:24492-24510)
  if (p->elastic_enabled && accel &&
      (p->elastic_iter_min <= 0 || p->elastic_iter_max < p->elastic_iter_min ||
       !(p->elastic_residual_tolerance > 0.0) || !std::isfinite(p->elastic_residual_tolerance) ||
       !(p->elastic_residual_absolute_floor > 0.0) || !std::isfinite(p->elastic_residual_absolute_floor) ||
       (p->elastic_fail_on_nonconvergence != 0 && p->elastic_fail_on_nonconvergence != 1)))
    return pf::fail(PF_E_PARAMS, "accelerated elastic iteration/residual contract is invalid");
  // zero-mode tolerance/iteration contract (main_cuda.cu:24468-24475)
  if (p->zero_mode && (!(p->zero_mode_tolerance_relative > 0.0) || !std::isfinite(p->zero_mode_tolerance_relative) ||
                       p->zero_mode_max_iterations <= 0))
    return pf::fail(PF_E_PARAMS, "PF zero-mode tolerance/iteration contract is invalid");
  if (p->elastic_warm_start_order < 1 || p->elastic_warm_start_order > 3)
    return pf::fail(PF_E_PARAMS, "elastic_warm_start_order must be 1, 2 or 3");
  if (p->elastic_anderson_depth < 0 || p->elastic_anderson_depth > 4)
    return pf::fail(PF_E_PARAMS, "elastic_anderson_depth must be in [0, 4]");
  return PF_OK;
}

int pf_create(const pf_params* p, int precision, pf_handle** out) {
  if (!p || !out) return pf::fail(PF_E_ARG, "pf_create: null");
  int rc = validate(p);
  if (rc) return rc;
  auto h = std::make_unique<pf_handle>();
  h->precision = precision;
  if (precision == PF_FP64)
    h->eng = pf::make_engine_f64();
  else if (precision == PF_MIXED)
    h->eng = pf::make_engine_f32();
  else
    return pf::fail(PF_E_ARG, "unknown precision");
  rc = h->eng->init(*p);
  if (rc) return rc;
  *out = h.release();
  return PF_OK;
}

int pf_set_fields(pf_handle* h, const double* phi, const double* xB, const double* dYdt, double xB_cap) {
  if (!h || !phi || !xB) return pf::fail(PF_E_ARG, "pf_set_fields: null");
  return h->eng->set_fields(phi, xB, dYdt, xB_cap);
}

int pf_step(pf_handle* h, int nsteps) {
  if (!h || nsteps < 0) return pf::fail(PF_E_ARG, "pf_step: bad argument");
  for (int s = 0; s < nsteps; ++s) {
    int rc = h->eng->step();
    if (rc) return rc;
  }
  return PF_OK;
}

int pf_get_field(pf_handle* h, int which, double* out) {
  if (!h || !out) return pf::fail(PF_E_ARG, "pf_get_field: null");
  return h->eng->get_field(which, out);
}

int pf_keep_stress(pf_handle* h, int enable) {
  if (!h) return pf::fail(PF_E_ARG, "pf_keep_stress: null");
  return h->eng->keep_stress(enable);
}

int pf_scalars_get(pf_handle* h, pf_scalars* out) {
  if (!h || !out) return pf::fail(PF_E_ARG, "pf_scalars_get: null");
  return h->eng->scalars(out);
}

long long pf_elastic_state_size(pf_handle* h) { return h ? h->eng->elastic_state_size() : -1; }

int pf_get_elastic_state(pf_handle* h, double* out) {
  if (!h || !out) return pf::fail(PF_E_ARG, "pf_get_elastic_state: null");
  return h->eng->get_elastic_state(out);
}

int pf_set_elastic_state(pf_handle* h, const double* in) {
  if (!h || !in) return pf::fail(PF_E_ARG, "pf_set_elastic_state: null");
  return h->eng->set_elastic_state(in);
}

int pf_time_kernel(pf_handle* h, int kernel_id, int reps, double* us, double* bytes) {
  if (!h || !us || !bytes || reps < 1 || kernel_id < 0 || kernel_id >= PF_K_COUNT)
    return pf::fail(PF_E_ARG, "pf_time_kernel: bad argument");
  return h->eng->time_kernel(kernel_id, reps, us, bytes);
}

int pf_probe_kernels(pf_handle* h, unsigned mask) {
  if (!h || (mask >> PF_K_COUNT) != 0u) return pf::fail(PF_E_ARG, "pf_probe_kernels: bad argument");
  return h->eng->probe_kernels(mask);
}

int pf_probe_read(pf_handle* h, int kernel_id, long long* launches, long long* worked, double* us_total,
                  double* bytes_total) {
  if (!h || !launches || !worked || !us_total || !bytes_total || kernel_id < 0 || kernel_id >= PF_K_COUNT)
    return pf::fail(PF_E_ARG, "pf_probe_read: bad argument");
  return h->eng->probe_read(kernel_id, launches, worked, us_total, bytes_total);
}

int pf_minimize_opts_default(pf_minimize_opts* o) {
  if (!o) return pf::fail(PF_E_ARG, "pf_minimize_opts_default: null");
  o->dt = 0.01;
  o->max_iter = 2000;
  o->V0 = 0.0;
  o->eta_lambda_vol = 0.2;
  o->post_projection_iters = 1;
  o->rms_dphi_threshold = 1.0e-6;
  o->rms_res_threshold = 1.0e-4;
  o->vol_err_rel_threshold = 1.0e-4;
  o->convergence_steps = 10;
  o->min_pseudo_time = 0.0;
  o->full_model = 0;
  o->rms_dY_threshold = 5.0e-5;
  o->energy_diff_rel_threshold = 1.0e-9;
  o->xB_max_safe = 0.07;
  o->mass_constraint = 0;
  o->mass_tolerance_relative = 1.0e-12;
  o->mass_max_iterations = 64;
  o->target_mass_code = 0.0;
  o->freeze_phi = 0;
  o->component_count = 0;
  o->component_labels = nullptr;
  o->component_target_h_sums = nullptr;
  return PF_OK;
}

int pf_minimize_begin(pf_handle* h, const pf_minimize_opts* o) {
  if (!h || !o) return pf::fail(PF_E_ARG, "pf_minimize_begin: null");
  if (!(o->dt > 0.0) || !std::isfinite(o->dt)) return pf::fail(PF_E_ARG, "minimize dt must be positive");
  if (o->max_iter < 0 || o->post_projection_iters < 0 || o->convergence_steps < 1)
    return pf::fail(PF_E_ARG, "minimize: bad iteration counts");
  // main_cuda.cu:24562-24636
  if (o->mass_constraint) {
    if (!o->full_model)
      return pf::fail(PF_E_ARG, "the minimize mass constraint requires the full-model minimize");
    if (!(o->mass_tolerance_relative > 0.0) || !std::isfinite(o->mass_tolerance_relative) ||
        o->mass_max_iterations <= 0 || !std::isfinite(o->target_mass_code))
      return pf::fail(PF_E_ARG, "minimize mass-constraint tolerance, iteration, or target contract is invalid");
  }
  if (o->freeze_phi) {
    if (!o->full_model || !o->mass_constraint)
      return pf::fail(PF_E_ARG, "freeze_phi requires the full-model minimize with the mass constraint");
    if (o->post_projection_iters != 0)
      return pf::fail(PF_E_ARG, "freeze_phi requires post_projection_iters 0 (phi stays byte-identical)");
    if (o->V0 > 0.0) return pf::fail(PF_E_ARG, "freeze_phi requires V0 <= 0 (the input h-volume is kept)");
  }
  if (o->component_count != 0) {
    if (!o->full_model || !o->mass_constraint || o->freeze_phi)
      return pf::fail(PF_E_ARG,
                      "component volume constraints require full-model constrained minimize with phi evolution "
                      "enabled");
    if (o->component_count < 2 || o->component_count > 64 || !o->component_labels || !o->component_target_h_sums)
      return pf::fail(PF_E_ARG,
                      "component constraints require a component count in [2,64], labels and one positive "
                      "target h sum per component");
    if (o->post_projection_iters != 0 || o->V0 > 0.0)
      return pf::fail(PF_E_ARG,
                      "component constraints own the phi volume projection: post_projection_iters 0 and V0 <= 0");
    for (int c = 0; c < o->component_count; ++c)
      if (!(o->component_target_h_sums[c] > 0.0) || !std::isfinite(o->component_target_h_sums[c]))
        return pf::fail(PF_E_ARG, "invalid component target h sum");
  }
  return h->eng->minimize_begin(*o);
}

int pf_minimize_iterate(pf_handle* h, pf_minimize_stats* st) {
  if (!h) return pf::fail(PF_E_ARG, "pf_minimize_iterate: null");
  return h->eng->minimize_iterate(st);
}

int pf_minimize(pf_handle* h, const pf_minimize_opts* o, pf_minimize_stats* st) {
  int rc = pf_minimize_begin(h, o);
  if (rc) return rc;
  pf_minimize_stats last{};
  for (int it = 0; it < o->max_iter; ++it) {
    if ((rc = h->eng->minimize_iterate(&last))) return rc;
    if (last.converged) break;
  }
  if (st) *st = last;
  return PF_OK;
}

int pf_eta_params_default(pf_eta_params* q) {
  if (!q) return pf::fail(PF_E_ARG, "pf_eta_params_default: null");
  std::memset(q, 0, sizeof(*q));
  q->gp_xB_fixed = 0.3529411764705882;
  q->gp_elastic_derivative_scale = 1.0;
  q->gp_reaction_nu_A = 0.65;
  q->gp_reaction_nu_B = 0.35;
  return PF_OK;
}

int pf_eta_update(pf_handle* h, const pf_eta_params* q, const double* eta, const double* phi,
                  const double* xB_alpha, const double* tr_sigma_or_null, double* eta_out) {
  if (!h || !q || !eta || !phi || !xB_alpha || !eta_out) return pf::fail(PF_E_ARG, "pf_eta_update: null");
  return h->eng->eta_update(*q, eta, phi, xB_alpha, tr_sigma_or_null, eta_out);
}

int pf_set_warm_start(pf_handle* h, const double* u_k_or_null) {
  if (!h) return pf::fail(PF_E_ARG, "pf_set_warm_start: null handle");
  return h->eng->set_warm_start(u_k_or_null);
}

int pf_set_warm_start_callback(pf_handle* h, pf_warm_start_fn fn, void* user) {
  if (!h) return pf::fail(PF_E_ARG, "pf_set_warm_start_callback: null handle");
  return h->eng->set_warm_start_callback(fn, user, h);
}

int pf_attach_learned_warm_start(pf_handle* h, pf_nn* net) {
  if (!h) return pf::fail(PF_E_ARG, "pf_attach_learned_warm_start: null handle");
  return h->eng->attach_nn(pf_nn_impl(net));
}

int pf_nn_collect(pf_nn* net, pf_handle* h) {
  if (!h || !net) return pf::fail(PF_E_ARG, "pf_nn_collect: null");
  int rc = h->eng->nn_collect(pf_nn_impl(net));
  return rc ? rc : pf_nn_dataset_size(net);
}

int pf_sync(pf_handle* h) {
  if (!h) return pf::fail(PF_E_ARG, "pf_sync: null");
  return h->eng->sync();
}

int pf_checkpoint_write(pf_handle* h, const char* path, const pf_checkpoint_provenance* prov) {
  if (!h || !path) return pf::fail(PF_E_ARG, "pf_checkpoint_write: null");
  return h->eng->checkpoint_write(path, prov);
}

int pf_checkpoint_read(pf_handle* h, const char* path, const pf_checkpoint_provenance* prov) {
  if (!h || !path) return pf::fail(PF_E_ARG, "pf_checkpoint_read: null");
  return h->eng->checkpoint_read(path, prov);
}

int pf_destroy(pf_handle* h) {
  delete h;
  return PF_OK;
}

static int fft_handle(int Nx, int Ny, int Nz, int precision, pf_handle** h) {
  pf_params p;
  pf_params_default(&p);
  p.Nx = Nx;
  p.Ny = Ny;
  p.Nz = Nz;
  p.W = p.kappa_phi = p.L_phi = p.D_alpha = p.D_compound = 1.0;
  p.mu_reference_scale = 1.0;
  return pf_create(&p, precision, h);
}

int pf_fft_forward(int Nx, int Ny, int Nz, int precision, const double* in, double* out) {
  if (!in || !out) return pf::fail(PF_E_ARG, "pf_fft_forward: null");
  pf_handle* h = nullptr;
  int rc = fft_handle(Nx, Ny, Nz, precision, &h);
  if (rc) return rc;
  rc = h->eng->fft_forward(in, out);
  pf_destroy(h);
  return rc;
}

int pf_fft_inverse(int Nx, int Ny, int Nz, int precision, const double* in, double* out) {
  if (!in || !out) return pf::fail(PF_E_ARG, "pf_fft_inverse: null");
  pf_handle* h = nullptr;
  int rc = fft_handle(Nx, Ny, Nz, precision, &h);
  if (rc) return rc;
  rc = h->eng->fft_inverse(in, out);
  pf_destroy(h);
  return rc;
}

}  // extern "C"
