// .params loader and defaults (host only).
//
// Mirrors params_default (main_cuda.cu:15781-16433) for the keys the time step
// reads, the key=value file syntax of load_pfparams_override_file
// (main_cuda.cu:22306-22360: '#' comments, surrounding whitespace trimmed,
// unknown keys warned and ignored), and the required-key contract of
// validate_required_pfparams_presence (main_cuda.cu:1089-1144).
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

#include "../../include/pfcore.h"
#include "pf_error.hpp"

namespace {

const char* const kVoigt21[21] = {"11", "12", "13", "14", "15", "16", "22", "23", "24", "25", "26",
                                  "33", "34", "35", "36", "44", "45", "46", "55", "56", "66"};
const char* const kEigen[6] = {"eps_xx00", "eps_yy00", "eps_zz00", "eps_yz00", "eps_xz00", "eps_xy00"};
const char* const kE0[6] = {"E0_xx", "E0_yy", "E0_zz", "E0_yz", "E0_xz", "E0_xy"};

// Keys of the schema-v1 file that the time step does not read; accepted
// without a warning (they drive unit reporting or non-default model modes).
const char* const kAcceptedOther[] = {
    "gp_D_ratio", "gp_L_eta_code", "gp_L_eta_phys", "gp_M_eta_ratio_to_crit", "gp_W_eta_code",
    "gp_W_eta_phys", "gp_gamma_alpha_gp", "gp_kappa_eta_code", "gp_kappa_eta_phys", "gp_l_eta_nm",
    "gp_reaction_nu_A", "gp_reaction_nu_B", "gp_xB_eq_alpha_for_eta", "t_real_unit_s", "dt_code",
    "model_mode", "pf_y_update_mode", "pf_composition_mode", nullptr};

void trim(std::string& s) {
  size_t b = 0, e = s.size();
  while (b < e && (s[b] == ' ' || s[b] == '\t' || s[b] == '\r' || s[b] == '\n')) ++b;
  while (e > b && (s[e - 1] == ' ' || s[e - 1] == '\t' || s[e - 1] == '\r' || s[e - 1] == '\n')) --e;
  s = s.substr(b, e - b);
}

bool parse_double(const char* v, double* out) {
  char* end = nullptr;
  errno = 0;
  double d = strtod(v, &end);
  if (end == v || errno == ERANGE) return false;
  while (*end == ' ' || *end == '\t') ++end;
  if (*end != '\0') return false;
  *out = d;
  return true;
}
bool parse_int(const char* v, int* out) {
  double d;
  if (!parse_double(v, &d)) return false;
  if (d != std::floor(d)) return false;
  *out = (int)d;
  return true;
}

struct Presence {
  unsigned long long named = 0;   // bit per required scalar key
  unsigned eigen = 0, s = 0, sp = 0;
};

const char* const kRequired[] = {"pf_params_schema_version", "dt", "dx", "dy", "dz", "t_real_unit",
                                 "temperature_C", "mu_reference_scale", "W", "kappa_phi", "L_phi",
                                 "D_alpha", "D_compound", "v_A", "v_B", "Vm_compound", "Vm_alpha_0",
                                 "dVm_alpha_dxB", "ic_vf_init_phi", "ic_vf_target_phi",
                                 "ic_phi_iface_w", "gamma_Jm2", "lambda_sm_m",
                                 "Vm_alpha_0_phys_m3mol", "elastic_shift_dimless",
                                 "eps_iso_over_vB", nullptr};

void mark(Presence& pr, const char* key) {
  std::string k = key;
  if (k == "dt_code") k = "dt";
  if (k == "t_real_unit_s") k = "t_real_unit";
  for (int i = 0; kRequired[i]; ++i)
    if (k == kRequired[i]) pr.named |= 1ull << i;
  for (int i = 0; i < 6; ++i)
    if (k == kEigen[i]) pr.eigen |= 1u << i;
  if (k.rfind("S_p_", 0) == 0) {
    for (int i = 0; i < 21; ++i)
      if (k.substr(4) == kVoigt21[i]) pr.sp |= 1u << i;
  } else if (k.rfind("S_", 0) == 0) {
    for (int i = 0; i < 21; ++i)
      if (k.substr(2) == kVoigt21[i]) pr.s |= 1u << i;
  }
}

}  // namespace

extern "C" int pf_params_default(pf_params* p) {
  if (!p) return pf::fail(PF_E_ARG, "pf_params_default: null");
  std::memset(p, 0, sizeof(*p));
  p->Nx = p->Ny = p->Nz = 64;
  p->dx = p->dy = p->dz = 1.0;
  p->dt = 0.01;
  p->nsteps = 100;
  p->out_every = 100;
  p->csv_out_every = 10;
  p->elastic_enabled = 1;
  p->pf_params_schema_version = 0;
  // physical inputs have no built-in values in the reference (-1 sentinels)
  p->temperature_C = 380.0;
  p->mu_reference_scale = 1.0;
  p->t_real_unit = 1.0;
  p->v_A = 0.0;
  p->v_B = 1.0;
  p->Vm_compound = 1.0;
  p->Vm_alpha_0 = 1.0;
  p->dVm_alpha_dxB = 0.0;
  p->W = -1.0;
  p->kappa_phi = -1.0;
  p->L_phi = -1.0;
  p->D_alpha = -1.0;
  p->D_compound = -1.0;
  p->eps_iso_over_vB = 0.0;
  p->elastic_shift_dimless = -1.0;
  p->Y_clip = 20.0;
  p->xB_eps = 1e-8;
  p->thermo_convex_extrapolation_enabled = 0;
  p->elastic_warm_start_enabled = 1;
  p->elastic_residual_control_enabled = 1;
  p->elastic_iter_min = 2;
  p->elastic_iter_max = 32;
  p->elastic_residual_tolerance = 1.0e-6;
  p->elastic_residual_absolute_floor = 1.0e-30;
  p->elastic_fail_on_nonconvergence = 1;
  p->zero_mode = 0;
  p->zero_mode_tolerance_relative = 1.0e-12;
  p->zero_mode_max_iterations = 24;
  p->enable_Y_rhs_previous_time_level = 0;
  p->disable_Y_rhs_gamma_term = 0;
  p->Y_rhs_term_h_scale = 1.0;
  p->mode = 0;
  p->elastic_warm_start_order = 1;
  p->elastic_anderson_depth = 0;
  p->ic_23d_xB_out = 0.03;
  return PF_OK;
}

extern "C" int pf_params_set(pf_params* p, const char* key, const char* value) {
  if (!p || !key || !value) return pf::fail(PF_E_ARG, "pf_params_set: null");
  std::string k = key;
  struct DKey {
    const char* name;
    double* dst;
  } dkeys[] = {{"dt", &p->dt},
               {"dt_code", &p->dt},
               {"dx", &p->dx},
               {"dy", &p->dy},
               {"dz", &p->dz},
               {"t_real_unit", &p->t_real_unit},
               {"t_real_unit_s", &p->t_real_unit},
               {"temperature_C", &p->temperature_C},
               {"mu_reference_scale", &p->mu_reference_scale},
               {"v_A", &p->v_A},
               {"v_B", &p->v_B},
               {"Vm_compound", &p->Vm_compound},
               {"Vm_alpha_0", &p->Vm_alpha_0},
               {"dVm_alpha_dxB", &p->dVm_alpha_dxB},
               {"W", &p->W},
               {"kappa_phi", &p->kappa_phi},
               {"L_phi", &p->L_phi},
               {"D_alpha", &p->D_alpha},
               {"D_compound", &p->D_compound},
               {"eps_iso_over_vB", &p->eps_iso_over_vB},
               {"elastic_shift_dimless", &p->elastic_shift_dimless},
               {"Y_clip", &p->Y_clip},
               {"xB_eps", &p->xB_eps},
               {"elastic_residual_tolerance", &p->elastic_residual_tolerance},
               {"elastic_residual_absolute_floor", &p->elastic_residual_absolute_floor},
               {"pf_zero_mode_tolerance_relative", &p->zero_mode_tolerance_relative},
               {"Y_rhs_term_h_scale", &p->Y_rhs_term_h_scale},
               {"gamma_Jm2", &p->gamma_Jm2},
               {"lambda_sm_m", &p->lambda_sm_m},
               {"ic_phi_iface_w", &p->ic_phi_iface_w},
               {"ic_vf_init_phi", &p->ic_vf_init_phi},
               {"ic_vf_target_phi", &p->ic_vf_target_phi},
               {"Vm_alpha_0_phys_m3mol", &p->Vm_alpha_0_phys_m3mol},
               {"ic_23d_xB_out", &p->ic_23d_xB_out},
               {nullptr, nullptr}};
  for (int i = 0; dkeys[i].name; ++i) {
    if (k == dkeys[i].name) {
      if (!parse_double(value, dkeys[i].dst))
        return pf::fail(PF_E_PARAMS, "bad value for key " + k + ": " + value);
      return 1;
    }
  }
  struct IKey {
    const char* name;
    int* dst;
  } ikeys[] = {{"pf_params_schema_version", &p->pf_params_schema_version},
               {"thermo_convex_extrapolation_enabled", &p->thermo_convex_extrapolation_enabled},
               {"elastic_enabled", &p->elastic_enabled},
               {"elastic_warm_start_enabled", &p->elastic_warm_start_enabled},
               {"elastic_residual_control_enabled", &p->elastic_residual_control_enabled},
               {"elastic_iter_min", &p->elastic_iter_min},
               {"elastic_iter_max", &p->elastic_iter_max},
               {"elastic_fail_on_nonconvergence", &p->elastic_fail_on_nonconvergence},
               {"pf_zero_mode_max_iterations", &p->zero_mode_max_iterations},
               {"enable_Y_rhs_previous_time_level", &p->enable_Y_rhs_previous_time_level},
               {"disable_Y_rhs_gamma_term", &p->disable_Y_rhs_gamma_term},
               {"elastic_warm_start_order", &p->elastic_warm_start_order},
               {"elastic_anderson_depth", &p->elastic_anderson_depth},
               {nullptr, nullptr}};
  for (int i = 0; ikeys[i].name; ++i) {
    if (k == ikeys[i].name) {
      if (!parse_int(value, ikeys[i].dst))
        return pf::fail(PF_E_PARAMS, "bad integer for key " + k + ": " + value);
      return 1;
    }
  }
  for (int i = 0; i < 6; ++i) {
    if (k == kEigen[i]) {
      if (!parse_double(value, &p->eps00[i])) return pf::fail(PF_E_PARAMS, "bad value for " + k);
      return 1;
    }
    if (k == kE0[i]) {
      if (!parse_double(value, &p->E0[i])) return pf::fail(PF_E_PARAMS, "bad value for " + k);
      return 1;
    }
  }
  if (k.rfind("S_p_", 0) == 0) {
    for (int i = 0; i < 21; ++i)
      if (k.substr(4) == kVoigt21[i]) {
        if (!parse_double(value, &p->S_p[i])) return pf::fail(PF_E_PARAMS, "bad value for " + k);
        return 1;
      }
    return 0;
  }
  if (k.rfind("S_", 0) == 0) {
    for (int i = 0; i < 21; ++i)
      if (k.substr(2) == kVoigt21[i]) {
        if (!parse_double(value, &p->S[i])) return pf::fail(PF_E_PARAMS, "bad value for " + k);
        return 1;
      }
    return 0;
  }
  for (int i = 0; kAcceptedOther[i]; ++i)
    if (k == kAcceptedOther[i]) return 1;
  return 0;
}

extern "C" int pf_params_load(const char* path, pf_params* p) {
  if (!path || !p) return pf::fail(PF_E_ARG, "pf_params_load: null");
  FILE* fp = std::fopen(path, "r");
  if (!fp) return pf::fail(PF_E_PARAMS, std::string("could not open pf-param file: ") + path);
  char line[4096];
  int line_no = 0;
  Presence pr;
  while (std::fgets(line, sizeof line, fp)) {
    ++line_no;
    char* c = std::strchr(line, '#');
    if (c) *c = '\0';
    std::string s = line;
    trim(s);
    if (s.empty()) continue;
    size_t eq = s.find('=');
    if (eq == std::string::npos) {
      std::fclose(fp);
      return pf::fail(PF_E_PARAMS, std::string(path) + ":" + std::to_string(line_no) +
                                       " expected key=value format");
    }
    std::string key = s.substr(0, eq), value = s.substr(eq + 1);
    trim(key);
    trim(value);
    int rc = pf_params_set(p, key.c_str(), value.c_str());
    if (rc < 0) {
      std::fclose(fp);
      return rc;
    }
    if (rc == 0) {
      std::fprintf(stderr, "[warn] %s:%d unknown key ignored: %s\n", path, line_no, key.c_str());
      continue;
    }
    mark(pr, key.c_str());
  }
  std::fclose(fp);
  std::string missing;
  for (int i = 0; kRequired[i]; ++i)
    if (!(pr.named & (1ull << i))) missing += std::string(" ") + kRequired[i];
  if (pr.eigen != 0x3fu) missing += " eps_xx00..eps_xy00";
  if (pr.s != (1u << 21) - 1u) missing += " S_*";
  if (pr.sp != (1u << 21) - 1u) missing += " S_p_*";
  if (!missing.empty())
    return pf::fail(PF_E_PARAMS, std::string(path) + " is missing required keys:" + missing);
  return PF_OK;
}
