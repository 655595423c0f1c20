// main_pf: drop-in for the reference `main_cuda` binary's dynamics path.
//
// Keeps the reference command line (main_cuda.cu:22457-22771):
//   main_pf Nx Ny Nz dt nsteps out_every csv_out_every elastic_enabled
//           --pf-param-file <file> [--mode=dynamics|minimize] [--init-mode raw_fields
//           --init-phi-raw f --init-xB-raw f --init-meta f [--init-dY-dt-prev-raw f]]
//           [--pf-zero-mode PF_CONSERVED_Y_ZERO_MODE_V1|OFF] [--init-case-tag tag]
//           [--precision fp64|mixed]
// and its Results/ layout (main_cuda.cu:25270-25660): run directory
// <ch|chel>_T<T>_cuda_<Nx>x<Ny>x<Nz>_dt<dt>_steps<n>_xB<xB>/<case_tag>/ holding
// pf_input.params, vf_precip_vs_time_<tag>.csv, elastic_solver_trace.csv,
// pf_zero_mode_trace.csv, performance_summary.{csv,json} and phi/xB VTKs
// (io_vtk_cuda.h:35-131; CUDA_STO_SUPPRESS_VTK_OUTPUT=1 suppresses them).
// --mode=minimize runs phi-only or (--minimize-full-model) full-model minimize
// (main_cuda.cu:28341-32600) and writes
// energy_minimize_T<T>_r<r>_xB<xB>_<tag>.csv and phi_<iter>_<tag>.vtk into
// <prefix>_T<T>_cuda_<Nx>x<Ny>x<Nz>_dt<dt>_steps<n>_r<R>nm_xB<xB>/<tag>/.
// Errors print "[fatal] ..." and exit 2, as the reference does.
#include <sys/stat.h>

#include <algorithm>
#include <cctype>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/pfcore.h"

namespace {

int fatal(const std::string& msg) {
  std::fprintf(stderr, "[fatal] %s\n", msg.c_str());
  return 2;
}

std::string lib_error() {
  char buf[1024];
  pf_last_error(buf, sizeof buf);
  return buf;
}

bool starts_with(const char* s, const char* p) { return std::strncmp(s, p, std::strlen(p)) == 0; }

bool copy_text_file(const std::string& src, const std::string& dst) {
  std::ifstream in(src, std::ios::binary);
  if (!in) return false;
  std::ofstream out(dst, std::ios::binary);
  out << in.rdbuf();
  return bool(out);
}

double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// Minimal reader for the flat init_meta.json of a raw_fields bundle
// (main_cuda.cu:2733-2763): numbers and strings at top level.
struct Meta {
  int Nx = -1, Ny = -1, Nz = -1;
  std::string dtype = "float64", order = "C";
  double xB_max_safe = NAN;
};

bool read_meta(const std::string& path, Meta* m) {
  std::ifstream in(path);
  if (!in) return false;
  std::stringstream ss;
  ss << in.rdbuf();
  std::string t = ss.str();
  auto find_val = [&](const std::string& key, std::string* out) {
    size_t k = t.find("\"" + key + "\"");
    if (k == std::string::npos) return false;
    size_t c = t.find(':', k);
    if (c == std::string::npos) return false;
    size_t b = t.find_first_not_of(" \t\r\n", c + 1);
    if (b == std::string::npos) return false;
    if (t[b] == '"') {
      size_t e = t.find('"', b + 1);
      *out = t.substr(b + 1, e - b - 1);
    } else {
      size_t e = t.find_first_of(",}\n", b);
      *out = t.substr(b, e - b);
    }
    return true;
  };
  std::string v;
  if (find_val("Nx", &v)) m->Nx = std::atoi(v.c_str());
  if (find_val("Ny", &v)) m->Ny = std::atoi(v.c_str());
  if (find_val("Nz", &v)) m->Nz = std::atoi(v.c_str());
  if (find_val("dtype", &v)) m->dtype = v;
  if (find_val("order", &v)) m->order = v;
  if (find_val("xB_max_safe", &v)) m->xB_max_safe = std::atof(v.c_str());
  return m->Nx > 0 && m->Ny > 0 && m->Nz > 0;
}

bool read_raw(const std::string& path, const std::string& dtype, size_t n, std::vector<double>* out) {
  std::ifstream in(path, std::ios::binary);
  if (!in) return false;
  out->resize(n);
  if (dtype == "float32" || dtype == "f4" || dtype == "<f4") {
    std::vector<float> f(n);
    in.read(reinterpret_cast<char*>(f.data()), n * sizeof(float));
    if ((size_t)in.gcount() != n * sizeof(float)) return false;
    for (size_t i = 0; i < n; ++i) (*out)[i] = f[i];
  } else {
    in.read(reinterpret_cast<char*>(out->data()), n * sizeof(double));
    if ((size_t)in.gcount() != n * sizeof(double)) return false;
  }
  return true;
}

// legacy ASCII VTK, %.5f, z-y-x order (io_vtk_cuda.h:35-131)
bool write_vtk(const std::string& path, const std::vector<double>& f, int Nx, int Ny, int Nz,
               const char* name, int step) {
  const char* sup = std::getenv("CUDA_STO_SUPPRESS_VTK_OUTPUT");
  if (sup && std::strcmp(sup, "1") == 0) {
    std::printf("VTK_OUTPUT_SUPPRESSED field=%s step=%d path=%s\n", name, step, path.c_str());
    return true;
  }
  FILE* fo = std::fopen(path.c_str(), "w");
  if (!fo) return false;
  const int ny_out = (Ny == 2) ? 1 : Ny;
  std::fprintf(fo, "# vtk DataFile Version 3.0\nPhase Field Simulation\nASCII\nDATASET STRUCTURED_POINTS\n");
  std::fprintf(fo, "DIMENSIONS %d %d %d\nASPECT_RATIO 1 1 1\nORIGIN 0 0 0\nPOINT_DATA %d\n", Nx, ny_out, Nz,
               Nx * ny_out * Nz);
  std::fprintf(fo, "SCALARS %s double 1\nLOOKUP_TABLE default\n", name);
  for (int k = 0; k < Nz; ++k)
    for (int j = 0; j < ny_out; ++j)
      for (int i = 0; i < Nx; ++i) std::fprintf(fo, "%.5f ", f[((size_t)i * Ny + j) * Nz + k]);
  std::fprintf(fo, "\n");
  std::fclose(fo);
  return true;
}

void usage() {
  std::printf(
      "Usage (positional, legacy):\n"
      "  ./main_pf Nx Ny Nz dt nsteps out_every csv_out_every elastic_enabled(0/1)\n"
      "    dt from --pf-param-file takes precedence over the positional dt.\n"
      "Flags:\n"
      "  --pf-param-file <path>   required: complete physical PF inputs (key=value)\n"
      "  --mode=dynamics|dynamics-continue|minimize|minimize-continue\n"
      "  --continue-phi-vtk <f> [--continue-xB-vtk <f>]   continue from VTK outputs (no zero mode)\n"
      "  --ic-23d-xB-out <x>      far-field x_B (run directory name, continuation rebuild, minimize baseline)\n"
      "  --minimize-full-model    chemical drive, composition and volume constraint (no zero mode)\n"
      "  --minimize-rms-dY-threshold <v> --minimize-energy-diff-rel-threshold <v> --minimize-xB-max-safe <x>\n"
      "  --minimize-mass-constraint [--minimize-target-mass-code <m>] [--minimize-mass-tolerance-relative <v>]\n"
      "  [--minimize-mass-max-iterations <n>] --minimize-freeze-phi\n"
      "  --minimize-component-volume-constraints --minimize-component-count <K>\n"
      "  --minimize-component-label-raw <int32 raw> --minimize-component-target-h-sums <v0,v1,...>\n"
      "  --minimize-max-iter <n> --minimize-dt <dt> --V0 <v> --eta-lambda-vol <v>\n"
      "  --minimize-post-projection-iters <n> --minimize-rms-dphi-threshold <v>\n"
      "  --minimize-rms-res-threshold <v> --minimize-vol-err-rel-threshold <v>\n"
      "  --minimize-convergence-steps <n> --minimize-min-pseudo-time <t> --radius <r>\n"
      "  --init-mode raw_fields --init-phi-raw <f> --init-xB-raw <f> --init-meta <f>\n"
      "  --init-dY-dt-prev-raw <f>\n"
      "  --pf-zero-mode PF_CONSERVED_Y_ZERO_MODE_V1|OFF\n"
      "  --init-case-tag <tag>\n"
      "  --pf-checkpoint-every <n>            0 disables periodic checkpoints\n"
      "  --pf-checkpoint-path <path>          atomic full-state checkpoint (+ <path>.state)\n"
      "  --pf-restart-from <path>             provenance-strict restart\n"
      "  --pf-initial-state-class <name>      checkpoint-pinned initial-state policy\n"
      "  --pf-fixture-manifest-sha256 <hex>   checkpoint-pinned fixture identity\n"
      "  --pf-profile-library-manifest-sha256 <hex> checkpoint-pinned profile-library identity\n"
      "  --precision fp64|mixed   (default fp64)\n");
}

// ---------------------------------------------------------------- VTK continuation
// The path up to its last '/', or "" (derive_continue_case_output_dir,
// main_cuda.cu:277-292); twice gives the run directory (derive_continue_output_root, 257-275).
std::string parent_dir(const std::string& p) {
  const size_t k = p.rfind('/');
  return k == std::string::npos ? std::string() : p.substr(0, k);
}
std::string base_name(const std::string& p) {
  const size_t k = p.rfind('/');
  return k == std::string::npos ? p : p.substr(k + 1);
}
bool file_nonempty(const std::string& p) {
  struct stat st;
  return stat(p.c_str(), &st) == 0 && st.st_size > 0;
}

// continue_dyn_<n+1> / continue_min_<n+1> with n the largest index already in
// the path (derive_next_continue_case_tag, main_cuda.cu:2492-2521)
std::string next_continue_tag(const std::string& vtk, bool minimize) {
  const std::string prefix = minimize ? "continue_min_" : "continue_dyn_";
  long best = 0;
  for (size_t k = vtk.find(prefix); k != std::string::npos; k = vtk.find(prefix, k + 1)) {
    size_t d = k + prefix.size();
    if (d < vtk.size() && std::isdigit((unsigned char)vtk[d])) {
      long v = 0;
      while (d < vtk.size() && std::isdigit((unsigned char)vtk[d])) v = v * 10 + (vtk[d++] - '0');
      best = std::max(best, v);
    }
  }
  return prefix + std::to_string(best + 1);
}

// An ASCII STRUCTURED_POINTS scalar file as main_pf writes it: the grid from
// DIMENSIONS, values after LOOKUP_TABLE in z-y-x order; a file of one y layer
// fills both layers of an Ny = 2 run (read_vtk_ascii_to_host, io_vtk_cuda.h:131-216).
bool read_vtk_scalars(const std::string& path, int Nx, int Ny, int Nz, std::vector<double>* out, const char* what) {
  FILE* fp = std::fopen(path.c_str(), "r");
  if (!fp) {
    std::fprintf(stderr, "[fatal] cannot open VTK file %s for reading\n", path.c_str());
    return false;
  }
  char line[4096];
  int vx = -1, vy = -1, vz = -1;
  bool dims = false, table = false;
  while (std::fgets(line, sizeof line, fp)) {
    if (!dims && std::strncmp(line, "DIMENSIONS", 10) == 0) {
      dims = std::sscanf(line, "DIMENSIONS %d %d %d", &vx, &vy, &vz) == 3;
      continue;
    }
    if (std::strncmp(line, "LOOKUP_TABLE", 12) == 0) {
      table = true;
      break;
    }
  }
  if (!dims || !table) {
    std::fprintf(stderr, "[fatal] invalid VTK file %s: missing DIMENSIONS or LOOKUP_TABLE header\n", path.c_str());
    std::fclose(fp);
    return false;
  }
  const bool flat = vy == 1 && Ny == 2;
  if (!(vx == Nx && vz == Nz && (vy == Ny || flat))) {
    std::fprintf(stderr, "[fatal] VTK grid mismatch for %s (%s): file=%dx%dx%d expected=%dx%dx%d\n", path.c_str(),
                 what, vx, vy, vz, Nx, Ny, Nz);
    std::fclose(fp);
    return false;
  }
  out->assign((size_t)Nx * Ny * Nz, 0.0);
  const int ny_file = flat ? 1 : Ny;
  for (int k = 0; k < Nz; ++k)
    for (int j = 0; j < ny_file; ++j)
      for (int i = 0; i < Nx; ++i) {
        double v = 0.0;
        if (std::fscanf(fp, "%lf", &v) != 1) {
          std::fprintf(stderr, "[fatal] unexpected end of VTK data in %s\n", path.c_str());
          std::fclose(fp);
          return false;
        }
        for (int jj = flat ? 0 : j; jj < (flat ? Ny : j + 1); ++jj) (*out)[((size_t)i * Ny + jj) * Nz + k] = v;
      }
  std::fclose(fp);
  return true;
}

// matrix solubility: Newton on RT ln x + L (1-x)^2 = 0 with L = 41212.9 - 18.05 T
// (solve_x_eq_device, thermo_utils.h:453-481)
double matrix_x_eq(double T) {
  const double L = 41212.9 - 18.05 * T;
  const double RT = 8.31446261815324 * T;
  double x = std::exp(-L / RT);
  if (x < 1e-9) x = 1e-9;
  if (x > 0.99) x = 0.5;
  for (int i = 0; i < 20; ++i) {
    const double f = RT * std::log(x) + L * (1.0 - x) * (1.0 - x);
    const double df = RT / x - 2.0 * L * (1.0 - x);
    const double delta = f / df;
    x -= delta;
    if (x < 1e-10) x = 1e-10;
    if (x > 0.999) x = 0.999;
    if (std::fabs(delta) < 1e-8) break;
  }
  return x;
}

double h_quintic(double p) { return p * p * p * (6.0 * p * p - 15.0 * p + 10.0); }
double clamp_frac(double x, double e) { return x < e ? e : (x > 1.0 - e ? 1.0 - e : x); }

// x_B of a continued dynamics or full-model minimize run without an x_B VTK:
// x_eq inside the particles and the far field ic_23d_xB_out outside, or,
// with ic_23d_xB_out <= 0, the far field that gives the seed-radius (or
// ic_vf_target_phi) composition budget (rebuild_full_model_composition_from_phi,
// main_cuda.cu:2523-2600; compute_effective_vf_target, 2458-2490)
void rebuild_composition(const pf_params& P, double seed_radius, std::vector<double>& phi, std::vector<double>* xB) {
  const size_t n = phi.size();
  const double xe = matrix_x_eq(P.temperature_C + 273.15);
  double sh = 0.0, sh2 = 0.0;
  for (size_t i = 0; i < n; ++i) {
    const double h = h_quintic(phi[i]);
    sh += h;
    sh2 += h * h;
  }
  const double mh = sh / (double)n, mh2 = sh2 / (double)n;
  double xout;
  if (P.ic_23d_xB_out > 0.0) {
    xout = clamp_frac(P.ic_23d_xB_out, P.xB_eps);
  } else {
    double vf = std::fmin(0.999, std::fmax(1e-6, P.ic_vf_target_phi));
    if (seed_radius > 0.0) {
      const bool three_d = P.Ny > 2;
      const double R = seed_radius;
      const double f = three_d ? (4.0 * M_PI / 3.0) * R * R * R / (P.Nx * P.dx * P.Ny * P.dy * P.Nz * P.dz)
                               : M_PI * R * R / (P.Nx * P.dx * P.Nz * P.dz);
      vf = std::fmin(0.999, std::fmax(1e-6, f));
    }
    const double A = 1.0 - 2.0 * mh + mh2, B = mh - mh2, C = P.v_B * mh;
    const double target = P.v_B * vf + (1.0 - vf) * xe;
    if (std::fabs(A) < 1e-12) {
      const double den = std::fmax(1.0 - mh, 1e-12);
      xout = ((vf - mh) * P.v_B + (1.0 - vf) * xe) / den;
    } else {
      xout = (target - B * xe - C) / A;
    }
    xout = clamp_frac(xout, P.xB_eps);
  }
  xB->resize(n);
  for (size_t i = 0; i < n; ++i) {
    const double h = h_quintic(phi[i]);
    (*xB)[i] = clamp_frac((1.0 - h) * xout + h * xe, P.xB_eps);
  }
  std::printf("[continue] x_B rebuilt from phi: x_eq=%.6e x_out=%.6e <h>=%.6f\n", xe, xout, mh);
}

// the last data row's leading integer and two reals of a CSV, if any
// (read_last_vf_csv_state / read_last_energy_iter, main_cuda.cu:311-352)
bool last_csv_row(const std::string& path, int* step, double* t0, double* t1) {
  FILE* fp = std::fopen(path.c_str(), "r");
  if (!fp) return false;
  char line[4096];
  bool found = false;
  while (std::fgets(line, sizeof line, fp)) {
    if (!std::isdigit((unsigned char)line[0]) && line[0] != '-') continue;
    int st = 0;
    double a = 0.0, b = 0.0;
    const int got = std::sscanf(line, "%d,%lf,%lf", &st, &a, &b);
    if (got >= 1 && (t0 == nullptr || got == 3)) {
      found = true;
      *step = st;
      if (t0) *t0 = a;
      if (t1) *t1 = b;
    }
  }
  std::fclose(fp);
  return found;
}

// "v0,v1,..." with exactly `count` positive finite entries
// (parse_component_target_h_sums, main_cuda.cu:2835-2873)
bool parse_target_h_sums(const std::string& text, int count, std::vector<double>* out) {
  out->clear();
  const char* p = text.c_str();
  for (int i = 0; i < count; ++i) {
    while (*p && std::isspace((unsigned char)*p)) ++p;
    char* end = nullptr;
    const double v = std::strtod(p, &end);
    if (end == p || !std::isfinite(v) || v <= 0.0) {
      std::fprintf(stderr, "[fatal] invalid component target h sum at entry %d in '%s'\n", i, text.c_str());
      return false;
    }
    out->push_back(v);
    p = end;
    while (*p && std::isspace((unsigned char)*p)) ++p;
    if (i + 1 < count) {
      if (*p != ',') {
        std::fprintf(stderr, "[fatal] expected comma after component target h sum %d in '%s'\n", i,
                     text.c_str());
        return false;
      }
      ++p;
    }
  }
  while (*p && std::isspace((unsigned char)*p)) ++p;
  if (*p) {
    std::fprintf(stderr, "[fatal] component target h sums has extra entries: '%s'\n", text.c_str());
    return false;
  }
  return true;
}

// exactly n little-endian int32 labels in [0, count)
// (read_raw_component_labels_i32, main_cuda.cu:2805-2833)
bool read_labels_i32(const std::string& path, size_t n, int count, std::vector<int>* out) {
  FILE* fp = std::fopen(path.c_str(), "rb");
  if (!fp) {
    std::fprintf(stderr, "[fatal] cannot open component label raw field: %s\n", path.c_str());
    return false;
  }
  std::vector<int32_t> tmp(n);
  const size_t got = std::fread(tmp.data(), sizeof(int32_t), n, fp);
  const int extra = std::fgetc(fp);
  std::fclose(fp);
  if (got != n || extra != EOF) {
    std::fprintf(stderr, "[fatal] component label raw size mismatch: got %zu int32 values, expected exactly %zu\n",
                 got, n);
    return false;
  }
  out->resize(n);
  for (size_t i = 0; i < n; ++i) {
    if (tmp[i] < 0 || tmp[i] >= count) {
      std::fprintf(stderr, "[fatal] component label out of range at idx=%zu: %d not in [0,%d)\n", i, (int)tmp[i],
                   count);
      return false;
    }
    (*out)[i] = tmp[i];
  }
  return true;
}

// Minimize loop with the reference's energy log (main_cuda.cu:27440-27466,
// 32455-32489), VTK outputs (build_case_vtk_path, main_cuda.cu:126-170) and
// convergence report (main_cuda.cu:32546-32600).
int run_minimize(pf_handle* h, const pf_params& P, const pf_minimize_opts& mo, const std::string& case_dir,
                 const std::string& tag, const std::string& csv_tag, bool continuing, double seed_radius) {
  const size_t N = (size_t)P.Nx * P.Ny * P.Nz;
  char name[512];
  std::snprintf(name, sizeof name, "/energy_minimize_T%.0f_r%.2f_xB%.3f_%s.csv", P.temperature_C, seed_radius,
                P.ic_23d_xB_out, csv_tag.c_str());
  // a continuation appends to an existing log of its case, numbering on
  // (main_cuda.cu:27436-27447)
  int iter_offset = 0;
  const bool append = continuing && file_nonempty(case_dir + name) &&
                      last_csv_row(case_dir + name, &iter_offset, nullptr, nullptr);
  FILE* efp = std::fopen((case_dir + name).c_str(), append ? "a" : "w");
  // main_cuda.cu:25667-25690
  FILE* tfp = nullptr;
  if (mo.mass_constraint) {
    const std::string tp = case_dir + "/minimize_mass_constraint_trace.csv";
    if (!(tfp = std::fopen(tp.c_str(), "w"))) {
      if (efp) std::fclose(efp);
      return fatal("cannot open minimize mass-constraint trace: " + tp);
    }
    std::fprintf(tfp,
                 "iter,target_mass_code,current_mass_code,lambda,residual_code,residual_relative,derivative_code,"
                 "iterations,newton_steps,bisection_steps,Y_star_min,Y_star_max,lambda_lower,lambda_upper,"
                 "accepted_constraint_steps\n");
  }
  auto close_all = [&]() {
    if (efp) std::fclose(efp);
    if (tfp) std::fclose(tfp);
    efp = tfp = nullptr;
  };
  if (efp && !append)
    std::fprintf(efp,
                 "iter,dt,mean_h,V0,lambda,F_surf_hat,F_el_hat,F_chem_excess_hat,F_total_excess_hat,"
                 "F_chem_CNT_hat,F_total_CNT_hat,total_interface_sum,total_el_core_sum,"
                 "rms_res,rms_dphi,rms_dY,rel_dF,vol_err_rel,mass_target_code,mass_current_code,mass_err_rel,"
                 "mass_lambda,post_proj_iters_used\n");
  auto dump = [&](const std::string& sfx, int step) {
    std::vector<double> f(N);
    pf_get_field(h, PF_PHI, f.data());
    write_vtk(case_dir + "/phi_" + sfx + "_" + tag + ".vtk", f, P.Nx, P.Ny, P.Nz, "phi", step);
  };
  dump("init", 0);
  if (pf_minimize_begin(h, &mo) != PF_OK) {
    close_all();
    return fatal(lib_error());
  }
  pf_minimize_stats st{};
  const double t0 = now_s();
  for (int it = 1; it <= mo.max_iter; ++it) {
    if (pf_minimize_iterate(h, &st) != PF_OK) {
      close_all();
      return fatal(lib_error() + " (minimize iteration " + std::to_string(it) + ")");
    }
    if (efp) {
      std::fprintf(efp,
                   "%d,%.8e,%.8e,%.8e,%.8e,%.8e,%.8e,%.8e,%.8e,%.8e,%.8e,%.8e,%.8e,%.8e,%.8e,%.8e,%.8e,%.8e,"
                   "%.17e,%.17e,%.17e,%.17e,%d\n",
                   it + iter_offset, mo.dt, st.mean_h, st.V0, st.lambda_vol, st.F_surf_hat, st.F_el_hat,
                   st.F_chem_excess_hat,
                   st.F_total_excess_hat, st.F_chem_CNT_hat, st.F_total_CNT_hat, st.interface_sum, st.el_core_sum,
                   st.rms_res, st.rms_dphi, st.rms_dY, st.energy_diff_rel, st.vol_err_rel, st.mass_target,
                   st.mass_current, st.mass_err_rel, st.mass_lambda, mo.post_projection_iters);
      std::fflush(efp);
    }
    if (tfp) {
      std::fprintf(tfp, "%d,%.17e,%.17e,%.17e,%.17e,%.17e,%.17e,%d,%d,%d,%.17e,%.17e,%.17e,%.17e,%d\n", it,
                   st.mass_target, st.mass_current, st.mass_lambda, st.mass_residual,
                   st.mass_residual / std::fmax(std::fabs(st.mass_target), 1.0), st.mass_derivative,
                   st.mass_iterations, st.mass_newton_steps, st.mass_bisection_steps, st.mass_Ystar_min,
                   st.mass_Ystar_max, st.mass_lambda_lower, st.mass_lambda_upper, it);
      std::fflush(tfp);
    }
    if (P.out_every > 0 && it % P.out_every == 0) dump(std::to_string(it), it);
    if (st.converged) {
      std::printf("[minimize] converged iter=%d rms_res=%.3e rms_dphi=%.3e vol_err_rel=%.3e\n", it, st.rms_res,
                  st.rms_dphi, st.vol_err_rel);
      if (mo.full_model)
        std::printf("[minimize] rms_dY=%.3e energy_diff_rel=%.3e F_total_excess_hat=%.6e\n", st.rms_dY,
                    st.energy_diff_rel, st.F_total_excess_hat);
      break;
    }
  }
  if (!st.converged)
    std::printf("[minimize] reached max_iter=%d without convergence (rms_res=%.3e rms_dphi=%.3e)\n", mo.max_iter,
                st.rms_res, st.rms_dphi);
  dump("final", st.iteration);
  close_all();
  FILE* pj = std::fopen((case_dir + "/minimize_summary.json").c_str(), "w");
  if (pj) {
    std::fprintf(pj,
                 "{\n  \"case_name\": \"%s\",\n  \"converged\": %d,\n  \"iterations\": %d,\n"
                 "  \"final_rms_res\": %.17e,\n  \"final_rms_dphi\": %.17e,\n  \"final_vol_err_rel\": %.17e,\n"
                 "  \"mean_h\": %.17e,\n  \"V0\": %.17e,\n  \"walltime_s\": %.9f\n}\n",
                 tag.c_str(), st.converged, st.iteration, st.rms_res, st.rms_dphi, st.vol_err_rel, st.mean_h, st.V0,
                 now_s() - t0);
    std::fclose(pj);
  }
  return 0;
}

}  // namespace

int main(int argc, char** argv) {
  const char* param_file = nullptr;
  std::string mode = "dynamics", init_mode, phi_raw, xB_raw, meta_path, dydt_raw, case_tag;
  std::string zero_mode_sel = "OFF";
  // checkpoint / restart (main_cuda.cu:22478-22500, 22720-22760)
  std::string ckpt_path, restart_from, init_class = "LEGACY_UNSPECIFIED", fixture_sha = "LEGACY_UNSPECIFIED",
                                       library_sha = "LEGACY_UNSPECIFIED";
  int ckpt_every = 0;
  int precision = PF_FP64;
  pf_minimize_opts mo;
  pf_minimize_opts_default(&mo);
  bool minimize_dt_set = false, max_iter_set = false;
  // V5 component constraints (main_cuda.cu:23740-23760)
  bool comp_enabled = false;
  int comp_count = 0;
  std::string comp_label_raw, comp_targets_text;
  double seed_radius = 10.0;   // ic_phi_seed_radius default (main_cuda.cu:16221)
  // VTK continuation (main_cuda.cu:22761-22771, 23770-23782)
  bool continuing = false;
  std::string continue_phi_vtk, continue_xB_vtk;
  double xB_out_flag = NAN;   // --ic-23d-xB-out
  for (int i = 1; i < argc; ++i) {
    std::string a = argv[i];
    auto next = [&](std::string* dst) {
      if (i + 1 >= argc) return false;
      *dst = argv[++i];
      return true;
    };
    if (a == "--help" || a == "-h") {
      usage();
      return 0;
    } else if (a == "--pf-param-file") {
      if (i + 1 >= argc) return fatal("--pf-param-file needs a path");
      param_file = argv[++i];
    } else if (starts_with(a.c_str(), "--mode=")) {
      mode = a.substr(7);
      if (mode == "minimize-continue" || mode == "dynamics-continue") {
        continuing = true;
        mode = mode == "minimize-continue" ? "minimize" : "dynamics";
      }
    } else if (a == "--continue-phi-vtk") {
      if (!next(&continue_phi_vtk)) return fatal(a + " needs a path");
      continuing = true;
    } else if (a == "--continue-xB-vtk") {
      if (!next(&continue_xB_vtk)) return fatal(a + " needs a path");
    } else if (a == "--ic-23d-xB-out") {
      std::string v;
      if (!next(&v)) return fatal(a + " needs a value");
      xB_out_flag = std::atof(v.c_str());
    } else if (a == "--init-mode") {
      if (!next(&init_mode)) return fatal("--init-mode needs a value");
    } else if (a == "--init-phi-raw") {
      if (!next(&phi_raw)) return fatal("--init-phi-raw needs a path");
    } else if (a == "--init-xB-raw") {
      if (!next(&xB_raw)) return fatal("--init-xB-raw needs a path");
    } else if (a == "--init-meta") {
      if (!next(&meta_path)) return fatal("--init-meta needs a path");
    } else if (a == "--init-dY-dt-prev-raw") {
      if (!next(&dydt_raw)) return fatal("--init-dY-dt-prev-raw needs a path");
    } else if (a == "--pf-zero-mode") {
      if (!next(&zero_mode_sel)) return fatal("--pf-zero-mode needs a value");
    } else if (a == "--init-case-tag") {
      if (!next(&case_tag)) return fatal("--init-case-tag needs a value");
    } else if (a == "--pf-checkpoint-every" || a == "--pf_checkpoint_every") {
      std::string v;
      if (!next(&v)) return fatal(a + " needs a value");
      ckpt_every = std::atoi(v.c_str());
    } else if (a == "--pf-checkpoint-path" || a == "--pf_checkpoint_path") {
      if (!next(&ckpt_path)) return fatal(a + " needs a path");
    } else if (a == "--pf-restart-from" || a == "--pf_restart_from") {
      if (!next(&restart_from)) return fatal(a + " needs a path");
    } else if (a == "--pf-initial-state-class" || a == "--pf_initial_state_class") {
      if (!next(&init_class)) return fatal(a + " needs a value");
    } else if (a == "--pf-fixture-manifest-sha256" || a == "--pf_fixture_manifest_sha256") {
      if (!next(&fixture_sha)) return fatal(a + " needs a value");
    } else if (a == "--pf-profile-library-manifest-sha256" || a == "--pf_profile_library_manifest_sha256") {
      if (!next(&library_sha)) return fatal(a + " needs a value");
    } else if (a == "--precision") {
      std::string v;
      if (!next(&v)) return fatal("--precision needs a value");
      if (v == "fp64")
        precision = PF_FP64;
      else if (v == "mixed")
        precision = PF_MIXED;
      else
        return fatal("--precision must be fp64 or mixed");
    } else if (a == "--minimize-max-iter" || a == "--minimize-dt" || a == "--V0" || a == "--eta-lambda-vol" ||
               a == "--eta_lambda_vol" || a == "--minimize-post-projection-iters" ||
               a == "--minimize-rms-dphi-threshold" || a == "--minimize-rms-res-threshold" ||
               a == "--minimize-vol-err-rel-threshold" || a == "--minimize-convergence-steps" ||
               a == "--minimize-min-pseudo-time" || a == "--radius" || a == "--minimize-rms-dY-threshold" ||
               a == "--minimize-rms-dy-threshold" || a == "--minimize-energy-diff-rel-threshold" ||
               a == "--minimize-xB-max-safe" || a == "--minimize-target-mass-code" ||
               a == "--minimize-mass-tolerance-relative" || a == "--minimize-mass-max-iterations") {
      // main_cuda.cu:23762-23862
      std::string v;
      if (!next(&v)) return fatal(a + " needs a value");
      const double x = std::atof(v.c_str());
      if (a == "--minimize-max-iter") {
        mo.max_iter = std::atoi(v.c_str());
        max_iter_set = true;
      } else if (a == "--minimize-dt") {
        mo.dt = x;
        minimize_dt_set = true;
      } else if (a == "--V0") {
        mo.V0 = x;
      } else if (a == "--eta-lambda-vol" || a == "--eta_lambda_vol") {
        mo.eta_lambda_vol = x;
      } else if (a == "--minimize-post-projection-iters") {
        mo.post_projection_iters = std::max(0, std::atoi(v.c_str()));
      } else if (a == "--minimize-rms-dphi-threshold") {
        mo.rms_dphi_threshold = x;
      } else if (a == "--minimize-rms-res-threshold") {
        mo.rms_res_threshold = x;
      } else if (a == "--minimize-vol-err-rel-threshold") {
        mo.vol_err_rel_threshold = x;
      } else if (a == "--minimize-convergence-steps") {
        mo.convergence_steps = std::atoi(v.c_str());
      } else if (a == "--minimize-min-pseudo-time") {
        mo.min_pseudo_time = x;
      } else if (a == "--minimize-rms-dY-threshold" || a == "--minimize-rms-dy-threshold") {
        mo.rms_dY_threshold = x;
      } else if (a == "--minimize-energy-diff-rel-threshold") {
        mo.energy_diff_rel_threshold = x;
      } else if (a == "--minimize-xB-max-safe") {
        mo.xB_max_safe = x;
      } else if (a == "--minimize-target-mass-code") {
        mo.target_mass_code = x;
      } else if (a == "--minimize-mass-tolerance-relative") {
        mo.mass_tolerance_relative = x;
      } else if (a == "--minimize-mass-max-iterations") {
        mo.mass_max_iterations = std::atoi(v.c_str());
      } else {
        seed_radius = x;
      }
    } else if (a == "--minimize-full-model" || a == "--minimize-full") {
      mo.full_model = 1;   // main_cuda.cu:23726-23730
    } else if (a == "--minimize-mass-constraint") {
      mo.mass_constraint = 1;   // main_cuda.cu:23732-23735
    } else if (a == "--minimize-freeze-phi") {
      mo.freeze_phi = 1;
    } else if (a == "--minimize-component-volume-constraints") {
      comp_enabled = true;
    } else if (a == "--minimize-component-count") {
      std::string v;
      if (!next(&v)) return fatal(a + " needs a value");
      comp_count = std::atoi(v.c_str());
    } else if (a == "--minimize-component-label-raw") {
      if (!next(&comp_label_raw)) return fatal(a + " needs a value");
    } else if (a == "--minimize-component-target-h-sums") {
      if (!next(&comp_targets_text)) return fatal(a + " needs a value");
    } else if (starts_with(a.c_str(), "--")) {
      std::fprintf(stderr, "[warn] flag not handled by main_pf: %s\n", a.c_str());
    }
  }
  // a continuation without --pf-param-file takes the parameters stored next to
  // the VTK it continues (main_cuda.cu:22543-22565)
  std::string stored_params;
  if (continuing && !continue_phi_vtk.empty() && !parent_dir(continue_phi_vtk).empty()) {
    const std::string cand = parent_dir(continue_phi_vtk) + "/pf_input.params";
    struct stat st;
    if (stat(cand.c_str(), &st) == 0) {
      if (!param_file) {
        stored_params = cand;
        param_file = stored_params.c_str();
        std::printf("[continue-param] loaded stored PF params from %s\n", cand.c_str());
      } else {
        std::printf("[continue-param] explicit --pf-param-file takes precedence; ignoring stored PF params at %s\n",
                    cand.c_str());
      }
    }
  }
  if (!param_file)
    return fatal("physical parameters have no built-in defaults: pass --pf-param-file <path>");
  if (mode != "dynamics" && mode != "minimize")
    return fatal("--mode=" + mode + " is not implemented (dynamics, minimize and their -continue forms)");
  // main_cuda.cu:24752-24770
  if (continuing && continue_phi_vtk.empty()) return fatal("continuation needs --continue-phi-vtk <path>");
  if (continuing && !init_mode.empty())
    return fatal("--init-mode raw_fields cannot be combined with --continue-phi-vtk/--continue-xB-vtk");
  const bool minimize = mode == "minimize";
  // main_cuda.cu:24562-24623
  if (mo.mass_constraint && !(minimize && mo.full_model))
    return fatal("--minimize-mass-constraint requires --mode=minimize --minimize-full-model");
  if (mo.freeze_phi && !(minimize && mo.full_model && mo.mass_constraint))
    return fatal("--minimize-freeze-phi requires --mode=minimize --minimize-full-model --minimize-mass-constraint");
  std::vector<double> comp_targets;
  if (comp_enabled && minimize && mo.full_model) {
    // main_cuda.cu:24638-24670
    if (!mo.mass_constraint || mo.freeze_phi)
      return fatal("--minimize-component-volume-constraints requires full-model constrained minimize with phi "
                   "evolution enabled");
    if (comp_count < 2 || comp_count > 64 || comp_label_raw.empty() || comp_targets_text.empty())
      return fatal("V5 component constraints require component count in [2,64], a label raw path, and exactly one "
                   "positive target h sum per component");
    if (mo.post_projection_iters != 0 || mo.V0 > 0.0)
      return fatal("V5 component constraints own the phi volume projection; require "
                   "--minimize-post-projection-iters 0 and V0<=0");
    if (!parse_target_h_sums(comp_targets_text, comp_count, &comp_targets)) return 2;
  } else if (comp_enabled) {
    return fatal("--minimize-component-volume-constraints requires full-model constrained minimize with phi "
                 "evolution enabled");
  }

  pf_params P;
  pf_params_default(&P);
  // positional arguments (main_cuda.cu:22527-22534); the param file is loaded
  // afterwards, so its dt wins (main_cuda.cu:22535-22542)
  auto pos = [&](int k) -> const char* { return (argc > k && argv[k][0] != '-') ? argv[k] : nullptr; };
  if (pos(1)) P.Nx = std::atoi(pos(1));
  if (pos(2)) P.Ny = std::atoi(pos(2));
  if (pos(3)) P.Nz = std::atoi(pos(3));
  if (pos(4)) P.dt = std::atof(pos(4));
  if (pos(5)) P.nsteps = std::atoi(pos(5));
  if (pos(6)) P.out_every = std::atoi(pos(6));
  if (pos(7)) P.csv_out_every = std::atoi(pos(7));
  if (pos(8)) P.elastic_enabled = std::atoi(pos(8));
  if (pf_params_load(param_file, &P) != PF_OK) return fatal(lib_error());
  if (std::isfinite(xB_out_flag)) P.ic_23d_xB_out = xB_out_flag;
  if (P.pf_params_schema_version != 1)
    return fatal("pf_params_schema_version mismatch: file=" + std::to_string(P.pf_params_schema_version) +
                 " expected=1");
  if (minimize) {
    // P.nsteps = minimize_max_iter, P.dt = minimize_dt, the latter defaulting
    // to the physical-input dt (main_cuda.cu:22540-22554, 24432-24435)
    if (!minimize_dt_set) mo.dt = P.dt;
    if (!max_iter_set) mo.max_iter = 2000;
    P.nsteps = mo.max_iter;
    P.dt = mo.dt;
    P.mode = 1;
  }
  if (zero_mode_sel == "PF_CONSERVED_Y_ZERO_MODE_V1")
    P.zero_mode = 1;
  else if (zero_mode_sel == "OFF")
    P.zero_mode = 0;
  else
    return fatal("unknown --pf-zero-mode selector: " + zero_mode_sel);
  // main_cuda.cu:24476-24491
  if (ckpt_every < 0) return fatal("--pf-checkpoint-every must be >= 0");
  if ((ckpt_every > 0 || !restart_from.empty()) && !P.zero_mode)
    return fatal("PF checkpoint/restart provenance is available only with PF_CONSERVED_Y_ZERO_MODE_V1");
  if (ckpt_every > 0 && ckpt_path.empty())
    return fatal("--pf-checkpoint-path is required when periodic checkpointing is enabled");
  if (minimize && (ckpt_every > 0 || !restart_from.empty()))
    return fatal("checkpoint/restart applies to dynamics runs");
  // main_cuda.cu:24706-24710
  if (continuing && P.zero_mode)
    return fatal("zero-mode restart cannot be mixed with VTK continuation; use --pf-restart-from for exact state "
                 "restoration");
  // initial-state identity (main_cuda.cu:24515-24555): either the complete
  // legacy triple, or the conditional-handoff class with two lowercase
  // SHA-256 manifests, the latter only for zero-mode dynamics from raw_fields
  // or a restart
  {
    auto is_lowercase_sha256 = [](const std::string& v) {
      if (v.size() != 64) return false;
      for (char c : v)
        if (!((c >= '0' && c <= '9') || (c >= 'a' && c <= 'f'))) return false;
      return true;
    };
    const std::string legacy = "LEGACY_UNSPECIFIED";
    const std::string handoff = "MASS_CONSERVING_LIBRARY_ASSEMBLED_CONDITIONAL_HANDOFF_V1";
    const bool legacy_id = init_class == legacy && fixture_sha == legacy && library_sha == legacy;
    const bool handoff_id = init_class == handoff && is_lowercase_sha256(fixture_sha) &&
                            is_lowercase_sha256(library_sha);
    if (!legacy_id && !handoff_id)
      return fatal("PF initial-state provenance must be either the complete LEGACY_UNSPECIFIED identity or " +
                   handoff + " with two lowercase SHA-256 identities");
    if (handoff_id) {
      if (!P.zero_mode || minimize)
        return fatal(handoff + " requires dynamics with PF_CONSERVED_Y_ZERO_MODE_V1 enabled");
      if (restart_from.empty() && init_mode != "raw_fields")
        return fatal("fresh " + handoff +
                     " runs require --init-mode raw_fields; continuation must use a provenance-matched checkpoint");
    }
  }
  // main_cuda.cu:24826-24841
  if (!(P.dt > 0.0)) return fatal("Invalid dt. Time step must be > 0.");
  if (P.nsteps <= 0) return fatal("Invalid nsteps=" + std::to_string(P.nsteps) + ". Total step count must be > 0.");
  if (P.out_every <= 0)
    return fatal("Invalid out_every=" + std::to_string(P.out_every) + ". VTK output interval must be > 0.");
  if (P.csv_out_every <= 0)
    return fatal("Invalid csv_out_every=" + std::to_string(P.csv_out_every) + ". CSV output interval must be > 0.");
  pf_checkpoint_provenance prov;
  pf_checkpoint_provenance_default(&P, &prov);
  auto set_identity = [](char* dst, size_t width, const std::string& v) {
    std::memset(dst, 0, width);
    std::strncpy(dst, v.c_str(), width - 1);
    return v.size() < width;
  };
  if (!set_identity(prov.initial_state_class, sizeof prov.initial_state_class, init_class) ||
      !set_identity(prov.fixture_manifest_sha256, sizeof prov.fixture_manifest_sha256, fixture_sha) ||
      !set_identity(prov.profile_library_manifest_sha256, sizeof prov.profile_library_manifest_sha256, library_sha))
    return fatal("checkpoint identity value is too long");
  // main_cuda.cu:24871-24878
  if (P.out_every > P.nsteps) {
    std::fprintf(stderr, "[warn] out_every (%d) exceeds nsteps (%d); output at the last step\n", P.out_every,
                 P.nsteps);
    P.out_every = P.nsteps;
  }
  if (P.csv_out_every > P.nsteps) {
    std::fprintf(stderr, "[warn] csv_out_every (%d) exceeds nsteps (%d); output at the last step\n",
                 P.csv_out_every, P.nsteps);
    P.csv_out_every = P.nsteps;
  }

  const size_t N = (size_t)P.Nx * P.Ny * P.Nz;
  std::vector<double> phi, xB, dydt;
  double xB_cap = NAN;
  if (init_mode == "raw_fields") {
    Meta m;
    if (meta_path.empty() || !read_meta(meta_path, &m)) return fatal("cannot read init meta: " + meta_path);
    if (m.Nx != P.Nx || m.Ny != P.Ny || m.Nz != P.Nz) return fatal("init meta grid does not match the run grid");
    if (!read_raw(phi_raw, m.dtype, N, &phi)) return fatal("cannot read phi raw field: " + phi_raw);
    if (minimize && mo.full_model && xB_raw.empty()) return fatal("full-model minimize needs --init-xB-raw");
    if (minimize && xB_raw.empty()) {
      xB.assign(N, 0.0);   // phi-only minimize allocates no composition (main_cuda.cu:2609)
    } else if (!read_raw(xB_raw, m.dtype, N, &xB)) {
      return fatal("cannot read xB raw field: " + xB_raw);
    }
    if (!dydt_raw.empty() && !read_raw(dydt_raw, m.dtype, N, &dydt))
      return fatal("cannot read dY_dt_prev raw field: " + dydt_raw);
    xB_cap = std::isfinite(m.xB_max_safe) ? m.xB_max_safe : mo.xB_max_safe;   // main_cuda.cu:2940
    if (case_tag.empty()) case_tag = "raw_fields_relax";
  } else if (continuing) {
    // load_continue_fields_from_vtk (main_cuda.cu:2602-2655): phi from its VTK;
    // with a composition (dynamics, full-model minimize) x_B from its VTK when
    // it exists, else rebuilt from phi
    if (!read_vtk_scalars(continue_phi_vtk, P.Nx, P.Ny, P.Nz, &phi, "phi")) return 2;
    for (double& v : phi) v = v < 0.0 ? 0.0 : (v > 1.0 ? 1.0 : v);
    const bool needs_x = !minimize || mo.full_model;
    if (!needs_x) {
      xB.assign(N, 0.0);
    } else if (!continue_xB_vtk.empty() && file_nonempty(continue_xB_vtk)) {
      if (!read_vtk_scalars(continue_xB_vtk, P.Nx, P.Ny, P.Nz, &xB, "xB")) return 2;
      for (double& v : xB) v = clamp_frac(v, P.xB_eps);
    } else {
      if (!continue_xB_vtk.empty())
        std::fprintf(stderr, "[warn] continuation x_B VTK missing or unreadable, rebuilding x_B from phi: %s\n",
                     continue_xB_vtk.c_str());
      rebuild_composition(P, seed_radius, phi, &xB);
    }
    if (case_tag.empty()) case_tag = next_continue_tag(continue_phi_vtk, minimize);
  } else {
    return fatal("main_pf initialises from --init-mode raw_fields or a VTK continuation only");
  }

  std::vector<int> comp_labels;
  if (!comp_targets.empty()) {
    if (!read_labels_i32(comp_label_raw, N, comp_count, &comp_labels)) return 2;
    mo.component_count = comp_count;
    mo.component_labels = comp_labels.data();
    mo.component_target_h_sums = comp_targets.data();
  }
  pf_handle* h = nullptr;
  if (pf_create(&P, precision, &h) != PF_OK) return fatal(lib_error());
  if (pf_set_fields(h, phi.data(), xB.data(), dydt.empty() ? nullptr : dydt.data(), xB_cap) != PF_OK)
    return fatal(lib_error());
  // restart: the checkpoint replaces the raw initial state (main_cuda.cu:27340-27395)
  long long first_step = 1;
  if (!restart_from.empty()) {
    if (pf_checkpoint_read(h, restart_from.c_str(), &prov) != PF_OK) return fatal(lib_error());
    pf_scalars sr;
    pf_scalars_get(h, &sr);
    if (sr.steps_done >= P.nsteps)
      return fatal("restart accepted_step=" + std::to_string(sr.steps_done) +
                   " must be less than requested final nsteps=" + std::to_string(P.nsteps));
    first_step = sr.steps_done + 1;
    std::printf("PF_ZERO_MODE_RESTART_LOADED checkpoint=%s accepted_step=%lld\n", restart_from.c_str(),
                sr.steps_done);
  }

  // Results tree (main_cuda.cu:25270-25462)
  const char* root_env = std::getenv("CUDA_STO_RESULTS_ROOT");
  const std::string root = (root_env && root_env[0]) ? root_env : "Results";
  char run_dir[512];
  std::string out_dir, csv_tag = case_tag;
  if (continuing) {
    // the continued run's own tree (main_cuda.cu:25363-25386, 25443-25456):
    // Results root + the source run directory's name when the root is set,
    // else the source run directory itself; CSVs keep the source case's tag
    const std::string src_case = parent_dir(continue_phi_vtk);
    const std::string src_run = parent_dir(src_case);
    if (root_env && root_env[0] && !base_name(src_run).empty() && !src_run.empty())
      out_dir = root + "/" + base_name(src_run);
    else if (!src_run.empty())
      out_dir = src_run;
    else
      out_dir = root;
    if (!src_case.empty()) csv_tag = base_name(src_case);
  }
  if (continuing) {
  } else if (minimize) {
    // the seed radius in nm from the physical dx (main_cuda.cu:25283-25396)
    const double dx_m = P.ic_phi_iface_w > 1e-30 ? P.lambda_sm_m / (2.0 * P.ic_phi_iface_w) : P.dx * 1e-9;
    const double unit_to_m = std::fabs(P.dx) > 1e-30 ? dx_m / P.dx : 1e-9;
    std::snprintf(run_dir, sizeof run_dir, "%s_T%.0f_cuda_%dx%dx%d_dt%.3g_steps%d_r%.3fnm_xB%.3f",
                  P.elastic_enabled ? "chel" : "ch", P.temperature_C, P.Nx, P.Ny, P.Nz, P.dt, P.nsteps,
                  seed_radius * unit_to_m * 1e9, P.ic_23d_xB_out);
  } else {
    std::snprintf(run_dir, sizeof run_dir, "%s_T%.0f_cuda_%dx%dx%d_dt%.3g_steps%d_xB%.3f",
                  P.elastic_enabled ? "chel" : "ch", P.temperature_C, P.Nx, P.Ny, P.Nz, P.dt, P.nsteps,
                  P.ic_23d_xB_out);
  }
  if (!continuing) out_dir = root + "/" + run_dir;
  const std::string case_dir = out_dir + "/" + case_tag;
  mkdir(root.c_str(), 0755);
  mkdir(out_dir.c_str(), 0755);
  mkdir(case_dir.c_str(), 0755);
  if (!minimize && !continuing) copy_text_file(param_file, out_dir + "/pf_input.params");
  copy_text_file(param_file, case_dir + "/pf_input.params");
  if (minimize) {
    const int rc = run_minimize(h, P, mo, case_dir, case_tag, csv_tag, continuing, seed_radius);
    pf_destroy(h);
    return rc;
  }

  // a continuation appends to an existing log of its case, numbering on
  // (main_cuda.cu:25571-25590)
  const std::string csv_path = case_dir + "/vf_precip_vs_time_" + csv_tag + ".csv";
  int csv_step0 = 0;
  double csv_t0 = 0.0, csv_treal0 = 0.0;
  const bool csv_append = continuing && file_nonempty(csv_path) && last_csv_row(csv_path, &csv_step0, &csv_t0,
                                                                                 &csv_treal0);
  if (!csv_append) csv_step0 = 0;
  FILE* csv = std::fopen(csv_path.c_str(), csv_append ? "a" : "w");
  if (csv && !csv_append) std::fprintf(csv, "step,t_code,t_real_s,vf_precip,R_avg\n");
  const bool accel = P.elastic_enabled && (P.elastic_warm_start_enabled || P.elastic_residual_control_enabled);
  // a restart appends to the traces of the run it continues (main_cuda.cu:25620-25650)
  const bool restarting = !restart_from.empty();
  auto open_trace = [&](const std::string& path, bool* fresh) -> FILE* {
    struct stat st;
    const bool empty = stat(path.c_str(), &st) != 0 || st.st_size == 0;
    std::FILE* f = std::fopen(path.c_str(), restarting ? "a" : "w");
    *fresh = f && (!restarting || empty);   // a header only at the start of a file
    return f;
  };
  bool etrace_fresh = false, ztrace_fresh = false;
  FILE* etrace = accel ? open_trace(case_dir + "/elastic_solver_trace.csv", &etrace_fresh) : nullptr;
  if (etrace && etrace_fresh)
    std::fprintf(etrace,
                 "step,warm_start_used,iterations,relative_residual,converged,iter_max,tolerance,"
                 "elastic_solver_wall_s,source_field_step\n");
  FILE* ztrace = P.zero_mode ? open_trace(case_dir + "/pf_zero_mode_trace.csv", &ztrace_fresh) : nullptr;
  if (ztrace && ztrace_fresh)
    std::fprintf(ztrace, "step,target_mass_code,lambda,iterations,mass_code\n");

  pf_scalars s0;
  pf_scalars_get(h, &s0);
  // the engine's frozen zero-mode target (after a restart, the checkpoint's
  // target_mass_code rather than the restored fields' own sum)
  const double target = P.zero_mode ? s0.zero_mode_target_mass : s0.mass_xBtot;
  auto write_row = [&](int step, const pf_scalars& sc) {
    double R_avg = 0.0;
    if (sc.n_if > 0.0) {
      const double ratio = sc.n_in / (sc.n_if / (P.ic_phi_iface_w * 2.0));
      R_avg = (P.Ny == 2 ? 2.0 : 3.0) * P.dx * ratio;
    }
    if (csv) {
      std::fprintf(csv, "%d,%.8f,%.8e,%.8e,%.8e\n", step + csv_step0, step * P.dt + csv_t0,
                   step * P.dt * P.t_real_unit + csv_treal0, sc.volume_fraction, R_avg);
      std::fflush(csv);
    }
    std::printf("[3D] step=%d <x_Btot>=%.10f vf_precip=%.6e R_avg=%.4f\n", step, sc.mean_xBtot,
                sc.volume_fraction, R_avg);
  };
  if (!restarting && !continuing) write_row(0, s0);   // main_cuda.cu:28309
  auto dump_vtk = [&](int step, bool final_) {
    std::vector<double> f(N);
    const std::string sfx = final_ ? "final" : std::to_string(step);
    pf_get_field(h, PF_PHI, f.data());
    write_vtk(case_dir + "/phi_" + sfx + ".vtk", f, P.Nx, P.Ny, P.Nz, "phi", step);
    pf_get_field(h, PF_XB, f.data());
    write_vtk(case_dir + "/xB_" + sfx + ".vtk", f, P.Nx, P.Ny, P.Nz, "xB", step);
    pf_get_field(h, PF_XBTOT, f.data());
    write_vtk(case_dir + "/xBtot_" + sfx + ".vtk", f, P.Nx, P.Ny, P.Nz, "xB_tot", step);
  };

  std::vector<double> wall;
  wall.reserve(P.nsteps);
  for (int step = (int)first_step; step <= P.nsteps; ++step) {
    const double t0 = now_s();
    const int rc = pf_step(h, 1);
    pf_sync(h);
    const double w = now_s() - t0;
    if (rc != PF_OK) {
      const std::string msg = lib_error();
      if (etrace) std::fclose(etrace);
      return fatal(msg + " (step " + std::to_string(step) + ")");
    }
    wall.push_back(w);
    pf_scalars sc;
    pf_scalars_get(h, &sc);
    if (etrace)
      std::fprintf(etrace, "%d,%d,%d,%.17e,%d,%d,%.17e,%.9f,%d\n", step, step > 1 ? 1 : 0,
                   sc.last_elastic_iterations, sc.last_elastic_residual,
                   sc.last_elastic_residual <= P.elastic_residual_tolerance ? 1 : 0, P.elastic_iter_max,
                   P.elastic_residual_tolerance, w, step - 1);
    if (ztrace)
      std::fprintf(ztrace, "%d,%.17e,%.17e,%d,%.17e\n", step, target, sc.zero_mode_lambda, sc.zero_mode_iterations,
                   sc.mass_xBtot);
    if (step % P.csv_out_every == 0 || step == P.nsteps) write_row(step, sc);
    if (step % P.out_every == 0) dump_vtk(step, false);
    // periodic checkpoint (main_cuda.cu:33494-33600)
    if (ckpt_every > 0 && (step % ckpt_every == 0 || step == P.nsteps)) {
      if (pf_checkpoint_write(h, ckpt_path.c_str(), &prov) != PF_OK) return fatal(lib_error());
      std::printf("PF_ZERO_MODE_CHECKPOINT_WRITTEN step=%d path=%s target_mass_code=%.17e lambda=%.17e "
                  "initial_state_class=%s fixture_manifest_sha256=%s profile_library_manifest_sha256=%s\n",
                  step, ckpt_path.c_str(), target, sc.zero_mode_lambda, prov.initial_state_class,
                  prov.fixture_manifest_sha256, prov.profile_library_manifest_sha256);
    }
  }
  dump_vtk(P.nsteps, true);
  if (csv) std::fclose(csv);
  if (etrace) std::fclose(etrace);
  if (ztrace) std::fclose(ztrace);

  // performance summary (main_cuda.cu:34245-34299)
  if (!wall.empty()) {
    std::vector<double> srt = wall;
    std::sort(srt.begin(), srt.end());
    double total = 0.0;
    for (double w : wall) total += w;
    const size_t n = wall.size();
    const double median = (n % 2 == 0) ? 0.5 * (srt[n / 2 - 1] + srt[n / 2]) : srt[n / 2];
    const size_t warm = n > 10 ? 10 : (n > 5 ? 5 : 0);
    double ws = 0.0;
    for (size_t i = warm; i < n; ++i) ws += wall[i];
    const double warm_avg = (n > warm) ? ws / (double)(n - warm) : total / (double)n;
    FILE* pc = std::fopen((case_dir + "/performance_summary.csv").c_str(), "w");
    if (pc) {
      std::fprintf(pc,
                   "case_name,picard_enabled,picard_iters,picard_omega,nsteps,dt,total_walltime_s,"
                   "avg_walltime_per_step_s,median_walltime_per_step_s,min_walltime_per_step_s,max_walltime_per_step_s,"
                   "warmup_excluded_avg_walltime_per_step_s,estimated_steps_per_hour,relative_slowdown_vs_baseline\n");
      std::fprintf(pc, "%s,%.17e,%.17e,%.17e,%.17e,%.17e,%.17e,%.17e,%.17e,%.17e,%.17e,%.17e,%.17e,%.17e\n",
                   case_tag.c_str(), 0.0, 1.0, 1.0, (double)n, P.dt, total, total / n, median, srt.front(),
                   srt.back(), warm_avg, 3600.0 / warm_avg, 1.0);
      std::fclose(pc);
    }
    FILE* pj = std::fopen((case_dir + "/performance_summary.json").c_str(), "w");
    if (pj) {
      std::fprintf(pj,
                   "{\n  \"case_name\": \"%s\",\n  \"nsteps\": %zu,\n  \"dt\": %.17e,\n  \"total_walltime_s\": %.17e,\n"
                   "  \"avg_walltime_per_step_s\": %.17e,\n  \"median_walltime_per_step_s\": %.17e,\n"
                   "  \"min_walltime_per_step_s\": %.17e,\n  \"max_walltime_per_step_s\": %.17e,\n"
                   "  \"warmup_excluded_avg_walltime_per_step_s\": %.17e,\n  \"estimated_steps_per_hour\": %.17e,\n"
                   "  \"precision\": \"%s\"\n}\n",
                   case_tag.c_str(), n, P.dt, total, total / n, median, srt.front(), srt.back(), warm_avg,
                   3600.0 / warm_avg, precision == PF_FP64 ? "fp64" : "mixed");
      std::fclose(pj);
    }
  }
  pf_destroy(h);
  return 0;
}
