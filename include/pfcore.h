/*
 * pfcore.h - C ABI of the MI355X phase-field time-step library (libpfcore.so).
 *
 * The reference has no library boundary: its drivers build argv and run the
 * `main_cuda` process (tools/analysis/run_pure_dynamics_mass_drift_benchmark.py:63-70,
 * scripts/run_pf_only_baseline_workstation.py:41-48), and inside that process
 * the step is ~150 `launch_*` C++ wrappers declared in cuda_kernels.h:167-1219,
 * called from the step loop main_cuda.cu:28427-33604.  This header is the seam
 * those wrappers and the loop collapse into.  Each entry point names the
 * reference code it replaces.
 *
 * Conventions: every function returns 0 on success and a negative PF_E_* code
 * on failure (the message is kept per thread, see pf_last_error).  Host
 * arrays are owned by the caller and copied; device memory is owned by the
 * handle.  Real fields are row-major [Nx][Ny][Nz] doubles (z fastest), the
 * reference's layout (cuda_kernels.cu:4157-4160).  One HIP stream per handle;
 * a handle is not thread-safe, distinct handles may be used from distinct
 * threads.
 */
#ifndef PFCORE_H
#define PFCORE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PFCORE_ABI_VERSION 1

enum pf_status {
  PF_OK = 0,
  PF_E_ARG = -1,            /* invalid argument / unsupported size */
  PF_E_PARAMS = -2,         /* .params file error or missing required key */
  PF_E_HIP = -3,            /* HIP runtime failure */
  PF_E_NONCONVERGED = -4,   /* elastic fixed point did not reach the tolerance */
  PF_E_ZERO_MODE = -5,      /* zero-mode shift rejected */
  PF_E_NONFINITE = -6,      /* non-finite residual or field */
  PF_E_IO = -7              /* checkpoint file: I/O, format, provenance or checksum */
};

enum pf_precision {
  PF_FP64 = 0,   /* all transforms and updates in FP64 */
  PF_MIXED = 1   /* FP32 transforms/updates; FP64 reductions, zero mode, residual */
};

enum pf_field {
  PF_PHI = 0, PF_XB = 1, PF_Y = 2, PF_DYDT = 3, PF_XBTOT = 4, PF_TRSIGMA = 5,
  /* stress components of the last step's elastic solution at its start-of-step
     fields (the reference's d_sigma_*_r arrays, main_cuda.cu:29138-29166);
     available for steps taken after pf_keep_stress(h, 1) */
  PF_SIGMA_XX = 6, PF_SIGMA_YY = 7, PF_SIGMA_ZZ = 8, PF_SIGMA_YZ = 9, PF_SIGMA_XZ = 10, PF_SIGMA_XY = 11
};

/* Code-unit parameters the time step reads (subset of PFParams, pf_params.h:6-702).
 * C and dC are the 21 upper-triangle Voigt entries 11,12,...,16,22,...,66
 * (keys S_ij and S_p_ij); strains are ordered xx, yy, zz, yz, xz, xy. */
typedef struct pf_params {
  int Nx, Ny, Nz;
  double dx, dy, dz, dt;
  int nsteps, out_every, csv_out_every, elastic_enabled;
  int pf_params_schema_version;
  double temperature_C, mu_reference_scale, t_real_unit;
  double v_A, v_B, Vm_compound, Vm_alpha_0, dVm_alpha_dxB;
  double W, kappa_phi, L_phi, D_alpha, D_compound;
  double eps_iso_over_vB, elastic_shift_dimless;
  double Y_clip, xB_eps;
  int thermo_convex_extrapolation_enabled;
  double S[21], S_p[21];
  double eps00[6], E0[6];
  int elastic_warm_start_enabled, elastic_residual_control_enabled;
  int elastic_iter_min, elastic_iter_max;
  double elastic_residual_tolerance, elastic_residual_absolute_floor;
  int elastic_fail_on_nonconvergence;
  int zero_mode;                       /* --pf-zero-mode PF_CONSERVED_Y_ZERO_MODE_V1 */
  double zero_mode_tolerance_relative; /* default 1e-12 (main_cuda.cu:22487) */
  int zero_mode_max_iterations;        /* default 24 (main_cuda.cu:22488) */
  int enable_Y_rhs_previous_time_level, disable_Y_rhs_gamma_term;
  double Y_rhs_term_h_scale;
  /* keys the loader records for the Results tree and unit reporting */
  double gamma_Jm2, lambda_sm_m, ic_phi_iface_w, ic_vf_init_phi, ic_vf_target_phi;
  double Vm_alpha_0_phys_m3mol;
  int mode;                            /* 0 dynamics, 1 minimize */
  /* Extensions of the accelerated elastic solver (no reference counterpart;
   * the defaults 1 and 0 reproduce main_cuda.cu:28760-29095): warm start from
   * a constant/linear/quadratic extrapolation of the last 1/2/3 converged
   * displacements, and Anderson mixing depth (0 = plain fixed point, <= 4).
   * The map, the residual and the tolerance test are unchanged. */
  int elastic_warm_start_order;
  int elastic_anderson_depth;
  /* far-field composition of the initial condition (key ic_23d_xB_out,
   * default 0.03, main_cuda.cu:16215); in full-model minimize the chemical
   * baseline x_ref when positive, else <x_B> of the initial fields
   * (main_cuda.cu:27400-27432) */
  double ic_23d_xB_out;
} pf_params;

typedef struct pf_scalars {
  double mass_xBtot;          /* sum (1-h) x_B + v_B h      (reduce_sum_xBtot_kernel) */
  double mean_xBtot;
  double volume_fraction;     /* <h(clamp01 phi)>           (reduce_sum_h_kernel) */
  double f_grad, f_dw, f_elastic, f_chem, f_total;   /* mean energy densities */
  double mean_DY;             /* last D-bar of the Y update */
  double last_elastic_residual;
  double zero_mode_lambda;
  int last_elastic_iterations;
  int zero_mode_iterations;
  long long steps_done;
  long long elastic_iterations_total;
  double n_in;                /* count of phi > 0.5          (reduce_sum_N_in_N_if_kernel) */
  double n_if;                /* count of 0.12 < phi < 0.88 */
  double zero_mode_target_mass; /* the frozen target of the zero mode (NaN without it): the mass at
                                   pf_set_fields, or the restarted checkpoint's target_mass_code
                                   (main_cuda.cu:28080-28105) */
  int kz_stride;              /* kz stride of the spectral arrays the plan tuner chose at pf_create
                                 (whole 128- or 64-byte lines or none; bitwise-neutral) */
  int zpass_rows;             /* rows per workgroup of the elastic solver's z pass, and the column-tile */
  int ypass_cols;             /* widths of the y and x passes, as the plan tuner chose them at pf_create */
  int xpass_cols;             /* (bitwise-neutral; PF_PLAN=<z>,<y>,<x> fixes them) */
} pf_scalars;

typedef struct pf_handle pf_handle;

/* Library ABI version (PFCORE_ABI_VERSION). */
int pf_version(void);

/* Copies the last error message of the calling thread. */
int pf_last_error(char* buf, size_t n);

/* Defaults of params_default (main_cuda.cu:15781-16433) for the hot-path keys. */
int pf_params_default(pf_params* p);

/* Apply one key=value (apply_pfparams_override_key, main_cuda.cu:21226-22305).
 * Returns 1 when the key is known, 0 when unknown (ignored), <0 on a bad value. */
int pf_params_set(pf_params* p, const char* key, const char* value);

/* Load a schema-v1 .params file (load_pfparams_override_file,
 * main_cuda.cu:22306-22360) including the required-key check
 * (validate_required_pfparams_presence, main_cuda.cu:1089-1144). */
int pf_params_load(const char* path, pf_params* p);

/* Allocate device state and plans (main_cuda.cu:27707-28214).
 * Sizes must factor into radices 2,3,5 and one further prime (e.g. 246). */
int pf_create(const pf_params* p, int precision, pf_handle** out);

/* Initial fields, raw_fields semantics (load_raw_init_fields,
 * main_cuda.cu:2905-2965): phi clamped to [0,1], x_B clamped to
 * [xB_eps, 1-xB_eps] and capped at xB_cap (pass NAN for none),
 * Y = logit(x_B); dY/dt defaults to 0.  Resets the elastic warm state and,
 * with the zero mode, freezes the target mass (main_cuda.cu:28080-28100). */
int pf_set_fields(pf_handle* h, const double* phi, const double* xB, const double* dYdt_or_null,
                  double xB_cap);

/* Advance nsteps time steps (the loop body main_cuda.cu:28427-31623). */
int pf_step(pf_handle* h, int nsteps);

/* Copy a real field to the host (N doubles).  PF_SIGMA_* fail with PF_E_ARG
 * unless the last step was taken with pf_keep_stress enabled. */
int pf_get_field(pf_handle* h, int which, double* out);

/* Keep (enable = 1) or stop keeping (0) the 6 stress components of each step
 * for pf_get_field; they are written by the pass that already evaluates the
 * stress (6 N values of the transform precision per step, allocated on the
 * first enable and held until pf_destroy).  Replaces reading the reference's
 * device stress arrays after a step (main_cuda.cu:29138-29166, the stress
 * that cuda_kernels.cu:7381-7480 evaluates). */
int pf_keep_stress(pf_handle* h, int enable);

/* Mass, volume fraction, energy densities and solver statistics. */
int pf_scalars_get(pf_handle* h, pf_scalars* out);

/* Elastic warm state u(k): 3 components over the retained (2/3-rule) modes,
 * [Ax][Ay][Az] in C order with no padding, interleaved (re,im) doubles.
 * pf_elastic_state_size gives the count of complex values per component. */
long long pf_elastic_state_size(pf_handle* h);
int pf_get_elastic_state(pf_handle* h, double* out);
int pf_set_elastic_state(pf_handle* h, const double* in);

/* Kernel ids for pf_time_kernel. */
enum pf_kernel_id {
  PF_K_ZPASS_HIJ = 0,      /* z pass: inverse 6 strains -> h_ij -> forward 6 (the increment form
                              where the solver iterates increments) */
  PF_K_YPASS_FWD6 = 1,     /* y pass, forward, 6 components */
  PF_K_YPASS_INV6 = 2,     /* y pass, inverse, 6 components */
  PF_K_XPASS_FWD6 = 3,     /* x pass, forward, 6 components */
  PF_K_XPASS_INV6 = 4,     /* x pass, inverse, 6 components */
  PF_K_GREEN = 5,          /* k-space Green update + residual (+ the Anderson history in the
                              accelerated solver, with a full history) */
  PF_K_ZPASS_PHIRHS = 6,   /* z pass: inverse 6 strains -> stress, dF/dphi -> forward 1 */
  PF_K_ZPASS_FLUX = 7,     /* z pass: inverse 3 gradients -> fluxes -> forward 5 */
  PF_K_AA_UPDATE = 8,      /* Anderson mixing update + strain of the increment (accelerated
                              solver with a history only; otherwise nothing is launched and
                              the bytes are 0) */
  PF_K_COUNT = 9
};

/* Time `reps` back-to-back launches of one kernel of the step on the
 * handle's stream with HIP events, using the handle's current buffers, and
 * report the mean duration (microseconds) and the bytes a launch must move
 * (each input read once, each output written once).  Measurement only:
 * the simulation state is not valid afterwards (call pf_set_fields). */
int pf_time_kernel(pf_handle* h, int kernel_id, int reps, double* us_per_launch, double* bytes_per_launch);

/* In-step launch probe.  pf_probe_kernels(h, mask) clears the probe and,
 * from the next step on, brackets every launch of the classes in `mask`
 * (bit 1 << PF_K_*; 0 switches the probe off) that the accelerated elastic
 * solve makes (the 6-component passes, the z pass, the Green and Anderson
 * updates) with HIP events on the handle's stream.  pf_probe_read
 * synchronizes the stream and reports, for one class since the last
 * pf_probe_kernels: launches, the launches that did their work (a launch
 * of an iteration after the one the device stopped the solve at does
 * nothing), their summed duration (microseconds, gated launches included)
 * and the algorithmic bytes of the working launches (as pf_time_kernel
 * counts them, with the history each iteration holds).  Launches of fused
 * batches are not probed.  Measurement only; no counterpart in the
 * reference, which has no kernel timing. */
int pf_probe_kernels(pf_handle* h, unsigned mask);
int pf_probe_read(pf_handle* h, int kernel_id, long long* launches, long long* worked, double* us_total,
                  double* bytes_total);

/* Minimize mode (--mode=minimize, phi-only or with --minimize-full-model).
 * Defaults of params_default (main_cuda.cu:16319-16347); CLI flags
 * main_cuda.cu:23762-23851.  Composition is absent: the eigenstrain is
 * h(phi) eps00 (eigenstrain_phi_only_point, cuda_kernels.cu:363-383). */
typedef struct pf_minimize_opts {
  double dt;                     /* --minimize-dt: pseudo-time step of the phi update */
  int max_iter;                  /* --minimize-max-iter (2000) */
  double V0;                     /* --V0: target <h>; <= 0 locks <h> of the first iteration */
  double eta_lambda_vol;         /* --eta-lambda-vol: multiplier under-relaxation (0.2) */
  int post_projection_iters;     /* --minimize-post-projection-iters (1) */
  double rms_dphi_threshold;     /* --minimize-rms-dphi-threshold (1e-6) */
  double rms_res_threshold;      /* --minimize-rms-res-threshold (1e-4) */
  double vol_err_rel_threshold;  /* --minimize-vol-err-rel-threshold (1e-4) */
  int convergence_steps;         /* --minimize-convergence-steps (10) */
  double min_pseudo_time;        /* minimize_min_pseudo_time (0) */
  /* --minimize-full-model (main_cuda.cu:23729): the chemical drive, the
   * composition eigenstrain and the composition update at the pseudo-time
   * step join the iteration (main_cuda.cu:29357-29700, 30283-30911) */
  int full_model;
  double rms_dY_threshold;          /* --minimize-rms-dY-threshold (5e-5) */
  double energy_diff_rel_threshold; /* --minimize-energy-diff-rel-threshold (1e-9) */
  double xB_max_safe;               /* --minimize-xB-max-safe (0.07): Y <= logit(xB_max_safe) */
  /* --minimize-mass-constraint (main_cuda.cu:32297-32366, full model only):
   * after the phi projections, the zero-mode root find shifts the clamped Y
   * of the iteration so that sum (1-h) x_B + h v_B equals the target, with Y
   * bounded to [-Y_clip, logit(xB_max_safe)]; a rejected search fails the
   * iteration with PF_E_ZERO_MODE.  The target is target_mass_code when
   * positive (it must match the initial fields), else the initial sum. */
  int mass_constraint;
  double mass_tolerance_relative;   /* --minimize-mass-tolerance-relative (1e-12) */
  int mass_max_iterations;          /* --minimize-mass-max-iterations (64) */
  double target_mass_code;          /* --minimize-target-mass-code (0: the initial ledger) */
  /* --minimize-freeze-phi (V4, main_cuda.cu:29357, 32508-32522): phi is left
   * untouched; only the composition update and the mass constraint run.
   * Needs mass_constraint, post_projection_iters 0 and V0 <= 0. */
  int freeze_phi;
  /* --minimize-component-volume-constraints (V5, main_cuda.cu:24638-24670,
   * 29446-29510, 29660-29728, 32207-32275): one h-volume constraint per
   * ownership label instead of the global one.  component_labels holds
   * Nx*Ny*Nz labels in 0..component_count-1 (the C-order int32 raw of
   * --minimize-component-label-raw, every label owning a cell) and
   * component_target_h_sums the target sum of h(phi) per label; both are
   * host arrays read by pf_minimize_begin.  Needs the mass constraint, phi
   * evolving, post_projection_iters 0 and V0 <= 0; component_count in
   * [2, 64], 0 = off.  vol_err_rel reports the largest relative error. */
  int component_count;
  const int* component_labels;
  const double* component_target_h_sums;
} pf_minimize_opts;

/* One row of energy_minimize_*.csv (main_cuda.cu:32455-32489) for phi-only
 * minimize, plus the convergence state (main_cuda.cu:32494-32560). */
typedef struct pf_minimize_stats {
  int iteration;
  int converged;                 /* condA held for convergence_steps iterations */
  int convergence_count;
  int elastic_iterations;
  double mean_h, V0, lambda_vol;
  double F_surf_hat, F_el_hat;   /* <grad + dw>, <0.5 sigma:eps_el> of phi at the iteration start */
  double interface_sum, el_core_sum;
  double rms_res, rms_dphi, vol_err_rel;
  double lambda_raw, lambda_kkt;
  /* full model (NaN in phi-only minimize): <g_bulk> - mu_mix(x_ref), the
   * total excess F_chem_excess + F_surf + F_el, the CNT summaries
   * <h>(mu0 - mu_B(x_out)) and F_surf + F_el + that (when ic_23d_xB_out > 0),
   * rms(Y^{n+1} - Y^n)/dt and the relative change of the total excess */
  double F_chem_excess_hat, F_total_excess_hat, F_chem_CNT_hat, F_total_CNT_hat;
  double rms_dY, energy_diff_rel;
  /* mass constraint (NaN and 0 without it): target, sum after the shift,
   * |sum - target| / max(|target|, 1), the shift and its search steps; the
   * search's residual and d sum/d lambda at the shift, the extremes of the
   * clamped Y it shifts and the final bracket (the columns of
   * minimize_mass_constraint_trace.csv, main_cuda.cu:25681-25685) */
  double mass_target, mass_current, mass_err_rel, mass_lambda;
  int mass_iterations, mass_newton_steps, mass_bisection_steps;
  double mass_residual, mass_derivative, mass_Ystar_min, mass_Ystar_max, mass_lambda_lower, mass_lambda_upper;
} pf_minimize_stats;

int pf_minimize_opts_default(pf_minimize_opts* o);

/* Start phi-only minimize from the handle's current phi (the minimize branch
 * of the step loop, main_cuda.cu:28341-28344, 29264-32600): resets the
 * multiplier, the volume target and the elastic warm state. */
int pf_minimize_begin(pf_handle* h, const pf_minimize_opts* o);

/* One minimize iteration; st may be NULL. */
int pf_minimize_iterate(pf_handle* h, pf_minimize_stats* st);

/* pf_minimize_begin, then iterations until convergence or max_iter; st (may
 * be NULL) receives the last iteration. */
int pf_minimize(pf_handle* h, const pf_minimize_opts* o, pf_minimize_stats* st);

/* GP-zone order parameter (not exercised by the two-phase workloads; kernel
 * level).  Keys of params_default (main_cuda.cu:15793-15828, 16164-16167). */
typedef struct pf_eta_params {
  double gp_xB_fixed, gp_delta_g0, gp_W_eta, gp_kappa_eta, gp_L_eta;
  double gp_eps_iso, gp_elastic_derivative_scale;
  int gp_elastic;                          /* gp elastic solver on and eta coupling active */
  int gp_raw_reaction_drive_only, gp_raw_reaction_drive_use_raw_units_debug;
  double gp_reaction_nu_A, gp_reaction_nu_B;
  double gp_mu_reference_raw;              /* compute_gp_mu_reference_mech_mix_host result */
} pf_eta_params;

int pf_eta_params_default(pf_eta_params* q);

/* One eta update on the handle's grid and dt (main_cuda.cu:29948-30006):
 * compute_eta_rhs_kernel (cuda_kernels.cu:845-946), the dealiased
 * semi-implicit update eta_semi_implicit_update_kernel (1045-1069) and
 * eta_normalize_and_clamp_kernel (1086-1098).  Host arrays of N doubles;
 * tr_sigma may be NULL (zero stress).  The simulation state is not touched. */
int pf_eta_update(pf_handle* h, const pf_eta_params* q, const double* eta, const double* phi,
                  const double* xB_alpha, const double* tr_sigma_or_null, double* eta_out);

/* ---------------------------------------------------------------------------
 * Warm start of the elastic fixed point.  The reference warm-starts from the
 * previous step's displacement and otherwise from the homogeneous solution
 * (main_cuda.cu:28760-28805).  These entry points let a caller supply the
 * start (the warm-start hook of the drop-in boundary); the iteration, its
 * residual and tolerance are unchanged. */

/* Set the warm state u(k) (3 compact components, interleaved re/im doubles,
 * pf_elastic_state_size values each) for the next solve; NULL forces the
 * homogeneous cold start. */
int pf_set_warm_start(pf_handle* h, const double* u_k_or_null);

/* Called at every cold start after the homogeneous start is formed; the
 * callback may replace it with pf_set_warm_start (return 0) or leave it
 * (return 0 without a call); a nonzero return aborts the step. */
typedef int (*pf_warm_start_fn)(pf_handle* h, void* user);
int pf_set_warm_start_callback(pf_handle* h, pf_warm_start_fn fn, void* user);

/* ---------------------------------------------------------------------------
 * Learned warm start (no reference counterpart): a 5-layer 3x3x3 convolution
 * network with 16 channels, BF16 weights and FP32 accumulation on the matrix
 * cores.  Input: phi and the 6 eigenstress components (C + h dC):eps0 of the
 * solve, scaled; output: 3 channels, a force density f whose Green
 * response Omega(k) f^(k) is the correction from the homogeneous cold start
 * to the converged displacement.  Training runs on the
 * GPU with the library's own forward and backward kernels and Adam. */
typedef struct pf_nn pf_nn;

/* Parameters: 5 layers of W[27][16][16] (tap = 9(dx+1)+3(dy+1)+(dz+1),
 * [cin][cout]) followed by bias[16] each, FP32. */
#define PF_NN_PARAMS 34640

/* Training geometry: batches of up to max_batch grids of Nx*Ny*Nz voxels and
 * a device-resident dataset of up to capacity snapshots of that grid. */
int pf_nn_create(int Nx, int Ny, int Nz, int max_batch, int capacity, pf_nn** out);
int pf_nn_destroy(pf_nn* net);

/* Feature scales: phi * s_phi, eigenstress * s_tau; network output times
 * s_out is the displacement correction. */
int pf_nn_set_scales(pf_nn* net, double s_phi, double s_tau, double s_out);
int pf_nn_get_scales(pf_nn* net, double* s_phi, double* s_tau, double* s_out);
int pf_nn_set_params(pf_nn* net, const float* params);   /* PF_NN_PARAMS values; resets Adam */
/* Resolution the network runs at inside a handle (pf_nn_collect, the cold-start
 * prediction): 1, the grid itself (default); 2, the grid averaged over 2x2x2
 * voxels (even Nx, Ny, Nz), whose prediction each of the 8 voxels reads back.
 * A pooled network is created with the pooled training grid (n/2)^3 and
 * collects from handles of n^3. */
int pf_nn_set_pool(pf_nn* net, int factor);
int pf_nn_get_pool(pf_nn* net, int* factor);
int pf_nn_get_params(pf_nn* net, float* params);

/* Converge the handle's elastic solve for its current phi (from its warm
 * state), then append (features, correction) to the dataset.  The handle's
 * grid must equal the training grid.  Returns the dataset size. */
int pf_nn_collect(pf_nn* net, pf_handle* h);
int pf_nn_dataset_size(pf_nn* net);

/* One Adam step on the snapshots idx[0..batch) (mean-square loss in output
 * units); loss may be NULL.  pf_nn_eval_loss evaluates without updating. */
int pf_nn_train_step(pf_nn* net, const int* idx, int batch, float lr, float* loss);
int pf_nn_eval_loss(pf_nn* net, const int* idx, int batch, float* loss);

/* Check entry points on host data for a grid of any size: in7 is
 * [batch][Nx*Ny*Nz][7] (already scaled), target [batch][3][Nx*Ny*Nz].
 * pf_nn_forward writes the output [batch][3][Nx*Ny*Nz]; pf_nn_gradients the
 * loss and the PF_NN_PARAMS gradients of the mean-square loss. */
int pf_nn_forward(pf_nn* net, int batch, int Nx, int Ny, int Nz, const float* in7, float* out3);
int pf_nn_gradients(pf_nn* net, int batch, int Nx, int Ny, int Nz, const float* in7, const float* target,
                    float* loss, float* grad);

/* Mean duration (microseconds, HIP events) of one inference over an
 * Nx*Ny*Nz grid: the five convolutions, on zero features. */
int pf_nn_time_inference(pf_nn* net, int Nx, int Ny, int Nz, int reps, double* us);

/* Use the network at every cold start of the handle's elastic solve (NULL
 * detaches).  The network must outlive the attachment. */
int pf_attach_learned_warm_start(pf_handle* h, pf_nn* net);

/* pf_nn_forward for one grid that allocates nothing.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * in7_dev [Nx*Ny*Nz][7] and out3_dev [3][Nx*Ny*Nz] are device arrays of the
 * caller (FP32, the layouts of pf_nn_forward); the packed input and the
 * activations live in a static device array of the library holding grids of
 * up to 262144 voxels (64^3; 24 MiB per GPU, held until the process ends).
 * Larger grids fail with PF_E_ARG.  The result equals pf_nn_forward's bit for
 * bit.  Every call on one GPU shares that array, so calls that overlap on one
 * GPU (from several threads or networks) race; the call returns after the
 * network's stream has finished. */
int pf_nn_forward_noalloc(pf_nn* net, int Nx, int Ny, int Nz, const float* in7_dev, float* out3_dev);

/* ---------------------------------------------------------------------------
 * Batched sweep: many independent instances (e.g. the 48 minimize runs of
 * uneven size of workload 4) driven concurrently.  Each handle keeps its own
 * stream; `threads` host workers take instances in the given order, so the
 * kernels of different instances overlap on the GPU.  Results equal those of
 * running each handle alone.  The handles must outlive the batch. */
typedef struct pf_batch pf_batch;
int pf_batch_create(int n, pf_handle* const* handles, int threads, pf_batch** out);
/* nsteps dynamics steps on every instance */
int pf_batch_step(pf_batch* b, int nsteps);
/* pf_minimize on every instance with opts[i]; stats (may be NULL) gets n entries */
int pf_batch_minimize(pf_batch* b, const pf_minimize_opts* opts, pf_minimize_stats* stats);
/* pf_minimize on every instance as one fused batch: the instances take their
 * iterations together, and each stage of an iteration (a transform pass, a
 * k-space update, a reduction, a controller) is one batched launch over all
 * instances that share the kernel (one per grid size for the transform
 * passes), driven from the calling thread on one stream with one wait per
 * phase for all instances.  For phi-only minimize with the accelerated
 * elastic solver (elastic_anderson_depth > 0) and no cold-start hook;
 * otherwise PF_E_ARG.  Results equal pf_minimize's on each handle bit for
 * bit.  No reference counterpart: the reference runs a sweep (e.g. 48 minimize
 * runs of 64^3 to 128^3 grids) as one process per instance. */
int pf_batch_minimize_fused(pf_batch* b, const pf_minimize_opts* opts, pf_minimize_stats* stats);
/* nsteps dynamics steps on every instance as one fused batch: each phase of a
 * step between the solver's and the zero mode's waits is recorded per
 * instance and replayed as one batched launch per kernel over the instances
 * that share it, recurring phases as HIP graphs (pf_batch_set_graphs); one
 * handle alone is one step as graph replays.  Accelerated elastic solver
 * (or no elasticity) and no cold-start hook; otherwise PF_E_ARG.  Results
 * equal pf_step's on each handle bit for bit in mixed precision and to
 * rounding in FP64 (a batched and a single compilation of one FP64 kernel
 * may fuse multiply-adds differently), with the same iteration counts.  No
 * reference counterpart. */
int pf_batch_step_fused(pf_batch* b, int nsteps);
/* launches of the last fused call: batched launches issued, the launches the
 * instances recorded (the count one-by-one execution issues), the HIP graphs
 * captured and the graph launches (any pointer may be NULL) */
int pf_batch_fused_launches(const pf_batch* b, long long* batched, long long* recorded, long long* graphs_built,
                            long long* graph_launches);
/* enable (1, the default) or disable (0) graph capture in fused calls: each
 * phase's batched launches are captured into a HIP graph the first time
 * their plan occurs and replayed as one graph launch when it recurs */
int pf_batch_set_graphs(pf_batch* b, int enable);
int pf_batch_destroy(pf_batch* b);

/* ---------------------------------------------------------------------------
 * Zero-mode checkpoint / restart in the reference's file format
 * (pf_zero_mode_checkpoint.h:9-96, pf_zero_mode_checkpoint.cpp:1-763):
 * V4 with the elastic warm state (904-byte header), V3 without (800 bytes),
 * V2 read-only; FNV-1a 64 payload checksum; written to <path>.tmp and
 * renamed.  A file is read only when the caller's expected provenance
 * matches it exactly.
 *
 * Provenance strings are NUL-padded; the fingerprints are those of
 * main_cuda.cu:4187-4267. */
typedef struct pf_checkpoint_provenance {
  char zero_mode[64], backend[64], composition_mode[64], y_update_mode[64];
  char explicit_context[64], reaction_discretization[64], elastic_solver_mode[64];
  char initial_state_class[96], fixture_manifest_sha256[96], profile_library_manifest_sha256[96];
  unsigned long long parameter_fingerprint, elastic_solver_fingerprint;
} pf_checkpoint_provenance;

/* The scalar part of a checkpoint (Checkpoint minus the arrays,
 * pf_zero_mode_checkpoint.h:39-76).  version is set by the reader. */
typedef struct pf_checkpoint_header {
  unsigned long long accepted_step;
  int nx, ny, nz, version;
  double dt_code, temperature_K;
  double target_mass_code, last_lambda, last_residual_code, last_derivative_code;
  unsigned long long last_iterations, accepted_zero_mode_steps;
  int elastic_present;
  unsigned long long elastic_source_field_step, elastic_last_iterations;
  double elastic_last_relative_residual;
  unsigned long long payload_checksum;
} pf_checkpoint_header;

/* pf_zero_mode_parameter_fingerprint, main_cuda.cu:4187-4221 */
unsigned long long pf_parameter_fingerprint(const pf_params* p);
/* elastic_solver_parameter_fingerprint, main_cuda.cu:4223-4267 */
unsigned long long pf_elastic_solver_fingerprint(const pf_params* p);
/* The provenance a reference run with these parameters writes
 * (main_cuda.cu:24721-24743 with the default selectors of 22478-22500):
 * zero mode V1, host Newton/bisection backend, "legacy" / "lagged_rhs",
 * LEGACY_UNSPECIFIED identities, and the elastic warm-state mode and
 * fingerprint when elastic_enabled and (warm start or residual control). */
int pf_checkpoint_provenance_default(const pf_params* p, pf_checkpoint_provenance* out);

/* File level: write_checkpoint / read_checkpoint (pf_zero_mode_checkpoint.h:79-91).
 * Arrays are [nx][ny][nz] doubles; displacement_k is ux, uy, uz over the full
 * half spectrum [nx][ny][nz/2+1] as interleaved float (re, im), present when
 * hdr->elastic_present.  The reader may be given NULL arrays. */
int pf_checkpoint_file_write(const char* path, const pf_checkpoint_provenance* prov,
                             const pf_checkpoint_header* hdr, const double* phi, const double* Y,
                             const double* xB, const double* dYdt, const float* displacement_k);
int pf_checkpoint_file_read(const char* path, const pf_checkpoint_provenance* expected,
                            pf_checkpoint_header* hdr, double* phi, double* Y, double* xB, double* dYdt,
                            float* displacement_k);

/* Handle level: the periodic checkpoint of main_cuda.cu:33494-33600 and the
 * restart of 27340-27395, 27997-28030, 28066-28105.  Requires the zero mode.
 * Write: phi, Y, x_B, dY/dt, the zero-mode runtime and, when the reference
 * would keep one, the elastic warm state after the last accepted step.  It
 * also writes <path>.state, this library's exact-resume companion (the FP64
 * warm state and extrapolation history), bound to the V4 file's checksum.
 * Read: checks grid, dt and temperature against the handle, restores the
 * fields as stored (no clamping), the zero-mode target (the field mass must
 * match it to 1e-12 relative) and the step count; the warm state comes from
 * <path>.state when it belongs to this file, else from the V4 FP32 state.
 * prov NULL means pf_checkpoint_provenance_default of the handle's params. */
int pf_checkpoint_write(pf_handle* h, const char* path, const pf_checkpoint_provenance* prov);
int pf_checkpoint_read(pf_handle* h, const char* path, const pf_checkpoint_provenance* prov);

/* Synchronise the handle's stream. */
int pf_sync(pf_handle* h);

int pf_destroy(pf_handle* h);

/* ---------------------------------------------------------------------------
 * FFT test entry points (the 3-D transforms that replace the cuFFT plans of
 * main_cuda.cu:28158-28214): forward R2C of a real [Nx][Ny][Nz] field to the
 * full half spectrum [Nx][Ny][Nz/2+1] with the 2/3 dealias applied (modes
 * outside the retained set are 0), and the matching unnormalised C2R.
 * Complex values are interleaved doubles. */
int pf_fft_forward(int Nx, int Ny, int Nz, int precision, const double* in, double* out_halfspec);
int pf_fft_inverse(int Nx, int Ny, int Nz, int precision, const double* in_halfspec, double* out);

#ifdef __cplusplus
}
#endif

#endif /* PFCORE_H */
