"""ctypes binding of libpfcore.so (the C ABI declared in include/pfcore.h).

The reference drivers start ``./main_cuda`` as a subprocess
(tools/analysis/run_pure_dynamics_mass_drift_benchmark.py:63-70); this module is
the in-process replacement they can call instead.  The product path is the
HIP library only: if ``libpfcore.so`` is missing, importing the binding raises.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("PFCORE_LIB", os.path.join(HERE, "libpfcore.so"))

PF_FP64 = 0
PF_MIXED = 1
PF_PHI, PF_XB, PF_Y, PF_DYDT, PF_XBTOT, PF_TRSIGMA = range(6)
# stress components xx, yy, zz, yz, xz, xy (kept while keep_stress is on)
PF_SIGMA_XX, PF_SIGMA_YY, PF_SIGMA_ZZ, PF_SIGMA_YZ, PF_SIGMA_XZ, PF_SIGMA_XY = range(6, 12)
PF_SIGMA = (PF_SIGMA_XX, PF_SIGMA_YY, PF_SIGMA_ZZ, PF_SIGMA_YZ, PF_SIGMA_XZ, PF_SIGMA_XY)
KERNELS = {"zpass_hij": 0, "ypass_fwd6": 1, "ypass_inv6": 2, "xpass_fwd6": 3, "xpass_inv6": 4,
           "green": 5, "zpass_phirhs": 6, "zpass_flux": 7, "aa_update": 8}

STATUS = {0: "PF_OK", -1: "PF_E_ARG", -2: "PF_E_PARAMS", -3: "PF_E_HIP", -4: "PF_E_NONCONVERGED",
          -5: "PF_E_ZERO_MODE", -6: "PF_E_NONFINITE", -7: "PF_E_IO"}

VOIGT21 = ["11", "12", "13", "14", "15", "16", "22", "23", "24", "25", "26", "33", "34", "35",
           "36", "44", "45", "46", "55", "56", "66"]


class PfParams(C.Structure):
    _fields_ = [
        ("Nx", C.c_int), ("Ny", C.c_int), ("Nz", C.c_int),
        ("dx", C.c_double), ("dy", C.c_double), ("dz", C.c_double), ("dt", C.c_double),
        ("nsteps", C.c_int), ("out_every", C.c_int), ("csv_out_every", C.c_int),
        ("elastic_enabled", C.c_int),
        ("pf_params_schema_version", C.c_int),
        ("temperature_C", C.c_double), ("mu_reference_scale", C.c_double), ("t_real_unit", C.c_double),
        ("v_A", C.c_double), ("v_B", C.c_double), ("Vm_compound", C.c_double),
        ("Vm_alpha_0", C.c_double), ("dVm_alpha_dxB", C.c_double),
        ("W", C.c_double), ("kappa_phi", C.c_double), ("L_phi", C.c_double),
        ("D_alpha", C.c_double), ("D_compound", C.c_double),
        ("eps_iso_over_vB", C.c_double), ("elastic_shift_dimless", C.c_double),
        ("Y_clip", C.c_double), ("xB_eps", C.c_double),
        ("thermo_convex_extrapolation_enabled", C.c_int),
        ("S", C.c_double * 21), ("S_p", C.c_double * 21),
        ("eps00", C.c_double * 6), ("E0", C.c_double * 6),
        ("elastic_warm_start_enabled", C.c_int), ("elastic_residual_control_enabled", C.c_int),
        ("elastic_iter_min", C.c_int), ("elastic_iter_max", C.c_int),
        ("elastic_residual_tolerance", C.c_double), ("elastic_residual_absolute_floor", C.c_double),
        ("elastic_fail_on_nonconvergence", C.c_int),
        ("zero_mode", C.c_int), ("zero_mode_tolerance_relative", C.c_double),
        ("zero_mode_max_iterations", C.c_int),
        ("enable_Y_rhs_previous_time_level", C.c_int), ("disable_Y_rhs_gamma_term", C.c_int),
        ("Y_rhs_term_h_scale", C.c_double),
        ("gamma_Jm2", C.c_double), ("lambda_sm_m", C.c_double), ("ic_phi_iface_w", C.c_double),
        ("ic_vf_init_phi", C.c_double), ("ic_vf_target_phi", C.c_double),
        ("Vm_alpha_0_phys_m3mol", C.c_double),
        ("mode", C.c_int),
        ("elastic_warm_start_order", C.c_int), ("elastic_anderson_depth", C.c_int),
        ("ic_23d_xB_out", C.c_double),
    ]


class PfScalars(C.Structure):
    _fields_ = [
        ("mass_xBtot", C.c_double), ("mean_xBtot", C.c_double), ("volume_fraction", C.c_double),
        ("f_grad", C.c_double), ("f_dw", C.c_double), ("f_elastic", C.c_double),
        ("f_chem", C.c_double), ("f_total", C.c_double), ("mean_DY", C.c_double),
        ("last_elastic_residual", C.c_double), ("zero_mode_lambda", C.c_double),
        ("last_elastic_iterations", C.c_int), ("zero_mode_iterations", C.c_int),
        ("steps_done", C.c_longlong), ("elastic_iterations_total", C.c_longlong),
        ("n_in", C.c_double), ("n_if", C.c_double), ("zero_mode_target_mass", C.c_double), ("kz_stride", C.c_int),
        ("zpass_rows", C.c_int), ("ypass_cols", C.c_int), ("xpass_cols", C.c_int),
    ]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


# every entry point declared in include/pfcore.h
class PfMinimizeOpts(C.Structure):
    _fields_ = [
        ("dt", C.c_double), ("max_iter", C.c_int), ("V0", C.c_double), ("eta_lambda_vol", C.c_double),
        ("post_projection_iters", C.c_int), ("rms_dphi_threshold", C.c_double),
        ("rms_res_threshold", C.c_double), ("vol_err_rel_threshold", C.c_double),
        ("convergence_steps", C.c_int), ("min_pseudo_time", C.c_double),
        ("full_model", C.c_int), ("rms_dY_threshold", C.c_double), ("energy_diff_rel_threshold", C.c_double),
        ("xB_max_safe", C.c_double),
        ("mass_constraint", C.c_int), ("mass_tolerance_relative", C.c_double), ("mass_max_iterations", C.c_int),
        ("target_mass_code", C.c_double), ("freeze_phi", C.c_int),
        ("component_count", C.c_int), ("component_labels", C.POINTER(C.c_int)),
        ("component_target_h_sums", C.POINTER(C.c_double)),
    ]


class PfMinimizeStats(C.Structure):
    _fields_ = [
        ("iteration", C.c_int), ("converged", C.c_int), ("convergence_count", C.c_int),
        ("elastic_iterations", C.c_int),
        ("mean_h", C.c_double), ("V0", C.c_double), ("lambda_vol", C.c_double),
        ("F_surf_hat", C.c_double), ("F_el_hat", C.c_double),
        ("interface_sum", C.c_double), ("el_core_sum", C.c_double),
        ("rms_res", C.c_double), ("rms_dphi", C.c_double), ("vol_err_rel", C.c_double),
        ("lambda_raw", C.c_double), ("lambda_kkt", C.c_double),
        ("F_chem_excess_hat", C.c_double), ("F_total_excess_hat", C.c_double),
        ("F_chem_CNT_hat", C.c_double), ("F_total_CNT_hat", C.c_double),
        ("rms_dY", C.c_double), ("energy_diff_rel", C.c_double),
        ("mass_target", C.c_double), ("mass_current", C.c_double), ("mass_err_rel", C.c_double),
        ("mass_lambda", C.c_double), ("mass_iterations", C.c_int), ("mass_newton_steps", C.c_int),
        ("mass_bisection_steps", C.c_int), ("mass_residual", C.c_double), ("mass_derivative", C.c_double),
        ("mass_Ystar_min", C.c_double), ("mass_Ystar_max", C.c_double), ("mass_lambda_lower", C.c_double),
        ("mass_lambda_upper", C.c_double),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class PfEtaParams(C.Structure):
    _fields_ = [
        ("gp_xB_fixed", C.c_double), ("gp_delta_g0", C.c_double), ("gp_W_eta", C.c_double),
        ("gp_kappa_eta", C.c_double), ("gp_L_eta", C.c_double), ("gp_eps_iso", C.c_double),
        ("gp_elastic_derivative_scale", C.c_double), ("gp_elastic", C.c_int),
        ("gp_raw_reaction_drive_only", C.c_int), ("gp_raw_reaction_drive_use_raw_units_debug", C.c_int),
        ("gp_reaction_nu_A", C.c_double), ("gp_reaction_nu_B", C.c_double), ("gp_mu_reference_raw", C.c_double),
    ]


class PfCheckpointProvenance(C.Structure):
    """pf_checkpoint_provenance: the restart identity of a checkpoint
    (pf_zero_mode_checkpoint.h:24-37)."""
    _fields_ = [("zero_mode", C.c_char * 64), ("backend", C.c_char * 64),
                ("composition_mode", C.c_char * 64), ("y_update_mode", C.c_char * 64),
                ("explicit_context", C.c_char * 64), ("reaction_discretization", C.c_char * 64),
                ("elastic_solver_mode", C.c_char * 64), ("initial_state_class", C.c_char * 96),
                ("fixture_manifest_sha256", C.c_char * 96),
                ("profile_library_manifest_sha256", C.c_char * 96),
                ("parameter_fingerprint", C.c_ulonglong), ("elastic_solver_fingerprint", C.c_ulonglong)]

    TEXT = ("zero_mode", "backend", "composition_mode", "y_update_mode", "explicit_context",
            "reaction_discretization", "elastic_solver_mode", "initial_state_class",
            "fixture_manifest_sha256", "profile_library_manifest_sha256")

    def as_dict(self):
        d = {k: getattr(self, k).decode() for k in self.TEXT}
        d["parameter_fingerprint"] = int(self.parameter_fingerprint)
        d["elastic_solver_fingerprint"] = int(self.elastic_solver_fingerprint)
        return d


class PfCheckpointHeader(C.Structure):
    _fields_ = [("accepted_step", C.c_ulonglong), ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
                ("version", C.c_int), ("dt_code", C.c_double), ("temperature_K", C.c_double),
                ("target_mass_code", C.c_double), ("last_lambda", C.c_double),
                ("last_residual_code", C.c_double), ("last_derivative_code", C.c_double),
                ("last_iterations", C.c_ulonglong), ("accepted_zero_mode_steps", C.c_ulonglong),
                ("elastic_present", C.c_int), ("elastic_source_field_step", C.c_ulonglong),
                ("elastic_last_iterations", C.c_ulonglong), ("elastic_last_relative_residual", C.c_double),
                ("payload_checksum", C.c_ulonglong)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


EXPORTS = {
    "pf_version": (C.c_int, []),
    "pf_last_error": (C.c_int, [C.c_char_p, C.c_size_t]),
    "pf_params_default": (C.c_int, [C.POINTER(PfParams)]),
    "pf_params_set": (C.c_int, [C.POINTER(PfParams), C.c_char_p, C.c_char_p]),
    "pf_params_load": (C.c_int, [C.c_char_p, C.POINTER(PfParams)]),
    "pf_create": (C.c_int, [C.POINTER(PfParams), C.c_int, C.POINTER(C.c_void_p)]),
    "pf_set_fields": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_double]),
    "pf_step": (C.c_int, [C.c_void_p, C.c_int]),
    "pf_get_field": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p]),
    "pf_keep_stress": (C.c_int, [C.c_void_p, C.c_int]),
    "pf_scalars_get": (C.c_int, [C.c_void_p, C.POINTER(PfScalars)]),
    "pf_elastic_state_size": (C.c_longlong, [C.c_void_p]),
    "pf_get_elastic_state": (C.c_int, [C.c_void_p, C.c_void_p]),
    "pf_set_elastic_state": (C.c_int, [C.c_void_p, C.c_void_p]),
    "pf_sync": (C.c_int, [C.c_void_p]),
    "pf_time_kernel": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.POINTER(C.c_double),
                                 C.POINTER(C.c_double)]),
    "pf_probe_kernels": (C.c_int, [C.c_void_p, C.c_uint]),
    "pf_probe_read": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_longlong), C.POINTER(C.c_longlong),
                                C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "pf_minimize_opts_default": (C.c_int, [C.POINTER(PfMinimizeOpts)]),
    "pf_minimize_begin": (C.c_int, [C.c_void_p, C.POINTER(PfMinimizeOpts)]),
    "pf_minimize_iterate": (C.c_int, [C.c_void_p, C.POINTER(PfMinimizeStats)]),
    "pf_minimize": (C.c_int, [C.c_void_p, C.POINTER(PfMinimizeOpts), C.POINTER(PfMinimizeStats)]),
    "pf_eta_params_default": (C.c_int, [C.POINTER(PfEtaParams)]),
    "pf_eta_update": (C.c_int, [C.c_void_p, C.POINTER(PfEtaParams), C.c_void_p, C.c_void_p, C.c_void_p,
                                C.c_void_p, C.c_void_p]),
    "pf_set_warm_start": (C.c_int, [C.c_void_p, C.c_void_p]),
    "pf_set_warm_start_callback": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "pf_nn_create": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_void_p)]),
    "pf_nn_destroy": (C.c_int, [C.c_void_p]),
    "pf_nn_set_scales": (C.c_int, [C.c_void_p, C.c_double, C.c_double, C.c_double]),
    "pf_nn_set_pool": (C.c_int, [C.c_void_p, C.c_int]),
    "pf_nn_get_pool": (C.c_int, [C.c_void_p, C.POINTER(C.c_int)]),
    "pf_nn_get_scales": (C.c_int, [C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_double),
                                   C.POINTER(C.c_double)]),
    "pf_nn_set_params": (C.c_int, [C.c_void_p, C.c_void_p]),
    "pf_nn_get_params": (C.c_int, [C.c_void_p, C.c_void_p]),
    "pf_nn_collect": (C.c_int, [C.c_void_p, C.c_void_p]),
    "pf_nn_dataset_size": (C.c_int, [C.c_void_p]),
    "pf_nn_train_step": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int, C.c_float, C.POINTER(C.c_float)]),
    "pf_nn_eval_loss": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int, C.POINTER(C.c_float)]),
    "pf_nn_forward": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_void_p]),
    "pf_nn_forward_noalloc": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_void_p]),
    "pf_nn_gradients": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_void_p,
                                  C.POINTER(C.c_float), C.c_void_p]),
    "pf_nn_time_inference": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_double)]),
    "pf_attach_learned_warm_start": (C.c_int, [C.c_void_p, C.c_void_p]),
    "pf_batch_create": (C.c_int, [C.c_int, C.c_void_p, C.c_int, C.POINTER(C.c_void_p)]),
    "pf_batch_step": (C.c_int, [C.c_void_p, C.c_int]),
    "pf_batch_minimize": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "pf_batch_minimize_fused": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "pf_batch_step_fused": (C.c_int, [C.c_void_p, C.c_int]),
    "pf_batch_fused_launches": (C.c_int, [C.c_void_p] + [C.POINTER(C.c_longlong)] * 4),
    "pf_batch_set_graphs": (C.c_int, [C.c_void_p, C.c_int]),
    "pf_batch_destroy": (C.c_int, [C.c_void_p]),
    "pf_destroy": (C.c_int, [C.c_void_p]),
    "pf_fft_forward": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_void_p]),
    "pf_fft_inverse": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_void_p]),
    "pf_parameter_fingerprint": (C.c_ulonglong, [C.POINTER(PfParams)]),
    "pf_elastic_solver_fingerprint": (C.c_ulonglong, [C.POINTER(PfParams)]),
    "pf_checkpoint_provenance_default": (C.c_int, [C.POINTER(PfParams), C.POINTER(PfCheckpointProvenance)]),
    "pf_checkpoint_file_write": (C.c_int, [C.c_char_p, C.POINTER(PfCheckpointProvenance),
                                           C.POINTER(PfCheckpointHeader), C.c_void_p, C.c_void_p, C.c_void_p,
                                           C.c_void_p, C.c_void_p]),
    "pf_checkpoint_file_read": (C.c_int, [C.c_char_p, C.POINTER(PfCheckpointProvenance),
                                          C.POINTER(PfCheckpointHeader), C.c_void_p, C.c_void_p, C.c_void_p,
                                          C.c_void_p, C.c_void_p]),
    "pf_checkpoint_write": (C.c_int, [C.c_void_p, C.c_char_p, C.POINTER(PfCheckpointProvenance)]),
    "pf_checkpoint_read": (C.c_int, [C.c_void_p, C.c_char_p, C.POINTER(PfCheckpointProvenance)]),
}


class PfError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{STATUS.get(code, code)}: {msg}")
        self.code = code


def load_library(path=LIB_PATH):
    if not os.path.exists(path):
        raise FileNotFoundError(
            f"{path} is not built; run `python -m cuda_sto_pf_amd.build` (there is no CPU fallback)")
    lib = C.CDLL(path)
    for name, (res, args) in EXPORTS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        _LIB = load_library()
    return _LIB


def _check(rc):
    if rc < 0:
        buf = C.create_string_buffer(1024)
        lib().pf_last_error(buf, 1024)
        raise PfError(rc, buf.value.decode(errors="replace"))
    return rc


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p)


def params_from_dict(d):
    """Fill a PfParams from a dict (keys of the reference .params file plus
    C / dC 6x6 Voigt matrices and eps00 / E0 6-tuples)."""
    p = PfParams()
    _check(lib().pf_params_default(C.byref(p)))
    for k, v in d.items():
        if k in ("C", "dC"):
            m = np.asarray(v, dtype=np.float64)
            arr = p.S if k == "C" else p.S_p
            for idx, suf in enumerate(VOIGT21):
                arr[idx] = float(m[int(suf[0]) - 1, int(suf[1]) - 1])
        elif k in ("eps00", "E0"):
            arr = getattr(p, k)
            for i in range(6):
                arr[i] = float(v[i])
        elif k == "zero_mode_tolerance_relative":
            p.zero_mode_tolerance_relative = float(v)
        elif hasattr(p, k):
            cur = getattr(p, k)
            setattr(p, k, type(cur)(v))
    return p


def load_params_file(path):
    p = PfParams()
    _check(lib().pf_params_default(C.byref(p)))
    _check(lib().pf_params_load(path.encode(), C.byref(p)))
    return p


class Simulation:
    """One device-resident simulation (a pf_handle)."""

    def __init__(self, params, precision=PF_FP64):
        if isinstance(params, dict):
            params = params_from_dict(params)
        self.params = params
        self.shape = (params.Nx, params.Ny, params.Nz)
        h = C.c_void_p()
        _check(lib().pf_create(C.byref(params), int(precision), C.byref(h)))
        self._h = h

    def close(self):
        if self._h:
            lib().pf_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_fields(self, phi, xB, dYdt=None, xB_cap=float("nan")):
        phi = np.ascontiguousarray(phi, dtype=np.float64).reshape(self.shape)
        xB = np.ascontiguousarray(xB, dtype=np.float64).reshape(self.shape)
        d = None if dYdt is None else np.ascontiguousarray(dYdt, dtype=np.float64).reshape(self.shape)
        _check(lib().pf_set_fields(self._h, _ptr(phi), _ptr(xB), None if d is None else _ptr(d),
                                   float(xB_cap)))

    def step(self, n=1):
        _check(lib().pf_step(self._h, int(n)))

    def field(self, which):
        out = np.empty(self.shape, dtype=np.float64)
        _check(lib().pf_get_field(self._h, int(which), _ptr(out)))
        return out

    def keep_stress(self, enable=True):
        """Keep the 6 stress components of each following step (PF_SIGMA_*)."""
        _check(lib().pf_keep_stress(self._h, 1 if enable else 0))

    def scalars(self):
        s = PfScalars()
        _check(lib().pf_scalars_get(self._h, C.byref(s)))
        return s.as_dict()

    def elastic_state(self):
        n = lib().pf_elastic_state_size(self._h)
        out = np.empty(3 * n * 2, dtype=np.float64)
        _check(lib().pf_get_elastic_state(self._h, _ptr(out)))
        return out.view(np.complex128).reshape(3, n)

    def set_elastic_state(self, u):
        u = np.ascontiguousarray(u, dtype=np.complex128)
        _check(lib().pf_set_elastic_state(self._h, _ptr(u)))

    def sync(self):
        _check(lib().pf_sync(self._h))

    def checkpoint_write(self, path, prov=None):
        """Zero-mode checkpoint (V4 with the elastic warm state, V3 without)
        plus the exact-resume companion <path>.state."""
        _check(lib().pf_checkpoint_write(self._h, str(path).encode(), None if prov is None else C.byref(prov)))

    def checkpoint_read(self, path, prov=None):
        _check(lib().pf_checkpoint_read(self._h, str(path).encode(), None if prov is None else C.byref(prov)))

    def minimize_begin(self, opts):
        self._opts = opts
        _check(lib().pf_minimize_begin(self._h, C.byref(opts)))

    def minimize_iterate(self):
        st = PfMinimizeStats()
        _check(lib().pf_minimize_iterate(self._h, C.byref(st)))
        return st.as_dict()

    def minimize(self, opts):
        st = PfMinimizeStats()
        _check(lib().pf_minimize(self._h, C.byref(opts), C.byref(st)))
        return st.as_dict()

    def eta_update(self, q, eta, phi, xB_alpha, tr_sigma=None):
        """One GP-zone eta update on this grid (pf_eta_update); q is a dict of
        pf_eta_params fields over the defaults."""
        qq = PfEtaParams()
        _check(lib().pf_eta_params_default(C.byref(qq)))
        for k, v in q.items():
            setattr(qq, k, v)
        arrs = [np.ascontiguousarray(a, dtype=np.float64).reshape(self.shape) for a in (eta, phi, xB_alpha)]
        t = None if tr_sigma is None else np.ascontiguousarray(tr_sigma, dtype=np.float64).reshape(self.shape)
        out = np.empty(self.shape, dtype=np.float64)
        _check(lib().pf_eta_update(self._h, C.byref(qq), _ptr(arrs[0]), _ptr(arrs[1]), _ptr(arrs[2]),
                                   None if t is None else _ptr(t), _ptr(out)))
        return out

    def set_warm_start(self, u=None):
        """Warm state for the next elastic solve (3 x pf_elastic_state_size
        complex values), or None for the homogeneous cold start."""
        if u is None:
            _check(lib().pf_set_warm_start(self._h, None))
        else:
            u = np.ascontiguousarray(u, dtype=np.complex128)
            _check(lib().pf_set_warm_start(self._h, _ptr(u)))

    def attach_net(self, net):
        """Use a learned network at every cold start (None detaches)."""
        self._net = net
        _check(lib().pf_attach_learned_warm_start(self._h, None if net is None else net._h))

    def time_kernel(self, kernel_id, reps=20):
        """Mean microseconds and bytes of one launch of a step kernel (HIP events)."""
        us = C.c_double()
        b = C.c_double()
        _check(lib().pf_time_kernel(self._h, int(kernel_id), int(reps), C.byref(us), C.byref(b)))
        return us.value, b.value

    def probe(self, kernel_ids):
        """Bracket the accelerated solve's launches of these kernel classes
        with HIP events from the next step on (an empty list stops it)."""
        mask = 0
        for k in kernel_ids:
            mask |= 1 << int(k)
        _check(lib().pf_probe_kernels(self._h, mask))

    def probe_read(self, kernel_id):
        """(launches, working launches, summed microseconds, bytes of the
        working launches) of one probed class since probe()."""
        n = C.c_longlong()
        w = C.c_longlong()
        us = C.c_double()
        b = C.c_double()
        _check(lib().pf_probe_read(self._h, int(kernel_id), C.byref(n), C.byref(w), C.byref(us), C.byref(b)))
        return n.value, w.value, us.value, b.value


NN_PARAMS = 34640


class Net:
    """The learned warm-start network (a pf_nn): training geometry `shape`,
    batches of up to `max_batch` grids, a device dataset of `capacity`
    snapshots."""

    def __init__(self, shape, max_batch=4, capacity=0):
        self.shape = tuple(shape)
        h = C.c_void_p()
        _check(lib().pf_nn_create(*[int(v) for v in self.shape], int(max_batch), int(capacity), C.byref(h)))
        self._h = h

    def close(self):
        if self._h:
            lib().pf_nn_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_scales(self, s_phi, s_tau, s_out):
        _check(lib().pf_nn_set_scales(self._h, float(s_phi), float(s_tau), float(s_out)))

    def set_pool(self, factor):
        """1: the network runs on the handle's grid; 2: on its 2x2x2 averages."""
        _check(lib().pf_nn_set_pool(self._h, int(factor)))

    def pool(self):
        v = C.c_int()
        _check(lib().pf_nn_get_pool(self._h, C.byref(v)))
        return v.value

    def scales(self):
        a, b, c = C.c_double(), C.c_double(), C.c_double()
        _check(lib().pf_nn_get_scales(self._h, C.byref(a), C.byref(b), C.byref(c)))
        return a.value, b.value, c.value

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        assert p.size == NN_PARAMS
        _check(lib().pf_nn_set_params(self._h, _ptr(p)))

    def params(self):
        out = np.empty(NN_PARAMS, dtype=np.float32)
        _check(lib().pf_nn_get_params(self._h, _ptr(out)))
        return out

    def collect(self, sim):
        rc = lib().pf_nn_collect(self._h, sim._h)
        _check(min(rc, 0))
        return rc

    def size(self):
        return lib().pf_nn_dataset_size(self._h)

    def train_step(self, idx, lr):
        idx = np.ascontiguousarray(idx, dtype=np.int32)
        loss = C.c_float()
        _check(lib().pf_nn_train_step(self._h, _ptr(idx), int(idx.size), float(lr), C.byref(loss)))
        return loss.value

    def eval_loss(self, idx):
        idx = np.ascontiguousarray(idx, dtype=np.int32)
        loss = C.c_float()
        _check(lib().pf_nn_eval_loss(self._h, _ptr(idx), int(idx.size), C.byref(loss)))
        return loss.value

    def forward(self, in7, shape):
        in7 = np.ascontiguousarray(in7, dtype=np.float32)
        B = in7.shape[0]
        out = np.empty((B, 3, int(np.prod(shape))), dtype=np.float32)
        _check(lib().pf_nn_forward(self._h, B, *[int(v) for v in shape], _ptr(in7), _ptr(out)))
        return out

    def time_inference(self, shape, reps=5):
        us = C.c_double()
        _check(lib().pf_nn_time_inference(self._h, *[int(v) for v in shape], int(reps), C.byref(us)))
        return us.value

    def gradients(self, in7, target, shape):
        in7 = np.ascontiguousarray(in7, dtype=np.float32)
        target = np.ascontiguousarray(target, dtype=np.float32)
        g = np.empty(NN_PARAMS, dtype=np.float32)
        loss = C.c_float()
        _check(lib().pf_nn_gradients(self._h, in7.shape[0], *[int(v) for v in shape], _ptr(in7), _ptr(target),
                                     C.byref(loss), _ptr(g)))
        return loss.value, g


class Batch:
    """Independent simulations driven concurrently by `threads` host workers
    (pf_batch_*); each keeps its own stream, results equal solo runs."""

    def __init__(self, sims, threads=8):
        self.sims = list(sims)
        arr = (C.c_void_p * len(self.sims))(*[s._h.value for s in self.sims])
        h = C.c_void_p()
        _check(lib().pf_batch_create(len(self.sims), arr, int(threads), C.byref(h)))
        self._h = h

    def close(self):
        if self._h:
            lib().pf_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def step(self, n=1):
        _check(lib().pf_batch_step(self._h, int(n)))

    def minimize(self, opts):
        """opts: one PfMinimizeOpts for all instances or a list of them."""
        n = len(self.sims)
        if isinstance(opts, PfMinimizeOpts):
            opts = [opts] * n
        arr = (PfMinimizeOpts * n)(*opts)
        st = (PfMinimizeStats * n)()
        _check(lib().pf_batch_minimize(self._h, arr, st))
        return [s.as_dict() for s in st]

    def step_fused(self, n=1):
        """pf_batch_step_fused: n steps of all instances together, each phase
        of a step one batched launch per kernel over the instances."""
        _check(lib().pf_batch_step_fused(self._h, int(n)))

    def minimize_fused(self, opts):
        """pf_batch_minimize_fused: all instances' iterations together, each
        stage one batched launch over the instances (phi-only minimize with
        the accelerated elastic solver)."""
        n = len(self.sims)
        if isinstance(opts, PfMinimizeOpts):
            opts = [opts] * n
        arr = (PfMinimizeOpts * n)(*opts)
        st = (PfMinimizeStats * n)()
        _check(lib().pf_batch_minimize_fused(self._h, arr, st))
        return [s.as_dict() for s in st]

    def fused_launches(self):
        """(batched launches issued, launches recorded, graphs captured, graph
        launches) of the last fused call"""
        v = [C.c_longlong() for _ in range(4)]
        _check(lib().pf_batch_fused_launches(self._h, *[C.byref(x) for x in v]))
        return tuple(x.value for x in v)

    def set_graphs(self, enable):
        _check(lib().pf_batch_set_graphs(self._h, 1 if enable else 0))


def minimize_opts(**kw):
    """pf_minimize_opts with the reference defaults, overridden by keywords."""
    o = PfMinimizeOpts()
    _check(lib().pf_minimize_opts_default(C.byref(o)))
    for k, v in kw.items():
        if not hasattr(o, k):
            raise KeyError(k)
        setattr(o, k, v)
    return o


def fft_forward(a, precision=PF_FP64):
    a = np.ascontiguousarray(a, dtype=np.float64)
    nx, ny, nz = a.shape
    out = np.empty((nx, ny, nz // 2 + 1), dtype=np.complex128)
    _check(lib().pf_fft_forward(nx, ny, nz, int(precision), _ptr(a), _ptr(out)))
    return out


def fft_inverse(ak, shape, precision=PF_FP64):
    nx, ny, nz = shape
    ak = np.ascontiguousarray(ak, dtype=np.complex128).reshape(nx, ny, nz // 2 + 1)
    out = np.empty(shape, dtype=np.float64)
    _check(lib().pf_fft_inverse(nx, ny, nz, int(precision), _ptr(ak), _ptr(out)))
    return out


# ---------------------------------------------------------------- checkpoint files
def _as_params(params):
    return params_from_dict(params) if isinstance(params, dict) else params


def parameter_fingerprint(params):
    return int(lib().pf_parameter_fingerprint(C.byref(_as_params(params))))


def elastic_solver_fingerprint(params):
    return int(lib().pf_elastic_solver_fingerprint(C.byref(_as_params(params))))


def checkpoint_provenance(params=None, **overrides):
    """The provenance a reference run with `params` writes, with optional
    field overrides (strings or fingerprints)."""
    q = PfCheckpointProvenance()
    if params is not None:
        _check(lib().pf_checkpoint_provenance_default(C.byref(_as_params(params)), C.byref(q)))
    for k, v in overrides.items():
        setattr(q, k, v.encode() if isinstance(v, str) else v)
    return q


def checkpoint_file_write(path, prov, header, phi, Y, xB, dYdt, displacement_k=None):
    """header: dict of PfCheckpointHeader fields; arrays [nx][ny][nz] float64;
    displacement_k: float32 of 6 * nx*ny*(nz/2+1) values when elastic_present."""
    h = PfCheckpointHeader()
    for k, v in header.items():
        setattr(h, k, v)
    arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (phi, Y, xB, dYdt)]
    dk = None if displacement_k is None else np.ascontiguousarray(displacement_k, dtype=np.float32)
    _check(lib().pf_checkpoint_file_write(str(path).encode(), C.byref(prov), C.byref(h),
                                          *[_ptr(a) for a in arrs], None if dk is None else _ptr(dk)))


def checkpoint_file_read(path, expected):
    """Returns (header dict, {phi, Y, xB, dYdt, displacement_k})."""
    h = PfCheckpointHeader()
    _check(lib().pf_checkpoint_file_read(str(path).encode(), C.byref(expected), C.byref(h),
                                         None, None, None, None, None))
    shape = (h.nx, h.ny, h.nz)
    out = {k: np.empty(shape, dtype=np.float64) for k in ("phi", "Y", "xB", "dYdt")}
    dk = np.empty(6 * h.nx * h.ny * (h.nz // 2 + 1), dtype=np.float32) if h.elastic_present else None
    _check(lib().pf_checkpoint_file_read(str(path).encode(), C.byref(expected), C.byref(h),
                                         _ptr(out["phi"]), _ptr(out["Y"]), _ptr(out["xB"]), _ptr(out["dYdt"]),
                                         None if dk is None else _ptr(dk)))
    out["displacement_k"] = dk
    return h.as_dict(), out
