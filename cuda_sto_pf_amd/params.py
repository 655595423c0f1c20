"""Schema-v1 ``.params`` files (key=value) as read by ``main_cuda``
(load_pfparams_override_file, main_cuda.cu:22306-22360) and written by the
reference's converter (Unit_Psedobinary.py:835)."""
import numpy as np

VOIGT21 = ["11", "12", "13", "14", "15", "16", "22", "23", "24", "25", "26", "33", "34", "35",
           "36", "44", "45", "46", "55", "56", "66"]
EIGEN = ["eps_xx00", "eps_yy00", "eps_zz00", "eps_yz00", "eps_xz00", "eps_xy00"]
E0KEYS = ["E0_xx", "E0_yy", "E0_zz", "E0_yz", "E0_xz", "E0_xy"]

# keys written for unit reporting (Unit_Psedobinary.py:741-818); the time step
# does not read them but the loader requires them (main_cuda.cu:1089-1144)
REPORTING_DEFAULTS = {
    "t_real_unit": 1.0, "ic_vf_init_phi": 0.0, "ic_vf_target_phi": 0.04, "ic_phi_iface_w": 2.0,
    "gamma_Jm2": 0.168, "lambda_sm_m": 4.0e-9, "Vm_alpha_0_phys_m3mol": 4.1009e-05,
}

SCALAR_KEYS = ["dt", "dx", "dy", "dz", "temperature_C", "mu_reference_scale", "v_A", "v_B",
               "Vm_compound", "Vm_alpha_0", "dVm_alpha_dxB", "W", "kappa_phi", "L_phi", "D_alpha",
               "D_compound", "eps_iso_over_vB", "elastic_shift_dimless"]
SOLVER_KEYS = ["elastic_warm_start_enabled", "elastic_residual_control_enabled", "elastic_iter_min",
               "elastic_iter_max", "elastic_residual_tolerance", "elastic_residual_absolute_floor",
               "elastic_fail_on_nonconvergence", "Y_clip", "xB_eps", "thermo_convex_extrapolation_enabled"]


def write_params_file(path, pd):
    """Write a code-unit parameter dict (see configs.base_params) as a
    schema-v1 .params file, %.16e like the reference converter."""
    lines = ["# code-unit parameters (schema v1)", "pf_params_schema_version=1"]
    for k in SCALAR_KEYS:
        lines.append(f"{k}={float(pd[k]):.16e}")
    for k, v in REPORTING_DEFAULTS.items():
        lines.append(f"{k}={float(pd.get(k, v)):.16e}")
    C = np.asarray(pd["C"])
    dC = np.asarray(pd["dC"])
    for s in VOIGT21:
        i, j = int(s[0]) - 1, int(s[1]) - 1
        lines.append(f"S_{s}={C[i, j]:.16e}")
        lines.append(f"S_p_{s}={dC[i, j]:.16e}")
    for k, v in zip(EIGEN, pd["eps00"]):
        lines.append(f"{k}={float(v):.16e}")
    for k, v in zip(E0KEYS, pd["E0"]):
        lines.append(f"{k}={float(v):.16e}")
    for k in SOLVER_KEYS:
        if k in pd:
            v = pd[k]
            lines.append(f"{k} = {v:.16e}" if isinstance(v, float) else f"{k} = {int(v)}")
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")
    return path
