"""The main_pf binary: the reference main_cuda command line and Results/ layout
(main_cuda.cu:22457-22771, 25270-25660).  CPU tests cover the fail-closed
paths (no device call); the GPU test runs a short raw_fields case."""
import json
import os
import subprocess

import numpy as np
import pytest

from cuda_sto_pf_amd import params as pfparams
from cuda_sto_pf_amd.configs import config

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "cuda_sto_pf_amd", "bin", "main_pf")


def _bundle(tmp_path, n=16, steps=4, zero_mode=True):
    pd, phi, xB, meta = config(2, n)
    prm = pfparams.write_params_file(str(tmp_path / "run.params"), pd)
    phi.astype(np.float64).tofile(tmp_path / "phi_init.raw")
    xB.astype(np.float64).tofile(tmp_path / "xB_init.raw")
    (tmp_path / "init_meta.json").write_text(json.dumps(
        {"Nx": n, "Ny": n, "Nz": n, "dx_nm": 2.0, "interface_width_nm": 4.0, "dtype": "float64",
         "order": "C", "xB_max_safe": meta["xB_max_safe"]}))
    args = [CLI, str(n), str(n), str(n), "0.01", str(steps), "2", "1", "1", "--pf-param-file", prm,
            "--init-mode", "raw_fields", "--init-phi-raw", str(tmp_path / "phi_init.raw"),
            "--init-xB-raw", str(tmp_path / "xB_init.raw"), "--init-meta", str(tmp_path / "init_meta.json")]
    if zero_mode:
        args += ["--pf-zero-mode", "PF_CONSERVED_Y_ZERO_MODE_V1"]
    return args, pd, phi, xB, meta


def _run(args, tmp_path, extra_env=None):
    env = dict(os.environ, CUDA_STO_RESULTS_ROOT=str(tmp_path / "Results"))
    env.update(extra_env or {})
    return subprocess.run(args, capture_output=True, text=True, env=env, timeout=300)


def test_missing_param_file_is_fatal(tmp_path):
    r = _run([CLI, "16", "16", "16", "0.01", "2", "1", "1", "1"], tmp_path)
    assert r.returncode == 2 and "[fatal]" in r.stderr


def test_incomplete_param_file_is_fatal(tmp_path):
    args, *_ = _bundle(tmp_path)
    prm = args[args.index("--pf-param-file") + 1]
    lines = [ln for ln in open(prm) if not ln.startswith("S_p_11")]
    open(prm, "w").writelines(lines)
    r = _run(args, tmp_path)
    assert r.returncode == 2 and "S_p_" in r.stderr


def test_unknown_mode_is_refused(tmp_path):
    args, *_ = _bundle(tmp_path)
    r = _run(args + ["--mode=nucleation"], tmp_path)
    assert r.returncode == 2 and "--mode=nucleation is not implemented" in r.stderr


@pytest.mark.gpu
def test_cli_raw_fields_run_matches_library(pf, tmp_path):
    args, pd, phi, xB, meta = _bundle(tmp_path, steps=4)
    r = _run(args, tmp_path, {"CUDA_STO_SUPPRESS_VTK_OUTPUT": "0"})
    print(r.stdout[-2000:], r.stderr[-2000:], flush=True)
    assert r.returncode == 0, r.stderr
    runs = os.listdir(tmp_path / "Results")
    assert len(runs) == 1 and runs[0].startswith("chel_T380_cuda_16x16x16_dt0.01_steps4_xB")
    case = tmp_path / "Results" / runs[0] / "raw_fields_relax"
    for f in ("pf_input.params", "vf_precip_vs_time_raw_fields_relax.csv", "elastic_solver_trace.csv",
              "pf_zero_mode_trace.csv", "performance_summary.csv", "performance_summary.json",
              "phi_2.vtk", "xB_4.vtk", "phi_final.vtk"):
        assert (case / f).exists(), f
    csv = (case / "vf_precip_vs_time_raw_fields_relax.csv").read_text().splitlines()
    assert csv[0] == "step,t_code,t_real_s,vf_precip,R_avg" and len(csv) == 1 + 5
    # the VTK holds the same phi as the library run to %.5f
    sim = pf.Simulation(pd, pf.PF_FP64)
    sim.set_fields(phi, xB, xB_cap=meta["xB_max_safe"])
    sim.step(4)
    ref = sim.field(pf.PF_PHI)
    txt = (case / "phi_final.vtk").read_text().split("LOOKUP_TABLE default\n")[1].split()
    vals = np.array([float(v) for v in txt]).reshape(16, 16, 16)   # z, y, x order
    assert np.abs(vals.transpose(2, 1, 0) - ref).max() <= 6e-6


def test_minimize_constraint_flags_fail_closed(tmp_path):
    """main_cuda.cu:24562-24623: the mass constraint needs the full-model
    minimize, freeze-phi needs the mass constraint; the per-component volume
    constraints are not built and refused."""
    args, *_ = _bundle(tmp_path, zero_mode=False)
    for extra, flag in ((["--minimize-mass-constraint"], "--minimize-mass-constraint"),
                        (["--mode=minimize", "--minimize-mass-constraint"], "--minimize-mass-constraint"),
                        (["--mode=minimize", "--minimize-full-model", "--minimize-freeze-phi"],
                         "--minimize-freeze-phi"),
                        (["--mode=minimize", "--minimize-full-model", "--minimize-component-volume-constraints"],
                         "--minimize-component-volume-constraints")):
        r = _run(args + extra, tmp_path)
        assert r.returncode == 2 and "[fatal]" in r.stderr and flag in r.stderr, (extra, r.stderr)


@pytest.mark.gpu
def test_cli_mass_constrained_minimize_matches_library(pf, tmp_path):
    """--minimize-full-model --minimize-mass-constraint: the energy log's mass
    columns and minimize_mass_constraint_trace.csv (main_cuda.cu:25667-25690,
    32368-32398) equal the library's own run."""
    args, pd, phi, xB, meta = _bundle(tmp_path, zero_mode=False)
    args = args + ["--mode=minimize", "--minimize-full-model", "--minimize-mass-constraint", "--minimize-max-iter",
                   "4", "--minimize-dt", "0.05", "--init-case-tag", "masscase"]
    r = _run(args, tmp_path)
    print(r.stdout[-2000:], r.stderr[-2000:], flush=True)
    assert r.returncode == 0, r.stderr
    runs = os.listdir(tmp_path / "Results")
    case = tmp_path / "Results" / runs[0] / "masscase"
    logs = [f for f in os.listdir(case) if f.startswith("energy_minimize_")]
    rows = (case / logs[0]).read_text().splitlines()
    head = rows[0].split(",")
    trace = (case / "minimize_mass_constraint_trace.csv").read_text().splitlines()
    assert trace[0].startswith("iter,target_mass_code,current_mass_code,lambda,") and len(trace) == 5
    sim = pf.Simulation(dict(pd, dt=0.05, zero_mode=0), pf.PF_FP64)
    sim.set_fields(phi, xB, xB_cap=meta["xB_max_safe"])
    # the fields load under the meta cap, Y under the minimize cap (0.07)
    sim.minimize_begin(pf.minimize_opts(dt=0.05, max_iter=4, full_model=1, mass_constraint=1))
    for i in range(4):
        st = sim.minimize_iterate()
        cols = dict(zip(head, rows[1 + i].split(",")))
        tr = trace[1 + i].split(",")
        assert float(cols["mass_target_code"]) == st["mass_target"] == float(tr[1])
        assert abs(float(cols["mass_lambda"]) - st["mass_lambda"]) <= 1e-14 + 1e-12 * abs(st["mass_lambda"])
        assert float(cols["mass_err_rel"]) < 1e-12
        assert int(tr[7]) == st["mass_iterations"] and int(tr[8]) + int(tr[9]) == int(tr[7])


@pytest.mark.gpu
def test_cli_minimize_matches_library(pf, tmp_path):
    """--mode=minimize: the energy log of main_cuda (energy_minimize_*.csv,
    main_cuda.cu:27440-27466) and phi VTKs with the case tag; the logged
    multiplier and volume equal the library's own minimize run."""
    args, pd, phi, xB, meta = _bundle(tmp_path, steps=4, zero_mode=False)
    V0 = 0.05
    args = args + ["--mode=minimize", "--minimize-max-iter", "5", "--minimize-dt", "0.05", "--V0", str(V0),
                   "--init-case-tag", "mincase"]
    r = _run(args, tmp_path, {"CUDA_STO_SUPPRESS_VTK_OUTPUT": "0"})
    print(r.stdout[-2000:], r.stderr[-2000:], flush=True)
    assert r.returncode == 0, r.stderr
    runs = os.listdir(tmp_path / "Results")
    assert len(runs) == 1 and "_steps5_r" in runs[0] and "_dt0.05_" in runs[0]
    case = tmp_path / "Results" / runs[0] / "mincase"
    logs = [f for f in os.listdir(case) if f.startswith("energy_minimize_T380_")]
    assert len(logs) == 1 and logs[0].endswith("_mincase.csv")
    rows = (case / logs[0]).read_text().splitlines()
    assert rows[0].startswith("iter,dt,mean_h,V0,lambda,F_surf_hat") and len(rows) == 1 + 5
    for f in ("phi_init_mincase.vtk", "phi_final_mincase.vtk", "phi_2_mincase.vtk", "minimize_summary.json"):
        assert (case / f).exists(), f
    sim = pf.Simulation(pd, pf.PF_FP64)
    sim.set_fields(phi, xB, xB_cap=meta["xB_max_safe"])
    sim.minimize_begin(pf.minimize_opts(dt=0.05, V0=V0, max_iter=5))
    for i in range(5):
        st = sim.minimize_iterate()
        cols = rows[1 + i].split(",")
        assert int(cols[0]) == i + 1
        assert abs(float(cols[2]) - st["mean_h"]) <= 1e-8 * abs(st["mean_h"])
        assert abs(float(cols[4]) - st["lambda_vol"]) <= 1e-8 * abs(st["lambda_vol"])


def test_checkpoint_flags_fail_closed(tmp_path):
    """main_cuda.cu:24476-24491: checkpoints need the zero mode, a path with a
    period, and a non-negative period."""
    args, *_ = _bundle(tmp_path, zero_mode=False)
    r = _run(args + ["--pf-checkpoint-every", "2", "--pf-checkpoint-path", str(tmp_path / "c.chk")], tmp_path)
    assert r.returncode == 2 and "only with PF_CONSERVED_Y_ZERO_MODE_V1" in r.stderr
    args, *_ = _bundle(tmp_path)
    r = _run(args + ["--pf-checkpoint-every", "2"], tmp_path)
    assert r.returncode == 2 and "--pf-checkpoint-path is required" in r.stderr
    r = _run(args + ["--pf-checkpoint-every", "-1"], tmp_path)
    assert r.returncode == 2 and "must be >= 0" in r.stderr


@pytest.mark.gpu
def test_cli_restart_reproduces_continuous_run(pf, tmp_path):
    """The reference's restart qualification through the binary: 6 steps in
    one run and 3 + restart + 3 end in byte-identical checkpoints, and the
    restart appends to the elastic and zero-mode traces."""
    ident = ["--pf-initial-state-class", "MASS_CONSERVING_LIBRARY_ASSEMBLED_CONDITIONAL_HANDOFF_V1",
             "--pf-fixture-manifest-sha256", "ab" * 32, "--pf-profile-library-manifest-sha256", "cd" * 32]
    args, *_ = _bundle(tmp_path, steps=6)
    cont = str(tmp_path / "continuous.chk")
    r = _run(args + ident + ["--pf-checkpoint-every", "6", "--pf-checkpoint-path", cont,
                             "--init-case-tag", "cont"], tmp_path)
    assert r.returncode == 0, r.stderr
    assert "PF_ZERO_MODE_CHECKPOINT_WRITTEN step=6" in r.stdout
    args3, *_ = _bundle(tmp_path, steps=3)
    mid = str(tmp_path / "mid.chk")
    r = _run(args3 + ident + ["--pf-checkpoint-every", "3", "--pf-checkpoint-path", mid,
                              "--init-case-tag", "half"], tmp_path)
    assert r.returncode == 0, r.stderr
    fin = str(tmp_path / "restart.chk")
    r = _run(args + ident + ["--pf-restart-from", mid, "--pf-checkpoint-every", "6", "--pf-checkpoint-path", fin,
                             "--init-case-tag", "half"], tmp_path)
    print(r.stdout[-1500:], r.stderr[-1500:], flush=True)
    assert r.returncode == 0, r.stderr
    assert "PF_ZERO_MODE_RESTART_LOADED" in r.stdout and "accepted_step=3" in r.stdout
    assert open(cont, "rb").read() == open(fin, "rb").read()
    # a restart without the identity it was written with is refused
    r = _run(args + ["--pf-restart-from", mid, "--init-case-tag", "bad"], tmp_path)
    assert r.returncode == 2 and "provenance mismatch" in r.stderr
    runs = [d for d in os.listdir(tmp_path / "Results") if "_steps6_" in d]
    trace = (tmp_path / "Results" / runs[0] / "half" / "elastic_solver_trace.csv").read_text().splitlines()
    assert trace[0].startswith("step,") and [ln.split(",")[0] for ln in trace[1:]] == ["4", "5", "6"]
    # the zero-mode trace of the continued run carries the checkpoint's frozen
    # target, the value the first half's rows carry
    half = [d for d in os.listdir(tmp_path / "Results") if "_steps3_" in d]
    z1 = (tmp_path / "Results" / half[0] / "half" / "pf_zero_mode_trace.csv").read_text().splitlines()
    z2 = (tmp_path / "Results" / runs[0] / "half" / "pf_zero_mode_trace.csv").read_text().splitlines()
    assert z1[0].startswith("step,target_mass_code")
    targets = {ln.split(",")[1] for ln in z1[1:] + z2[1:]}
    assert len(targets) == 1, targets


@pytest.mark.gpu
def test_front_end_to_run(pf, tmp_path):
    """Physical inputs -> .params (units.py) and a uniform matrix bundle
    (raw_init.py) drive main_pf: a matrix-only run keeps phi = 0 and the
    composition uniform, with the mass held by the zero mode."""
    from cuda_sto_pf_amd import raw_init, units
    prm = str(tmp_path / "phys.params")
    units.write_params(prm, units.load_physical_inputs(None, gamma=0.168, lambda_sm=4e-9, pf_dx=1e-9,
                                                       phys_dx_ref=1e-9, dt=0.02))
    b = raw_init.uniform(tmp_path / "bundle", 16, 16, 16, 1.0, 4.0, 0.03)
    args = [CLI, "16", "16", "16", "0.02", "3", "3", "1", "1", "--pf-param-file", prm, "--init-mode", "raw_fields",
            "--init-phi-raw", str(b / "phi_init.raw"), "--init-xB-raw", str(b / "xB_init.raw"),
            "--init-meta", str(b / "init_meta.json"), "--pf-zero-mode", "PF_CONSERVED_Y_ZERO_MODE_V1",
            "--init-case-tag", "uniform"]
    r = _run(args, tmp_path, {"CUDA_STO_SUPPRESS_VTK_OUTPUT": "1"})
    print(r.stdout[-1500:], r.stderr[-1500:], flush=True)
    assert r.returncode == 0, r.stderr
    runs = os.listdir(tmp_path / "Results")
    assert len(runs) == 1 and runs[0].startswith("chel_T380_cuda_16x16x16_dt0.02_steps3_xB0.030")
    rows = (tmp_path / "Results" / runs[0] / "uniform" / "vf_precip_vs_time_uniform.csv").read_text().splitlines()
    assert len(rows) == 1 + 4 and all(float(r.split(",")[3]) == 0.0 for r in rows[1:])


@pytest.mark.gpu
def test_cli_nonconvergence_is_fatal(pf, tmp_path):
    """A step whose elastic solve stops at elastic_iter_max above tolerance
    ends the run with [fatal] and status 2, as the reference does."""
    args, *_ = _bundle(tmp_path, steps=2)
    prm = args[args.index("--pf-param-file") + 1]
    lines = []
    for ln in open(prm):
        if ln.startswith("elastic_iter_max"):
            ln = "elastic_iter_max = 3\n"
        elif ln.startswith("elastic_residual_tolerance"):
            ln = "elastic_residual_tolerance = 1e-15\n"
        lines.append(ln)
    open(prm, "w").writelines(lines)
    r = _run(args, tmp_path, {"CUDA_STO_SUPPRESS_VTK_OUTPUT": "1"})
    print(r.stdout[-800:], r.stderr[-800:], flush=True)
    assert r.returncode == 2 and "[fatal]" in r.stderr and "did not converge" in r.stderr


HANDOFF = "MASS_CONSERVING_LIBRARY_ASSEMBLED_CONDITIONAL_HANDOFF_V1"


@pytest.mark.parametrize("flags,match", [
    # a mixed triple: legacy class with a real manifest (main_cuda.cu:24527-24545)
    (["--pf-initial-state-class", "LEGACY_UNSPECIFIED", "--pf-fixture-manifest-sha256", "a" * 64],
     "complete LEGACY_UNSPECIFIED identity"),
    # handoff class with an upper-case or a short digest
    (["--pf-initial-state-class", HANDOFF, "--pf-fixture-manifest-sha256", "A" * 64,
      "--pf-profile-library-manifest-sha256", "b" * 64], "two lowercase SHA-256"),
    (["--pf-initial-state-class", HANDOFF, "--pf-fixture-manifest-sha256", "a" * 63,
      "--pf-profile-library-manifest-sha256", "b" * 64], "two lowercase SHA-256"),
    (["--pf-initial-state-class", "SOMETHING_ELSE"], "complete LEGACY_UNSPECIFIED identity"),
])
def test_checkpoint_identity_rule_is_fail_closed(tmp_path, flags, match):
    args, *_ = _bundle(tmp_path)
    r = _run(args + flags, tmp_path)
    assert r.returncode == 2 and match in r.stderr, r.stderr


def test_handoff_identity_needs_zero_mode_dynamics_and_raw_fields(tmp_path):
    """A well-formed handoff identity is accepted only for zero-mode dynamics
    (main_cuda.cu:24546-24553) started from raw_fields or a restart (24554-24561)."""
    ident = ["--pf-initial-state-class", HANDOFF, "--pf-fixture-manifest-sha256", "a" * 64,
             "--pf-profile-library-manifest-sha256", "b" * 64]
    args, *_ = _bundle(tmp_path, zero_mode=False)
    r = _run(args + ident, tmp_path)
    assert r.returncode == 2 and "requires dynamics with PF_CONSERVED_Y_ZERO_MODE_V1" in r.stderr
    args, *_ = _bundle(tmp_path)
    r = _run(args + ident + ["--mode=minimize"], tmp_path)
    assert r.returncode == 2 and "requires dynamics" in r.stderr
    i = args.index("--init-mode")
    no_raw = args[:i] + args[i + 2:]
    r = _run(no_raw + ident, tmp_path)
    assert r.returncode == 2 and "require --init-mode raw_fields" in r.stderr


@pytest.mark.parametrize("pos,value,match", [(5, "0", "Invalid out_every=0"), (6, "0", "Invalid csv_out_every=0"),
                                             (4, "0", "Invalid nsteps=0")])
def test_nonpositive_intervals_are_fatal(tmp_path, pos, value, match):
    """out_every, csv_out_every and nsteps <= 0 exit [fatal] (main_cuda.cu:24830-24841)
    rather than being clamped."""
    args, *_ = _bundle(tmp_path)
    args = list(args)
    args[pos + 1] = value
    r = _run(args, tmp_path)
    assert r.returncode == 2 and match in r.stderr, r.stderr


def test_component_constraint_flags_fail_closed(tmp_path):
    """V5 command line (main_cuda.cu:2805-2873, 24638-24670): count in [2,64],
    exactly one positive target per component, an int32 label raw of the grid's
    size with labels in range, the mass constraint on and no post-projection."""
    args, pd, phi, *_ = _bundle(tmp_path, zero_mode=False)
    lab = np.zeros(phi.shape, dtype=np.int32)
    lab[8:] = 1
    lab.tofile(tmp_path / "labels.raw")
    lab[:2].tofile(tmp_path / "short.raw")
    bad = lab.copy()
    bad[0, 0, 0] = 2
    bad.tofile(tmp_path / "range.raw")
    base = args + ["--mode=minimize", "--minimize-full-model", "--minimize-post-projection-iters", "0",
                   "--minimize-component-volume-constraints"]
    good = ["--minimize-component-count", "2", "--minimize-component-label-raw", str(tmp_path / "labels.raw"),
            "--minimize-component-target-h-sums", "10.0,12.5"]
    cases = [
        ([], "requires full-model constrained minimize"),
        (["--minimize-mass-constraint", "--minimize-component-count", "1"] + good[2:], "component count in [2,64]"),
        (["--minimize-mass-constraint"] + good[:4] + ["--minimize-component-target-h-sums", "10.0"],
         "expected comma"),
        (["--minimize-mass-constraint"] + good[:4] + ["--minimize-component-target-h-sums", "10.0,12.5,3"],
         "extra entries"),
        (["--minimize-mass-constraint"] + good[:4] + ["--minimize-component-target-h-sums", "10.0,-1"],
         "invalid component target"),
        (["--minimize-mass-constraint", "--minimize-component-count", "2", "--minimize-component-label-raw",
          str(tmp_path / "short.raw")] + good[4:], "size mismatch"),
        (["--minimize-mass-constraint", "--minimize-component-count", "2", "--minimize-component-label-raw",
          str(tmp_path / "range.raw")] + good[4:], "out of range"),
    ]
    for extra, match in cases:
        r = _run(base + extra, tmp_path)
        assert r.returncode == 2 and "[fatal]" in r.stderr and match in r.stderr, (extra, r.stderr)
    r = _run(args + ["--mode=minimize", "--minimize-full-model", "--minimize-mass-constraint",
                     "--minimize-component-volume-constraints"] + good, tmp_path)
    assert r.returncode == 2 and "post-projection-iters 0" in r.stderr


def _read_vtk(path, shape):
    txt = path.read_text().split("LOOKUP_TABLE default\n")[1].split()
    return np.array([float(v) for v in txt]).reshape(shape[2], shape[1], shape[0]).transpose(2, 1, 0)


def _write_vtk(path, f, name="phi"):
    nx, ny, nz = f.shape
    head = ("# vtk DataFile Version 3.0\nPhase Field Simulation\nASCII\nDATASET STRUCTURED_POINTS\n"
            f"DIMENSIONS {nx} {ny} {nz}\nASPECT_RATIO 1 1 1\nORIGIN 0 0 0\nPOINT_DATA {nx * ny * nz}\n"
            f"SCALARS {name} double 1\nLOOKUP_TABLE default\n")
    path.write_text(head + " ".join(f"{v:.5f}" for v in f.transpose(2, 1, 0).ravel()) + "\n")


def test_continuation_flags_fail_closed(tmp_path):
    """main_cuda.cu:24706-24770: a continuation needs --continue-phi-vtk,
    cannot start from raw fields, refuses the zero mode, and checks the VTK
    grid; without --pf-param-file it takes the stored pf_input.params next
    to the VTK (main_cuda.cu:22543-22565)."""
    args, pd, phi, *_ = _bundle(tmp_path, zero_mode=False)
    prm = args[args.index("--pf-param-file") + 1]
    pos = args[:9]
    src = tmp_path / "Results" / "run" / "case"
    src.mkdir(parents=True)
    _write_vtk(src / "phi_final.vtk", phi)
    _write_vtk(tmp_path / "small.vtk", phi[:8])
    cases = [(pos + ["--pf-param-file", prm, "--mode=dynamics-continue"], "--continue-phi-vtk"),
             (args + ["--mode=dynamics-continue", "--continue-phi-vtk", str(src / "phi_final.vtk")], "raw_fields"),
             (pos + ["--pf-param-file", prm, "--mode=dynamics-continue", "--continue-phi-vtk",
                     str(src / "phi_final.vtk"), "--pf-zero-mode", "PF_CONSERVED_Y_ZERO_MODE_V1"], "zero-mode"),
             (pos + ["--pf-param-file", prm, "--mode=dynamics-continue", "--continue-phi-vtk",
                     str(tmp_path / "small.vtk")], "VTK grid mismatch")]
    for a, match in cases:
        r = _run(a, tmp_path)
        assert r.returncode == 2 and match in r.stderr, (a[-3:], r.stderr)
    # no parameter file at all, then the one stored beside the VTK
    r = _run(pos + ["--mode=dynamics-continue", "--continue-phi-vtk", str(src / "phi_final.vtk")], tmp_path)
    assert r.returncode == 2 and "no built-in defaults" in r.stderr
    import shutil
    shutil.copy(prm, src / "pf_input.params")
    r = _run(pos + ["--mode=dynamics-continue", "--continue-phi-vtk", str(src / "phi_final.vtk"),
                    "--pf-zero-mode", "PF_CONSERVED_Y_ZERO_MODE_V1"], tmp_path)
    assert "[continue-param] loaded stored PF params from" in r.stdout and "zero-mode" in r.stderr


@pytest.mark.gpu
def test_cli_vtk_continuation_matches_library(pf, tmp_path):
    """--mode=dynamics-continue from the VTKs of a finished run (main_cuda.cu:
    2602-2655, 25363-25456): the continued case is continue_dyn_1 beside the
    source case, its CSV carries the source case's tag, and its fields equal
    a library run from the same %.5f fields; without an x_B VTK the
    composition is rebuilt from phi with x_eq(T) inside and ic_23d_xB_out
    outside; --mode=minimize-continue logs under the source tag."""
    from oracle import thermo as th
    args, pd, phi, xB, meta = _bundle(tmp_path, steps=4, zero_mode=False)
    env = {"CUDA_STO_SUPPRESS_VTK_OUTPUT": "0"}
    r = _run(args, tmp_path, env)
    assert r.returncode == 0, r.stderr
    run = tmp_path / "Results" / os.listdir(tmp_path / "Results")[0]
    src = run / "raw_fields_relax"
    pos = args[:9]
    pos[5] = "3"
    prm = args[args.index("--pf-param-file") + 1]
    cont = pos + ["--pf-param-file", prm, "--mode=dynamics-continue", "--continue-phi-vtk",
                  str(src / "phi_final.vtk"), "--continue-xB-vtk", str(src / "xB_final.vtk")]
    r = _run(cont, tmp_path, env)
    print(r.stdout[-1500:], r.stderr[-1500:], flush=True)
    assert r.returncode == 0, r.stderr
    case = run / "continue_dyn_1"
    rows = (case / "vf_precip_vs_time_raw_fields_relax.csv").read_text().splitlines()
    assert rows[0].startswith("step,") and [int(x.split(",")[0]) for x in rows[1:]] == [1, 2, 3]
    shape = phi.shape
    p0 = np.clip(_read_vtk(src / "phi_final.vtk", shape), 0.0, 1.0)
    x0 = np.clip(_read_vtk(src / "xB_final.vtk", shape), pd["xB_eps"], 1.0 - pd["xB_eps"])
    sim = pf.Simulation(dict(pd, zero_mode=0), pf.PF_FP64)
    sim.set_fields(p0, x0)
    sim.step(3)
    assert np.abs(_read_vtk(case / "phi_final.vtk", shape) - sim.field(pf.PF_PHI)).max() <= 6e-6
    assert np.abs(_read_vtk(case / "xB_final.vtk", shape) - sim.field(pf.PF_XB)).max() <= 6e-6
    # a second continuation numbers on
    r = _run(pos + ["--pf-param-file", prm, "--mode=dynamics-continue", "--continue-phi-vtk",
                    str(case / "phi_final.vtk")], tmp_path, env)
    assert r.returncode == 0, r.stderr
    case2 = run / "continue_dyn_2"
    assert (case2 / "vf_precip_vs_time_continue_dyn_1.csv").exists()
    # that one rebuilt x_B from phi: x_eq inside, the far field outside
    p1 = np.clip(_read_vtk(case / "phi_final.vtk", shape), 0.0, 1.0)
    h = th.h_of_phi(p1)
    xe = th.solve_x_eq(pd.get("temperature_C", 380.0) + 273.15)
    xr = np.clip((1.0 - h) * 0.03 + h * xe, pd["xB_eps"], 1.0 - pd["xB_eps"])
    sim2 = pf.Simulation(dict(pd, zero_mode=0), pf.PF_FP64)
    sim2.set_fields(p1, xr)
    sim2.step(3)
    assert np.abs(_read_vtk(case2 / "xB_final.vtk", shape) - sim2.field(pf.PF_XB)).max() <= 6e-6
    # phi-only minimize continued from the VTK
    r = _run(pos + ["--pf-param-file", prm, "--mode=minimize-continue", "--continue-phi-vtk",
                    str(src / "phi_final.vtk"), "--minimize-max-iter", "3", "--minimize-dt", "0.05"], tmp_path, env)
    assert r.returncode == 0, r.stderr
    logs = [f for f in os.listdir(run / "continue_min_1") if f.startswith("energy_minimize_")]
    assert len(logs) == 1 and logs[0].endswith("_raw_fields_relax.csv")
    assert len((run / "continue_min_1" / logs[0]).read_text().splitlines()) == 1 + 3
