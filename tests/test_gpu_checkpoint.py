"""Checkpoint / restart on the GPU (pf_checkpoint_write / pf_checkpoint_read).

The reference qualifies restarts by byte equality: a run of N+M steps and a
run of N steps, checkpoint, restart and M more steps end in identical V4
checkpoints (reports/pf_elastic_warm_start_residual_v1/qualification_report.md:99-111).
The same criterion is applied here to the reference iteration and to the
accelerated solver in both precisions.  The exact-resume companion file is
what makes it hold for the FP64 warm state and the extrapolation history; a
restart from the V4 file alone (the reference's FP32 warm state) is checked
to continue within the solver tolerance.
"""
import os
import subprocess

import numpy as np
import pytest

from cuda_sto_pf_amd.configs import config

HERE = os.path.dirname(os.path.abspath(__file__))
TOOL = os.path.join(os.path.dirname(HERE), "oracle", "_ref", "ckpt_ref_tool")

FIELDS = ("PF_PHI", "PF_Y", "PF_XB", "PF_DYDT")


def _case(accel):
    pd, phi, xB, meta = config(2, 16)
    if accel:
        pd = dict(pd, elastic_warm_start_order=3, elastic_anderson_depth=3)
    return pd, phi, xB, meta["xB_max_safe"]


def _sim(pf, pd, phi, xB, cap, prec):
    s = pf.Simulation(pd, prec)
    s.set_fields(phi, xB, xB_cap=cap)
    return s


def _fields(pf, s):
    return [s.field(getattr(pf, f)) for f in FIELDS]


CASES = [("fp64_reference", False, "PF_FP64"), ("fp64_accel", True, "PF_FP64"), ("mixed_accel", True, "PF_MIXED")]


@pytest.mark.gpu
@pytest.mark.parametrize("name,accel,prec", CASES, ids=[c[0] for c in CASES])
def test_continuous_equals_restart_bitwise(pf, tmp_path, name, accel, prec):
    pd, phi, xB, cap = _case(accel)
    prec = getattr(pf, prec)
    n, m = 4, 3
    a = _sim(pf, pd, phi, xB, cap, prec)
    a.step(n + m)
    a.checkpoint_write(tmp_path / "continuous.chk")
    b = _sim(pf, pd, phi, xB, cap, prec)
    b.step(n)
    b.checkpoint_write(tmp_path / "mid.chk")
    b.close()
    c = pf.Simulation(pd, prec)
    c.checkpoint_read(tmp_path / "mid.chk")
    assert c.scalars()["steps_done"] == n
    c.step(m)
    c.checkpoint_write(tmp_path / "restart.chk")
    for f, fa, fc in zip(FIELDS, _fields(pf, a), _fields(pf, c)):
        assert np.array_equal(fa, fc), f
    sa, sc = a.scalars(), c.scalars()
    assert sa["steps_done"] == sc["steps_done"] == n + m
    assert sa["zero_mode_lambda"] == sc["zero_mode_lambda"]
    assert sa["last_elastic_iterations"] == sc["last_elastic_iterations"]
    da = (tmp_path / "continuous.chk").read_bytes()
    dc = (tmp_path / "restart.chk").read_bytes()
    print(f"{name}: checkpoint {len(da)} bytes, identical={da == dc}", flush=True)
    assert da[:8] == b"PFZMCHK4"
    assert da == dc
    assert (tmp_path / "continuous.chk.state").read_bytes() == (tmp_path / "restart.chk.state").read_bytes()


@pytest.mark.gpu
@pytest.mark.parametrize("prec", ["PF_FP64", "PF_MIXED"])
def test_restart_from_v4_alone(pf, tmp_path, prec):
    """Without the companion file the warm state comes from the V4 file's
    FP32 u(k), as in the reference: the run continues and stays within the
    elastic tolerance's reach of the continuous one."""
    pd, phi, xB, cap = _case(True)
    prec = getattr(pf, prec)
    a = _sim(pf, pd, phi, xB, cap, prec)
    a.step(6)
    b = _sim(pf, pd, phi, xB, cap, prec)
    b.step(3)
    b.checkpoint_write(tmp_path / "mid.chk")
    os.remove(tmp_path / "mid.chk.state")
    c = pf.Simulation(pd, prec)
    c.checkpoint_read(tmp_path / "mid.chk")
    c.step(3)
    dx = np.abs(a.field(pf.PF_XB) - c.field(pf.PF_XB)).max()
    dphi = np.abs(a.field(pf.PF_PHI) - c.field(pf.PF_PHI)).max()
    print(f"V4-only restart: max|dxB|={dx:.2e} max|dphi|={dphi:.2e}", flush=True)
    # FP64: the FP32 warm start only changes where a 1e-8 solve starts;
    # mixed: FP32 transforms make the paths differ at their own 1e-7 level
    tol = 1e-8 if prec == pf.PF_FP64 else 1e-6
    assert dx < tol and dphi < tol
    assert abs(c.scalars()["mass_xBtot"] - a.scalars()["mass_xBtot"]) <= 1e-12 * abs(a.scalars()["mass_xBtot"])


@pytest.mark.gpu
def test_checkpoint_contents_and_reference_reader(pf, tmp_path):
    """The V4 file holds the handle's fields, the zero-mode runtime and the
    warm state at the retained modes (zero elsewhere).  Its bytes equal what
    the host writer (pf_checkpoint_file_write) makes of the same content, and
    that writer is byte-identical to the reference's own writer built from
    the reference sources (tests/test_checkpoint.py, CPU), so the GPU-written
    file is the reference's file.  Where that build is present here too, the
    reference reader and writer also round-trip the file directly."""
    pd, phi, xB, cap = _case(True)
    s = _sim(pf, pd, phi, xB, cap, pf.PF_FP64)
    s.step(3)
    path = str(tmp_path / "s.chk")
    s.checkpoint_write(path)
    prov = pf.checkpoint_provenance(s.params)
    h, data = pf.checkpoint_file_read(path, prov)
    sc = s.scalars()
    assert h["version"] == 4 and h["accepted_step"] == 3 and h["elastic_source_field_step"] == 2
    assert h["accepted_zero_mode_steps"] == 3
    assert h["last_lambda"] == sc["zero_mode_lambda"] and h["last_iterations"] == sc["zero_mode_iterations"]
    assert h["elastic_last_iterations"] == sc["last_elastic_iterations"]
    assert h["temperature_K"] == pd.get("temperature_C", 380.0) + 273.15
    for f, name in zip(FIELDS, ("phi", "Y", "xB", "dYdt")):
        assert np.array_equal(data[name], s.field(getattr(pf, f))), name
    u = s.elastic_state()                          # compact retained modes, FP64
    d = data["displacement_k"].reshape(3, 16, 16, 9, 2)
    assert np.count_nonzero(d) > 0
    nz_full = np.count_nonzero(np.abs(d).sum(axis=-1))
    assert nz_full <= np.count_nonzero(u)          # only retained modes are populated
    assert np.allclose(np.sort(np.abs(d[..., 0] + 1j * d[..., 1]).ravel())[-nz_full:],
                       np.sort(np.abs(u).astype(np.float32).ravel())[-nz_full:], rtol=1e-6)
    # the chain to the reference writer: re-serialise the content with the host writer
    hdr = {k: v for k, v in h.items() if k not in ("version", "payload_checksum")}
    again = str(tmp_path / "again.chk")
    pf.checkpoint_file_write(again, prov, hdr, data["phi"], data["Y"], data["xB"], data["dYdt"],
                             data["displacement_k"])
    assert open(path, "rb").read() == open(again, "rb").read()
    if os.path.exists(TOOL):
        args = [prov.as_dict()[k] for k in pf.PfCheckpointProvenance.TEXT]
        args += [format(prov.parameter_fingerprint, "x"), format(prov.elastic_solver_fingerprint, "x")]
        back = str(tmp_path / "back.chk")
        r = subprocess.run([TOOL, "roundtrip", path, back] + args, capture_output=True, text=True, timeout=60)
        assert r.returncode == 0, r.stdout
        assert open(path, "rb").read() == open(back, "rb").read()


@pytest.mark.gpu
def test_restart_rejections(pf, tmp_path):
    pd, phi, xB, cap = _case(True)
    s = _sim(pf, pd, phi, xB, cap, pf.PF_FP64)
    with pytest.raises(pf.PfError, match="refusing to checkpoint"):
        s.checkpoint_write(tmp_path / "early.chk")    # no accepted step yet
    s.step(2)
    path = tmp_path / "s.chk"
    s.checkpoint_write(path)
    # a different grid
    other = pf.Simulation(dict(pd, Nx=32), pf.PF_FP64)
    with pytest.raises(pf.PfError, match="provenance mismatch|grid/dt/temperature"):
        other.checkpoint_read(path)
    # a different identity than the one written
    q = pf.checkpoint_provenance(s.params, fixture_manifest_sha256="f" * 64)
    fresh = pf.Simulation(pd, pf.PF_FP64)
    with pytest.raises(pf.PfError, match="provenance mismatch") as ei:
        fresh.checkpoint_read(path, q)
    assert ei.value.code == -7
    # zero mode off: no checkpoint provenance
    nz = _sim(pf, dict(pd, zero_mode=0), phi, xB, cap, pf.PF_FP64)
    nz.step(1)
    with pytest.raises(pf.PfError, match="only with PF_CONSERVED_Y_ZERO_MODE_V1"):
        nz.checkpoint_write(tmp_path / "nz.chk")
    # the handle that failed to restore still works from set_fields
    fresh.set_fields(phi, xB, xB_cap=cap)
    fresh.step(1)
