"""Zero-mode checkpoint files (no GPU): this library's writer and reader
against the reference's own implementation and its unit test.

The reference's host-only checkpoint code (pf_zero_mode_checkpoint.cpp) and
its unit test (tests/test_pf_zero_mode_checkpoint.cpp) compile here; the
recipe is oracle/Makefile, which builds them from the reference tree into
oracle/_ref/ together with a small driver (oracle/ckpt_ref_tool.cpp).  The
tests that need those binaries skip when neither they nor the reference tree
are present.

Pins:
- the reference's unit test passes (it pins the reference implementation);
- a file written by this library is read by the reference reader, and the
  reference writer turns it back into the same bytes (V4 and V3);
- files written by the reference writer (V4, V3, and a V2 converted by it)
  are read by this library with the same content;
- the two fingerprints (main_cuda.cu:4187-4267) of the reference's committed
  qualification parameters equal the values its run printed
  (reports/pf_elastic_warm_start_residual_v1/accelerated_continuous_stdout.log:70,100);
- the rejections the reference test checks: provenance, time level, reaction
  discretization, fixture, library, initial-state class, elastic solver
  fingerprint, V4 state read by a solver without elastic state, a corrupted
  byte.
"""
import os
import shutil
import struct
import subprocess

import numpy as np
import pytest

from cuda_sto_pf_amd import pfcore as pf

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
REF_DIR = os.path.join(ROOT, "oracle", "_ref")
TOOL = os.path.join(REF_DIR, "ckpt_ref_tool")
REF_TEST = os.path.join(REF_DIR, "test_pf_zero_mode_checkpoint")
REFERENCE = "/root/reference"

FIX_SHA = "3df27d489abeb808f3e72e75d4910ee53a7f490d702af9f1ed17f5d20d03f82b"
LIB_SHA = "58803a8bc6679b823e45e7a7b85df16ae68efa55338d52d4c4151b414a5ef0fe"


@pytest.fixture(scope="module")
def lib():
    try:
        pf.load_library()
    except FileNotFoundError:
        pytest.skip("libpfcore.so is not built")
    return pf


@pytest.fixture(scope="module")
def ref_tools():
    if not (os.path.exists(TOOL) and os.path.exists(REF_TEST)):
        if not os.path.isdir(REFERENCE) or shutil.which("make") is None:
            pytest.skip("reference checkpoint code not built and the reference tree is absent")
        subprocess.run(["make", "-C", os.path.join(ROOT, "oracle")], check=True, capture_output=True)
    return TOOL


def _prov(elastic, **kw):
    d = dict(zero_mode="PF_CONSERVED_Y_ZERO_MODE_V1", backend="HOST_NEWTON_BISECTION_V1",
             composition_mode="legacy", y_update_mode="lagged_rhs", explicit_context="SM_EXPLICIT_CONTEXT_N_V1",
             reaction_discretization="SM_TANGENT_N_V1",
             elastic_solver_mode="ELASTIC_WARM_START_RESIDUAL_V1" if elastic else "ELASTIC_RUNTIME_STATE_NOT_REQUIRED",
             initial_state_class="MASS_CONSERVING_LIBRARY_ASSEMBLED_CONDITIONAL_HANDOFF_V1",
             fixture_manifest_sha256=FIX_SHA, profile_library_manifest_sha256=LIB_SHA,
             parameter_fingerprint=0x123456789ABCDEF0, elastic_solver_fingerprint=0xFEDCBA9876543210 if elastic else 0)
    d.update(kw)
    return d


def _prov_args(d):
    return [d[k] for k in pf.PfCheckpointProvenance.TEXT] + [format(d["parameter_fingerprint"], "x"),
                                                             format(d["elastic_solver_fingerprint"], "x")]


def _prov_struct(d):
    return pf.checkpoint_provenance(**d)


# the deterministic content of oracle/ckpt_ref_tool.cpp (make_value)
def _make_value(a, i):
    i = np.asarray(i, dtype=np.int64)
    if a == 0:
        return 0.5 + 0.5 * ((i * 37) % 101) / 101.0 - 0.25
    if a == 1:
        return -4.0 + 0.01 * ((i * 13) % 97)
    if a == 2:
        return 0.02 + 0.0001 * ((i * 7) % 89)
    if a == 3:
        return 1.0e-6 * (i.astype(np.float64) - 10.0)
    return 0.001 * ((i * 11) % 83) - 0.04


def _content(nx, ny, nz, elastic):
    n = nx * ny * nz
    arrays = [_make_value(a, np.arange(n)).reshape(nx, ny, nz) for a in range(4)]
    k = nx * ny * (nz // 2 + 1)
    disp = _make_value(4, np.arange(6 * k)).astype(np.float32) if elastic else None
    header = dict(accepted_step=9, nx=nx, ny=ny, nz=nz, dt_code=0.01, temperature_K=653.15,
                  target_mass_code=123.25, last_lambda=-2.5e-7, last_residual_code=1.5e-14,
                  last_derivative_code=7.75, last_iterations=3, accepted_zero_mode_steps=9,
                  elastic_present=1 if elastic else 0, elastic_source_field_step=8 if elastic else 0,
                  elastic_last_iterations=5 if elastic else 0,
                  elastic_last_relative_residual=4.5e-9 if elastic else 0.0)
    return header, arrays, disp


def _write_mine(path, elastic, dims=(3, 4, 5)):
    header, arrays, disp = _content(*dims, elastic)
    pf.checkpoint_file_write(path, _prov_struct(_prov(elastic)), header, *arrays, displacement_k=disp)
    return header, arrays, disp


def _check_content(path, elastic, dims=(3, 4, 5), prov=None):
    header, arrays, disp = _content(*dims, elastic)
    h, data = pf.checkpoint_file_read(path, _prov_struct(prov or _prov(elastic)))
    for k, v in header.items():
        assert h[k] == v, k
    for name, ref in zip(("phi", "Y", "xB", "dYdt"), arrays):
        assert np.array_equal(data[name], ref), name
    if elastic:
        assert np.array_equal(data["displacement_k"], disp)
    return h


def test_fingerprints_match_the_reference_run(lib):
    """warm_start_residual.params (reports/pf_elastic_warm_start_residual_v1/provenance/)
    at the run's 246^3 grid with elastic_enabled=1 (the CLI's positional flag)."""
    p = pf.load_params_file(os.path.join(HERE, "golden", "warm_start_residual.params"))
    p.Nx = p.Ny = p.Nz = 246
    p.elastic_enabled = 1
    assert pf.parameter_fingerprint(p) == 0xAA304B81A6B1DBC7
    assert pf.elastic_solver_fingerprint(p) == 0x941613C0FF53008C
    q = pf.checkpoint_provenance(p).as_dict()
    assert q["elastic_solver_mode"] == "ELASTIC_WARM_START_RESIDUAL_V1"
    assert q["parameter_fingerprint"] == 0xAA304B81A6B1DBC7
    assert q["initial_state_class"] == "LEGACY_UNSPECIFIED"
    # one input changed changes the fingerprint; an input outside it does not
    p.dt = 0.01
    assert pf.parameter_fingerprint(p) != 0xAA304B81A6B1DBC7
    p.dt = 0.02
    p.mu_reference_scale = 1.0
    assert pf.parameter_fingerprint(p) == 0xAA304B81A6B1DBC7


def test_reference_unit_test_passes(ref_tools):
    r = subprocess.run([REF_TEST], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0 and "PASS_PF_ZERO_MODE_CHECKPOINT_PROVENANCE_V3_V4" in r.stdout


@pytest.mark.parametrize("elastic", [True, False], ids=["v4", "v3"])
def test_my_file_through_the_reference_is_byte_identical(lib, ref_tools, tmp_path, elastic):
    mine, back = str(tmp_path / "mine.chk"), str(tmp_path / "back.chk")
    _write_mine(mine, elastic)
    with open(mine, "rb") as f:
        head = f.read(16)
    assert head[:8] == (b"PFZMCHK4" if elastic else b"PFZMCHK3")
    assert struct.unpack("<II", head[8:])[1] == (904 if elastic else 800)
    r = subprocess.run([ref_tools, "roundtrip", mine, back] + _prov_args(_prov(elastic)), capture_output=True,
                       text=True, timeout=60)
    assert r.returncode == 0, r.stdout
    assert open(mine, "rb").read() == open(back, "rb").read()


@pytest.mark.parametrize("elastic", [True, False], ids=["v4", "v3"])
def test_reference_file_read_here(lib, ref_tools, tmp_path, elastic):
    ref = str(tmp_path / "ref.chk")
    r = subprocess.run([ref_tools, "make", ref, "3", "4", "5", "1" if elastic else "0"] + _prov_args(_prov(elastic)),
                       capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout
    h = _check_content(ref, elastic)
    assert h["version"] == (4 if elastic else 3)


def _v2_bytes(dims=(2, 3, 4)):
    """A V2 file (DiskHeaderV2, pf_zero_mode_checkpoint.cpp:27-53) built by
    hand: the reference writer no longer produces V2."""
    header, arrays, _ = _content(*dims, False)
    nx, ny, nz = dims
    sel = [b"PF_CONSERVED_Y_ZERO_MODE_V1", b"HOST_NEWTON_BISECTION_V1", b"legacy", b"lagged_rhs",
           b"SM_EXPLICIT_CONTEXT_N_V1", b"SM_TANGENT_N_V1"]
    body = struct.pack("<8sIIQQiiii6d3Q", b"PFZMCHK2", 2, 512, nx * ny * nz, header["accepted_step"], nx, ny, nz, 0,
                       header["dt_code"], header["temperature_K"], header["target_mass_code"], header["last_lambda"],
                       header["last_residual_code"], header["last_derivative_code"], header["last_iterations"],
                       header["accepted_zero_mode_steps"], 0x123456789ABCDEF0)
    body += b"".join(s.ljust(64, b"\0") for s in sel)
    payload = b"".join(np.ascontiguousarray(a).tobytes() for a in arrays)

    def fnv(data, h=1469598103934665603):
        for b in data:
            h = ((h ^ b) * 1099511628211) & 0xFFFFFFFFFFFFFFFF
        return h
    s = fnv(payload, fnv(body + b"\0" * 8))
    return body + struct.pack("<Q", s) + payload


def test_v2_file_read_here_and_by_the_reference(lib, ref_tools, tmp_path):
    legacy = _prov(False, initial_state_class="LEGACY_UNSPECIFIED", fixture_manifest_sha256="LEGACY_UNSPECIFIED",
                   profile_library_manifest_sha256="LEGACY_UNSPECIFIED")
    v2 = tmp_path / "v2.chk"
    v2.write_bytes(_v2_bytes())
    h = _check_content(str(v2), False, dims=(2, 3, 4), prov=legacy)
    assert h["version"] == 2
    # a V2 file needs the legacy identities
    with pytest.raises(pf.PfError, match="legacy checkpoint"):
        pf.checkpoint_file_read(str(v2), _prov_struct(_prov(False)))
    # the reference reads it too and rewrites it as V3, which reads back the same
    v3 = str(tmp_path / "v3.chk")
    r = subprocess.run([ref_tools, "roundtrip", str(v2), v3] + _prov_args(legacy), capture_output=True, text=True,
                       timeout=60)
    assert r.returncode == 0, r.stdout
    assert _check_content(v3, False, dims=(2, 3, 4), prov=legacy)["version"] == 3


@pytest.mark.parametrize("field,value,match", [
    ("backend", "UNQUALIFIED_BACKEND", "provenance mismatch"),
    ("explicit_context", "LEGACY_MIXED_PHI_NP1_V1", "provenance mismatch"),
    ("reaction_discretization", "FINITE_DELTA_H_REACTION_V1", "provenance mismatch"),
    ("fixture_manifest_sha256", "0" + FIX_SHA[1:], "provenance mismatch"),
    ("profile_library_manifest_sha256", "0" + LIB_SHA[1:], "provenance mismatch"),
    ("initial_state_class", "LEGACY_UNSPECIFIED", "provenance mismatch"),
    ("parameter_fingerprint", 0x123456789ABCDEF1, "provenance mismatch"),
])
@pytest.mark.parametrize("elastic", [True, False], ids=["v4", "v3"])
def test_provenance_mismatch_rejected(lib, tmp_path, elastic, field, value, match):
    path = str(tmp_path / "c.chk")
    _write_mine(path, elastic)
    with pytest.raises(pf.PfError, match=match) as ei:
        pf.checkpoint_file_read(path, _prov_struct(_prov(elastic, **{field: value})))
    assert ei.value.code == -7


def test_elastic_state_rules(lib, tmp_path):
    v4, v3 = str(tmp_path / "e.chk"), str(tmp_path / "n.chk")
    _write_mine(v4, True)
    _write_mine(v3, False)
    # elastic solver fingerprint
    with pytest.raises(pf.PfError, match="V4 elastic/zero-mode provenance mismatch"):
        pf.checkpoint_file_read(v4, _prov_struct(_prov(True, elastic_solver_fingerprint=0xFEDCBA9876543211)))
    # a V4 elastic state offered to a solver without elastic state
    with pytest.raises(pf.PfError, match="provenance mismatch"):
        pf.checkpoint_file_read(v4, _prov_struct(_prov(False)))
    # a V3 file offered to a solver that needs the elastic state
    with pytest.raises(pf.PfError, match="V3 lacks required elastic runtime state"):
        pf.checkpoint_file_read(v3, _prov_struct(_prov(True)))
    # the writer refuses an elastic state whose source step is not accepted_step - 1
    header, arrays, disp = _content(3, 4, 5, True)
    header["elastic_source_field_step"] = 7
    with pytest.raises(pf.PfError, match="elastic runtime state is invalid"):
        pf.checkpoint_file_write(str(tmp_path / "bad.chk"), _prov_struct(_prov(True)), header, *arrays,
                                 displacement_k=disp)
    assert not os.path.exists(str(tmp_path / "bad.chk.tmp"))


@pytest.mark.parametrize("where", ["last", "header", "field"])
def test_corruption_rejected(lib, tmp_path, where):
    path = str(tmp_path / "c.chk")
    _write_mine(path, True)
    data = bytearray(open(path, "rb").read())
    at = {"last": len(data) - 1, "header": 100, "field": 904 + 8 * 7}[where]
    data[at] ^= 0x1
    open(path, "wb").write(bytes(data))
    with pytest.raises(pf.PfError, match="checksum mismatch|provenance mismatch"):
        pf.checkpoint_file_read(path, _prov_struct(_prov(True)))
    # a trailing byte is rejected as well
    _write_mine(path, True)
    with open(path, "ab") as f:
        f.write(b"\0")
    with pytest.raises(pf.PfError, match="checksum mismatch"):
        pf.checkpoint_file_read(path, _prov_struct(_prov(True)))
