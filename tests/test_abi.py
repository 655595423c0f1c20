"""C ABI of libpfcore.so (CPU only: no device calls)."""
import ctypes as C
import os
import re

import pytest

from cuda_sto_pf_amd import params as pfparams
from cuda_sto_pf_amd.configs import base_params

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "pfcore.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:int|long long)\s+(pf_\w+)\s*\(", text, flags=re.M)))


def test_every_declared_symbol_is_exported(pf):
    names = declared_functions()
    assert len(names) >= 15
    raw = C.CDLL(pf.LIB_PATH)
    for n in names:
        assert hasattr(raw, n), n
        assert n in pf.EXPORTS, f"{n} missing from the ctypes binding"


def test_abi_version_and_struct_sizes(pf):
    assert pf.lib().pf_version() == 1
    p = pf.PfParams()
    assert pf.lib().pf_params_default(C.byref(p)) == 0
    assert p.Y_clip == 20.0 and p.xB_eps == 1e-8
    assert p.elastic_iter_min == 2 and p.elastic_iter_max == 32
    assert p.elastic_residual_tolerance == 1e-6 and p.zero_mode_max_iterations == 24


def test_params_file_roundtrip(pf, tmp_path):
    pd = base_params(32, hetero=True)
    path = pfparams.write_params_file(str(tmp_path / "run.params"), pd)
    p = pf.load_params_file(path)
    assert p.dt == pd["dt"] and p.kappa_phi == pd["kappa_phi"]
    assert abs(p.S[0] - pd["C"][0, 0]) == 0.0 and abs(p.S_p[15] - pd["dC"][3, 3]) == 0.0
    assert p.elastic_residual_tolerance == pd["elastic_residual_tolerance"]
    assert tuple(p.eps00) == tuple(pd["eps00"])


def test_params_missing_required_key_is_rejected(pf, tmp_path):
    pd = base_params(32)
    path = pfparams.write_params_file(str(tmp_path / "bad.params"), pd)
    lines = [ln for ln in open(path) if not ln.startswith("pf_params_schema_version")]
    open(path, "w").writelines(lines)
    with pytest.raises(pf.PfError) as e:
        pf.load_params_file(path)
    assert "pf_params_schema_version" in str(e.value)


def test_params_bad_value_and_syntax(pf, tmp_path):
    path = tmp_path / "syntax.params"
    path.write_text("dt 0.1\n")
    with pytest.raises(pf.PfError):
        pf.load_params_file(str(path))
    path.write_text("dt=abc\n")
    with pytest.raises(pf.PfError):
        pf.load_params_file(str(path))


def test_unknown_key_is_ignored(pf):
    p = pf.PfParams()
    pf.lib().pf_params_default(C.byref(p))
    assert pf.lib().pf_params_set(C.byref(p), b"no_such_key", b"1") == 0
    assert pf.lib().pf_params_set(C.byref(p), b"S_p_46", b"2.5") == 1
    assert p.S_p[17] == 2.5


@pytest.mark.parametrize("key,value,match", [
    ("elastic_iter_min", 0, "accelerated elastic iteration/residual contract"),
    ("elastic_iter_max", 1, "accelerated elastic iteration/residual contract"),   # below iter_min = 2
    ("elastic_residual_tolerance", 0.0, "accelerated elastic iteration/residual contract"),
    ("elastic_residual_tolerance", float("inf"), "accelerated elastic iteration/residual contract"),
    ("elastic_residual_absolute_floor", 0.0, "accelerated elastic iteration/residual contract"),
    ("elastic_fail_on_nonconvergence", 2, "accelerated elastic iteration/residual contract"),
    ("zero_mode_tolerance_relative", -1e-12, "zero-mode tolerance/iteration contract"),
    ("zero_mode_max_iterations", 0, "zero-mode tolerance/iteration contract"),
])
def test_create_refuses_invalid_solver_contracts(pf, key, value, match):
    """pf_create fails with PF_E_PARAMS before any device call on the
    reference's contracts (main_cuda.cu:24468-24475, 24492-24510)."""
    pd = dict(base_params(16, hetero=True), **{key: value})
    p = pf.params_from_dict(pd)
    h = C.c_void_p()
    rc = pf.lib().pf_create(C.byref(p), 0, C.byref(h))
    assert rc == -2
    buf = C.create_string_buffer(1024)
    pf.lib().pf_last_error(buf, 1024)
    assert match in buf.value.decode()
