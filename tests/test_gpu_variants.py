"""Opt-in implementation variants selected by environment at library load:
PF_ZPASS=reg (the register-resident z pass, pf_regfft.hpp) and PF_KZ_PAD=0
(spectral rows not padded to whole cache lines).  Each runs in a child
process (the selection is read once per process) and must reproduce the FP64
oracle step to the same 1e-10 as the default path."""
import json
import os
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

CHILD = r"""
import json, sys
import numpy as np
sys.path.insert(0, sys.argv[1])
from cuda_sto_pf_amd import pfcore as pf
from cuda_sto_pf_amd.configs import base_params, composition, spheres_phi
from oracle.step_ref import Oracle, Params
shape = tuple(json.loads(sys.argv[2]))
pd = base_params(shape, hetero=True, zero_mode=True, elastic=True)
rng = np.random.default_rng(11)
phi = spheres_phi(shape, [(shape[0] / 2, shape[1] / 2, shape[2] / 2)], [3.0], width=1.5)
xB = composition(phi, rng)
o = Oracle(Params(**pd), phi, xB, xB_cap=0.5)
s = pf.Simulation(pd, pf.PF_FP64)
s.set_fields(phi, xB, xB_cap=0.5)
worst = 0.0
for _ in range(3):
    o.step()
    s.step()
    for f, ref in ((pf.PF_PHI, o.phi), (pf.PF_Y, o.Y), (pf.PF_XB, o.xB)):
        worst = max(worst, float(np.abs(s.field(f) - ref).max()))
print(json.dumps({"worst": worst, "iters": s.scalars()["last_elastic_iterations"],
                  "oracle_iters": o.last_elastic_iterations}))
"""

CASES = [("zpass_reg_32", {"PF_ZPASS": "reg"}, (32, 32, 32)),
         ("zpass_reg_kzpad", {"PF_ZPASS": "reg"}, (24, 20, 128)),
         ("no_kz_padding", {"PF_KZ_PAD": "0"}, (24, 20, 128))]


@pytest.mark.gpu
@pytest.mark.parametrize("name,env,shape", CASES, ids=[c[0] for c in CASES])
def test_variant_matches_oracle(name, env, shape):
    print(f"variant {name} {env} shape={shape}", flush=True)
    r = subprocess.run([sys.executable, "-c", CHILD, ROOT, json.dumps(shape)], capture_output=True, text=True,
                       timeout=240, env=dict(os.environ, **env))
    print(r.stdout[-800:], r.stderr[-1500:], flush=True)
    assert r.returncode == 0
    res = json.loads(r.stdout.strip().splitlines()[-1])
    assert res["worst"] < 1e-10
    assert res["iters"] == res["oracle_iters"]


KZ_CHILD = r"""
import hashlib, json, sys
import numpy as np
sys.path.insert(0, sys.argv[1])
from cuda_sto_pf_amd import pfcore as pf
from cuda_sto_pf_amd.configs import base_params, composition, spheres_phi
shape = (24, 20, 128)
pd = dict(base_params(shape, hetero=True, zero_mode=True, elastic=True), elastic_warm_start_order=3,
          elastic_anderson_depth=3)
rng = np.random.default_rng(12)
phi = spheres_phi(shape, [(12.0, 10.0, 64.0)], [4.0], width=1.5)
xB = composition(phi, rng)
out = {}
for name, prec in (("fp64", pf.PF_FP64), ("mixed", pf.PF_MIXED)):
    s = pf.Simulation(pd, prec)
    s.set_fields(phi, xB, xB_cap=0.5)
    s.step(4)
    h = hashlib.sha256()
    for f in (pf.PF_PHI, pf.PF_Y, pf.PF_XB):
        h.update(s.field(f).tobytes())
    sc = s.scalars()
    out[name] = {"hash": h.hexdigest(), "kz_stride": sc["kz_stride"],
                 "plan": [sc["zpass_rows"], sc["ypass_cols"], sc["xpass_cols"]]}
print(json.dumps(out))
"""


@pytest.mark.gpu
def test_kz_stride_choice_is_bitwise_neutral():
    """The plan tuner's kz strides (whole lines, half lines, none) move only
    where spectra sit: four steps give bitwise the same fields with each
    forced stride and with the tuned one (tile plan fixed)."""
    res = {}
    for pad in ("16", "4", "0", None):
        env = dict(os.environ, PF_PLAN="1,0,0")   # the tile plan fixed (it changes rounding)
        env.pop("PF_KZ_PAD", None)
        if pad is not None:
            env["PF_KZ_PAD"] = pad
        print(f"kz stride child PF_KZ_PAD={pad}", flush=True)
        r = subprocess.run([sys.executable, "-c", KZ_CHILD, ROOT], capture_output=True, text=True, timeout=240,
                           env=env)
        print(r.stdout[-600:], r.stderr[-1200:], flush=True)
        assert r.returncode == 0
        res[pad] = json.loads(r.stdout.strip().splitlines()[-1])
    for prec in ("fp64", "mixed"):
        strides = {res[p][prec]["kz_stride"] for p in ("16", "4", "0")}
        assert len(strides) >= 2, strides          # the forced strides differ
        assert len({res[p][prec]["hash"] for p in res}) == 1, prec
        assert res[None][prec]["kz_stride"] in {48, 44, 43}


PLAN_CHILD = r"""
import hashlib, json, sys
import numpy as np
sys.path.insert(0, sys.argv[1])
from cuda_sto_pf_amd import pfcore as pf
from cuda_sto_pf_amd.configs import base_params, composition, spheres_phi
mode, out_path, ck = sys.argv[2], sys.argv[3], sys.argv[4]
shape = (24, 20, 128)
pd = dict(base_params(shape, hetero=True, zero_mode=True, elastic=True), elastic_warm_start_order=3,
          elastic_anderson_depth=3)
rng = np.random.default_rng(12)
phi = spheres_phi(shape, [(12.0, 10.0, 64.0)], [4.0], width=1.5)
xB = composition(phi, rng)
res = {}
for name, prec in (("fp64", pf.PF_FP64), ("mixed", pf.PF_MIXED)):
    s = pf.Simulation(pd, prec)
    if mode == "restart":
        s.checkpoint_read(f"{ck}_{name}")
        s.step(3)
    else:
        s.set_fields(phi, xB, xB_cap=0.5)
        s.step(3)
        if mode == "checkpoint":
            s.checkpoint_write(f"{ck}_{name}")
        else:
            s.step(3)
    f = {k: s.field(v) for k, v in (("phi", pf.PF_PHI), ("Y", pf.PF_Y), ("xB", pf.PF_XB))}
    np.savez(f"{out_path}_{name}.npz", **f)
    h = hashlib.sha256()
    for k in ("phi", "Y", "xB"):
        h.update(f[k].tobytes())
    sc = s.scalars()
    res[name] = {"hash": h.hexdigest(), "plan": [sc["zpass_rows"], sc["ypass_cols"], sc["xpass_cols"]]}
print(json.dumps(res))
"""


def _plan_child(tmp_path, plan, mode, tag, ck):
    env = dict(os.environ)
    env.pop("PF_PLAN", None)
    if plan is not None:
        env["PF_PLAN"] = plan
    print(f"plan child PF_PLAN={plan} {mode}", flush=True)
    r = subprocess.run([sys.executable, "-c", PLAN_CHILD, ROOT, mode, str(tmp_path / tag), str(ck)],
                       capture_output=True, text=True, timeout=240, env=env)
    print(r.stdout[-600:], r.stderr[-1200:], flush=True)
    assert r.returncode == 0
    return json.loads(r.stdout.strip().splitlines()[-1])


@pytest.mark.gpu
def test_tile_plans_agree(tmp_path):
    """The plan tuner's tiles (rows per workgroup of the solver's z pass,
    column-tile widths of the y and x passes) change the compiled transform
    code, and with it the rounding of some operations: six steps with forced
    plans and with the tuned one agree to rounding in FP64 and to the FP32
    transforms' rounding in mixed precision."""
    ck = tmp_path / "unused"
    plans = ("1,0,0", "0,1,1", "2,2,2", "3,3,3", None)
    res = {p: _plan_child(tmp_path, p, "run", f"run{i}", ck) for i, p in enumerate(plans)}
    for prec, tol in (("fp64", 1e-12), ("mixed", 1e-6)):
        assert len({tuple(res[p][prec]["plan"]) for p in plans[:4]}) >= 3   # the forced plans differ
        ref = np.load(tmp_path / f"run0_{prec}.npz")
        for i in range(1, len(plans)):
            other = np.load(tmp_path / f"run{i}_{prec}.npz")
            for k in ("phi", "Y", "xB"):
                d = float(np.abs(other[k] - ref[k]).max())
                assert d < tol * max(1.0, float(np.abs(ref[k]).max())), (plans[i], prec, k, d)


@pytest.mark.gpu
def test_restart_continues_with_the_checkpointed_plan(tmp_path):
    """The exact-resume companion carries the tile plan: a restart in a
    process whose own plan differs continues bitwise like the run that wrote
    the checkpoint."""
    ck = tmp_path / "ck"
    cont = _plan_child(tmp_path, "3,3,3", "run", "cont", ck)
    _plan_child(tmp_path, "3,3,3", "checkpoint", "first", ck)
    rest = _plan_child(tmp_path, "0,1,1", "restart", "rest", ck)
    for prec in ("fp64", "mixed"):
        assert rest[prec]["plan"] == cont[prec]["plan"]
        assert rest[prec]["hash"] == cont[prec]["hash"], prec
