"""Physical inputs -> code-unit .params (cuda_sto_pf_amd/units.py, the
reference's Unit_Psedobinary.py converter).

Pin: the qualification run's parameter file
(reports/pf_elastic_warm_start_residual_v1/provenance/warm_start_residual.params,
copied to tests/golden/) is the converter's output for the built-in inputs
with gamma = 0.168 J/m^2, lambda_sm = 4 nm, pf_dx = phys_dx_ref = 1 nm and
dt = 0.02 (its values fix these: S_11 = 108 GPa / (12 gamma / lambda) = 214.29,
ic_phi_iface_w = lambda / (2 pf_dx) = 2, D_alpha = (5 lambda / dx)^2 = 400).
All 95 key=value lines must come out identical; the lines the qualification
driver appended after them (the elastic solver controls) are excluded.
"""
import json
import math
import os

import pytest

from cuda_sto_pf_amd import units

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "warm_start_residual.params")
QUAL = dict(gamma=0.168, lambda_sm=4e-9, pf_dx=1e-9, phys_dx_ref=1e-9, dt=0.02)


def _golden_block():
    return open(GOLDEN).read().split("\n\n")[0] + "\n"


def test_reproduces_the_qualification_params_bitwise():
    o = units.main_cuda_overrides(units.load_physical_inputs(None, **QUAL))
    for k in units.ELASTIC_SOLVER_DEFAULTS:
        o.pop(k)
    assert units.format_params(o) == _golden_block()


def test_command_line_and_loader(tmp_path):
    phys = tmp_path / "phys.json"
    phys.write_text(json.dumps(dict(QUAL, _comment="SI units")))
    out = tmp_path / "run.params"
    units.main(["--input", str(phys), "--out", str(out)])
    text = out.read_text()
    assert text.startswith(_golden_block()[:200])
    assert "elastic_warm_start_enabled=1" in text and "elastic_residual_tolerance=9.9999999999999995e-07" in text
    from cuda_sto_pf_amd import pfcore as pf
    try:
        pf.load_library()
    except FileNotFoundError:
        pytest.skip("libpfcore.so is not built")
    p = pf.load_params_file(str(out))    # schema-checked like main_cuda's loader
    assert p.pf_params_schema_version == 1 and p.elastic_iter_max == 32
    assert p.S[0] == pytest.approx(108.0e9 / (12 * 0.168 / 4e-9))


def test_unknown_input_rejected(tmp_path):
    phys = tmp_path / "phys.json"
    phys.write_text(json.dumps({"gamma": 0.1, "gama": 0.2}))
    with pytest.raises(KeyError, match="gama"):
        units.load_physical_inputs(str(phys))


def test_physics_invariants():
    inp = units.load_physical_inputs(None, **QUAL)
    s = units.convert(inp)
    T = inp.temperature_C + 273.15
    # the matrix solubility solves the tangent condition RT ln x + L (1-x)^2 = 0
    x = s["xAg2Te_eq"]
    L = 41212.9 - 18.05 * T
    assert abs(units.R_GAS * T * math.log(x) + L * (1 - x) ** 2) < 1e-6 * L
    # diffusivity ratio, and the precipitate stiffness as matrix + S_p
    assert s["D_compound"] / s["D_alpha"] == pytest.approx(inp.D_ratio, rel=1e-14)
    o = units.main_cuda_overrides(inp)
    w = 12 * inp.gamma / inp.lambda_sm
    assert o["S_11"] + o["S_p_11"] == pytest.approx(inp.C_tensor_Ag2Te_GPa[0][0] * 1e9 / w, rel=1e-13)
    # a 90-degree rotation about z swaps the xx and yy eigenstrains
    rot = units.load_physical_inputs(None, eigenstrain_rotation_matrix=[[0, -1, 0], [1, 0, 0], [0, 0, 1]])
    r = units.main_cuda_overrides(rot)
    assert r["eps_xx00"] == pytest.approx(-0.022) and r["eps_yy00"] == pytest.approx(0.046)
    with pytest.raises(ValueError, match="orthogonal"):
        units.main_cuda_overrides(units.load_physical_inputs(None, eigenstrain_rotation_matrix=[[1, 1, 0], [0, 1, 0],
                                                                                                [0, 0, 1]]))


def test_uniform_raw_bundle(tmp_path):
    """The matrix-only bundle of tools/analysis/write_uniform_matrix_raw_init.py:
    float32 zeros and a constant x_B, and the same meta keys and values."""
    import numpy as np
    from cuda_sto_pf_amd import raw_init
    raw_init.main(["uniform", "--out-dir", str(tmp_path / "b"), "--nx", "4", "--ny", "3", "--nz", "2",
                   "--dx-nm", "1", "--interface-width-nm", "4", "--xb-out", "0.03"])
    d = tmp_path / "b"
    assert np.array_equal(np.fromfile(d / "phi_init.raw", dtype=np.float32), np.zeros(24, np.float32))
    assert np.array_equal(np.fromfile(d / "xB_init.raw", dtype=np.float32), np.full(24, 0.03, np.float32))
    meta = json.loads((d / "init_meta.json").read_text())
    assert meta == {"Nx": 4, "Ny": 3, "Nz": 2, "dx_nm": 1.0, "interface_width_nm": 4.0, "dtype": "float32",
                    "order": "C", "dt_recommended": None, "mean_xBtot": 0.03, "xB_max_safe": 0.03,
                    "reference_type": "matrix_only_same_strain", "mean_phi": 0.0, "mean_h": 0.0, "voxel_count": 0}


def test_calphad_functions_agree_with_the_step_thermodynamics():
    """The converter's SGTE/CALPHAD functions (restated from
    Unit_Psedobinary.py:441-485) and the time step's (oracle/thermo.py, from
    thermo_utils.h:316-377) are two independent transcriptions of two
    reference sources: they agree on both sides of every SGTE breakpoint
    (600.61 K Pb, 722.66 K Te, 1234.93 K Ag), so a dropped term or a wrong
    coefficient in either fails here."""
    import numpy as np
    from oracle import thermo
    for T in (400.0, 600.0, 600.62, 653.15, 700.0, 722.65, 722.67, 900.0, 1234.92, 1234.94, 1400.0):
        assert units._ghser_pb(T) == pytest.approx(thermo.GHSER_Pb(T), rel=1e-14, abs=1e-9)
        assert units._ghser_ag(T) == pytest.approx(thermo.GHSER_Ag(T), rel=1e-14, abs=1e-9)
        assert units._ghser_te(T) == pytest.approx(thermo.GHSER_Te(T), rel=1e-14, abs=1e-9)
        assert units._g_pbte(T) == pytest.approx(thermo.G_PbTe_Solid(T), rel=1e-14)
        assert units._g_ag2te(T) == pytest.approx(thermo.G_Ag2Te_Solid(T), rel=1e-14)
        assert units._l0(T) == pytest.approx(thermo.get_L_param(T), rel=1e-15)
        for x in (1e-6, 0.004665, 0.03, 0.5, 0.97):
            assert units.mu_pbte(T, x) == pytest.approx(float(thermo.mu_PbTe_calphad(T, np.float64(x))), rel=1e-13)
            assert units.mu_ag2te(T, x) == pytest.approx(float(thermo.mu_Ag2Te_calphad(T, np.float64(x))), rel=1e-13)
