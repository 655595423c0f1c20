"""raw_fields init bundles for ``main_pf --init-mode raw_fields``.

A bundle is ``phi_init.raw`` and ``xB_init.raw`` (C-order [Nx][Ny][Nz],
float32 or float64) with ``init_meta.json`` (grid, spacing, dtype, order,
xB_max_safe, ...), the input of the reference's ``load_raw_init_fields``
(main_cuda.cu:2905-2965).  ``uniform`` restates the reference's
matrix-only writer (tools/analysis/write_uniform_matrix_raw_init.py:1-60):
phi = 0 and x_B = xb_out everywhere, float32, with the same meta keys.

    python -m cuda_sto_pf_amd.raw_init uniform --out-dir D --nx 64 --ny 64 --nz 64 \\
        --dx-nm 1 --interface-width-nm 4 --xb-out 0.03
"""
import argparse
import json
from pathlib import Path

import numpy as np


def write_bundle(out_dir, phi, xB, dx_nm, interface_width_nm, dtype="float64", **meta):
    """Write phi/xB as raw arrays of `dtype` and the meta JSON; extra meta
    keys (xB_max_safe, mean_xBtot, ...) are written as given."""
    out = Path(out_dir).expanduser().resolve()
    out.mkdir(parents=True, exist_ok=True)
    phi = np.ascontiguousarray(phi, dtype=dtype)
    xB = np.ascontiguousarray(xB, dtype=dtype)
    if phi.ndim != 3 or phi.shape != xB.shape:
        raise ValueError("phi and xB must be 3-D arrays of one shape")
    phi.tofile(out / "phi_init.raw")
    xB.tofile(out / "xB_init.raw")
    nx, ny, nz = phi.shape
    payload = {"Nx": nx, "Ny": ny, "Nz": nz, "dx_nm": float(dx_nm), "interface_width_nm": float(interface_width_nm),
               "dtype": np.dtype(dtype).name, "order": "C"}
    payload.update(meta)
    (out / "init_meta.json").write_text(json.dumps(payload, indent=2), encoding="utf-8")
    return out


def ownership_labels(shape, centres):
    """Periodic nearest-centre ownership (the V5 ownership map of
    scripts/materialize_pf_elastic_multi_particle_component_constraint_v5.py:
    every cell goes to the closest centre on the periodic grid, ties to the
    lower index); centres in grid units.  int32, C order."""
    coords = np.indices(shape, dtype=np.float64)
    best = np.full(shape, np.inf)
    owner = np.zeros(shape, dtype=np.int32)
    for k, c in enumerate(centres):
        d2 = np.zeros(shape)
        for ax, n in enumerate(shape):
            d = (coords[ax] - c[ax] + 0.5 * n) % n - 0.5 * n
            d2 += d * d
        closer = d2 < best
        owner[closer] = k
        best[closer] = d2[closer]
    return owner


def component_h_sums(phi, labels, count):
    """The per-label sums of h(phi) = phi^3 (6 phi^2 - 15 phi + 10) over
    clamp01(phi) that --minimize-component-target-h-sums takes."""
    p = np.clip(phi, 0.0, 1.0)
    h = p * p * p * (6.0 * p * p - 15.0 * p + 10.0)
    return np.bincount(labels.ravel(), weights=h.ravel(), minlength=count)


def uniform(out_dir, nx, ny, nz, dx_nm, interface_width_nm, xb_out):
    """Matrix-only bundle (write_uniform_matrix_raw_init.py:28-57)."""
    shape = (nx, ny, nz)
    return write_bundle(out_dir, np.zeros(shape, dtype=np.float32), np.full(shape, float(xb_out), dtype=np.float32),
                        dx_nm, interface_width_nm, dtype="float32", dt_recommended=None,
                        mean_xBtot=float(xb_out), xB_max_safe=float(xb_out),
                        reference_type="matrix_only_same_strain", mean_phi=0.0, mean_h=0.0, voxel_count=0)


def main(argv=None):
    ap = argparse.ArgumentParser(description="raw_fields init bundles for main_pf")
    sub = ap.add_subparsers(dest="cmd", required=True)
    u = sub.add_parser("uniform", help="matrix-only bundle: phi = 0, x_B = xb_out (float32)")
    u.add_argument("--out-dir", type=Path, required=True)
    u.add_argument("--nx", type=int, required=True)
    u.add_argument("--ny", type=int, required=True)
    u.add_argument("--nz", type=int, required=True)
    u.add_argument("--dx-nm", type=float, required=True)
    u.add_argument("--interface-width-nm", type=float, required=True)
    u.add_argument("--xb-out", type=float, required=True)
    u.add_argument("--dtype", default="float32", choices=("float32",))
    a = ap.parse_args(argv)
    print(uniform(a.out_dir, a.nx, a.ny, a.nz, a.dx_nm, a.interface_width_nm, a.xb_out))
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
