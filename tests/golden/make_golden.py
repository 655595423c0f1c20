"""Write the golden fixtures of the oracle runs (CPU, numpy FP64).

  cfg1_64_100.npz   workload 1 (64^3 sphere, homogeneous elasticity, zero mode),
                    100 steps: central x-, y- and z-planes of phi, Y, x_B, plus
                    field checksums (sum, sum of squares, min, max), mass and
                    the energy densities of the last step.
  cfg2_16_10.npz    workload 2 geometry at 16^3 (heterogeneous, tol 1e-8,
                    zero mode), 10 steps: full phi, Y, x_B, dY/dt fields and the
                    per-step elastic iteration counts.
  cfg2_32_100.npz   workload 2 geometry at 32^3 (heterogeneous, tol 1e-8, zero
                    mode), 100 steps: full phi and Y fields, x_B checksums, the
                    per-step elastic iteration counts, mass and energy densities.

Run:  python tests/golden/make_golden.py [names...]   (a few minutes)
"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from cuda_sto_pf_amd.configs import config  # noqa: E402
from oracle.step_ref import Oracle, Params  # noqa: E402


def checksums(a):
    return np.array([a.sum(), (a * a).sum(), a.min(), a.max()])


def planes(a):
    n = a.shape
    return a[n[0] // 2].copy(), a[:, n[1] // 2].copy(), a[:, :, n[2] // 2].copy()


def cfg1():
    pd, phi, xB, meta = config(1)
    o = Oracle(Params(**pd), phi, xB, xB_cap=meta["xB_max_safe"])
    iters = []
    t = time.time()
    for s in range(100):
        o.step()
        iters.append(o.last_elastic_iterations)
    out = {"steps": 100, "iters": np.array(iters), "mass": o.mass(), "target_mass": o.target_mass,
           "energies": np.array([o.last_energies[k] for k in ("grad", "dw", "elastic", "chem")])}
    for name, f in (("phi", o.phi), ("Y", o.Y), ("xB", o.xB)):
        px, py, pz = planes(f)
        out[f"{name}_px"], out[f"{name}_py"], out[f"{name}_pz"] = px, py, pz
        out[f"{name}_sums"] = checksums(f)
    np.savez_compressed(os.path.join(HERE, "cfg1_64_100.npz"), **out)
    print(f"cfg1_64_100: {time.time() - t:.1f}s, mass drift {o.mass() - o.target_mass:.3e}")


def cfg2_small():
    pd, phi, xB, meta = config(2, 16)
    o = Oracle(Params(**pd), phi, xB, xB_cap=meta["xB_max_safe"])
    iters = []
    for s in range(10):
        o.step()
        iters.append(o.last_elastic_iterations)
    np.savez_compressed(os.path.join(HERE, "cfg2_16_10.npz"), phi=o.phi, Y=o.Y, xB=o.xB, dYdt=o.dYdt,
                        iters=np.array(iters), mass=o.mass(), target_mass=o.target_mass)
    print("cfg2_16_10 iterations", iters)


def cfg2_32():
    pd, phi, xB, meta = config(2, 32)
    o = Oracle(Params(**pd), phi, xB, xB_cap=meta["xB_max_safe"])
    iters = []
    t = time.time()
    for s in range(100):
        o.step()
        iters.append(o.last_elastic_iterations)
    np.savez_compressed(os.path.join(HERE, "cfg2_32_100.npz"), phi=o.phi, Y=o.Y, xB_sums=checksums(o.xB),
                        iters=np.array(iters), mass=o.mass(), target_mass=o.target_mass,
                        energies=np.array([o.last_energies[k] for k in ("grad", "dw", "elastic", "chem")]))
    print(f"cfg2_32_100: {time.time() - t:.1f}s, iterations {iters[:5]}...{iters[-3:]}")


JOBS = {"cfg2_16_10": cfg2_small, "cfg1_64_100": cfg1, "cfg2_32_100": cfg2_32}

if __name__ == "__main__":
    for name in (sys.argv[1:] or list(JOBS)):
        JOBS[name]()
