"""Seeded synthetic workloads (initial fields + code-unit parameters).

The reference launches ``main_cuda`` with raw_fields fixtures
(``main_cuda.cu:2733-3037``) and a schema-v1 ``.params`` file
(``main_cuda.cu:1089-1144``).  This module builds both for the five benchmark
workloads, with every random draw seeded so that the oracle and the HIP path
see identical inputs.

Stiffness is given in GPa and non-dimensionalised by ``12 gamma / lambda``
(``Unit_Psedobinary.py:592-712``) with gamma = 0.168 J/m^2 and lambda = 4 nm,
i.e. by 0.504 GPa.
"""
import math

import numpy as np

GPA_SCALE = 12.0 * 0.168 / 4.0e-9 / 1.0e9   # 0.504 GPa per code unit

# cubic PbTe matrix (GPa) and the Ag2Te - PbTe modulus difference (GPa)
C_MATRIX_GPA = (108.0, 7.7, 13.4)
DC_HETERO_GPA = (-48.0, 22.0, 1.6)
EPS00 = (0.046, -0.022, -0.017, 0.0, 0.0, 0.0)   # xx yy zz yz xz xy


def cubic_voigt(c11, c12, c44):
    m = np.zeros((6, 6))
    for i in range(3):
        for j in range(3):
            m[i, j] = c12
        m[i, i] = c11
        m[i + 3, i + 3] = c44
    return m


def stiffness(hetero: bool):
    C = cubic_voigt(*C_MATRIX_GPA) / GPA_SCALE
    dC = cubic_voigt(*DC_HETERO_GPA) / GPA_SCALE if hetero else np.zeros((6, 6))
    return C, dC


def base_params(n, dt=0.01, hetero=False, tol=1e-8, zero_mode=True, elastic=True):
    """Code-unit parameter dict shared by the oracle and the HIP library.

    Keys are the reference ``.params`` keys (``pf_params.h``); C and dC are
    6x6 Voigt matrices.
    """
    C, dC = stiffness(hetero)
    nx, ny, nz = (n, n, n) if np.isscalar(n) else n
    return dict(
        Nx=nx, Ny=ny, Nz=nz, dx=1.0, dy=1.0, dz=1.0, dt=dt,
        temperature_C=380.0, mu_reference_scale=20668.536,
        v_A=0.0, v_B=1.0, Vm_compound=1.0, Vm_alpha_0=1.0, dVm_alpha_dxB=0.0,
        W=1.0, kappa_phi=2.0, L_phi=1.0, D_alpha=1.0, D_compound=0.01,
        eps_iso_over_vB=0.00233, elastic_shift_dimless=0.0,
        Y_clip=20.0, xB_eps=1e-8, thermo_convex_extrapolation_enabled=0,
        elastic_enabled=1 if elastic else 0,
        C=C, dC=dC, eps00=EPS00, E0=(0.0,) * 6,
        elastic_warm_start_enabled=1, elastic_residual_control_enabled=1,
        elastic_iter_min=2, elastic_iter_max=32, elastic_residual_tolerance=tol,
        elastic_residual_absolute_floor=1e-30, elastic_fail_on_nonconvergence=1,
        zero_mode=1 if zero_mode else 0,
    )


def spheres_phi(shape, centres, radii, width=4.0, aspects=None, rotations=None, tail=10.0):
    """phi = max_p 1/2 (1 + tanh((R_p - r_p)/w)) with periodic distances.

    Each particle is evaluated inside a periodic box of half-width
    R/aspect + tail*w around its centre (the profile there is below
    e^(-2 tail) ~ 2e-9); outside the box its contribution is taken as 0.
    """
    nx, ny, nz = shape
    phi = np.zeros(shape)
    for p, (c, R) in enumerate(zip(centres, radii)):
        a = 1.0 if aspects is None else aspects[p]
        half = int(math.ceil(R / min(a, 1.0) + tail * width))
        idx = []
        for ax, n in enumerate(shape):
            if 2 * half + 1 >= n:
                idx.append(np.arange(n))
            else:
                c0 = int(math.floor(c[ax]))
                idx.append(np.arange(c0 - half, c0 + half + 2) % n)
        X = idx[0].astype(np.float64)[:, None, None]
        Yv = idx[1].astype(np.float64)[None, :, None]
        Z = idx[2].astype(np.float64)[None, None, :]
        dx = X - c[0]
        dy = Yv - c[1]
        dz = Z - c[2]
        dx = dx - nx * np.round(dx / nx)
        dy = dy - ny * np.round(dy / ny)
        dz = dz - nz * np.round(dz / nz)
        if rotations is not None:
            Rm = rotations[p]
            vx = Rm[0, 0] * dx + Rm[0, 1] * dy + Rm[0, 2] * dz
            vy = Rm[1, 0] * dx + Rm[1, 1] * dy + Rm[1, 2] * dz
            vz = Rm[2, 0] * dx + Rm[2, 1] * dy + Rm[2, 2] * dz
            dx, dy, dz = vx, vy, vz
        r = np.sqrt(dx * dx + dy * dy + (dz / a) ** 2)
        val = 0.5 * (1.0 + np.tanh((R - r) / width))
        sub = np.ix_(idx[0], idx[1], idx[2])
        phi[sub] = np.maximum(phi[sub], val)
    return phi


# x_B inside a precipitate.  The reference initialises the matrix composition
# field inside beta to x_eq(T) (init_xB_Y_from_phi_kernel, main_cuda.cu:2316-2330).
# A literal 0.33 puts x_B inside the regular-solution spinodal
# (0.103 < x < 0.897 at 380 C), where Gamma floors at 1e-4, M_eff hits the
# 1000 D cap and the explicit flux diverges at dt = 0.01 within five steps in
# the reference arithmetic; x_eq(380 C) = 0.004665 is used instead.
X_EQ_380C = 0.0046649518214523


def composition(phi, rng, inside=X_EQ_380C, outside=0.02, noise=1e-3):
    h = phi * phi * phi * (6.0 * phi * phi - 15.0 * phi + 10.0)
    return (1.0 - h) * outside + h * inside + rng.uniform(-noise, noise, size=phi.shape)


def place_nonoverlapping(rng, n, shape, radii, gap=2.0, max_tries=20000):
    centres = []
    for R in radii:
        for _ in range(max_tries):
            c = rng.uniform(0, 1, size=3) * np.array(shape)
            ok = True
            for c2, R2 in zip(centres, radii):
                d = np.abs(c - c2)
                d = np.minimum(d, np.array(shape) - d)
                if np.sqrt((d * d).sum()) < R + R2 + gap:
                    ok = False
                    break
            if ok:
                centres.append(c)
                break
        else:
            raise RuntimeError("could not place particle")
    return centres


def config(idx: int, n=None):
    """Return (params, phi0, xB0, meta) for benchmark workload ``idx`` (1-5).

    ``n`` optionally overrides the grid edge (used for reduced-size parity
    runs of the same geometry family).
    """
    if idx == 1:
        n = n or 64
        p = base_params(n, dt=0.01, hetero=False)
        rng = np.random.default_rng(1)
        c = (n / 2.0, n / 2.0, n / 2.0)
        phi = spheres_phi((n, n, n), [c], [10.0 * n / 64.0 if n != 64 else 10.0])
        xB = composition(phi, rng)
        meta = dict(workload="sphere_r10_homogeneous", steps=100)
    elif idx == 2:
        n = n or 128
        p = base_params(n, dt=0.01, hetero=True)
        rng = np.random.default_rng(2)
        radii = rng.uniform(6, 14, size=8) * (n / 128.0)
        centres = place_nonoverlapping(rng, 8, (n, n, n), radii)
        phi = spheres_phi((n, n, n), centres, radii)
        xB = composition(phi, rng)
        meta = dict(workload="eight_spheres_hetero", steps=500)
    elif idx == 3:
        n = n or 246
        p = base_params(n, dt=0.01, hetero=True)
        rng = np.random.default_rng(3)
        radii = np.exp(rng.normal(math.log(8.0), 0.3, size=60)) * (n / 246.0)
        centres = place_nonoverlapping(rng, 60, (n, n, n), radii, gap=1.0)
        phi = spheres_phi((n, n, n), centres, radii)
        xB = composition(phi, rng)
        meta = dict(workload="psd60_coarsening_246", steps=2000)
    elif idx == 5:
        n = n or 256
        p = base_params(n, dt=0.01, hetero=True)
        rng = np.random.default_rng(5)
        radii = rng.uniform(5, 12, size=40) * (n / 256.0)
        centres = place_nonoverlapping(rng, 40, (n, n, n), radii, gap=1.0)
        phi = spheres_phi((n, n, n), centres, radii)
        xB = composition(phi, rng)
        meta = dict(workload="forty_particles_hetero_256", steps=None)
    else:
        raise ValueError(f"unknown workload {idx}")
    meta["xB_max_safe"] = 0.5
    meta["centres"] = [tuple(float(v) for v in c) for c in (centres if idx != 1 else [c])]
    return p, phi, xB, meta


# ---------------------------------------------------------------- workload 4
def haar_rotation(rng):
    """Uniformly distributed rotation (QR of a Gaussian matrix, sign-fixed)."""
    Q, Rr = np.linalg.qr(rng.standard_normal((3, 3)))
    Q = Q * np.sign(np.diag(Rr))
    if np.linalg.det(Q) < 0:
        Q[:, 0] = -Q[:, 0]
    return Q


def rotated_eigenstrain(Q, principal=(0.046, -0.022, -0.017)):
    """Q diag(principal) Q^T in the order xx, yy, zz, yz, xz, xy."""
    E = Q @ np.diag(principal) @ Q.T
    return (E[0, 0], E[1, 1], E[2, 2], E[1, 2], E[0, 2], E[0, 1])


def sweep_instances(seed=4, counts=((128, 8), (96, 16), (64, 24))):
    """Workload 4: 48 minimize instances of uneven size, each one sphere or
    oblate spheroid (radius U(10, 30), aspect 1 or U(0.5, 1)) whose principal
    eigenstrains (0.046, -0.022, -0.017) are rotated by a seeded Haar-random
    rotation; the spheroid's axes follow the same rotation.  Heterogeneous
    elasticity; largest grids first."""
    rng = np.random.default_rng(seed)
    out = []
    for n, k in counts:
        for i in range(k):
            R = float(rng.uniform(10.0, 30.0))
            a = float(rng.uniform(0.5, 1.0)) if rng.random() < 0.5 else 1.0
            Q = haar_rotation(rng)
            pd = base_params(n, hetero=True, zero_mode=False)
            pd["eps00"] = rotated_eigenstrain(Q)
            c = (n / 2.0, n / 2.0, n / 2.0)
            phi = spheres_phi((n, n, n), [c], [R], aspects=[a], rotations=[Q.T])
            out.append(dict(name=f"n{n}_{i}", shape=(n, n, n), params=pd, phi=phi, radius=R, aspect=a))
    return out
