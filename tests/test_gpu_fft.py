"""Pruned 3-D FFT passes (libpfcore) against numpy's rfftn/irfftn with the 2/3
dealias of dealias_23_kernel (cuda_kernels.cu:103-161)."""
import numpy as np
import pytest

from oracle.step_ref import KSpace, Params

# 256 and 512 run two radix-16 stages (the second with 16x16 twiddles), 192 and
# 384 mix radix 16 with 3 and 8; 256 along each axis is the headline grid's plan
SHAPES = [(8, 8, 8), (16, 12, 8), (12, 20, 24), (32, 32, 32), (6, 8, 246), (246, 4, 6), (64, 48, 40),
          (96, 16, 128), (256, 6, 8), (8, 256, 6), (6, 8, 256), (4, 6, 512), (192, 4, 384)]


def _mask(shape):
    nx, ny, nz = shape
    return KSpace(Params(Nx=nx, Ny=ny, Nz=nz)).mask


@pytest.mark.gpu
@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("prec,tol", [(0, 1e-13), (1, 2e-6)])
def test_forward_matches_numpy(pf, shape, prec, tol):
    rng = np.random.default_rng(sum(shape) + prec)
    a = rng.standard_normal(shape)
    ref = np.fft.rfftn(a) * _mask(shape)
    got = pf.fft_forward(a, prec)
    scale = np.abs(ref).max()
    err = np.abs(got - ref).max() / scale
    print(f"forward {shape} prec={prec} rel_err={err:.3e}", flush=True)
    assert err < tol


@pytest.mark.gpu
@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("prec,tol", [(0, 1e-13), (1, 2e-6)])
def test_inverse_matches_numpy(pf, shape, prec, tol):
    rng = np.random.default_rng(7 + sum(shape) + prec)
    nx, ny, nz = shape
    spec = (rng.standard_normal((nx, ny, nz // 2 + 1))
            + 1j * rng.standard_normal((nx, ny, nz // 2 + 1))) * _mask(shape)
    ref = np.fft.irfftn(spec, s=shape, axes=(0, 1, 2)) * (nx * ny * nz)
    got = pf.fft_inverse(spec, shape, prec)
    err = np.abs(got - ref).max() / np.abs(ref).max()
    print(f"inverse {shape} prec={prec} rel_err={err:.3e}", flush=True)
    assert err < tol
