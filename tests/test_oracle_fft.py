"""Pins the oracle's 3D FFTs (fft_dir / fft_inv) against numpy's pocketfft and
the DFT definition."""
import numpy as np
import pytest

from gturbo3d_amd.params import Params
import oracle_lib as ol


@pytest.mark.parametrize("shape", [(8, 8, 8), (16, 16, 16), (16, 24, 32), (96, 12, 20), (30, 10, 6)])
def test_fft_dir_matches_numpy(shape):
    nz, ny, nx = shape
    p = Params(nx=nx, ny=ny, nz=nz)
    o = ol.Oracle(p)
    rng = np.random.default_rng(1)
    x = rng.standard_normal(shape)
    got = o.fft_dir(x)
    expect = np.fft.rfftn(x) / x.size  # fft_dir divides by NY, NX, NZ (src/fft_dir.f90:21)
    assert np.allclose(got, expect, rtol=0, atol=1e-15 * np.abs(expect).max() * 10)


@pytest.mark.parametrize("shape", [(8, 8, 8), (16, 24, 32), (12, 18, 10)])
def test_fft_inv_matches_numpy(shape):
    nz, ny, nx = shape
    p = Params(nx=nx, ny=ny, nz=nz)
    o = ol.Oracle(p)
    rng = np.random.default_rng(2)
    X = np.fft.rfftn(rng.standard_normal(shape))
    got = o.fft_inv(X)
    expect = np.fft.irfftn(X, s=shape, axes=(0, 1, 2)) * (nx * ny * nz)  # unnormalised Z2D
    assert np.allclose(got, expect, rtol=0, atol=1e-13 * np.abs(expect).max())


def test_fft_inv_ignores_imag_of_real_planes():
    # standard C2R: only Re X[kx=0] / Re X[kx=NX/2] enter after the y/z passes
    shape = (8, 8, 8)
    p = Params(nx=8, ny=8, nz=8)
    o = ol.Oracle(p)
    X = np.zeros((8, 8, 5), np.complex128)
    X[0, 0, 0] = 1.0 + 0.7j
    X[0, 0, 4] = -2.0 + 0.3j
    got = o.fft_inv(X)
    n = np.arange(8)
    expect = 1.0 + (-2.0) * np.cos(np.pi * n)[None, None, :] * np.ones(shape)
    assert np.allclose(got, expect, atol=1e-14)


def test_fft_direct_dft_small():
    nz, ny, nx = 6, 4, 10
    p = Params(nx=nx, ny=ny, nz=nz)
    o = ol.Oracle(p)
    rng = np.random.default_rng(3)
    x = rng.standard_normal((nz, ny, nx))
    got = o.fft_dir(x)
    z, y, xx = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    for kz in range(nz):
        for ky in range(ny):
            for kx in range(nx // 2 + 1):
                ph = np.exp(-2j * np.pi * (kx * xx / nx + ky * y / ny + kz * z / nz))
                assert abs(got[kz, ky, kx] - np.sum(x * ph) / x.size) < 1e-15


@pytest.mark.parametrize("n", [16, 32, 96])
def test_round_trip(n):
    p = Params(nx=n, ny=n, nz=n)
    o = ol.Oracle(p)
    x = np.random.default_rng(4).standard_normal((n, n, n))
    assert np.allclose(o.fft_inv(o.fft_dir(x)), x, atol=1e-13)
