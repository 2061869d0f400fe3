"""Device FFTs (replacing cufftExecR2C/C2R, pm_periodic.c:140,201) against
numpy golden vectors and numpy at the mesh sizes the workloads use."""
import numpy as np
import pytest

from cuco_amd import engine as E

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("M", [8, 12, 16])
def test_r2c_matches_golden(M, golden_dir):
    g = np.load(f"{golden_dir}/fft_golden.npz")
    a, ref = g[f"r3d_{M}_in"], g[f"r3d_{M}_rfftn"]
    c = E.fft3d_r2c(a, E.FP64)
    assert np.abs(c - ref).max() <= 1e-12 * np.abs(ref).max()
    c32 = E.fft3d_r2c(a, E.FP32)
    assert np.abs(c32 - ref).max() <= 2e-6 * np.abs(ref).max()


@pytest.mark.parametrize("M", [32, 64, 96, 128, 160, 192, 256])
def test_r2c_c2r_match_numpy(M):
    rng = np.random.default_rng(M)
    a = rng.standard_normal((M, M, M))
    ref = np.fft.rfftn(a)
    c = E.fft3d_r2c(a, E.FP64)
    assert np.abs(c - ref).max() <= 1e-12 * np.abs(ref).max()
    back = E.fft3d_c2r(ref, E.FP64)
    assert np.abs(back / M ** 3 - a).max() <= 1e-12 * np.abs(a).max()
    c32 = E.fft3d_r2c(a, E.FP32)
    assert np.abs(c32 - ref).max() <= 3e-6 * np.abs(ref).max()
    back32 = E.fft3d_c2r(ref, E.FP32)
    assert np.abs(back32 / M ** 3 - a).max() <= 3e-6 * np.abs(a).max()


def test_c2r_ignores_imaginary_dc_and_nyquist():
    M = 16
    rng = np.random.default_rng(2)
    a = rng.standard_normal((M, M, M))
    c = np.fft.rfftn(a)
    c[..., 0] += 0.5j * (np.abs(c[..., 0]) < np.inf)     # junk imaginary parts
    c[..., -1] += 0.25j
    ref = np.fft.irfftn(c, s=(M, M, M), axes=(0, 1, 2)) * M ** 3
    back = E.fft3d_c2r(c, E.FP64)
    assert np.abs(back - ref).max() <= 1e-11 * np.abs(ref).max()
