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


@pytest.mark.parametrize("M", [512, 1024])
def test_fp32_large_separable(M):
    """FP32 transforms at the production meshes (the register-resident
    kernels) on a separable input f(x) g(y) h(z), whose transform is the
    outer product F(kx) G(ky) H(kz) of 1D numpy transforms; checked plane by
    plane to keep host memory small."""
    rng = np.random.default_rng(M + 7)
    f, g, h = (rng.standard_normal(M) for _ in range(3))
    a = (np.outer(f, g).astype(np.float32)[:, :, None] * h.astype(np.float32)[None, None, :])
    F, G, H = np.fft.fft(f), np.fft.fft(g), np.fft.rfft(h)
    scale = np.abs(F).max() * np.abs(G).max() * np.abs(H).max()
    c = E.fft3d_r2c(a, E.FP32)
    err = 0.0
    for x in range(M):
        ref = F[x] * np.outer(G, H)
        err = max(err, float(np.abs(c[x] - ref).max()))
    assert err <= 3e-6 * scale, err / scale
    back = E.fft3d_c2r(c, E.FP32)
    del c
    amax = float(np.abs(a).max())
    berr = 0.0
    for x in range(0, M, 7):
        berr = max(berr, float(np.abs(back[x] / M ** 3 - a[x]).max()))
    assert berr <= 3e-6 * amax, berr / amax
