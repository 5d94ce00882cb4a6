"""The mean-field focus series on the GPU (iscg_wf_conv3d: one 3D FFT convolution) against the oracle's direct
per-focus convolution (cli/widefield_check.cpp:287-343 DirectImage, PlanFocus's blend and clamp): FP32 path vs FP64,
relative L2 <= 1e-5 per focus; the fused camera noise bit-exact against the oracle's chain on the same frames."""
import numpy as np
import pytest

import oracle as O
from insiliscope_amd import host

pytestmark = pytest.mark.gpu

TOL = 1e-5  # relative L2 per focus image, FP32 transforms against the FP64 direct sum


def rel_l2(a, b):
    return float(np.sqrt(((np.asarray(a, np.float64) - b) ** 2).sum() / (np.asarray(b, np.float64) ** 2).sum()))


def small_case(golden, seed=3, nx=48, ny=40, nzD=6):
    P = golden("psf_ref.npz")
    planes = P["small_planes"]  # 9 planes, 65^2, os 4, 50 nm apart
    rng = np.random.default_rng(seed)
    dens = np.where(rng.random((nzD, ny, nx)) < 0.05, 1.0 + np.floor(rng.random((nzD, ny, nx)) * 3), 0.0)
    return planes, dens.astype(np.float32)


@pytest.mark.parametrize("delta_um,fov", [(0.0, (0, 0, 0, 0)), (0.013, (5, 3, 30, 25)), (-0.031, (0, 7, 48, 20))])
def test_wf_focus_series_vs_direct(ctx, orc, golden, delta_um, fov):
    planes, dens = small_case(golden)
    nzD, ny, nx = dens.shape
    dz, k0, nF = 0.05, -3, 7
    focus0 = (k0 + 0.5) * dz + delta_um
    print("delta", delta_um, "fov", fov, flush=True)
    cache = host.PsfKernelCache(planes, oversampling=4, zStepNm=50.0)
    wf = host.WidefieldStack(ctx)
    assert wf.Plan(cache, nx, ny, nzD, dz, k0, focus0, nF, fov=fov), wf.err
    nF_, ch, cw = wf.image_shape
    img = np.zeros((nF, ch, cw), np.float32)
    assert wf.FocusSeries(dens, outImages=img), wf.err
    R = orc.wf_radius(65, 4, 1)
    K = np.stack([orc.wf_kernel(p, 4, 1, R) for p in planes])
    foci = focus0 + dz * np.arange(nF)
    want = orc.wf_direct(dens, dz, k0, K, 50.0, foci)
    x0, y0 = fov[0], fov[1]
    want = np.maximum(want[:, y0:y0 + ch, x0:x0 + cw], 0.0)
    for f in range(nF):
        e = rel_l2(img[f], want[f])
        assert e <= TOL, f"focus {f}: rel L2 {e:.3g}"


def test_wf_noise_epilogue_bit_exact(ctx, orc, golden):
    planes, dens = small_case(golden, seed=4)
    nzD, ny, nx = dens.shape
    dens *= 40.0
    cache = host.PsfKernelCache(planes, oversampling=4, zStepNm=50.0)
    wf = host.WidefieldStack(ctx)
    assert wf.Plan(cache, nx, ny, nzD, 0.05, 0, 0.025, 5), wf.err
    shp = wf.image_shape
    off = np.full(shp[1:], 100.0, np.float32)
    cam = host.CameraNoiseParams(quantumEfficiency=1.0, gainPhotonsPerAdu=2.0, readNoiseElectrons=1.5)
    img = np.zeros(shp, np.float32)
    adu = np.zeros(shp, np.uint16)
    assert wf.FocusSeries(dens, cam=cam, offset=off, background=20.0, seed=1, frameBase=100, outImages=img,
                          outAdu=adu), wf.err
    ocam = O.camera(quantumEfficiency=1.0, gainPhotonsPerAdu=2.0, readNoiseElectrons=1.5)
    for f in range(shp[0]):
        assert np.array_equal(adu[f], orc.noise(img[f], ocam, 1, 100 + f, offset=off))


def test_wf_delta_density_reproduces_kernel(ctx, orc, golden):
    """One dye at a plane centre seen from that plane: the kernel plane at zero defocus, centred on the dye."""
    P = golden("psf_ref.npz")
    planes = P["small_planes"]
    nx = ny = 32
    dens = np.zeros((3, ny, nx), np.float32)
    dens[1, 16, 15] = 1.0
    cache = host.PsfKernelCache(planes, oversampling=4, zStepNm=50.0)
    wf = host.WidefieldStack(ctx)
    assert wf.Plan(cache, nx, ny, 3, 0.05, 0, 0.075, 3), wf.err
    img = np.zeros((3, ny, nx), np.float32)
    assert wf.FocusSeries(dens, outImages=img), wf.err
    R = orc.wf_radius(65, 4, 1)
    K = orc.wf_kernel(planes[4], 4, 1, R)
    want = np.zeros((ny, nx))
    for dy in range(-R, R + 1):
        for dx in range(-R, R + 1):
            y, x = 16 + dy, 15 + dx
            if 0 <= y < ny and 0 <= x < nx:
                want[y, x] = K[dy + R, dx + R]
    assert rel_l2(img[0], want) < TOL
    assert abs(img[0].astype(np.float64).sum() - want.sum()) < 1e-5
