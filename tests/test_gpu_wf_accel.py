"""The per-focus mean-field seam on the GPU (iscg_wf_images / has_plane / fits behind host.WidefieldGpuHost, the
reference's WidefieldAccelerator): the reference scenes' own jobs against their CPU images, and synthetic jobs at
the C2 grid against the FP64 restatement of JobReferenceHost, to the reference's JobReferenceHost bar (1e-5)."""
from types import SimpleNamespace

import numpy as np
import pytest

import oracle as O
from insiliscope_amd import host

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


@pytest.fixture()
def accel(ctx):
    return host.WidefieldGpuHost(ctx)


def test_reference_jobs_match_reference_scene_images(accel, golden):
    G = golden("wf_jobs_ref.npz")
    for pre in ("a_", "b_"):
        jobs = O.unpack_wf_jobs(G, pre)
        print("scene", pre, "jobs", len(jobs), "grid", jobs[0].NX, jobs[0].NY, flush=True)
        outs = []
        for j in jobs:
            got = []
            assert accel.Images(j, got), accel.err
            outs.append(got[0])
        for key, _, _ in jobs[0].planes:
            assert accel.HasPlane(jobs[0].geometry, key)
        for f, want in enumerate(G[pre + "images"]):
            assert rel_l2(outs[f + 1], want) < 1e-5, (pre, f, rel_l2(outs[f + 1], want))


def random_job(rng, NX, NY, nx, ny, nPlanes, nKernels, nChannels, geometry, fov, frac, keys=None, carry=True):
    kernels = []
    for p in range(nKernels):
        k = np.zeros((NY, NX))
        r = 6
        g = rng.random((2 * r + 1, 2 * r + 1))
        for dy in range(-r, r + 1):
            for dx in range(-r, r + 1):
                k[dy % NY, dx % NX] = g[dy + r, dx + r]
        kernels.append((p, np.fft.rfft2(k).astype(np.complex64)))
    keys = keys if keys is not None else [1000 + i for i in range(nPlanes)]
    planes = []
    if carry:
        for key in keys:
            n = int(rng.integers(1, nx * ny // 20))
            cells = rng.choice(nx * ny, n, replace=False).astype(np.uint32)
            planes.append((key, cells, (rng.random(n) * 3 + 0.1).astype(np.float32)))
    channels = []
    for c in range(nChannels):
        deps = []
        for key in keys:
            p0 = int(rng.integers(0, nKernels - 1))
            two = bool(rng.random() < 0.7)
            w1 = float(rng.random()) if two else 0.0
            deps.append((key, p0, p0 + 1 if two else p0, 1.0 - w1, w1, two))
        channels.append(deps)
    return SimpleNamespace(NX=NX, NY=NY, nx=nx, ny=ny, fovX0=fov[0], fovY0=fov[1], cw=fov[2], ch=fov[3],
                           fracX=frac[0], fracY=frac[1], geometry=geometry, kernels=kernels, planes=planes,
                           channels=channels, hasPersistent=True)


@pytest.mark.parametrize("NX,NY,nx,ny,fov", [(512, 512, 256, 256, (0, 0, 256, 256)),
                                             (1024, 256, 700, 200, (37, 11, 600, 180)),
                                             (64, 128, 40, 90, (3, 5, 30, 70)),
                                             (1024, 2048, 900, 1800, (50, 100, 800, 1600)),
                                             (2048, 512, 1500, 400, (10, 20, 1400, 360))])
def test_synthetic_jobs_vs_oracle(ctx, NX, NY, nx, ny, fov):
    """Two jobs per geometry: the first carries the planes, the second only refers to them (resident), with
    another sub-cell shift and other blends; three channels."""
    rng = np.random.default_rng(NX + NY)
    accel = host.WidefieldGpuHost(ctx)
    ref = O.JobReferenceHost()
    geometry = 0x5000 + NX + NY
    j1 = random_job(rng, NX, NY, nx, ny, 6, 8, 3, geometry, fov, (0.31, -0.77))
    keys = [pl[0] for pl in j1.planes]
    j2 = random_job(rng, NX, NY, nx, ny, 6, 8, 3, geometry, fov, (-0.5, 0.5), keys=keys, carry=False)
    j2.kernels = j1.kernels
    for j in (j1, j2):
        print("grid", NX, NY, "planes carried", len(j.planes), flush=True)
        got = []
        assert accel.Images(j, got), accel.err
        want = ref.Images(j)
        assert len(got) == 3
        for c in range(3):
            assert got[c].shape == (fov[3], fov[2])
            assert rel_l2(got[c], want[c]) < 1e-5, (c, rel_l2(got[c], want[c]))


def test_plane_residency_errors_and_recarry(ctx):
    rng = np.random.default_rng(9)
    accel = host.WidefieldGpuHost(ctx)
    ref = O.JobReferenceHost()
    fov = (0, 0, 48, 40)
    j1 = random_job(rng, 64, 64, 48, 40, 3, 4, 1, 77, fov, (0.2, 0.1))
    assert accel.Images(j1, []), accel.err
    ref.Images(j1)
    assert accel.HasPlane(77, 1000) and not accel.HasPlane(77, 5) and not accel.HasPlane(78, 1000)
    # a dep on a plane neither carried nor resident fails, as the reference's host does
    bad = random_job(rng, 64, 64, 48, 40, 1, 4, 1, 77, fov, (0.0, 0.0), keys=[424242], carry=False)
    bad.kernels = j1.kernels
    assert not accel.Images(bad, [])
    assert "does not carry" in accel.err
    # a resident plane carried again (new values): transformed into its slot, out of order with a new one
    j3 = random_job(rng, 64, 64, 48, 40, 2, 4, 1, 77, fov, (0.0, 0.0), keys=[1001, 2000])
    j3.kernels = j1.kernels
    got = []
    assert accel.Images(j3, got), accel.err
    assert rel_l2(got[0], ref.Images(j3)[0]) < 1e-5


def test_fits_and_budget(ctx):
    accel = host.WidefieldGpuHost(ctx)
    assert accel.Fits(91, 512, 512, 100)
    assert not accel.Fits(92, 2048, 2048, 10 ** 7)
    # grids the passes do not take do not fit (the scene then renders on the CPU instead of failing in Images):
    # powers of two up to WidefieldGpuJob::kMaxGpuFft = 2048 fit
    for geom, NX, NY in ((96, 96, 64), (97, 8, 64), (99, 4096, 64), (100, 64, 4096)):
        assert not accel.Fits(geom, NX, NY, 1), (NX, NY)
    for geom, NX, NY in ((94, 2048, 512), (95, 1024, 1024), (98, 2048, 2048)):
        assert accel.Fits(geom, NX, NY, 1), (NX, NY)
    accel.SetBudget(3 * 64 * 48 * 8)  # room for three 64-point plane spectra
    try:
        rng = np.random.default_rng(2)
        j = random_job(rng, 64, 64, 40, 40, 8, 4, 1, 93, (0, 0, 40, 40), (0.0, 0.0))
        assert not accel.Images(j, [])
        assert "budget" in accel.err
    finally:
        accel.SetBudget(0)


def test_device_output_and_kernel_change(ctx):
    """Images into a device tensor; a kernel spectrum at a new address within the geometry is taken again."""
    import torch
    rng = np.random.default_rng(5)
    accel = host.WidefieldGpuHost(ctx)
    ref = O.JobReferenceHost()
    fov = (0, 0, 100, 90)
    j = random_job(rng, 128, 128, 100, 90, 4, 5, 2, 501, fov, (0.4, 0.25))
    out = torch.empty((2, 90, 100), dtype=torch.float32, device="cuda")
    assert accel.Images(j, outArray=out), accel.err
    torch.cuda.synchronize()
    want = ref.Images(j)
    for c in range(2):
        assert rel_l2(out[c].cpu().numpy(), want[c]) < 1e-5
    j2 = SimpleNamespace(**vars(j))
    j2.planes = []
    j2.kernels = [(p, (spec * 2).astype(np.complex64)) for p, spec in j.kernels]
    got = []
    assert accel.Images(j2, got), accel.err
    for c in range(2):
        assert rel_l2(got[c], 2 * want[c]) < 1e-5


def test_widefield_images_render_bit_exact_vs_reference(ctx, golden):
    """WidefieldImages::Render on the GPU: bleach coefficients (zeros bound the terms), 1x1 and 2x2 binning, with and
    without the persistent channel, onto a base image: the compiled reference's frames bit for bit."""
    G = golden("wf_jobs_ref.npz")
    ch, base, a = G["render_channels"], G["render_base"], G["render_coeffs"]
    ph, _ = host.WidefieldImagesRender(ch, True, 1, a, ctx=ctx)
    assert np.array_equal(ph, G["render_u1"])
    ph, _ = host.WidefieldImagesRender(ch, True, 2, a, base=np.stack([base] * len(a)), ctx=ctx)
    assert np.array_equal(ph, G["render_u2"])
    ph, _ = host.WidefieldImagesRender(ch[1:], False, 2, a, ctx=ctx)
    assert np.array_equal(ph, G["render_u2_nop"])


def test_widefield_render_noise_on_accelerator_output(ctx, orc):
    """The per-focus seam end to end on the device: job -> channel images (device) -> frames with bleach
    coefficients and sCMOS noise; the ADU equal the noise chain of the rendered photons (oracle), bit for bit."""
    import torch
    rng = np.random.default_rng(31)
    accel = host.WidefieldGpuHost(ctx)
    fov = (0, 0, 64, 48)
    j = random_job(rng, 128, 64, 64, 48, 3, 4, 3, 9001, fov, (0.2, -0.4))
    imgs = torch.empty((3, 48, 64), dtype=torch.float32, device="cuda")
    assert accel.Images(j, outArray=imgs), accel.err
    a = np.array([[0.8, 0.1], [0.0, 0.5], [0.0, 0.0]])
    cam = host.CameraNoiseParams(1.0, 0.0, 2.0, 1.5, False)
    off = torch.full((48, 64), 100.0, dtype=torch.float32, device="cuda")
    frames = np.array([7, 8, 9], np.uint32)
    ph, _ = host.WidefieldImagesRender(imgs, True, 1, torch.from_numpy(a).cuda(), ctx=ctx)
    _, adu = host.WidefieldImagesRender(imgs, True, 1, torch.from_numpy(a).cuda(), cam=cam, offset=off, seed=3,
                                        frames=torch.from_numpy(frames.view(np.int32)).cuda(), ctx=ctx)
    torch.cuda.synchronize()
    ph = ph.cpu().numpy()
    adu = adu.cpu().numpy().view(np.uint16)
    want = orc.wf_images_render(imgs.cpu().numpy(), True, 1, a[0])
    assert np.array_equal(ph[0], want)
    oc = O.camera(quantumEfficiency=1.0, gainPhotonsPerAdu=2.0, readNoiseElectrons=1.5, emccd=0)
    for k in range(3):
        assert np.array_equal(adu[k], orc.noise(ph[k], oc, 3, int(frames[k]), offset=np.full((48, 64), 100.0,
                                                                                                    np.float32)))
