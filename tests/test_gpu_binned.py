"""The binned blink regime on the GPU (iscg_binned_setup / iscg_binned_render behind host.BinnedBlinkRenderer, the
reference's BinnedBlinkRenderer, Sim/BinnedBlinks.cpp): the compiled reference's frames (kernel planes at upscale 1,
2, 4, a halo cut, the Gaussian; edge-case emitters; added to non-zero images) and the oracle's FP64 direct
convolution. FP32 transforms on a power-of-two grid against the reference's FastSize grid: relative L2 of the added
image <= 1e-5 per frame; frames nothing reaches are left bit for bit."""
import numpy as np
import pytest

from conftest import record_accuracy
from insiliscope_amd import host

pytestmark = pytest.mark.gpu

TOL = 1e-5
CASES = ("k1", "k2", "k4", "halo", "gauss")


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


def emitters(rows):
    e = np.zeros(len(rows), host.FRAME_EMITTER_DTYPE)
    if len(rows):
        e["xPx"], e["yPx"], e["photons"], e["zIndex"] = rows[:, 0], rows[:, 1], rows[:, 2], rows[:, 3].astype(int)
    return e


def setup(ctx, G, P, name):
    u_ref, R_ref, up = (int(v) for v in G[name + "_uR"])
    r = host.BinnedBlinkRenderer(ctx)
    if name == "gauss":
        kernel = None
    else:
        kernel = host.PsfKernelCache(P, oversampling=4, zStepNm=50.0)
        r2 = int(G[name + "_haloR2"])
        if r2 >= 0:
            kernel.halo = host.PsfHaloSpans(lo=np.zeros((9, 1), np.int16), hi=np.zeros((9, 1), np.int16),
                                            radius2=np.array([r2], np.int32), camRad=0)
    assert r.Setup(kernel, float(G[name + "_sigma"]), int(G["W"]), int(G["H"]), up), r.err
    assert (r.Upscale(), r.RadiusCells()) == (u_ref, R_ref)
    return r


@pytest.mark.parametrize("name", CASES)
def test_binned_vs_reference(ctx, golden, name):
    G = golden("binned_ref.npz")
    P = golden("psf_ref.npz")["small_planes"]
    print("case", name, flush=True)
    r = setup(ctx, G, P, name)
    em, st = G[name + "_em"], G[name + "_start"]
    frames = [emitters(em[st[f]:st[f + 1]]) for f in range(len(st) - 1)]
    img = G[name + "_img0"].copy()
    assert r.RenderFrames(frames, img), r.err
    errs = [0.0]
    for f in range(len(frames)):
        if not len(frames[f]):
            assert np.array_equal(img[f], G[name + "_img0"][f])
            continue
        e = rel_l2(img[f] - G[name + "_img0"][f], G[name + "_img"][f] - G[name + "_img0"][f])
        errs.append(e)
        assert e <= TOL, (name, f, e)
    record_accuracy("binned_" + name + "_vs_compiled_reference", rel_l2_max=max(errs))


def test_binned_single_frame_render_alone_and_device_images(ctx, golden, orc):
    """Render (one frame, with the ground-truth image `alone`) and RenderFrames into a device tensor agree with the
    oracle's direct convolution."""
    import torch
    G = golden("binned_ref.npz")
    P = golden("psf_ref.npz")["small_planes"]
    r = setup(ctx, G, P, "k2")
    W, H = int(G["W"]), int(G["H"])
    em, st = G["k2_em"], G["k2_start"]
    rows = em[st[0]:st[1]]
    img = np.full((H, W), 3.0, np.float32)
    alone = np.zeros((H, W), np.float32)
    assert r.Render(emitters(rows), img, alone), r.err
    u, R = r.Upscale(), r.RadiusCells()
    K = orc.binned_kernels(P, 4, 2, R)
    want = orc.binned_render(K, u, R, W, H, rows)
    assert rel_l2(alone, want) <= TOL
    assert rel_l2(img - 3.0, want) <= TOL
    dev = torch.zeros((2, H, W), dtype=torch.float32, device="cuda")
    assert r.RenderFrames([emitters(rows), emitters(rows[:7])], dev), r.err
    torch.cuda.synchronize()
    d = dev.cpu().numpy()
    assert rel_l2(d[0], want) <= TOL
    assert rel_l2(d[1], orc.binned_render(K, u, R, W, H, rows[:7])) <= TOL


def test_binned_many_planes_large_frames_in_chunks(ctx, golden, orc):
    """512 x 512 frames (a 1024 x 1024 grid) with 800 emitters over 9 planes each: a 40-frame batch spans several
    device chunks; sampled frames against the oracle, and the batch equal to frame-by-frame renders."""
    P = golden("psf_ref.npz")["small_planes"]
    W = H = 512
    r = host.BinnedBlinkRenderer(ctx)
    assert r.Setup(host.PsfKernelCache(P, oversampling=4, zStepNm=50.0), 0.0, W, H, 1), r.err
    u, R = r.Upscale(), r.RadiusCells()
    rng = np.random.default_rng(21)
    frames = []
    for f in range(40):
        n = 800
        rows = np.stack([rng.uniform(-4, W + 4, n), rng.uniform(-4, H + 4, n), rng.uniform(100, 5000, n),
                         rng.integers(0, 9, n).astype(np.float64)], 1)
        frames.append(rows)
    imgs = np.zeros((40, H, W), np.float32)
    print("batch of 40 frames", flush=True)
    assert r.RenderFrames([emitters(x) for x in frames], imgs), r.err
    K = orc.binned_kernels(P, 4, 1, R)
    for f in (0, 17, 39):
        want = orc.binned_render(K, u, R, W, H, frames[f])
        assert rel_l2(imgs[f], want) <= TOL, f
    one = np.zeros((H, W), np.float32)
    assert r.Render(emitters(frames[17]), one), r.err
    assert np.array_equal(one, imgs[17])


def test_binned_errors(ctx):
    r = host.BinnedBlinkRenderer(ctx)
    img = np.zeros((1, 16, 16), np.float32)
    assert not r.Setup(None, 0.0, 64, 64, 1)          # neither a kernel nor a Gaussian
    assert not r.RenderFrames([np.zeros(0, host.FRAME_EMITTER_DTYPE)], img)  # no valid setup
    assert "setup" in r.err
    assert not r.Setup(None, 2.0, 4000, 64, 1)        # beyond a 1024-point grid
    assert r.Setup(None, 1.0, 16, 16, 1), r.err
    assert r.RenderFrames([np.zeros(0, host.FRAME_EMITTER_DTYPE)], img), r.err
    assert not img.any()
