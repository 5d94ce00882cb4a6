"""Camera noise on the GPU (iscg_noise) against the compiled reference's ADU fixtures and the oracle: bit-exact."""
import numpy as np
import pytest

import oracle as O
from insiliscope_amd import host
from test_oracle import NOISE_CASES, noise_case_args

pytestmark = pytest.mark.gpu


def cam_of(c: O.Camera) -> host.CameraNoiseParams:
    return host.CameraNoiseParams(c.quantumEfficiency, c.darkCurrentElectrons, c.gainPhotonsPerAdu,
                                  c.readNoiseElectrons, bool(c.emccd), c.emGain, c.cicElectrons, c.bitDepth)


@pytest.mark.parametrize("case", NOISE_CASES, ids=[c[0] for c in NOISE_CASES])
def test_noise_bit_exact_vs_reference_fixture(ctx, golden, case):
    G = golden("noise_ref.npz")
    ph, cam, seed, frame, kw, base = noise_case_args(G, case)
    print("case", case[0], flush=True)
    h, w = ph.shape
    got = host.ApplyNoiseChain(ph, w, h, cam_of(cam), offsetMap=kw.get("offset"), gainMap=kw.get("gain"),
                               readNoiseMap=kw.get("read"), noiseSeed=seed, frame=frame, pixelBase=base, ctx=ctx)
    assert np.array_equal(got, G[case[0]])


def test_noise_batch_device_tensors_vs_oracle(ctx, orc):
    import torch
    rng = np.random.default_rng(11)
    n, h, w = 6, 128, 160
    ph = (rng.random((n, h, w)) * 120.0).astype(np.float32)
    ph[:, :4, :4] = 0.0
    frames = np.array([0, 1, 2, 17, 1000, 4294967295], np.uint32)
    off = np.full((h, w), 100.0, np.float32)
    for emccd in (False, True):
        cam = host.CameraNoiseParams(1.0, 0.0, 1.0 / 300.0 if emccd else 2.0, 50.0 if emccd else 1.5, emccd, 300.0,
                                     0.002, 16)
        print("emccd", emccd, flush=True)
        d_ph = torch.from_numpy(ph).cuda()
        d_fr = torch.from_numpy(frames.view(np.int32)).cuda()
        d_off = torch.from_numpy(off).cuda()
        out = host.ApplyNoiseChainFrames(d_ph, cam, 42, d_fr, offsetMap=d_off, ctx=ctx)
        assert out.is_cuda
        ctx.synchronize()
        got = out.cpu().numpy().view(np.uint16)
        c = O.camera(quantumEfficiency=1.0, darkCurrentElectrons=0.0, gainPhotonsPerAdu=cam.gainPhotonsPerAdu,
                     readNoiseElectrons=cam.readNoiseElectrons, emccd=int(emccd), emGain=300.0, cicElectrons=0.002,
                     bitDepth=16)
        for k in range(n):
            want = orc.noise(ph[k], c, 42, int(frames[k]), offset=off)
            assert np.array_equal(got[k], want), f"frame {k} emccd {emccd}: {(got[k] != want).sum()} mismatches"


def test_noise_failure_injection_leaves_caller_in_charge(ctx):
    """An injected device-side failure returns an error (the caller's CPU fallback path), then the next call works."""
    ph = np.full((8, 8), 10.0, np.float32)
    cam = host.CameraNoiseParams(quantumEfficiency=1.0)
    ctx.inject_failure(1, "test: device lost")
    with pytest.raises(host.IscgError, match="device lost"):
        host.ApplyNoiseChain(ph, 8, 8, cam, ctx=ctx)
    out = host.ApplyNoiseChain(ph, 8, 8, cam, ctx=ctx)
    assert out.shape == (8, 8)


def test_noise_fp32_path_verified_on_device(ctx):
    """The FP32 filter's error bounds, measured on the device over every input it can receive, leave the 2x margin."""
    b = ctx.noise_bounds()
    print(b, flush=True)
    assert b["fp32_path"]
    assert 0.0 < b["gauss_bound"] <= 5e-6
    assert 3.6e-6 < b["threshold_rel_bound"] <= 1e-5


def _both_paths(ctx, fn):
    ctx.set_noise_path(True)
    try:
        exact = fn()
    finally:
        ctx.set_noise_path(False)
    fast = fn()
    return exact, fast


@pytest.mark.parametrize("emccd", [False, True])
@pytest.mark.parametrize("read", [1.6, 300.0, 1e4])
def test_noise_fp32_path_equals_fp64_chain_full_size(ctx, emccd, read):
    """64 frames of 256 x 256 (the mean-field stack's size) through both paths: identical ADU. Photon levels span the
    Knuth branch, the normal approximation (l > 60), zero and the 16-bit clip; read = 1e4 puts a tenth of the pixels
    within the margin, so lanes of one wave take both paths."""
    import torch
    print("emccd", emccd, "read", read, flush=True)
    g = torch.Generator(device="cuda").manual_seed(5)
    n, h, w = 64, 256, 256
    lvl = torch.tensor([0.0, 0.3, 3.0, 20.0, 59.0, 61.0, 400.0, 5e4, 2e5], device="cuda")
    idx = torch.randint(0, len(lvl), (n, h, w), device="cuda", generator=g)
    ph = (lvl[idx] * (0.5 + torch.rand((n, h, w), device="cuda", generator=g))).contiguous()
    off = (100.0 + 5.0 * torch.rand((h, w), device="cuda", generator=g)).contiguous()
    gain = (1.8 + 0.4 * torch.rand((h, w), device="cuda", generator=g)).contiguous()
    frames = torch.arange(n, dtype=torch.int32, device="cuda") * 3 + 1
    cam = host.CameraNoiseParams(0.82, 0.05, 1.0 / 200.0 if emccd else 2.0, read, emccd, 200.0, 0.003, 16)
    exact, fast = _both_paths(
        ctx, lambda: host.ApplyNoiseChainFrames(ph, cam, 77, frames, offsetMap=off,
                                                gainMap=None if emccd else gain, ctx=ctx).cpu())
    assert torch.equal(exact, fast), f"{(exact != fast).sum().item()} of {exact.numel()} differ"


def test_noise_set_path_rejects_bad_value(ctx):
    with pytest.raises(host.IscgError, match="exact must be 0 or 1"):
        ctx.check(ctx.L.iscg_noise_set_path(ctx.h, 2))


def test_tiff_frames_from_device_memory(ctx, tmp_path):
    """A stack of ADU frames rendered into device memory goes to the movie file through iscg_tiff_frames_u16 (the
    copy from the device inside the library): the same bytes as the pages written from host memory."""
    import torch
    from insiliscope_amd import configs
    ph = np.random.default_rng(4).gamma(2.0, 30.0, (4, 33, 47)).astype(np.float32)
    cam = configs.scmos_c1()
    adu = torch.zeros((4, 33, 47), dtype=torch.int16, device="cuda")
    host.ApplyNoiseChainFrames(torch.from_numpy(ph).cuda(), cam, 7, np.arange(4, dtype=np.uint32), out=adu, ctx=ctx)
    torch.cuda.synchronize()
    a = host.TiffWriter(str(tmp_path / "dev.tif"))
    assert a.Frames(adu, "gpu frames") and a.Close(), a.error()
    b = host.TiffWriter(str(tmp_path / "host.tif"))
    assert b.Frames(adu.cpu().numpy().view(np.uint16), "gpu frames") and b.Close(), b.error()
    assert open(tmp_path / "dev.tif", "rb").read() == open(tmp_path / "host.tif", "rb").read()


def test_noise_setup_maps_feed_gpu_noise(ctx, orc, golden):
    """NoiseSetup's pixel maps (iscg_noise_setup: Pixel*Map::Generate, Sim/SMLMNoise.cpp:89-140, equal to the
    compiled reference's maps in maps_ref.npz) as the GPU noise's inputs (Sim/ScopeMovie.cpp:973-996 feeds them to
    ApplyNoiseChain): a whole sensor (sCMOS, offset + gain + read-noise maps) and a row window (EMCCD, pixelBase of
    the window), through iscg_noise and through the splat's fused noise, ADU bit-exact against the oracle's chain
    on the same maps."""
    G = golden("maps_ref.npz")
    rng = np.random.default_rng(23)
    cases = [
        (host.NoiseSetup.Make(1234567, 48, 40, 100.0, 2.5, 2.0, 0.05, 1.6, 0.2), G["maps_full"],
         host.CameraNoiseParams(1.0, 0.0, 2.0, 1.6, False, 300.0, 0.002, 16)),
        (host.NoiseSetup.Make(-77, 48, 12, 100.0, 2.5, 1.0 / 300.0, 3.0, 50.0, 1.5, sensorHeight=40, rowY0=17),
         G["maps_window"], host.CameraNoiseParams(1.0, 0.0, 1.0 / 300.0, 50.0, True, 300.0, 0.002, 16)),
    ]
    for s, want_maps, cam in cases:
        for got, want in zip((s.offsetMap, s.gainMap, s.readNoiseMap), want_maps):
            assert np.array_equal(got, want)
        h, w = s.offsetMap.shape
        print("maps", (h, w), "pixelBase", s.pixelBase, "emccd", cam.emccd, flush=True)
        ocam = O.camera(quantumEfficiency=cam.quantumEfficiency, gainPhotonsPerAdu=cam.gainPhotonsPerAdu,
                        readNoiseElectrons=cam.readNoiseElectrons, emccd=int(cam.emccd), emGain=cam.emGain,
                        cicElectrons=cam.cicElectrons)
        ph = (rng.gamma(0.8, 60.0, (3, h, w))).astype(np.float32)
        frames = np.array([0, 5, 9], np.uint32)
        adu = host.ApplyNoiseChainFrames(ph, cam, s.noiseSeed, frames, s.offsetMap, s.gainMap, s.readNoiseMap,
                                         pixelBase=s.pixelBase, ctx=ctx)
        for k, f in enumerate(frames):
            want = orc.noise(ph[k], ocam, s.noiseSeed, int(f), offset=s.offsetMap, gain=s.gainMap,
                             read=s.readNoiseMap, pixel_base=s.pixelBase)
            assert np.array_equal(adu[k], want), f"iscg_noise frame {f}: {(adu[k] != want).sum()} ADU differ"
        # the splat's fused noise with the same maps: an empty frame's flat background, noised
        sp = host.SplatGpuHost(ctx)
        assert sp.SetStatic(w, h, pixelBase=s.pixelBase, offset=s.offsetMap, gain=s.gainMap,
                            readNoise=s.readNoiseMap, flatBackground=33.0), sp.err
        out = np.zeros((1, h, w), np.uint16)
        assert sp.RenderBatch(np.zeros(0, host.FRAME_EMITTER_DTYPE), np.zeros(2, np.uint32),
                              np.array([4], np.uint32), cam, s.noiseSeed, outAdu=out), sp.err
        want = orc.noise(np.full((h, w), 33.0, np.float32), ocam, s.noiseSeed, 4, offset=s.offsetMap,
                         gain=s.gainMap, read=s.readNoiseMap, pixel_base=s.pixelBase)
        assert np.array_equal(out[0], want), f"fused noise: {(out[0] != want).sum()} ADU differ"
