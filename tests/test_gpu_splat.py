"""The splat (+ fused noise) on the GPU through the C ABI, against the compiled reference's RenderPhotonImage frames
and the oracle: bit-exact photon images and ADU."""
import numpy as np
import pytest

import oracle as O
from insiliscope_amd import configs, host

pytestmark = pytest.mark.gpu


def fixture(golden):
    G, P = golden("splat_ref.npz"), golden("psf_ref.npz")
    ev = G["events"].view(host.BLINK_EVENT_DTYPE).reshape(-1)
    w, h = (int(x) for x in G["w_h"])
    return G, P, ev, w, h


def batch(G, frames):
    ems = [G[f"emitters_{f}"].view(host.FRAME_EMITTER_DTYPE) for f in frames]
    start = np.zeros(len(ems) + 1, np.uint32)
    start[1:] = np.cumsum([len(e) for e in ems])
    return np.concatenate(ems), start


@pytest.mark.parametrize("interp,name", [(0, "nearest"), (1, "linear"), (2, "cubic")])
def test_splat_photons_bit_exact_vs_reference(ctx, golden, interp, name):
    G, P, ev, w, h = fixture(golden)
    print("interp", name, flush=True)
    sp = host.SplatGpuHost(ctx)
    cache = host.PsfKernelCache(P["small_planes"], oversampling=4, zStepNm=50.0, interpMode=interp)
    assert sp.SetKernels([cache]), sp.err
    assert sp.SetStatic(w, h, flatBackground=20.0), sp.err
    em, start = batch(G, [0, 1, 2])
    ph = np.zeros((3, h, w), np.float32)
    assert sp.RenderBatch(em, start, None, None, 0, outPhotons=ph), sp.err
    for f in range(3):
        want = G[f"frame_{name}_{f}"]
        assert np.array_equal(ph[f], want), f"frame {f}: {(ph[f] != want).sum()} pixels differ"


def test_splat_halo_and_min_photons(ctx, golden):
    G, P, ev, w, h = fixture(golden)
    sp = host.SplatGpuHost(ctx)
    halo = host.PsfHaloSpans(G["halo_lo"], G["halo_hi"], G["halo_r2"], camRad=8, threshold=3e-4)
    cache = host.PsfKernelCache(P["small_planes"], oversampling=4, zStepNm=50.0, interpMode=2, halo=halo)
    assert sp.SetKernels([cache]) and sp.SetStatic(w, h, flatBackground=20.0), sp.err
    em, start = batch(G, [1])
    ph = np.zeros((1, h, w), np.float32)
    assert sp.RenderBatch(em, start, None, None, 0, outPhotons=ph), sp.err
    assert np.array_equal(ph[0], G["frame_cubic_halo_1"])
    cache2 = host.PsfKernelCache(P["small_planes"], oversampling=4, zStepNm=50.0, interpMode=2)
    assert sp.SetKernels([cache2]), sp.err
    assert sp.RenderBatch(em, start, None, None, 0, minPhotons=900.0, outPhotons=ph), sp.err
    assert np.array_equal(ph[0], G["frame_cubic_min_1"])


def test_splat_gpu_emitter_records(ctx, golden):
    """The reference's 64-byte wire format (float weights): within float rounding of the CPU frames."""
    G, P, ev, w, h = fixture(golden)
    sp = host.SplatGpuHost(ctx)
    cache = host.PsfKernelCache(P["small_planes"], oversampling=4, zStepNm=50.0, interpMode=2)
    assert sp.SetKernels([cache]) and sp.SetStatic(w, h, flatBackground=20.0), sp.err
    recs = [G[f"gpu_emitters_{f}"].view(host.SPLAT_EMITTER_DTYPE) for f in range(3)]
    start = np.zeros(4, np.uint32)
    start[1:] = np.cumsum([len(r) for r in recs])
    ph = np.zeros((3, h, w), np.float32)
    assert sp.RenderBatch(np.concatenate(recs), start, None, None, 0, outPhotons=ph, records=True), sp.err
    for f in range(3):
        want = G[f"frame_cubic_{f}"].astype(np.float64)
        err = np.sqrt(((ph[f] - want) ** 2).sum() / (want ** 2).sum())
        assert err < 1e-6


def test_splat_fused_noise_bit_exact(ctx, orc, golden):
    """Splat + EMCCD noise in one pass = the reference frames through the oracle's ApplyNoiseChain; also through
    RenderFrames (SplatGpuHost's per-frame lists)."""
    G, P, ev, w, h = fixture(golden)
    sp = host.SplatGpuHost(ctx)
    cache = host.PsfKernelCache(P["small_planes"], oversampling=4, zStepNm=50.0, interpMode=2)
    off = np.full((h, w), 100.0, np.float32)
    assert sp.SetKernels([cache]) and sp.SetStatic(w, h, offset=off, flatBackground=20.0), sp.err
    cam = configs.emccd_c4()
    ocam = O.camera(quantumEfficiency=1.0, gainPhotonsPerAdu=1.0 / 300.0, readNoiseElectrons=50.0, emccd=1,
                    emGain=300.0, cicElectrons=0.002)
    em, start = batch(G, [0, 1, 2])
    frames = np.array([10, 11, 12], np.uint32)
    adu = np.zeros((3, h, w), np.uint16)
    assert sp.RenderBatch(em, start, frames, cam, 1, outAdu=adu), sp.err
    for f in range(3):
        want = orc.noise(G[f"frame_cubic_{f}"], ocam, 1, int(frames[f]), offset=off)
        assert np.array_equal(adu[f], want)
    recs = [G[f"gpu_emitters_{f}"].view(host.SPLAT_EMITTER_DTYPE) for f in range(3)]
    outs = [np.zeros((h, w), np.uint16) for _ in range(3)]
    assert sp.RenderFrames(recs, frames, None, cam, 1, outs), sp.err
    assert all(o.any() for o in outs)


def test_splat_uneven_batch_c4_like_vs_oracle(ctx, orc):
    """512x512 frames with Poisson(800) blinks, one empty frame, C4's kernel geometry (hw 8, os 4) on a short stack
    at 15 nm, Catmull-Rom, EMCCD: every frame bit-exact against oracle render + noise."""
    req = configs.c4_psf(seed=1)
    req.update(nz=11, pupilSamples=64)
    planes = orc.psf_planes(O.psf_request(**req), normalize=True)
    cache = host.PsfKernelCache(planes, oversampling=4, zStepNm=15.0, interpMode=2)
    ev, st = configs.c4_events(seed=1, n_frames=3)
    ems = [host.CollectFrameEmitters(ev[st[f]:st[f + 1]], f, 100.0, 1.0, psfCache=cache) for f in range(3)]
    ems.insert(1, np.zeros(0, host.FRAME_EMITTER_DTYPE))
    start = np.zeros(len(ems) + 1, np.uint32)
    start[1:] = np.cumsum([len(e) for e in ems])
    w = h = 512
    off = np.full((h, w), 100.0, np.float32)
    sp = host.SplatGpuHost(ctx)
    assert sp.SetKernels([cache]) and sp.SetStatic(w, h, offset=off, flatBackground=50.0), sp.err
    frames = np.arange(len(ems), dtype=np.uint32)
    adu = np.zeros((len(ems), h, w), np.uint16)
    ph = np.zeros((len(ems), h, w), np.float32)
    cam = configs.emccd_c4()
    assert sp.RenderBatch(np.concatenate(ems), start, frames, cam, 1, outAdu=adu, outPhotons=ph), sp.err
    bs = np.stack([orc.block_sums(p, 4) for p in planes])
    ocam = O.camera(quantumEfficiency=1.0, gainPhotonsPerAdu=1.0 / 300.0, readNoiseElectrons=50.0, emccd=1,
                    emGain=300.0, cicElectrons=0.002)
    for f, em in enumerate(ems):
        want = orc.render_frame(w, h, 50.0, em, bs, planes.shape[1], 4, 2)
        assert np.array_equal(ph[f], want), f"frame {f}: {(ph[f] != want).sum()} photon pixels differ"
        want_adu = orc.noise(want, ocam, 1, f, offset=off)
        assert np.array_equal(adu[f], want_adu), f"frame {f}: {(adu[f] != want_adu).sum()} ADU differ"


def test_splat_extras_and_bg_scale(ctx, orc, golden):
    """Blinks added onto a caller's extra image replace the background; a background map with a per-frame scale
    follows RenderPhotonImage's float(map * scale)."""
    G, P, ev, w, h = fixture(golden)
    cache = host.PsfKernelCache(P["small_planes"], oversampling=4, zStepNm=50.0, interpMode=2)
    rng = np.random.default_rng(5)
    bg = (rng.random((h, w)) * 30).astype(np.float32)
    sp = host.SplatGpuHost(ctx)
    assert sp.SetKernels([cache]) and sp.SetStatic(w, h, background=bg, flatBackground=0.0), sp.err
    em, start = batch(G, [0, 1])
    extras = np.stack([np.zeros((h, w), np.float32), (rng.random((h, w)) * 7).astype(np.float32)])
    ph = np.zeros((2, h, w), np.float32)
    assert sp.RenderBatch(em, start, None, None, 0, bgScales=np.array([0.7, 1.0], np.float32), outPhotons=ph), sp.err
    bs = np.stack([orc.block_sums(p, 4) for p in P["small_planes"]])
    e0 = G["emitters_0"].view(O.EMITTER_DTYPE)
    want0 = orc.render_frame(w, h, 0.0, e0, bs, 65, 4, 2, bg_map=bg, bg_scale=float(np.float32(0.7)))
    assert np.array_equal(ph[0], want0)
    assert sp.RenderBatch(em, start, None, None, 0, extras=extras, outPhotons=ph), sp.err
    # onto extras[1]: the oracle frame with a background map equal to extras[1] and scale 1 starts from the same
    e1 = G["emitters_1"].view(O.EMITTER_DTYPE)
    want1 = orc.render_frame(w, h, 0.0, e1, bs, 65, 4, 2, bg_map=extras[1], bg_scale=1.0)
    assert np.array_equal(ph[1], want1)


def test_splat_fft_placement_vs_reference(ctx, golden):
    """The Fft placement (interp 3): one FP64 Fourier shift of the raw plane per emitter, the reference's radix-2
    operations; the phase cos/sin are the device's, so a pixel may differ by a float rounding where one of them
    differs by an ulp (none so far); required: at most 1 pixel in 10^4 and those within 1e-6 relative."""
    G, P, ev, w, h = fixture(golden)
    sp = host.SplatGpuHost(ctx)
    cache = host.PsfKernelCache(P["small_planes"], oversampling=4, zStepNm=50.0, interpMode=3)
    assert sp.SetKernels([cache]) and sp.SetStatic(w, h, flatBackground=20.0), sp.err
    em, start = batch(G, [0, 1, 2])
    ph = np.zeros((3, h, w), np.float32)
    assert sp.RenderBatch(em, start, None, None, 0, outPhotons=ph), sp.err
    for f in range(3):
        want = G[f"frame_fft_{f}"]
        diff = ph[f] != want
        print("frame", f, "pixels differing", int(diff.sum()), flush=True)
        assert diff.sum() <= max(1, want.size // 10000)
        assert np.allclose(ph[f], want, rtol=1e-6, atol=1e-6)
    halo = host.PsfHaloSpans(G["halo_fft_lo"], G["halo_fft_hi"], None, camRad=8, threshold=3e-4)
    cache_h = host.PsfKernelCache(P["small_planes"], oversampling=4, zStepNm=50.0, interpMode=3, halo=halo)
    assert sp.SetKernels([cache_h]), sp.err
    em, start = batch(G, [1])
    ph1 = np.zeros((1, h, w), np.float32)
    assert sp.RenderBatch(em, start, None, None, 0, outPhotons=ph1), sp.err
    assert np.allclose(ph1[0], G["frame_fft_halo_1"], rtol=1e-6, atol=1e-6)
    assert (ph1[0] != G["frame_fft_halo_1"]).sum() <= 1


def test_splat_fft_placement_many_emitters_vs_oracle(ctx, orc, golden):
    """200 emitters over 4 frames (noise fused too), against the FP64 restatement, and the records path refusing
    Fft stacks as the reference's GPU host does."""
    P = golden("psf_ref.npz")
    planes = P["small_planes"]
    w, h = 64, 56
    ev, st = configs.blink_events(11, 4, w, h, 100.0, 50.0, 200.0, 1500.0)
    cache = host.PsfKernelCache(planes, oversampling=4, zStepNm=50.0, interpMode=3)
    ems = [host.CollectFrameEmitters(ev[st[f]:st[f + 1]], f, 100.0, 1.0, psfCache=cache) for f in range(4)]
    start = np.zeros(5, np.uint32)
    start[1:] = np.cumsum([len(e) for e in ems])
    em = np.concatenate(ems)
    sp = host.SplatGpuHost(ctx)
    assert sp.SetKernels([cache]) and sp.SetStatic(w, h, flatBackground=5.0), sp.err
    ph = np.zeros((4, h, w), np.float32)
    assert sp.RenderBatch(em, start, None, None, 0, outPhotons=ph), sp.err
    bad = 0
    for f in range(4):
        want = orc.render_frame_fft(w, h, 5.0, ems[f].view(O.EMITTER_DTYPE), planes, 4)
        assert np.allclose(ph[f], want, rtol=1e-6, atol=1e-6)
        bad += int((ph[f] != want).sum())
    print("pixels differing from the FP64 restatement", bad, flush=True)
    assert bad <= max(1, ph.size // 10000)
    rec = host.CollectGpuEmitters(ems[0], host.PsfKernelCache(planes, oversampling=4, zStepNm=50.0, interpMode=2))
    assert not sp.RenderBatch(rec, np.array([0, len(rec)], np.uint32), np.zeros(1, np.uint32), None, 0,
                              outPhotons=ph[:1], records=True)
    assert "Fft" in sp.err


@pytest.mark.parametrize("interp", [0, 1, 2, 3])
def test_halo_spans_bit_exact_vs_reference(ctx, orc, interp):
    """WithHaloCut's spans from the GPU (block sums, then per-pixel window maxima) equal the compiled reference's
    lo, hi and radius2 at C4's kernel geometry (hw 8, os 4) and at os 1 and 3, over thresholds from a full to an
    empty cut; a new cut gets a new splat serial, the same cut again the same spans."""
    ref = O.Reference()
    req = configs.c4_psf(seed=1)
    req.update(nz=9, pupilSamples=64)
    for os_, hw in ((4, 8), (1, 12), (3, 5)):
        req.update(oversampling=os_, kernelHalfWidthPx=hw)
        planes = orc.psf_planes(O.psf_request(**req), normalize=True)
        cache = host.PsfKernelCache(planes, oversampling=os_, zStepNm=15.0, interpMode=interp)
        serials = set()
        for thr in (1e-7, 3e-4, 1e-3, 1e-2, 0.3):
            c = host.WithHaloCut(cache, thr, ctx=ctx)
            lo, hi, r2 = ref.halo_spans(planes, os_, interp, thr)
            assert np.array_equal(c.halo.lo, lo) and np.array_equal(c.halo.hi, hi), (os_, thr)
            assert np.array_equal(c.halo.radius2, r2), (os_, thr)
            assert c.serial == cache.serial and c.SplatSerial == c.halo.serial
            serials.add(c.halo.serial)
            assert host.WithHaloCut(cache, thr, ctx=ctx).halo is c.halo
        assert len(serials) == 5
    assert host.WithHaloCut(cache, 0.0, ctx=ctx).halo is None


def test_collect_frame_emitters_on_gpu_vs_reference():
    """CollectFrameEmitters for a batch of frames on the GPU (iscg_collect_frame_emitters) against the compiled
    reference's, frame by frame: events spanning frames, ending exactly on a frame edge, NaN and infinite times,
    brightness and excitation factors, per-frame drift, an illumination field (read at sites off the sensor too) and
    z beyond the stack (clamped); the records bit for bit, in event order, and the clamped / total z counts."""
    R = O.Reference()
    rng = np.random.default_rng(21)
    n, nF, w, h = 3000, 12, 90, 70
    ev = np.zeros(n, host.DEVICE_EVENT_DTYPE)
    ev["xUm"] = rng.uniform(-0.5, 9.5, n)
    ev["yUm"] = rng.uniform(-0.5, 7.5, n)
    ev["zNm"] = rng.uniform(-900.0, 900.0, n)
    ts = rng.uniform(-1.0, nF + 1.0, n)
    ev["tStart"] = ts
    ev["tEnd"] = ts + rng.exponential(0.8, n)
    ev["brightness"] = np.where(rng.random(n) < 0.3, 1.0, rng.lognormal(0.0, 0.4, n))
    ev["excite"] = np.where(rng.random(n) < 0.5, 1.0, rng.uniform(0.2, 1.5, n))
    ev["tStart"][:4] = [3.0, np.nan, 2.0, -np.inf]
    ev["tEnd"][:4] = [3.0, 5.0, np.inf, 1.5]
    illum = rng.uniform(0.5, 1.5, (h, w)).astype(np.float32)
    drift = rng.normal(0.0, 0.7, (nF, 2))
    frames = np.arange(nF, dtype=np.int64)
    bst, bidx = host.BucketEventsByFrame(ev[["xUm", "yUm", "zNm", "tStart", "tEnd", "brightness"]].astype(
        host.BLINK_EVENT_DTYPE), nF)
    cache = host.PsfKernelCache(np.zeros((41, 5, 5), np.float32), zStepNm=40.0)
    for il in (None, illum):
        recs, start, zc = host.CollectFrameEmittersBatch(ev, bst, bidx, frames, w, h, 110.0, 1500.0, drift=drift,
                                                         psfCache=cache, globalZOffsetUm=0.1, kernel=2,
                                                         illumination=il)
        zt = [0, 0]
        for f in range(nF):
            sub = ev[bidx[bst[f]:bst[f + 1]]]
            e7 = np.stack([sub[k] for k in host.DEVICE_EVENT_DTYPE.names], axis=1)
            want, (c, t) = R.collect_emitters_full(e7, f, w, h, 110.0, 1500.0, tuple(drift[f]), 41, 40.0, 0.1, il)
            zt[0] += c
            zt[1] += t
            got = recs[start[f]:start[f + 1]]
            assert len(got) == len(want), f
            for k in ("xPx", "yPx", "photons", "zIndex"):
                assert np.array_equal(got[k], want[k]), (f, k)
            assert (got["kernel"] == 2).all()
        assert zc == tuple(zt) and zt[0] > 0


def test_bucket_events_on_gpu_vs_reference():
    """BucketEventsByFrame on the GPU (iscg_bucket_events: stable radix sort of the (frame, event) pairs) against the
    compiled reference's lists: events spanning frames, before and after the movie, NaN and infinite times; then
    events -> buckets -> frame records entirely on the device, equal to the reference's CollectFrameEmitters."""
    R = O.Reference()
    rng = np.random.default_rng(13)
    n = 20000
    ev = np.zeros(n, host.BLINK_EVENT_DTYPE)
    ts = rng.uniform(-5.0, 300.0, n)
    ev["tStart"] = ts
    ev["tEnd"] = ts + rng.exponential(1.5, n)
    ev["xUm"] = rng.uniform(0.0, 5.0, n)
    ev["yUm"] = rng.uniform(0.0, 5.0, n)
    ev["brightness"] = 1.0
    special = [(-1e300, 2.5), (3.0, np.inf), (np.nan, 4.0), (2.0, np.nan), (-np.inf, 0.0), (49.999, 50.0),
               (50.0, 51.0), (0.0, 0.0), (-0.5, 1e-300), (7.0, 3.0), (-2.0, 1e30)]
    for i, (a, b) in enumerate(special):
        ev["tStart"][i], ev["tEnd"][i] = a, b
    for nF in (0, 1, 50, 297):
        start, idx = host.BucketEventsByFrameGpu(ev, nF)
        counts, want = R.bucket_events(ev, nF)
        assert np.array_equal(np.diff(start), counts) and np.array_equal(idx, want), nF
    nF = 297
    start, idx = host.BucketEventsByFrameGpu(ev, nF)
    recs, rst, _ = host.CollectFrameEmittersBatch(ev, start, idx, np.arange(nF, dtype=np.int64), 50, 50, 100.0,
                                                  800.0)
    for f in (0, 1, 150, 296):
        e7 = np.zeros((int(start[f + 1] - start[f]), 7))
        sub = ev[idx[start[f]:start[f + 1]]]
        for j, k in enumerate(("xUm", "yUm", "zNm", "tStart", "tEnd", "brightness")):
            e7[:, j] = sub[k]
        e7[:, 6] = 1.0
        want, _ = R.collect_emitters_full(e7, f, 50, 50, 100.0, 800.0)
        got = recs[rst[f]:rst[f + 1]]
        assert len(got) == len(want) and all(np.array_equal(got[k], want[k]) for k in ("xPx", "yPx", "photons")), f


def test_collect_frame_emitters_device_outputs_without_waiting():
    """The collect with device outputs and no z counts returns without waiting for the device; the records and
    offsets it leaves there equal those of the waiting (host-output) call."""
    import ctypes as C
    import torch
    from insiliscope_amd import host as H
    ctx = H.default_context()
    rng = np.random.default_rng(8)
    n, nF = 5000, 40
    ev = np.zeros(n, H.DEVICE_EVENT_DTYPE)
    ts = rng.uniform(0.0, nF, n)
    ev["tStart"], ev["tEnd"] = ts, ts + rng.exponential(0.7, n)
    ev["xUm"], ev["yUm"], ev["zNm"] = rng.uniform(0, 6, n), rng.uniform(0, 6, n), rng.uniform(-500, 500, n)
    ev["brightness"], ev["excite"] = 1.0, 1.0
    start, idx = H.BucketEventsByFrameGpu(ev, nF, ctx=ctx)
    cache = H.PsfKernelCache(np.zeros((21, 5, 5), np.float32), zStepNm=50.0)
    want, wst, _ = H.CollectFrameEmittersBatch(ev, start, idx, np.arange(nF, dtype=np.int64), 64, 64, 100.0, 900.0,
                                               psfCache=cache, ctx=ctx)
    d_ev = torch.from_numpy(ev.view(np.uint8).reshape(n, -1)).cuda()
    d_st = torch.from_numpy(start.astype(np.uint32).view(np.int32)).cuda()
    d_idx = torch.from_numpy(idx.view(np.int32)).cuda()
    d_fi = torch.arange(nF, dtype=torch.int64, device="cuda")
    d_out = torch.zeros((len(idx), H.FRAME_EMITTER_DTYPE.itemsize), dtype=torch.uint8, device="cuda")
    d_ost = torch.zeros(nF + 1, dtype=torch.int32, device="cuda")
    ctx.follow_torch()
    ctx.check(ctx.L.iscg_collect_frame_emitters(ctx.h, d_ev.data_ptr(), n, d_st.data_ptr(), d_idx.data_ptr(),
                                                d_fi.data_ptr(), nF, 64, 64, 100.0, 900.0, None, cache.nz, 50.0, 0.0,
                                                0, None, d_out.data_ptr(), d_out.shape[0], d_ost.data_ptr(), None))
    torch.cuda.synchronize()
    ost = d_ost.cpu().numpy().view(np.uint32)
    assert np.array_equal(ost, wst)
    got = d_out.cpu().numpy()[: ost[-1]].copy().view(H.FRAME_EMITTER_DTYPE).reshape(-1)
    assert got.tobytes() == want.tobytes()
