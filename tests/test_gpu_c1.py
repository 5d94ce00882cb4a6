"""The C1 SMLM movie as the benchmark renders it (host.SmlmMovieGpu: 128 x 128 px, 50 frames, ~200 blinks per frame,
the Gibson-Lanni + Zernike PSF at 129 x 129 (hw 16, os 4, M 128, 41 planes at 50 nm), the events bucketed and
collected on the device, Catmull-Rom splats whose camera radius 16 spans up to 3 x 3 splat tiles, bg 20, sCMOS):
the PSF within 1e-5 relative L2 of the FP64 oracle per plane, the device's frame emitters byte-identical to
CollectFrameEmitters' host mirror, and every frame's photons and ADU bit-exact against the oracle's RenderPhotonImage
+ ApplyNoiseChain on the device's own planes."""
import numpy as np
import pytest

import oracle as O
from insiliscope_amd import configs, host
from conftest import record_accuracy

pytestmark = pytest.mark.gpu


def test_c1_movie_end_to_end_vs_oracle(ctx, orc):
    import torch
    req = configs.c1_psf(seed=1, oversampling=4)
    ev, st = configs.c1_events(seed=1, n_frames=50)
    w = h = 128
    off = np.full((h, w), 100.0, np.float32)
    cam = configs.scmos_c1()
    print("C1 movie: %d blinks over 50 frames" % len(ev), flush=True)
    movie = host.SmlmMovieGpu(ev, 50, w, h, 100.0, 1.0, 20.0, offset=off, ctx=ctx)
    adu = torch.empty((50, h, w), dtype=torch.int16, device="cuda")
    ph = torch.empty((50, h, w), dtype=torch.float32, device="cuda")
    assert movie.Render(req, 4, 50.0, 2, cam, 1, outAdu=adu, outPhotons=ph), movie.splat.err
    planes = movie.planes[0].cpu().numpy()
    adu = adu.cpu().numpy().view(np.uint16)
    ph = ph.cpu().numpy()

    # the PSF: 41 sum-1 planes of 129^2 against the FP64 oracle
    want = orc.psf_planes(O.psf_request(**req), normalize=True)
    assert want.shape == planes.shape == (41, 129, 129)
    err = np.array([np.sqrt(((planes[k] - want[k]) ** 2).sum() / (want[k] ** 2).sum()) for k in range(41)])
    print("PSF max rel L2 per plane", err.max(), flush=True)
    record_accuracy("psf_c1_129x129_os4_m128_vs_fp64_oracle", rel_l2_max=err.max())
    assert err.max() <= 1e-5

    # the frames' emitters collected on the device = CollectFrameEmitters (host mirror, pinned to the reference)
    ems, start = movie.Emitters()
    cache = host.PsfKernelCache(planes, oversampling=4, zStepNm=50.0, interpMode=2)
    for f in range(50):
        host_em = host.CollectFrameEmitters(ev[st[f]:st[f + 1]], f, 100.0, 1.0, psfCache=cache)
        got = ems[start[f]:start[f + 1]]
        assert got.tobytes() == np.ascontiguousarray(host_em).tobytes(), f"frame {f}: emitters differ"

    # photons and ADU: the oracle's splat on the block sums of the device's planes, then its noise chain
    bs = np.stack([orc.block_sums(p, 4) for p in planes])
    ocam = O.camera(quantumEfficiency=1.0, gainPhotonsPerAdu=2.0, readNoiseElectrons=1.5, emccd=0)
    hits3 = 0
    for f in range(50):
        em = ems[start[f]:start[f + 1]]
        want_ph = orc.render_frame(w, h, 20.0, em, bs, 129, 4, 2)
        assert np.array_equal(ph[f], want_ph), f"frame {f}: {(ph[f] != want_ph).sum()} photon pixels differ"
        want_adu = orc.noise(want_ph, ocam, 1, f, offset=off)
        assert np.array_equal(adu[f], want_adu), f"frame {f}: {(adu[f] != want_adu).sum()} ADU differ"
        # blinks whose 33 x 33 patch, clipped to the frame, reaches three 16-px tiles along both axes
        x0 = np.floor(em["xPx"] + 0.5).astype(int)
        y0 = np.floor(em["yPx"] + 0.5).astype(int)
        span = lambda c: np.minimum(c + 16, 127) // 16 - np.maximum(c - 16, 0) // 16
        hits3 += int(((span(x0) == 2) & (span(y0) == 2)).sum())
    print("blinks spanning 3 x 3 tiles:", hits3, flush=True)
    assert hits3 > 0

    # the whole movie captured as one HIP graph (the benchmark's step): every replay bit-identical to the eager
    # render, with eager work between replays
    want_adu = torch.from_numpy(adu.view(np.int16)).cuda()
    out = torch.zeros((50, h, w), dtype=torch.int16, device="cuda")
    movie.Capture(req, 4, 50.0, 2, cam, 1, outAdu=out)
    for k in range(3):
        out.zero_()
        movie.Replay()
        torch.cuda.synchronize()
        assert torch.equal(out, want_adu), f"replay {k}"
        other = torch.empty_like(out)
        assert movie.Render(req, 4, 50.0, 2, cam, 1, outAdu=other), movie.splat.err
        torch.cuda.synchronize()
        assert torch.equal(other, want_adu)
