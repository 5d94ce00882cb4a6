"""Batches captured in HIP graphs (torch.cuda.CUDAGraph over the library's launches on the capturing stream): a
mean-field stack on the slab path with sCMOS noise, a splat + EMCCD batch of frames with uneven blink counts, and
the 512-PSF library.
Each graph is replayed several times, with the program's own eager calls and synchronisations between replays,
and every replay's output is compared bit for bit with the eager result. The calls are warmed up on the capture
stream first, so that nothing is allocated and the library does not switch streams while capturing."""
import numpy as np
import pytest

import oracle as O
from insiliscope_amd import configs, host

pytestmark = pytest.mark.gpu


def small_density(golden, seed=5, nx=48, ny=40, nzD=6):
    rng = np.random.default_rng(seed)
    d = np.where(rng.random((nzD, ny, nx)) < 0.05, 1.0 + np.floor(rng.random((nzD, ny, nx)) * 3), 0.0)
    return (40.0 * d).astype(np.float32)


def test_graph_wf_stack_slab_with_noise(ctx, golden):
    import torch
    P = golden("psf_ref.npz")["small_planes"]
    dens_np = small_density(golden)
    nzD, ny, nx = dens_np.shape
    s = torch.cuda.Stream()
    cam = host.CameraNoiseParams(quantumEfficiency=1.0, gainPhotonsPerAdu=2.0, readNoiseElectrons=1.5)
    with torch.cuda.stream(s):
        wf = host.WidefieldStack(ctx)
        assert wf.Plan(host.PsfKernelCache(P, oversampling=4, zStepNm=50.0), nx, ny, nzD, 0.05, 0, 0.025, 7,
                       fft=(64, 512, 128)), wf.err
        shp = wf.image_shape
        dens = torch.from_numpy(dens_np).cuda()
        off = torch.full(shp[1:], 100.0, dtype=torch.float32, device="cuda")
        adu = torch.zeros(shp, dtype=torch.int16, device="cuda")
        eager = torch.zeros(shp, dtype=torch.int16, device="cuda")

        def call(out):
            assert wf.FocusSeries(dens, cam=cam, offset=off, background=5.0, seed=7, frameBase=3, outAdu=out), wf.err

        call(eager)
        call(adu)
    torch.cuda.synchronize()
    want = eager.cpu().numpy()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=s):
        call(adu)
    for k in range(3):
        adu.zero_()
        g.replay()
        torch.cuda.synchronize()
        print("wf replay", k, flush=True)
        assert np.array_equal(adu.cpu().numpy(), want), f"replay {k}"
        with torch.cuda.stream(s):
            call(eager)  # eager work between replays
        torch.cuda.synchronize()
        assert np.array_equal(eager.cpu().numpy(), want)


def test_graph_splat_batch_with_noise(ctx, orc):
    import torch
    req = configs.c4_psf(seed=1)
    req.update(nz=5, pupilSamples=64)
    planes = orc.psf_planes(O.psf_request(**req), normalize=True)
    cache = host.PsfKernelCache(planes, oversampling=4, zStepNm=15.0, interpMode=2)
    w = h = 96
    nF = 4
    ev, st = configs.blink_events(3, nF, w, h, 100.0, 25.0, 25.0, 3000.0)
    ems = [host.CollectFrameEmitters(ev[st[f]:st[f + 1]], f, 100.0, 1.0, psfCache=cache) for f in range(nF)]
    start = np.zeros(nF + 1, np.uint32)
    start[1:] = np.cumsum([len(e) for e in ems])
    flat = np.concatenate(ems)
    s = torch.cuda.Stream()
    cam = configs.emccd_c4()
    with torch.cuda.stream(s):
        sp = host.SplatGpuHost(ctx)
        assert sp.SetKernels([cache]) and sp.SetStatic(w, h, offset=np.full((h, w), 100.0, np.float32),
                                                       flatBackground=50.0), sp.err
        d_em = torch.from_numpy(flat.view(np.uint8).reshape(len(flat), -1)).cuda()
        d_st = torch.from_numpy(start.view(np.int32)).cuda()
        d_fr = torch.arange(nF, dtype=torch.int32, device="cuda")
        adu = torch.zeros((nF, h, w), dtype=torch.int16, device="cuda")
        eager = torch.zeros_like(adu)

        def call(out):
            assert sp.RenderBatch(d_em, d_st, d_fr, cam, 11, outAdu=out), sp.err

        call(eager)
        call(adu)
    torch.cuda.synchronize()
    want = eager.cpu().numpy()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=s):
        call(adu)
    for k in range(3):
        adu.zero_()
        g.replay()
        torch.cuda.synchronize()
        print("splat replay", k, flush=True)
        assert np.array_equal(adu.cpu().numpy(), want), f"replay {k}"
        with torch.cuda.stream(s):
            call(eager)
        torch.cuda.synchronize()
        assert np.array_equal(eager.cpu().numpy(), want)


def test_graph_psf_library_c3(ctx):
    """The whole C3 library (8 wavelengths x 64 Zernike sets, 41 planes of 65x65, M 128) as one captured graph:
    replays equal the eager run and the one-shot batch call bit for bit; after Update with other aberrations, a
    replay computes the new library."""
    import torch
    reqs = configs.c3_requests(seed=1)
    reqs2 = configs.c3_requests(seed=2)
    pick = [0, 100, 511]
    want = host.ComputePsfPlanesBatch([reqs[i] for i in pick], normalize=True, ctx=ctx)
    want2 = host.ComputePsfPlanesBatch([reqs2[i] for i in pick], normalize=True, ctx=ctx)
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        lib = host.PsfLibrary(reqs, ctx)
        assert lib.shape == (512, 41, 65, 65)
        out = torch.zeros(lib.shape, dtype=torch.float32, device="cuda")
        eager = torch.zeros_like(out)
        lib.Run(eager, normalize=True)
        lib.Run(out, normalize=True)
    torch.cuda.synchronize()
    for j, i in enumerate(pick):
        assert np.array_equal(eager[i].cpu().numpy(), want[j]), i
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=s):
        lib.Run(out, normalize=True)
    for k in range(3):
        out.zero_()
        g.replay()
        torch.cuda.synchronize()
        print("psf replay", k, flush=True)
        assert torch.equal(out, eager), f"replay {k}"
        with torch.cuda.stream(s):
            lib.Run(eager, normalize=True)  # eager work between replays
        torch.cuda.synchronize()
    with torch.cuda.stream(s):
        lib.Update(reqs2)
    out.zero_()
    g.replay()
    torch.cuda.synchronize()
    for j, i in enumerate(pick):
        assert np.array_equal(out[i].cpu().numpy(), want2[j]), i
    lib.close()
