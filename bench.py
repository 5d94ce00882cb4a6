"""Throughput of the image-formation path on one MI355X, against the reference CPU path.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload wf|smlm|psf|dyes|adjoint|c1] [--no-cpu]
                    [--no-secondary]

Default workload (`wf`): the widefield mean field of a 256 x 256 x 64 dye density (100 x 100 x 50 nm voxels, seeded)
convolved with a 127-plane Gibson-Lanni PSF (257 x 257 at 100 nm, 50 nm apart; NA 1.45, ni 1.518, ns 1.33, 670 nm,
Noll 5-15 +-0.05 waves) by a zero-padded 512 x 512 x 128 FP32 R2C/C2R 3D FFT, sampled at the 64 dye-plane foci, then
sCMOS noise (read 1.5 e-, gain 2 e-/ADU, offset 100) -> 64 uint16 frames. A step is one such focus stack; the metric
is frames (focus images) per second with the density already in HBM. By default 1000 stacks are timed after 100
warm-up stacks (about 0.15 s of GPU work), the sustained rate: the first few milliseconds of work on an idle GPU run
slower (20 stacks after 3 read about 7 % lower).

Secondary lines in the same JSON: the SMLM DNA-PAINT movie (512 x 512, Poisson(800) blinks/frame, Catmull-Rom splat
from a 65 x 65 x 81 table, EMCCD noise) in frames/s, and the PSF library (512 PSFs x 41 planes of 65 x 65, M 128) in
PSFs/s.

roofline: the dominant kernel's algorithmic bytes per launch / its mean launch time (HIP events on the library's
stream during the timed steps) against 8 TB/s HBM. cpu_baseline: the reference's own CPU code (compiled from its
sources into oracle/_ref) on a bounded sample, all host cores.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "frames/s (SMLM 512² & WF 256²×64z) + HBM GB/s; fwd+adjoint at ≤1e-5 rel-L2"
HBM_PEAK_GBS = 8000.0
# host threads for the CPU baseline: the host's CPU share per GPU (16 on an MI355X node), or fewer where it has fewer
CORES = max(1, min(16, int(os.environ.get("OMP_NUM_THREADS", "16")), os.cpu_count() or 1))
MFMA_FP16_PEAK_TFS = 2500.0
FP32_PEAK_TFS = 157.3
# bounded CPU samples: the headline baseline repeats its sample for at least this many seconds, the secondary ones
# for at least CPU_MIN_S_SECONDARY (so the default run still ends within a few minutes)
CPU_MIN_S = 10.0
CPU_MIN_S_SECONDARY = 3.0


def timed_repeat(fn, min_s):
    """Runs fn() until at least min_s seconds have passed (at least once); returns (repetitions, seconds)."""
    reps, t0 = 0, time.time()
    while True:
        fn()
        reps += 1
        el = time.time() - t0
        if el >= min_s:
            return reps, el


def log(*a):
    print(*a, file=sys.stderr, flush=True)


class Timer:
    """HIP events on the library context's stream."""

    def __init__(self, ctx):
        import torch
        self.torch = torch
        self.ctx = ctx

    def __enter__(self):
        # the wrappers enqueue on torch's current stream (Context.follow_torch): time that stream
        self.ctx.follow_torch()
        self.stream = self.torch.cuda.current_stream(self.ctx.device)
        self.ctx.synchronize()
        self.a = self.torch.cuda.Event(enable_timing=True)
        self.b = self.torch.cuda.Event(enable_timing=True)
        self.a.record(self.stream)
        return self

    def __exit__(self, *exc):
        self.b.record(self.stream)
        self.b.synchronize()
        self.ms = self.a.elapsed_time(self.b)


def profiled(ctx, step, n):
    """Per-kernel (launches, total ms) over n more steps with the library's event pairs around each launch (kept out
    of the timed steps: the events add their own gaps)."""
    L = ctx.L
    L.iscg_profile_enable(ctx.h, 1)
    profile(ctx, L)
    for k in range(n):
        step(k)
    prof = profile(ctx, L)
    L.iscg_profile_enable(ctx.h, 0)
    return prof


def profile(ctx, L):
    import ctypes as C
    cap = 64
    names = C.create_string_buffer(64 * cap)
    launches = (C.c_int64 * cap)()
    total = (C.c_double * cap)()
    n = L.iscg_profile_read(ctx.h, names, launches, total, cap, 1)
    out = {}
    for i in range(min(n, cap)):
        nm = names.raw[64 * i:64 * (i + 1)].split(b"\0")[0].decode()
        out[nm] = (int(launches[i]), float(total[i]))
    return out


def traffic_for(kernel):
    p = os.path.join(ROOT, "profiles", f"{kernel}_traffic.json")
    if os.path.exists(p):
        with open(p) as f:
            return json.load(f).get("bytes_per_launch")
    return None


# ---- widefield mean field (C2) -----------------------------------------------------------------------------

def wf_psf_request():
    from insiliscope_amd import configs
    return configs.psf_request_fields(wavelengthNm=670.0, na=1.45, immersionIndex=1.518, sampleIndex=1.33,
                                      workingDistanceUm=170.0, pixelSizeNm=100.0, zStepNm=50.0, oversampling=1,
                                      kernelHalfWidthPx=128, nz=127, pupilSamples=256,
                                      zernike=configs.zernike_noll(1, range(5, 16), amplitude=0.05))


def wf_bytes(nx, ny, nzD, nF, NX, NY, NZ):
    """Algorithmic HBM bytes per launch of each pass: each pass reads its input and writes its output once (half
    spectra of NX/2 + 1 complex64 columns; zero padding never read). wf_slab (the fused y / z / inverse-y pass of
    the 512 x 512 x 128 grid) reads the x-transformed dye rows and the kernel spectrum and writes the FOV rows; the
    five-pass names are those of the other grids."""
    h = NX // 2 + 1
    c = 8
    return {
        "wf_rows_r2c": nzD * ny * nx * 4 + nzD * ny * h * c,
        "wf_cols_fwd": nzD * ny * h * c + nzD * NY * h * c,
        "wf_zconv": nzD * NY * h * c + NZ * NY * h * c + nF * NY * h * c,
        "wf_cols_inv": nF * NY * h * c + nF * ny * h * c,
        "wf_slab": nzD * ny * h * c + NZ * NY * h * c + nF * ny * h * c,
        "wf_rows_c2r_noise": nF * ny * h * c + nF * ny * nx * 2,
    }


def run_wf(ctx, steps, warmup):
    import torch
    from insiliscope_amd import configs, host
    L = ctx.L
    nx = ny = 256
    nzD = nF = 64
    t0 = time.time()
    psf = host.ComputePsfPlanesBatch([wf_psf_request()], normalize=True,
                                     out=torch.empty((1, 127, 257, 257), dtype=torch.float32, device="cuda"),
                                     ctx=ctx)[0]
    ctx.synchronize()
    cache = host.PsfKernelCache(psf, oversampling=1, zStepNm=50.0)
    wf = host.WidefieldStack(ctx)
    if not wf.Plan(cache, nx, ny, nzD, 0.05, 0, 0.025, nF, fft=(512, 512, 128)):
        raise RuntimeError(wf.err)
    dens = torch.from_numpy(configs.c2_density(seed=1)).cuda()
    off = torch.full((ny, nx), 100.0, dtype=torch.float32, device="cuda")
    adu = torch.empty((nF, ny, nx), dtype=torch.int16, device="cuda")
    cam = configs.scmos_c1()
    ctx.synchronize()
    log(f"wf setup {time.time() - t0:.1f} s (PSF 127 x 257^2 on the GPU, plan)")

    def step(k):
        if not wf.FocusSeries(dens, cam=cam, offset=off, seed=1, frameBase=k * nF, outAdu=adu):
            raise RuntimeError(wf.err)

    for k in range(warmup):
        step(k)
    with Timer(ctx) as t:
        for k in range(steps):
            step(warmup + k)
    ms = t.ms / steps
    prof = profiled(ctx, lambda k: step(warmup + steps + k), min(steps, 10))
    nbytes = wf_bytes(nx, ny, nzD, nF, 512, 512, 128)
    dom = max(prof, key=lambda k: prof[k][1])
    avg_ms = prof[dom][1] / prof[dom][0]
    achieved = nbytes[dom] / (avg_ms * 1e-3) / 1e9
    total_bytes = sum(nbytes[k] for k in prof if k in nbytes)
    kernels = {k: {"launches": v[0], "avg_ms": round(v[1] / v[0], 5),
                   "GB_s": round(nbytes.get(k, 0) / (v[1] / v[0] * 1e-3) / 1e9, 1)} for k, v in prof.items()}
    traffic = traffic_for(dom)
    return {
        "value": nF * 1000.0 / ms, "ms_per_step": ms, "dominant": dom, "_psf": psf,
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "kernel": dom, "algorithmic_bytes": nbytes[dom]},
        "stack_GB_s": round(total_bytes / (ms * 1e-3) / 1e9, 1),
        "kernels": kernels,
    }


def cpu_wf(planes):
    """The reference's WidefieldScene (UpdateFromGrid + FocusSeries: per dye plane a 2D real FFT, per focus the
    spectra re-paired with the PSF plane spectra and one inverse FFT) on the same density and PSF planes, timed on
    the host cores (ParallelFor, ISC_THREADS = all cores): one whole 64-focus stack after a warm-up stack."""
    import oracle as O
    from insiliscope_amd import configs
    ref = O.load_reference()
    if ref is None:
        return None
    dens = configs.c2_density(seed=1)
    foci = (np.arange(64) + 0.5) * 0.05
    cam = configs.scmos_c1()
    oc = O.camera(quantumEfficiency=1.0, gainPhotonsPerAdu=cam.gainPhotonsPerAdu,
                  readNoiseElectrons=cam.readNoiseElectrons, emccd=0)
    off = np.full((256, 256), 100.0, np.float32)

    def stack():
        # the same work as the GPU step: the 64 focus images, then ApplyNoiseChain (sCMOS, offset map) on each,
        # its rows in parallel as the reference runs it (ScopeMovie.cpp:7642-7650)
        imgs = ref.wf_focus_series(dens, 50.0, planes, 1, 50.0, foci, kernel_cap_um=12.8)
        for f in range(len(foci)):
            ref.noise(imgs[f], oc, 1, f, offset=off, parallel=True)

    ref.wf_focus_series(dens, 50.0, planes, 1, 50.0, foci[:2], kernel_cap_um=12.8)
    reps, el = timed_repeat(stack, CPU_MIN_S)
    return {"value": round(reps * len(foci) / el, 2), "unit": "frames/s", "cores": CORES, "kind": "reference",
            "sample": f"WidefieldScene::UpdateFromGrid + FocusSeries of the 256x256x64 density at all 64 foci, then "
                      f"ApplyNoiseChain (sCMOS, offset map, rows in parallel) on each image, {reps} stacks in "
                      f"{el:.2f} s (kernel radius 128 cells)"}


# ---- SMLM DNA-PAINT (C4) ---------------------------------------------------------------------------------------

# VALU wave instructions of one splat_tiles_kernel<32> launch over C4's 2000 frames (SQ_INSTS_VALU per dispatch,
# rocprofv3 --pmc on bench.py --workload smlm), and the VALU issue peak in G wave-instructions/s
SPLAT_C4_VALU_INSTS = 14_033_300_000  # mean of 21 dispatches (they vary by about 2e6)
VALU_ISSUE_PEAK_G = 256 * 2.4


def run_smlm(ctx, steps, warmup, n_frames=2000):
    import torch
    from insiliscope_amd import configs, host
    L = ctx.L
    t0 = time.time()
    planes = host.ComputePsfPlanesBatch([configs.c4_psf(seed=1)], normalize=True,
                                        out=torch.empty((1, 81, 65, 65), dtype=torch.float32, device="cuda"),
                                        ctx=ctx)[0]
    cache = host.PsfKernelCache(planes, oversampling=4, zStepNm=15.0, interpMode=2)
    ev, st = configs.c4_events(seed=1, n_frames=n_frames)
    w = h = 512
    off = torch.full((h, w), 100.0, dtype=torch.float32, device="cuda")
    # the blink events resident in HBM and bucketed by frame once; each step collects every frame's emitters on the
    # device (as the reference's RenderPhotonImage does per frame), then splats and noises the 2000 frames
    tc = time.time()
    movie = host.SmlmMovieGpu(ev, n_frames, w, h, 100.0, 1.0, 50.0, offset=off, ctx=ctx)
    bucket_s = time.time() - tc
    adu = torch.empty((n_frames, h, w), dtype=torch.int16, device="cuda")
    cam = configs.emccd_c4()
    ctx.synchronize()
    log(f"smlm setup {time.time() - t0:.1f} s ({len(ev)} blinks over {n_frames} frames; events uploaded and "
        f"bucketed on the GPU in {bucket_s * 1e3:.1f} ms)")

    def step():
        if not movie.RenderFrames(cache, cam, 1, outAdu=adu):
            raise RuntimeError(movie.splat.err)

    for _ in range(warmup):
        step()
    with Timer(ctx) as t:
        for _ in range(steps):
            step()
    ms = t.ms / steps
    prof = profiled(ctx, lambda k: step(), min(steps, 3))
    # algorithmic bytes of the splat+noise pass: ADU out, emitter records and setup in, offset map
    tiles = prof.get("splat_tiles_noise", (1, ms))
    b = n_frames * w * h * 2 + len(ev) * (112 + 8) + w * h * 4
    t_ms = tiles[1] / tiles[0]
    return {"value": n_frames * 1000.0 / ms, "unit": "frames/s", "ms_per_step": ms,
            "config": {"workload": "SMLM DNA-PAINT 512x512 x 2000 frames, Poisson(800) blinks/frame, Catmull-Rom "
                                   "splat from 65x65x81 (os 4, 15 nm), EMCCD gain 300 read 50 e-",
                       "blinks": int(len(ev))},
            "roofline": {"bound": "hbm", "achieved": round(b / (t_ms * 1e-3) / 1e9, 1), "peak": HBM_PEAK_GBS,
                         "unit": "GB/s", "frac": round(b / (t_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4), "traffic": None,
                         "kernel": "splat_tiles_noise", "algorithmic_bytes": b,
                         "note": "bound by the noise chain's VALU work (pcg4d counter draws, the Knuth Poisson "
                                 "loop and the EMCCD Gamma per pixel), not by HBM"},
            "splat_noise_kernel": {"avg_ms": round(t_ms, 4), "GB_s": round(b / (t_ms * 1e-3) / 1e9, 1),
                                   "bound": "fp64 valu (counter RNG + Knuth/Gamma noise)"},
            # the kernel's real roof: VALU issue. Work = the VALU wave instructions one launch over these 2000
            # frames issues (SQ_INSTS_VALU, rocprofv3 --pmc; the inputs are seeded, so the count is fixed), over the
            # launch time measured here; peak = one wave instruction per CU per cycle (4 SIMDs, 4 cycles per wave64
            # instruction) x 256 CUs x 2.4 GHz
            "valu_roofline": {"bound": "valu", "instructions": SPLAT_C4_VALU_INSTS,
                              "achieved": round(SPLAT_C4_VALU_INSTS / (t_ms * 1e-3) / 1e9, 1),
                              "peak": VALU_ISSUE_PEAK_G, "unit": "G wave-instructions/s",
                              "frac": round(SPLAT_C4_VALU_INSTS / (t_ms * 1e-3) / 1e9 / VALU_ISSUE_PEAK_G, 4),
                              "lane_instructions_per_pixel": round(SPLAT_C4_VALU_INSTS * 64 / (n_frames * w * h),
                                                                   1)},
            "_events": (ev, st, np.asarray(planes.cpu()))}


def run_binned(ctx, ev, st, planes, n_frames=64):
    """The binned blink regime (BinnedBlinkRenderer) on C4's frames: the same emitters as the splat, deposited per
    z plane on the 512 + 2R grid (u = 1) and FFT-convolved with the 81 plane kernels (1024 x 1024 transforms)."""
    import torch
    from insiliscope_amd import host
    cache = host.PsfKernelCache(planes, oversampling=4, zStepNm=15.0, interpMode=2)
    frames = [host.CollectFrameEmitters(ev[st[f]:st[f + 1]], f, 100.0, 1.0, psfCache=cache) for f in range(n_frames)]
    r = host.BinnedBlinkRenderer(ctx)
    if not r.Setup(cache, 0.0, 512, 512, 1):
        raise RuntimeError(r.err)
    imgs = torch.zeros((n_frames, 512, 512), dtype=torch.float32, device="cuda")

    def step():
        if not r.RenderFrames(frames, imgs):
            raise RuntimeError(r.err)

    step()
    with Timer(ctx) as t:
        step()
        step()
    ms = t.ms / 2
    planes_per_frame = float(np.mean([len(np.unique(np.clip(f["zIndex"], 0, 80))) for f in frames]))
    return {"value": n_frames * 1000.0 / ms, "unit": "frames/s", "ms_per_step": ms,
            "config": {"workload": f"binned blinks: C4's first {n_frames} frames (512x512, u 1, R {r.RadiusCells()}, "
                                   f"81 kernel planes, {planes_per_frame:.1f} occupied per frame), 1024x1024 FFT grid"},
            "_frames": [np.stack([f["xPx"], f["yPx"], f["photons"], f["zIndex"].astype(np.float64)], 1)
                        for f in frames[:4]]}


def cpu_binned(frames, planes):
    import oracle as O
    ref = O.load_reference()
    if ref is None:
        return None
    ref.binned_render(planes, 4, 1, 512, 512, frames[:1])
    reps, el = timed_repeat(lambda: ref.binned_render(planes, 4, 1, 512, 512, frames), CPU_MIN_S_SECONDARY)
    return {"value": round(reps * len(frames) / el, 2), "unit": "frames/s", "cores": CORES, "kind": "reference",
            "sample": f"BinnedBlinkRenderer::Render of {len(frames)} frames, {reps}x in {el:.2f} s (FFT lines on all "
                      f"cores)"}


def cpu_smlm(ev, st, planes, frames=96):
    import oracle as O
    from insiliscope_amd import configs
    ref = O.load_reference()
    if ref is None:
        return None
    cam = configs.emccd_c4()
    oc = O.camera(quantumEfficiency=1.0, gainPhotonsPerAdu=cam.gainPhotonsPerAdu, readNoiseElectrons=50.0,
                  emccd=1, emGain=300.0, cicElectrons=0.002)
    off = np.full((512, 512), 100.0, np.float32)
    sub_ev = ev[: st[frames]]
    reps, el = timed_repeat(lambda: ref.render_noise_frames(sub_ev, st[: frames + 1], 512, 512, 100.0, 50.0, planes,
                                                            4, 15.0, 2, oc, off, 1), CPU_MIN_S_SECONDARY)
    return {"value": round(reps * frames / el, 2), "unit": "frames/s", "cores": CORES, "kind": "reference",
            "sample": f"RenderPhotonImage (Cubic) + ApplyNoiseChain (EMCCD), {frames} of the 2000 frames, {reps}x in "
                      f"{el:.2f} s, frames in parallel as FluorescenceMovie batches them"}


# ---- PSF library (C3) -------------------------------------------------------------------------------------------

def run_psf(ctx, steps, warmup):
    import torch
    from insiliscope_amd import configs, host
    L = ctx.L
    reqs = host.PsfRequestArray(configs.c3_requests(seed=1))  # the requests as the C structs the seam takes
    out = torch.empty((len(reqs), 41, 65, 65), dtype=torch.float32, device="cuda")

    def step():
        host.ComputePsfPlanesBatch(reqs, out=out, ctx=ctx)

    # the one-shot call (requests uploaded by each call)
    for _ in range(warmup):
        step()
    with Timer(ctx) as t:
        for _ in range(steps):
            step()
    ms_call = t.ms / steps
    # the library held on the device, its run captured once into a HIP graph: a step is one graph launch
    side = torch.cuda.Stream()
    with torch.cuda.stream(side):
        lib = host.PsfLibrary(reqs, ctx)
        for _ in range(max(1, warmup)):
            lib.Run(out)
    torch.cuda.synchronize()
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph, stream=side):
        lib.Run(out)
    with torch.cuda.stream(side):
        graph.replay()
        with Timer(ctx) as t:
            for _ in range(steps):
                graph.replay()
        ms = t.ms / steps
        prof = profiled(ctx, lambda k: lib.Run(out), min(steps, 3))
    torch.cuda.synchronize()
    ctx.set_stream(None)
    lib.close()
    planes = len(reqs) * 41
    flops = planes * 8.0 * 65 * 128 * (128 + 65)
    k = prof.get("psf_planes_mfma", (1, ms))
    kms = k[1] / k[0]
    tf = flops / (kms * 1e-3) / 1e12
    return {"value": len(reqs) * 1000.0 / ms, "unit": "PSFs/s", "ms_per_step": ms,
            "planes_per_s": planes * 1000.0 / ms, "ms_per_step_one_shot_call": round(ms_call, 4),
            "config": {"workload": "PSF library: 8 wavelengths x 64 Zernike sets (Noll 4-21, sigma 0.04), 41 planes "
                                   "of 65x65, M 128; a step = one replay of the library's captured HIP graph"},
            "roofline": {"bound": "mfma", "achieved": round(3 * tf, 1), "peak": MFMA_FP16_PEAK_TFS,
                         "unit": "TFLOP/s", "frac": round(3 * tf / MFMA_FP16_PEAK_TFS, 4), "traffic": None,
                         "kernel": "psf_planes_mfma",
                         "note": "FP16 flops of the split products (3 per algorithmic complex-FP32 product, P 65, "
                                 "M 128); dense FP16 peak"},
            "mfma_kernel": {"avg_ms": round(kms, 4), "TFLOP_s_complex_fp32_equiv": round(tf, 1),
                            "TFLOP_s_executed_fp16": round(3 * tf, 1),
                            "frac_of_fp16_dense_peak": round(3 * tf / MFMA_FP16_PEAK_TFS, 4)},
            "kernels": {k: {"launches": v[0], "avg_ms": round(v[1] / v[0], 4)} for k, v in prof.items()}}


def cpu_psf(n=64):
    import oracle as O
    from insiliscope_amd import configs
    ref = O.load_reference()
    if ref is None:
        return None
    reqs = [O.psf_request(**r) for r in configs.c3_requests(seed=1)[:n]]
    reps, el = timed_repeat(lambda: ref.psf_library(reqs), CPU_MIN_S_SECONDARY)
    return {"value": round(reps * n / el, 2), "unit": "PSFs/s", "cores": CORES, "kind": "reference",
            "sample": f"ComputeZernikePsfPlanes x {n} requests (41 planes each, planes in parallel), {reps}x in "
                      f"{el:.2f} s"}


# ---- SMLM 128x128 x 50 frames end to end (C1) -------------------------------------------------------------------

def run_c1(ctx, steps, warmup):
    """One step = the C1 PSF on the GPU (41 planes, 129x129 at os 4, M 128), its splat kernel (block sums) set up on
    the device, the movie's blink events (resident in HBM, bucketed by frame once beforehand) collected into frame
    emitters on the device, and 50 frames of 128x128 rendered with Catmull-Rom splats, bg 20 and sCMOS noise: the
    movie end to end, as the reference's CPU timing covers it (RenderPhotonImage collects each frame's emitters)."""
    import torch
    from insiliscope_amd import configs, host
    req = configs.c1_psf(seed=1, oversampling=4)
    ev, st = configs.c1_events(seed=1, n_frames=50)
    w = h = 128
    off = torch.full((h, w), 100.0, dtype=torch.float32, device="cuda")
    adu = torch.empty((50, h, w), dtype=torch.int16, device="cuda")
    cam = configs.scmos_c1()
    # the blink events resident in HBM, bucketed by frame once (the reference buckets a movie's events once,
    # ScopeMovie.cpp:6183); each step computes the PSF, collects every frame's emitters on the device (as
    # RenderPhotonImage does per frame on the CPU), splats and noises the 50 frames
    movie = host.SmlmMovieGpu(ev, 50, w, h, 100.0, 1.0, 20.0, offset=off, ctx=ctx)

    def render():
        if not movie.Render(req, 4, 50.0, 2, cam, 1, outAdu=adu):
            raise RuntimeError(movie.splat.err)

    # the whole movie (PSF, block sums, collect, splat + noise) captured once as a HIP graph; a step is one replay
    movie.Capture(req, 4, 50.0, 2, cam, 1, outAdu=adu)

    def step():
        movie.Replay()

    for _ in range(warmup):
        step()
    with Timer(ctx) as t:
        for _ in range(steps):
            step()
    ms = t.ms / steps
    prof = profiled(ctx, lambda k: render(), min(steps, 3))  # per-kernel times from eager renders
    return {"value": 50 * 1000.0 / ms, "unit": "frames/s", "ms_per_step": ms,
            "config": {"workload": "SMLM 128x128 x 50 frames end to end: Gibson-Lanni PSF (41 planes 129x129, os 4, "
                                   "M 128, Noll 5-15 +-0.05) + frame emitters collected from the events on the GPU + "
                                   "Catmull-Rom splats of ~200 blinks/frame + sCMOS; a step = one replay of the "
                                   "movie's captured HIP graph",
                       "blinks": int(len(ev))},
            "kernels": {k: {"launches": v[0], "avg_ms": round(v[1] / v[0], 4)} for k, v in prof.items()},
            "_events": (ev, st)}


def cpu_c1(ev, st):
    """The reference on the host cores: BuildZernikePsfKernelCache (ComputeZernikePsfPlanes, sum-1, block sums) of
    the C1 request, then RenderPhotonImage (Cubic) + ApplyNoiseChain for the 50 frames (frames in parallel)."""
    import oracle as O
    from insiliscope_amd import configs
    ref = O.load_reference()
    if ref is None:
        return None
    req = O.psf_request(**configs.c1_psf(seed=1, oversampling=4))
    cam = configs.scmos_c1()
    oc = O.camera(quantumEfficiency=1.0, gainPhotonsPerAdu=cam.gainPhotonsPerAdu,
                  readNoiseElectrons=cam.readNoiseElectrons, emccd=0)
    off = np.full((128, 128), 100.0, np.float32)
    def movie():
        planes, _ = ref.psf_kernel_cache(req, 129, 41, 4)
        ref.render_noise_frames(ev, st, 128, 128, 100.0, 20.0, planes, 4, 50.0, 2, oc, off, 1)

    reps, el = timed_repeat(movie, CPU_MIN_S_SECONDARY)
    return {"value": round(reps * 50 / el, 2), "unit": "frames/s", "cores": CORES, "kind": "reference",
            "sample": f"the whole 50-frame movie, {reps}x in {el:.2f} s: PSF kernel cache + 50 rendered, noised "
                      f"frames"}


# ---- ROI forward + adjoint (C5) ----------------------------------------------------------------------------------

def run_c5(ctx, steps, warmup):
    """One step = the Poisson NLL and its gradient (x, y, z, photons, background per ROI; the 28 shared Zernike
    terms) of 4096 ROIs of 15x15 (pupil M 64, 4x4 sub-samples per pixel), data drawn once from the counter stream."""
    import torch
    from insiliscope_amd import configs, host
    z, p = configs.c5_rois(seed=1, n=4096)
    prm = np.zeros(len(p), host.ROI_PARAMS_DTYPE)
    for i, k in enumerate(("x", "y", "z", "photons", "background")):
        prm[k] = p[:, i]
    m = host.roi_model()
    mu = torch.from_numpy(host.RoiForward(m, z, prm, ctx=ctx)).cuda()
    cam = host.CameraNoiseParams(1.0, 0.0, 1.0, 0.0, False)  # ADU = Poisson(mu) counts
    frames = torch.arange(len(p), dtype=torch.int32, device="cuda")
    counts = host.ApplyNoiseChainFrames(mu, cam, 5, frames, ctx=ctx)
    data = counts.to(torch.int32).bitwise_and(0xFFFF).to(torch.float32)
    d_prm = torch.from_numpy(prm.view(np.uint8).reshape(len(prm), -1)).cuda()
    grad = torch.empty((len(p), 5), dtype=torch.float64, device="cuda")

    def step():
        host.RoiBackward(m, z, d_prm, data, gradParams=grad, ctx=ctx)

    for _ in range(warmup):
        step()
    with Timer(ctx) as t:
        for _ in range(steps):
            step()
    ms = t.ms / steps
    prof = profiled(ctx, lambda k: step(), min(steps, 3))
    # the work roi_kernel<true> executes per ROI: the forward field recomputed (2 complex matrix products) and the
    # adjoint (2 more), each a 64^3 complex product (M 64; the 60 sub-samples per axis padded to 64) = 4 real
    # 64^3 products on FP32 MFMA: 4 x 8 x 64^3 flops. rocprofv3 counts the same 3.44e10 FP32 MFMA flops per
    # 4096-ROI dispatch (SQ_INSTS_VALU_MFMA_MOPS_F32 x 512).
    P, M = 15 * 4, 64
    executed = len(p) * 4 * 8.0 * M ** 3
    algorithmic = len(p) * 2 * 8.0 * P * M * (M + P)  # what the adjoint needs: 2 x 8 P M (M + P), P unpadded
    k = prof.get("roi_backward", (1, ms))
    kms = k[1] / k[0]  # roi_kernel<true> alone (the reductions are timed as roi_reduce)
    tf = executed / (kms * 1e-3) / 1e12
    return {"value": len(p) * 1000.0 / ms, "unit": "ROI gradients/s", "ms_per_step": ms,
            "config": {"workload": "C5 adjoint: 4096 ROIs 15x15, Gibson-Lanni + 21 Zernike (Noll 4-24), pupil M 64, "
                                   "os 4, Poisson NLL gradient wrt x, y, z, N, b and the shared coefficients"},
            "roofline": {"bound": "mfma", "achieved": round(tf, 2), "peak": FP32_PEAK_TFS,
                         "unit": "TFLOP/s", "frac": round(tf / FP32_PEAK_TFS, 4),
                         "traffic": None, "kernel": "roi_backward",
                         "note": "executed FP32 MFMA flops of roi_kernel<true> (4 complex 64^3 products per ROI: "
                                 "the forward recomputed and the adjoint) over its own launch time; FP32 MFMA peak"},
            "roi_kernel": {"avg_ms": round(kms, 4), "TFLOP_s_fp32_executed": round(tf, 2),
                           "TFLOP_s_fp32_algorithmic": round(algorithmic / (kms * 1e-3) / 1e12, 2),
                           "flops_executed_per_roi": 4 * 8 * M ** 3,
                           "flops_algorithmic_per_roi": 2 * 8 * P * M * (M + P)},
            "kernels": {k: {"launches": v[0], "avg_ms": round(v[1] / v[0], 4)} for k, v in prof.items()},
            "_case": (z, prm, data.cpu().numpy())}


def cpu_c5(z, prm, data):
    """The FP64 C restatement of the ROI model and its adjoint (oracle.c, one thread) over the whole batch."""
    import oracle as O
    orc = O.load_oracle()
    p = np.zeros(len(prm), O.ROI_PARAMS_DTYPE)
    for k in ("x", "y", "z", "photons", "background"):
        p[k] = prm[k]
    t0 = time.time()
    orc.roi_backward(O.roi_model(), z, p, data.astype(np.float64))
    el = time.time() - t0
    return {"value": round(len(p) / el, 2), "unit": "ROI gradients/s", "cores": 1, "kind": "port",
            "sample": f"all {len(p)} ROIs ({el:.2f} s, FP64, one thread; the reference has no adjoint)"}


# ---- dye schedules (LabelSchedule over a batch of dye addresses) ----------------------------------------------------

DYE_KINETICS = dict(activationRatePerSec=0.2, onSec=0.05, offSec=1.0, bleachProb=0.05, photonCV=0.4,
                    initialOnSec=3.0)
DYE_T_MAX = 300.0


def dye_addresses(n, seed=17):
    rng = np.random.default_rng(seed)
    d = np.zeros((n, 3), np.uint32)
    d[:, 0] = rng.integers(0, 2 ** 32, n, dtype=np.uint64).astype(np.uint32)
    d[:, 1] = rng.integers(0, 64, n).astype(np.uint32)
    d[:, 2] = rng.integers(0, 1 << 24, n).astype(np.uint32)
    return d


def run_dyes(ctx, steps, warmup, n=1 << 21):
    """isc::LabelSchedule (dSTORM: initial ON, then activation / ON / brightness / bleach / OFF draws, about 20
    blinks per dye over 300 s) for n dye addresses resident in HBM, blinks written to HBM: one step = count, scan
    and fill launches of one iscg_label_blinks call."""
    import torch
    from insiliscope_amd import host
    d = dye_addresses(n)
    dd = torch.from_numpy(d.view(np.int32).reshape(-1)).cuda()
    kin = host.Kinetics(**DYE_KINETICS)
    offsets = torch.zeros(n + 1, dtype=torch.int64, device="cuda")
    total = torch.zeros(1, dtype=torch.int64, device="cuda")
    ctx.follow_torch()
    L = ctx.L
    ctx.check(L.iscg_label_blinks(ctx.h, dd.data_ptr(), n, 0, host.C.byref(kin), None, DYE_T_MAX, offsets.data_ptr(),
                                  None, 0, total.data_ptr()), "iscg_label_blinks")
    torch.cuda.synchronize()
    cap = int(total.item())
    out = torch.empty(cap * 3, dtype=torch.float64, device="cuda")

    def step(_=0):
        ctx.check(L.iscg_label_blinks(ctx.h, dd.data_ptr(), n, 0, host.C.byref(kin), None, DYE_T_MAX,
                                      offsets.data_ptr(), out.data_ptr(), cap, total.data_ptr()), "iscg_label_blinks")

    for _ in range(warmup):
        step()
    with Timer(ctx) as t:
        for _ in range(steps):
            step()
    ms = t.ms / steps
    prof = profiled(ctx, step, 3)
    kernels = {k: round(v[1] / max(v[0], 1) * 1000, 2) for k, v in prof.items() if k.startswith("dye_")}
    fill_us = kernels.get("dye_fill")
    # the fill launch's algorithmic HBM bytes: addresses 12 B + counts 4 B + offsets 8 B per dye, 24 B per blink
    fill_bytes = 24 * n + 24 * cap
    roof = None
    if fill_us:
        ach = fill_bytes / (fill_us * 1e-6) / 1e9
        roof = {"bound": "hbm", "achieved": round(ach, 1), "peak": 8000.0, "unit": "GB/s",
                "frac": round(ach / 8000.0, 4), "traffic": None, "kernel": "dye_fill",
                "note": "FP64 VALU-bound (a log per draw; log, cos, exp per brightness), not HBM-bound"}
    return {"value": n * 1000.0 / ms, "unit": "dyes/s", "ms_per_step": ms, "blinks_per_s": cap * 1000.0 / ms,
            "kernels_us": kernels, "roofline": roof,
            "config": {"workload": f"LabelSchedule dSTORM blinks of {n} dye addresses, tMax {DYE_T_MAX} s, "
                                   f"{cap / n:.2f} blinks per dye ({cap} blinks, 24 B each)",
                       "kinetics": DYE_KINETICS}}


def cpu_dyes(n=200000):
    import ctypes as C
    import oracle as O
    ref = O.load_reference()
    if ref is None:
        return None
    d = np.ascontiguousarray(dye_addresses(n, 23))
    kin = O.kinetics(**DYE_KINETICS)
    counts = np.zeros(n, np.int32)
    total = []
    reps, el = timed_repeat(lambda: total.append(ref.L.ref_label_blinks(
        d.ctypes.data, n, 0, C.byref(kin), DYE_T_MAX, counts.ctypes.data_as(C.POINTER(C.c_int32)), None, 0, 0, None,
        None)),
        CPU_MIN_S_SECONDARY)
    return {"value": round(reps * n / el, 1), "unit": "dyes/s", "cores": 1, "kind": "reference",
            "sample": f"isc::LabelSchedule of {n} dyes ({total[0]} blinks) in one thread, {reps}x in {el:.2f} s"}


WORLD_WINDOW = (0.0, 0.0, 25.6, 25.6, 0.0, 3.2)


def run_world(ctx, steps, warmup):
    """World::Density3dInWindow on the device over a synthetic cell field the size of the seed-1 field's 25.6 um
    window (5 cells, 2200 microtubules, 296,606 centreline points, 6300 pores; configs.world_field), microtubules and
    pores labelled (DNA-PAINT, density 0.7): one step = the walk's selection, every walked block's dyes generated
    (DyesInBlock / NpcDyesInBlock, LocalToWorld, the 0.1 nm lattice) and binned into the widefield workload's
    256 x 256 x 64 density (100 x 100 x 50 nm voxels), written to HBM."""
    import torch
    from insiliscope_amd import configs, host
    g = host.WorldGeometry.from_rows(**configs.world_field(1))
    W = host.World(1, g, ctx)
    out = torch.zeros((64, 256, 256), dtype=torch.float32, device="cuda")
    n = [0]

    def step(_=0):
        n[0], _o = W.Density3dInWindow(*WORLD_WINDOW, 256, 256, 64, 0xF, out=out)

    for _ in range(warmup):
        step()
    with Timer(ctx) as t:
        for _ in range(steps):
            step()
    ms = t.ms / steps
    prof = profiled(ctx, step, 3)
    kernels = {k: round(v[1] / max(v[0], 1) * 1000, 2) for k, v in prof.items() if k.startswith("world_")}
    W.close()
    return {"value": n[0] * 1000.0 / ms, "unit": "dyes/s", "ms_per_step": ms, "dyes_per_step": n[0],
            "kernels_us": kernels,
            "config": {"workload": "Density3dInWindow of a 25.6 um synthetic cell field (5 cells, 2200 microtubules, "
                                   "6300 pores; DNA-PAINT 0.7), 256x256x64 voxels, dyes generated every query"}}


def cpu_world():
    import oracle as O
    from insiliscope_amd import configs
    ref = O.load_reference()
    if ref is None:
        return None
    f = configs.world_field(1)
    cells = np.stack([f["cells"][k] for k in ("cx", "cy", "x", "y", "packRot")], 1)[:1]
    lk = (12.0, 0.0, 2.0, 5.0)
    res = []
    reps, el = timed_repeat(lambda: res.append(ref.mt_density(1, cells, f["mtStart"][:2], f["ptStart"], f["pts"], 0.7,
                                                              1.0, lk, WORLD_WINDOW, 256, 256, 64)[0]),
                            CPU_MIN_S_SECONDARY)
    return {"value": round(reps * res[0] / el, 1), "unit": "dyes/s", "cores": 1, "kind": "reference",
            "sample": f"DyesInBlock + LocalToWorld + lattice + binning (World::GenerateDyes / Density3dInWindow) "
                      f"over the microtubules of the first cell ({res[0]} dyes in the window) in one thread, "
                      f"{reps}x in {el:.2f} s"}


def measure_peaks():
    """The rates this box attains, beside the datasheet peaks the rooflines use: a 1 GiB device-to-device copy
    (read + write bytes / time) and square GEMMs through torch (hipBLASLt / rocBLAS) in FP32, FP16 and BF16."""
    import torch

    def timed(fn, reps):
        fn()
        torch.cuda.synchronize()
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record()
        for _ in range(reps):
            fn()
        b.record()
        torch.cuda.synchronize()
        return a.elapsed_time(b) / reps * 1e-3

    out = {"metric": "attainable peaks", "datasheet": {"hbm_GB_s": HBM_PEAK_GBS, "fp32_TFLOP_s": FP32_PEAK_TFS,
                                                       "fp16_dense_TFLOP_s": 2500.0}}
    n = 1 << 28  # 1 GiB of float32
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    t = timed(lambda: dst.copy_(src), 20)
    out["copy_GB_s"] = round(2 * n * 4 / t / 1e9, 1)
    del src, dst
    for name, dt, m in (("fp32", torch.float32, 8192), ("fp16", torch.float16, 16384), ("bf16", torch.bfloat16, 16384)):
        a = torch.randn(m, m, device="cuda").to(dt)
        b = torch.randn(m, m, device="cuda").to(dt)
        t = timed(lambda: torch.matmul(a, b), 10)
        out[f"gemm_{name}_TFLOP_s"] = round(2 * m ** 3 / t / 1e12, 1)
        del a, b
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--workload", default="wf", choices=["wf", "smlm", "psf", "dyes", "adjoint", "c1", "world"])
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-secondary", action="store_true")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="first copy MIB MiB device to device (a kernel of known traffic for PMC byte counters)")
    ap.add_argument("--peaks", action="store_true",
                    help="only measure the box's attainable rates (device copy, FP32/FP16/BF16 GEMM) and print them")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the image-formation path has no multi-GPU mode in this build "
              "(frames and PSFs shard across GPUs without exchange; run one process per GPU).", file=sys.stderr)
        sys.exit(2)
    os.environ["ISC_THREADS"] = str(CORES)
    import torch
    from insiliscope_amd import host
    if not torch.cuda.is_available():
        print("no GPU", file=sys.stderr)
        sys.exit(1)
    if args.peaks:
        print(json.dumps(measure_peaks()))
        return
    ctx = host.Context(0)
    if args.calibrate_copy:
        n = args.calibrate_copy << 18  # float32 elements
        src = torch.ones(n, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        log(f"calibration copy: {args.calibrate_copy} MiB read + {args.calibrate_copy} MiB written")
        del src, dst
    secondary = {}
    if args.workload == "wf":
        r = run_wf(ctx, args.steps, args.warmup)
        config = {"workload": "WF mean field 256x256x64 density (100x100x50 nm), 127-plane Gibson-Lanni PSF 257x257, "
                              "512x512x128 FP32 R2C/C2R 3D FFT, 64 foci, sCMOS noise",
                  "frames_per_step": 64, "fft": [512, 512, 128], "kernels": r["kernels"],
                  "stack_GB_s_algorithmic": r["stack_GB_s"]}
        unit, dtype = "frames/s", "f32"
    elif args.workload == "smlm":
        r = run_smlm(ctx, args.steps, args.warmup)
        r.pop("_events")
        config, unit, dtype = r["config"], "frames/s", "f64"
    elif args.workload == "dyes":
        r = run_dyes(ctx, args.steps, args.warmup)
        config, unit, dtype = r["config"], "dyes/s", "f64"
    elif args.workload == "world":
        r = run_world(ctx, args.steps, args.warmup)
        config, unit, dtype = r["config"], "dyes/s", "f64"
        r["roofline"] = None
    elif args.workload == "c1":
        r = run_c1(ctx, args.steps, args.warmup)
        r.pop("_events")
        config, unit, dtype = dict(r["config"], kernels=r["kernels"]), "frames/s", "f64"
        r["roofline"] = None
    elif args.workload == "adjoint":
        r = run_c5(ctx, args.steps, args.warmup)
        r.pop("_case")
        config, unit, dtype = r["config"], "ROI gradients/s", "f32"
    else:
        r = run_psf(ctx, args.steps, args.warmup)
        config, unit, dtype = r["config"], "PSFs/s", "f16x3/f32"
    cpu = None
    if args.workload == "wf" and not args.no_secondary:
        ev = None
        try:
            s = run_smlm(ctx, 2, 1)
            ev, st, planes = s.pop("_events")
            secondary["smlm_c4"] = s
            if not args.no_cpu:
                c = cpu_smlm(ev, st, planes)
                if c:
                    s["cpu_baseline"] = c
                    s["speedup_vs_cpu"] = round(s["value"] / c["value"], 1)
        except Exception as e:  # a secondary line must not hide the primary one
            secondary["smlm_c4"] = {"error": str(e)}
        if ev is not None:
            try:
                b = run_binned(ctx, ev, st, planes)
                frames = b.pop("_frames")
                secondary["binned_c4"] = b
                if not args.no_cpu:
                    c = cpu_binned(frames, planes)
                    if c:
                        b["cpu_baseline"] = c
                        b["speedup_vs_cpu"] = round(b["value"] / c["value"], 1)
            except Exception as e:
                secondary["binned_c4"] = {"error": str(e)}
        try:
            c = run_c1(ctx, 5, 1)
            ev, st = c.pop("_events")
            secondary["smlm_c1"] = c
            if not args.no_cpu:
                b = cpu_c1(ev, st)
                if b:
                    c["cpu_baseline"] = b
                    c["speedup_vs_cpu"] = round(c["value"] / b["value"], 1)
        except Exception as e:
            secondary["smlm_c1"] = {"error": str(e)}
        try:
            c = run_c5(ctx, 10, 2)
            case = c.pop("_case")
            secondary["adjoint_c5"] = c
            if not args.no_cpu:
                b = cpu_c5(*case)
                c["cpu_baseline"] = b
                c["speedup_vs_cpu"] = round(c["value"] / b["value"], 1)
        except Exception as e:
            secondary["adjoint_c5"] = {"error": str(e)}
        try:
            wr = run_world(ctx, 10, 2)
            secondary["world_density"] = wr
            if not args.no_cpu:
                c = cpu_world()
                if c:
                    wr["cpu_baseline"] = c
                    wr["speedup_vs_cpu"] = round(wr["value"] / c["value"], 1)
        except Exception as e:
            secondary["world_density"] = {"error": str(e)}
        try:
            p = run_psf(ctx, 3, 1)
            secondary["psf_library_c3"] = p
            if not args.no_cpu:
                c = cpu_psf()
                if c:
                    p["cpu_baseline"] = c
                    p["speedup_vs_cpu"] = round(p["value"] / c["value"], 1)
        except Exception as e:
            secondary["psf_library_c3"] = {"error": str(e)}
        try:
            dy = run_dyes(ctx, 5, 1)
            secondary["dye_schedules"] = dy
            if not args.no_cpu:
                c = cpu_dyes()
                if c:
                    dy["cpu_baseline"] = c
                    dy["speedup_vs_cpu"] = round(dy["value"] / c["value"], 1)
        except Exception as e:
            secondary["dye_schedules"] = {"error": str(e)}
    if not args.no_cpu and args.workload == "wf":
        try:
            cpu = cpu_wf(np.ascontiguousarray(r["_psf"].cpu().numpy()))
        except Exception as e:
            cpu = {"error": str(e)}
    r.pop("_psf", None)
    # launch parameters the autotuner chose during the warm-up calls (shape key -> value)
    config["autotuned"] = ctx.autotune_choices()
    line = {
        "metric": METRIC, "value": round(r["value"], 2), "unit": unit, "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(r["ms_per_step"], 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": dtype, "data": "synthetic (pcg4d-seeded)",
        "config": config, "roofline": r.get("roofline"), "cpu_baseline": cpu,
    }
    if cpu and cpu.get("value"):
        line["speedup_vs_cpu"] = round(r["value"] / cpu["value"], 1)
    if secondary:
        line["secondary"] = secondary
    print(json.dumps(line))


if __name__ == "__main__":
    main()
