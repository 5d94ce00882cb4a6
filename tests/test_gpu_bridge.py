"""The drop-in boundary, compiled: the reference's own engine (oracle/_ref/libiscbridge.so, built from its sources in
place with unmodified headers) running through this build's link-time drop-ins over libiscg.so --
insiliscope_amd/integration/SplatGpuIscg.cpp as sim::SplatGpuHost, GpuSimIscg.cpp as sim::GpuSimulator, and
WidefieldIscgHost.h as its WidefieldAccelerator -- against the same engine's CPU path.

The splat hosts take the reference's GPU wire format (GpuSplatEmitter: float weights), so their frames equal the
CPU's up to float rounding: an ADU differs only where a noise draw lands on a count boundary (the reference's own
GPU hosts accept 90 % of the pixels equal; here at least 99.9 %). The mean-field scenes' images are within 1e-5
relative L2 of the CPU's."""
import os

import numpy as np
import pytest

import oracle as O
from insiliscope_amd import configs

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def bridge():
    if not os.path.exists(O.BRIDGE_PATH):
        pytest.skip("libiscbridge.so not built (needs the reference checkout at build time)")
    return O.Bridge()


def equal_fraction(a, b):
    return float((a == b).mean())


@pytest.mark.parametrize("via,name", [(0, "SplatGpuHost"), (1, "GpuSimulator")])
@pytest.mark.parametrize("emccd", [False, True])
def test_splat_dropins_vs_engine_cpu(bridge, orc, via, name, emccd):
    """The reference's CollectGpuEmitters records of seeded blinks (C4's kernel geometry: hw 8, os 4, Catmull-Rom,
    11 planes at 15 nm; 96 x 80 px, 6 frames with Poisson(60) blinks) through the drop-in host, against
    RenderPhotonImage + ApplyNoiseChain of the same engine."""
    req = configs.c4_psf(seed=1)
    req.update(nz=11, pupilSamples=64)
    planes = orc.psf_planes(O.psf_request(**req), normalize=True)
    ev, st = configs.blink_events(4, 6, 96, 80, 100.0, 60.0, 75.0, 3000.0)
    cam = [1.0, 0.0, 1.0 / 300.0, 50.0, 1.0, 300.0, 0.002, 16] if emccd else [1.0, 0.0, 2.0, 1.5, 0.0, 300.0, 0.0, 16]
    print(name, "emccd" if emccd else "scmos", len(ev), "blinks", flush=True)
    gpu, cpu = bridge.splat_frames(via, ev, 6, 96, 80, 100.0, 1.0, 50.0, planes, 4, 15.0, 2, cam, 100.0, 7)
    frac = equal_fraction(gpu, cpu)
    print("equal ADU fraction", frac, flush=True)
    assert cpu.mean() > 100.0 and frac >= 0.999


def test_scope_movie_through_splat_dropin(bridge):
    """RenderScopeMovie (the engine insiliscope_cli runs) on the seed-1 cell field, 96 x 96 px, 12 frames: with this
    build's SplatGpuHost installed as the process-wide blink splat (SetDefaultBlinkSplatGpu, which makes it through
    MakeSplatGpuHost and its self-check) the frames equal the CPU movie's up to float rounding."""
    opts = {"size": 96, "frames": 12, "world-seed": 1, "seed": 1, "disk-cache": 0, "blink-min-photons": 0}
    cpu, ci = bridge.movie(opts, False, 96 * 96 * 16)
    gpu, gi = bridge.movie(opts, True, 96 * 96 * 16)
    print("movie", cpu.shape, "blinks", ci["blinks"], "host:", gi["host"], flush=True)
    assert "libiscg" in gi["host"] and gi["blinks"] == ci["blinks"] > 0
    assert gpu.shape == cpu.shape == (12, 96, 96)
    frac = equal_fraction(gpu, cpu)
    print("equal ADU fraction", frac, "max |diff|", int(np.abs(gpu.astype(int) - cpu.astype(int)).max()), flush=True)
    assert frac >= 0.999


def test_widefield_scene_through_accelerator(bridge, golden):
    """The reference's GPU-mode WidefieldScene with WidefieldIscgHost as its accelerator: FocusSeries of a 48 x 40 x 6
    persistent population (foci on and between the dye planes, the FOV off the cell edges) served by the host, each
    image within 1e-5 relative L2 of the same scene's CPU images."""
    W = golden("wf_ref.npz")
    planes = golden("psf_ref.npz")["small_planes"]
    foci = np.array([0.125, 0.02, 0.21, 0.2003, 0.31])
    gpu, cpu, used = bridge.wf_scene(W["scene_density"], 50.0, planes, 4, 50.0, foci, dx_um=0.0123, dy_um=-0.031)
    assert used, "the scene did not use the accelerator"
    for f in range(len(foci)):
        e = np.sqrt(((gpu[f] - cpu[f]).astype(np.float64) ** 2).sum() / (cpu[f].astype(np.float64) ** 2).sum())
        print("focus", foci[f], "rel L2", e, flush=True)
        assert e <= 1e-5
