"""The binned blink regime against the splat on one set of emitters, on the GPU: the reference's own regime check
(BinnedVsSplat, cli/blink_regimes_check.cpp:170-209) restated on this library's two renderers. 3000 blinks in a 64 x 64
frame (x, y uniform in [-0.5, 6.9] um, z uniform in [-500, 2500] nm against a 0.5 um stage, each on for 0.6 s from a
start uniform in [-0.3, 1.3) s), 1000 photons per blink, a 3000 nm Catmull-Rom kernel at the reference's default
oversampling (6 samples per pixel, so that every upscale divides it): the same photons within 1 %, the image
within the reference's relative L2 bounds (0.25 at upscale 1, 0.08 at 2, 0.06 at 3; the snap to the binned grid moves
each spot by up to half a cell). The binned frame is also the same on a second render (the reference checks serial =
parallel)."""
import numpy as np
import pytest

from insiliscope_amd import configs, host

pytestmark = pytest.mark.gpu

W = H = 64
PIXEL_NM, PER_BLINK, Z_STAGE_UM = 100.0, 1000.0, 0.5
REL_MAX = {1: 0.25, 2: 0.08, 3: 0.06}


@pytest.fixture(scope="module")
def case(ctx):
    req = configs.psf_request_fields(wavelengthNm=670.0, na=1.45, immersionIndex=1.518, sampleIndex=1.33,
                                     workingDistanceUm=170.0, pixelSizeNm=PIXEL_NM, zStepNm=100.0, oversampling=6,
                                     kernelHalfWidthPx=30, nz=31, pupilSamples=0, interpMode=2)
    cache = host.BuildZernikePsfKernelCache(req, ctx=ctx)
    rng = np.random.default_rng(7)
    ev = np.zeros(3000, host.BLINK_EVENT_DTYPE)
    ev["xUm"] = rng.uniform(-0.5, 6.9, len(ev))
    ev["yUm"] = rng.uniform(-0.5, 6.9, len(ev))
    ev["zNm"] = rng.uniform(-500.0, 2500.0, len(ev))
    ev["tStart"] = rng.uniform(-0.3, 1.3, len(ev))
    ev["tEnd"] = ev["tStart"] + 0.6
    ev["brightness"] = 1.0
    ems = host.CollectFrameEmitters(ev, 0, PIXEL_NM, PER_BLINK, psfCache=cache, globalZOffsetUm=Z_STAGE_UM)
    sp = host.SplatGpuHost(ctx)
    assert sp.SetKernels([cache]) and sp.SetStatic(W, H), sp.err
    splat = np.zeros((1, H, W), np.float32)
    assert sp.RenderBatch(ems, np.array([0, len(ems)], np.uint32), np.zeros(1, np.uint32), None, 1,
                          outPhotons=splat), sp.err
    return cache, ems, splat[0].astype(np.float64)


@pytest.mark.parametrize("upscale", [1, 2, 3])
def test_binned_equals_splat_within_the_reference_bounds(ctx, case, upscale):
    cache, ems, splat = case
    r = host.BinnedBlinkRenderer(ctx)
    assert r.Setup(cache, 1.0, W, H, upscale), r.err
    a = np.zeros((1, H, W), np.float32)
    b = np.zeros((1, H, W), np.float32)
    assert r.RenderFrames([ems], a) and r.RenderFrames([ems], b), r.err
    a = a[0].astype(np.float64)
    photons = a.sum() / splat.sum() - 1.0
    rel = float(np.linalg.norm(a - splat) / np.linalg.norm(splat))
    print("upscale %d (R %d cells): %d emitters, photons %+.4f %%, image rel L2 %.4f" % (
        r.Upscale(), r.RadiusCells(), len(ems), 100 * photons, rel), flush=True)
    assert r.Upscale() == upscale
    assert len(ems) > 1000 and splat.sum() > 0
    assert abs(photons) < 0.01
    assert rel < REL_MAX[upscale]
    assert np.array_equal(a, b[0].astype(np.float64))
