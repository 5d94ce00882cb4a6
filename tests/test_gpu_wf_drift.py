"""Drifting samples on the mean field (iscg_wf_shifted_frames behind host.WidefieldDriftFrames: RenderShiftedImages
for a batch of frames) against the compiled reference's WidefieldDriftFrames frames (wf_drift_ref.npz) and the FP64
restatement: 1e-5 relative L2 per frame; bleach channels and binning by linearity and against the restatement."""
import numpy as np
import pytest

from insiliscope_amd import host
from conftest import record_accuracy
from test_oracle import drift_case, rel_l2, shifted_frames_restated

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", ["a", "b"])
def test_drift_frames_vs_reference(ctx, golden, name):
    G = golden("wf_drift_ref.npz")
    desc, frames = drift_case(G, name)
    spectra = G[name + "_spectra"][:, None]  # (nImages, 1 channel, NY, W): the persistent population
    df = host.WidefieldDriftFrames(ctx)
    df.Begin(host.DriftRange(G[name + "_drift"]), [float(G[name + "_base"])])
    assert len(df.Foci()) == spectra.shape[0]
    df.SetSeries(dict(desc, nImages=spectra.shape[0], nChannels=1), spectra, float(G[name + "_frac_pitch"][2]))
    base = np.full((12, desc["ch"], desc["cw"]), 0.25, np.float32)  # added to
    got = df.Render(0, G[name + "_drift"], cam=base.copy()) - base
    want = G[name + "_frames"]
    errs = [rel_l2(got[f], want[f]) for f in range(len(want))]
    print(name, "max rel L2", max(errs), flush=True)
    record_accuracy("wf_drift_frames_%s_vs_compiled_reference" % name, rel_l2_max=max(errs))
    assert max(errs) <= 1e-5


def test_drift_frames_bleach_channels_and_binning(ctx, golden):
    """Three channels (the persistent one and two bleach maps) with per-frame coefficients equal one pre-summed
    channel (the sum is linear; 1e-6), and 2 x 2 binning equals the restatement's (1e-5)."""
    G = golden("wf_drift_ref.npz")
    desc, frames = drift_case(G, "a")
    P = G["a_spectra"]
    rng = np.random.default_rng(4)
    B1 = (P * rng.uniform(0.2, 0.8)).astype(np.complex64)
    B2 = np.roll(P, 1, axis=0).astype(np.complex64)
    a = np.stack([rng.uniform(0.0, 1.0, len(frames)), rng.uniform(0.0, 0.5, len(frames))], axis=1)
    a[3, 0] = 0.0  # a zero coefficient drops its term
    three = np.stack([P, B1, B2], axis=1)
    d3 = dict(desc, nImages=P.shape[0], nChannels=3)
    got = host.WidefieldShiftedFrames(d3, three, frames, a, ctx=ctx)
    for f, (i0, i1, w, dx, dy) in enumerate(frames):
        one = (P + np.float32(a[f, 0]) * B1 + np.float32(a[f, 1]) * B2).astype(np.complex64)
        want = host.WidefieldShiftedFrames(dict(desc, nImages=P.shape[0], nChannels=1), one[:, None],
                                           [(i0, i1, w, dx, dy)], ctx=ctx)[0]
        assert rel_l2(got[f], want) < 1e-6, f
    d2 = dict(desc, upscale=2)
    got2 = host.WidefieldShiftedFrames(dict(d2, nImages=P.shape[0], nChannels=1), P[:, None], frames, ctx=ctx)
    want2 = shifted_frames_restated(d2, P, frames)
    assert got2.shape == (len(frames), desc["ch"] // 2, desc["cw"] // 2)
    assert max(rel_l2(got2[f], want2[f]) for f in range(len(frames))) < 1e-5
