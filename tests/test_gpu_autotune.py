"""The launch-parameter autotuner (iscg_autotune_*): every candidate of a tuned parameter gives bit-identical results,
the tuned choice is one of the candidates, and kept choices survive a dump/load round trip."""
import numpy as np
import pytest

import oracle as O
from insiliscope_amd import configs, host

pytestmark = pytest.mark.gpu


def _only_key(choices, prefix):
    keys = [k for k in choices if k.startswith(prefix)]
    assert len(keys) == 1, choices
    return keys[0]


def test_psf_small_kernel_every_candidate_identical():
    """16 C3 requests (65x65, 41 planes, M 128, a row-block shape): the workgroups per request (1 to 32) only change
    which workgroup computes a plane, so the planes are the same bits for every choice, tuned or fixed."""
    reqs = configs.c3_requests(seed=1)[:16]
    fixed = host.Context(0)
    fixed.set_autotune(False)
    want = host.ComputePsfPlanesBatch(reqs, ctx=fixed)
    assert fixed.autotune_choices() == {}
    fixed.close()
    c = host.Context(0)
    got = host.ComputePsfPlanesBatch(reqs, ctx=c)
    ch = c.autotune_choices()
    key = _only_key(ch, "psf_small")
    cands = [zg + 64 for zg in (1, 2, 4, 8, 16, 32)]
    assert ch[key] in cands, ch
    assert np.array_equal(got, want)
    for c_ in cands:
        c.load_autotune({key: c_})
        again = host.ComputePsfPlanesBatch(reqs, ctx=c)
        assert np.array_equal(again, want), f"choice {c_} (workgroups per request + 64: the row-block kernel)"
    c.close()


@pytest.mark.parametrize("shape", ["auto_m96_33", "m128_69"])
def test_psf_chunked_kernel_vs_oracle_every_candidate_identical(orc, shape):
    """The chunked small-plane kernel (psf_small_kernel: pupil sizes or plane sizes the row-block form does not
    take): C1's optics with the automatic pupil rule (pupilSamples 0 -> M 96) at 33 x 33, and M 128 at 69 x 69
    (padded 80): every plane within 1e-5 relative L2 of the FP64 oracle, and every workgroups-per-request
    candidate (1 to 32) bit-identical."""
    base = configs.c1_psf(seed=1)
    reqs = []
    for i in range(4):
        r = dict(base, zernike=configs.zernike_noll(10 + i, range(5, 16), amplitude=0.05))
        if shape == "auto_m96_33":
            r.update(pupilSamples=0, nz=21)
        else:
            r.update(kernelHalfWidthPx=34, nz=9)
        reqs.append(r)
    size, nz, M = host.PsfGeometry(reqs[0])
    assert (size, M) == ((33, 96) if shape == "auto_m96_33" else (69, 128)), (size, M)
    fixed = host.Context(0)
    fixed.set_autotune(False)
    want = host.ComputePsfPlanesBatch(reqs, ctx=fixed)
    fixed.close()
    errs = []
    for i, r in enumerate(reqs):
        o = orc.psf_planes(O.psf_request(**r))
        for k in range(nz):
            errs.append(np.sqrt(((want[i, k] - o[k]) ** 2).sum() / (o[k] ** 2).sum()))
    print(shape, "max rel L2 vs oracle", max(errs), flush=True)
    assert max(errs) <= 1e-5
    c = host.Context(0)
    got = host.ComputePsfPlanesBatch(reqs, ctx=c)
    key = _only_key(c.autotune_choices(), "psf_small")
    cands = [zg for zg in (1, 2, 4, 8, 16, 32) if zg <= nz]
    assert c.autotune_choices()[key] in cands
    assert np.array_equal(got, want)
    for c_ in cands:
        c.load_autotune({key: c_})
        assert np.array_equal(host.ComputePsfPlanesBatch(reqs, ctx=c), want), f"choice {c_}"
    c.close()


def test_splat_grid_every_candidate_identical(orc):
    """16 frames of 512x512 with Poisson(800) blinks, Catmull-Rom, EMCCD: the persistent grid's workgroups per CU
    (4, 8, 16, 32), the tile edge (32, or 16: + 64) and the form (per hit, or per pixel: + 128) give the same photons
    and ADU; frame 0 equals the oracle's."""
    req = configs.c4_psf(seed=1)
    req.update(nz=11, pupilSamples=64)
    planes = orc.psf_planes(O.psf_request(**req), normalize=True)
    cache = host.PsfKernelCache(planes, oversampling=4, zStepNm=15.0, interpMode=2)
    n = 16
    ev, st = configs.c4_events(seed=1, n_frames=n)
    ems = [host.CollectFrameEmitters(ev[st[f]:st[f + 1]], f, 100.0, 1.0, psfCache=cache) for f in range(n)]
    start = np.zeros(n + 1, np.uint32)
    start[1:] = np.cumsum([len(e) for e in ems])
    allems = np.concatenate(ems)
    w = h = 512
    off = np.full((h, w), 100.0, np.float32)
    cam = configs.emccd_c4()
    frames = np.arange(n, dtype=np.uint32)

    def render(c):
        sp = host.SplatGpuHost(c)
        assert sp.SetKernels([cache]) and sp.SetStatic(w, h, offset=off, flatBackground=50.0), sp.err
        adu = np.zeros((n, h, w), np.uint16)
        ph = np.zeros((n, h, w), np.float32)
        assert sp.RenderBatch(allems, start, frames, cam, 1, outAdu=adu, outPhotons=ph), sp.err
        return adu, ph

    fixed = host.Context(0)
    fixed.set_autotune(False)
    want_adu, want_ph = render(fixed)
    fixed.close()
    bs = np.stack([orc.block_sums(p, 4) for p in planes])
    assert np.array_equal(want_ph[0], orc.render_frame(w, h, 50.0, ems[0], bs, planes.shape[1], 4, 2))
    c = host.Context(0)
    adu, ph = render(c)
    ch = c.autotune_choices()
    key = _only_key(ch, "splat_tiles")
    cands = [per_cu + 64 * t16 + 128 * pix for pix in (0, 1) for t16 in (0, 1) for per_cu in (4, 8, 16, 32)]
    assert ch[key] in cands, ch
    assert np.array_equal(adu, want_adu) and np.array_equal(ph, want_ph)
    for cand in cands:
        c.load_autotune({key: cand})
        adu, ph = render(c)
        assert np.array_equal(adu, want_adu) and np.array_equal(ph, want_ph), f"choice {cand}"
    c.close()


def test_autotune_dump_load_round_trip():
    c = host.Context(0)
    assert c.load_autotune("a key=3\nanother key 2=16\n\n") == 2
    assert c.autotune_choices() == {"a key": 3, "another key 2": 16}
    d = host.Context(0)
    assert d.load_autotune(c.autotune_choices()) == 2
    assert d.autotune_choices() == c.autotune_choices()
    for bad in ("no value\n", "k=\n", "=3\n", "k=-1\n", "k=3x\n"):
        with pytest.raises(host.IscgError):
            d.load_autotune(bad)
    assert d.L.iscg_autotune_set(d.h, 2) != 0
    c.close()
    d.close()
