"""PSF planes on the GPU (iscg_psf_planes: FP64 phase, split-FP16 MFMA matrix DFT) against the reference's fixture
(tests/psf/zernike_ref.bin), the compiled reference's planes at the C1/C3 geometries and the FP64 oracle:
relative L2 <= 1e-5 per plane."""
import os

import numpy as np
import pytest

import oracle as O
from insiliscope_amd import configs, host
from conftest import record_accuracy
from test_oracle import ZREF_CASES, GOLDEN

pytestmark = pytest.mark.gpu

TOL = 1e-5  # relative L2 per plane, split-precision MFMA path against FP64


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.sqrt(((a - b) ** 2).sum() / (b ** 2).sum()))


def test_psf_vs_zernike_ref_bin(ctx):
    b = open(os.path.join(GOLDEN, "zernike_ref.bin"), "rb").read()
    cases, nx, ny, nz = np.frombuffer(b[4:20], np.int32)
    ref = np.frombuffer(b[20:], np.float32).reshape(cases, nz, ny, nx)
    reqs = []
    for name, ns, depth, zern, mask in ZREF_CASES:
        reqs.append(dict(wavelengthNm=660.0, na=1.4, immersionIndex=1.518, sampleIndex=ns, workingDistanceUm=150.0,
                         sampleDepthNm=depth, pixelSizeNm=100.0, oversampling=4, kernelHalfWidthPx=8, nz=3,
                         zStepNm=100.0, maskModes=5, maskWaist=1.0, pupilSamples=64, zernike=zern, maskType=mask))
    got = host.ComputePsfPlanesBatch(reqs, ctx=ctx)
    errs, mx = [], []
    for c in range(cases):
        for z in range(nz):
            e = rel_l2(got[c, z], ref[c, z])
            print(ZREF_CASES[c][0], z, e, flush=True)
            errs.append(e)
            mx.append(np.abs(got[c, z] - ref[c, z]).max() / np.abs(ref[c, z]).max())
            assert e <= TOL and np.argmax(got[c, z]) == np.argmax(ref[c, z])
    record_accuracy("psf_zernike_ref_bin_4_cases_vs_reference_fixture", rel_l2_max=max(errs),
                    max_abs_rel_to_peak=max(mx))


def test_psf_c1_all_planes_vs_oracle(ctx, orc, golden):
    G = golden("psf_ref.npz")
    c1 = configs.c1_psf(seed=1)
    got = host.ComputeZernikePsfPlanes(c1, ctx=ctx)
    want = orc.psf_planes(O.psf_request(**c1))
    assert got.shape == want.shape == (41, 33, 33)
    errs = [rel_l2(got[z], want[z]) for z in range(41)]
    print("C1 max rel L2", max(errs), flush=True)
    record_accuracy("psf_c1_41_planes_33x33_vs_fp64_oracle", rel_l2_max=max(errs),
                    max_abs_rel_to_peak=max(np.abs(got[z] - want[z]).max() / np.abs(want[z]).max()
                                            for z in range(41)))
    assert max(errs) <= TOL
    for j, z in enumerate(G["c1_planes_idx"]):
        assert rel_l2(got[z], G["c1_planes"][j]) <= TOL


def test_psf_c3_batch_vs_reference(ctx, golden):
    G = golden("psf_ref.npz")
    reqs = configs.c3_requests(seed=1)
    pick = [int(i) for i in G["c3_pick"]]
    got = host.ComputePsfPlanesBatch([reqs[i] for i in pick], ctx=ctx)
    errs = []
    for j in range(len(pick)):
        for k, z in enumerate((0, 20, 40)):
            e = rel_l2(got[j, z], G["c3_planes"][j, k])
            errs.append(e)
            assert e <= TOL, (j, z, e)
    record_accuracy("psf_c3_65x65_vs_compiled_reference_fp64", rel_l2_max=max(errs))


def test_psf_normalized_and_wide_kernel(ctx, orc):
    """M 256 and a 129-wide kernel (several output-row blocks per plane); sum-1 planes."""
    req = dict(wavelengthNm=600.0, na=1.3, immersionIndex=1.518, sampleIndex=1.45, sampleDepthNm=800.0,
               pixelSizeNm=100.0, oversampling=1, kernelHalfWidthPx=64, nz=5, zStepNm=250.0, pupilSamples=256,
               zernike={3: 0.04, 7: -0.03, 12: 0.05})
    got = host.ComputePsfPlanesBatch([req], normalize=True, ctx=ctx)[0]
    want = orc.psf_planes(O.psf_request(**req), normalize=True)
    assert got.shape == want.shape == (5, 129, 129)
    for z in range(5):
        assert rel_l2(got[z], want[z]) <= TOL
        assert abs(got[z].astype(np.float64).sum() - 1.0) < 1e-5


def test_psf_mixed_batch_rejected(ctx):
    a = configs.c1_psf(seed=1)
    b = dict(a, kernelHalfWidthPx=8)
    with pytest.raises(host.IscgError, match="share"):
        host.ComputePsfPlanesBatch([a, b], ctx=ctx)
    with pytest.raises(host.IscgError, match="must be > 0"):
        host.ComputePsfPlanesBatch([dict(a, na=0.0)], ctx=ctx)


def test_compute_psf_kernel_cache_on_gpu(ctx, tmp_path):
    """ComputePsfKernelCache with the planes computed on the GPU: sum-1 planes within 1e-5 relative L2 of the
    compiled reference's BuildZernikePsfKernelCache, stored to disk (the file reads back bit for bit), and a repeat
    (another interpMode) is a memo hit with the same planes and serial."""
    import time
    from insiliscope_amd import psf_cache as P
    text = "0,0,0,0.05,-0.03,0.02,0,0,0,0,0.01,0,0,0,%.17g" % (1e-3 * (time.time() % 1.0))
    fields = dict(wavelengthNm=670.0, na=1.45, immersionIndex=1.518, sampleIndex=1.33, workingDistanceUm=170.0,
                  sampleDepthNm=500.0, pixelSizeNm=100.0, zStepNm=50.0, oversampling=2, kernelHalfWidthPx=6, nz=9,
                  pupilSamples=64, interpMode=2)
    req = P.PsfGeneratorRequest(model=P.GIBSON_LANNI_ZERNIKE, zernikeCoefficients=text, **fields)
    P.SetPsfKernelDiskCacheDir(str(tmp_path))
    ok, c, err = P.ComputePsfKernelCache(req, ctx=ctx)
    assert ok, err
    size = 2 * 6 * 2 + 1
    assert c.planes.shape == (9, size, size) and c.interpMode == 2 and c.oversampling == 2
    z, zok = P.ParseZernikeCoefficients(text)
    assert zok
    want, _ = O.Reference().psf_kernel_cache(O.psf_request(zernike=z, **fields), size, 9, 2)
    for k in range(9):
        assert rel_l2(c.planes[k], want[k]) <= TOL, k
    back = P.LoadKernelFromDisk(req)
    assert back is not None and np.array_equal(back.planes, c.planes)
    log = []
    ok, c2, _ = P.ComputePsfKernelCache(P.PsfGeneratorRequest(**{**req.__dict__, "interpMode": 1}), log.append,
                                        ctx=ctx)
    assert ok and c2.serial == c.serial and c2.planes is c.planes and c2.interpMode == 1
    assert log == ["PSF: kernel unchanged, reusing the one already computed."]
    P.SetPsfKernelDiskCacheDir("")


@pytest.mark.parametrize("name", ["m1024", "m288", "auto"])
def test_psf_large_pupils_vs_reference(ctx, golden, name):
    """Pupils above 256 samples on the chunked kernel (psf_planes_big_kernel): explicit M 1024 (65 x 65) and 288
    (129 x 129), and the automatic rule's M 416 (97 x 97): each plane within 1e-5 relative L2 of the compiled
    reference's (psf_big_ref.npz); two requests in one call give the same planes as one."""
    from test_oracle import PSF_BIG
    G = golden("psf_big_ref.npz")
    fields = dict(configs.c1_psf(seed=1), **PSF_BIG[name])
    assert host.PsfGeometry(fields)[2] == int(G[name + "_M"])
    got = host.ComputePsfPlanesBatch([fields, fields], ctx=ctx)
    want = G[name + "_planes"]
    errs = [rel_l2(got[0, k], want[k]) for k in range(len(want))]
    print(name, "M", int(G[name + "_M"]), "max rel L2", max(errs), flush=True)
    record_accuracy("psf_large_pupil_%s_vs_compiled_reference" % name, rel_l2_max=max(errs))
    assert max(errs) <= 1e-5 and np.array_equal(got[0], got[1])
