"""The PSF request's coefficient text and the kernel caches around the PSF seam (insiliscope_amd.psf_cache) against
the compiled reference: ParseZernikeCoefficients / FormatZernikeCoefficients on edge and random inputs, the disk
store's file and fingerprint in both directions (a file the reference wrote is read here; a file written here is
read by the reference's ComputePsfKernelCache), and the memo's rules (SameKernel, most recent first, at least two
entries, one computation for concurrent requesters) with the computation replaced on the CPU."""
import math
import threading
import time

import numpy as np
import pytest

import oracle as O
from insiliscope_amd import host, psf_cache as P

Z15 = ",".join(["0"] * 14)
EDGE = ["1e-310", "4.9406564584124654e-324", "2.2250738585072014e-308", "2.225073858507201e-308", "1e-400",
        "1e400", "inf", "-INFINITY", "nan", "nan(abc)", "0x1.8p3", "0x", "0x.8", "0x1p-1030", "1.5abc", "1e", "1e+",
        "+.5", ".e1", "\n7", "-0", "abc", "0.0e-999", "00012", "1.7976931348623157e308", "1.8e308", "-", "."]
LISTS = [",1,2,3,4,5,6,7,8,9,10,11,12,13,14,15", "1,2,3,4,5,6,7,8,9,10,11,12,13,14,15,",
         "1;2 3\t4,5,6,7,8,9,10,11,12,13,14,15", "1,,2,3,4,5,6,7,8,9,10,11,12,13,14,15", ",," + Z15 + ",0",
         "1, ,2,3,4,5,6,7,8,9,10,11,12,13,14", "  1 2 3 4 5 6 7 8 9 10 11 12 13 14 15  ", ",".join(["0.25"] * 28),
         ",".join(["1"] * 16), "", ";", "1;;2" + ",0" * 13, ",".join(["-1e-3"] * 27) + ";;"]


@pytest.fixture(scope="module")
def ref():
    return O.Reference()


def _same(a, b):
    return all((math.isnan(x) and math.isnan(y)) or (x == y and math.copysign(1, x) == math.copysign(1, y))
               for x, y in zip(a, b))


def test_parse_zernike_edge_cases_vs_reference(ref):
    for t in EDGE:
        text = t + "," + Z15
        got, ok = P.ParseZernikeCoefficients(text)
        want, wok = ref.parse_zernike(text)
        assert ok == wok and _same(got, want), (t, ok, wok, got[0], want[0])
    for text in LISTS:
        got, ok = P.ParseZernikeCoefficients(text)
        want, wok = ref.parse_zernike(text)
        assert ok == wok and _same(got, want), text


def test_parse_and_format_random_vs_reference(ref):
    rng = np.random.default_rng(3)
    seps = [",", ";", " ", "\t", ", ", " ;"]
    for _ in range(300):
        n = int(rng.choice([14, 15, 16, 27, 28, 29]))
        vals = rng.normal(0, 10.0 ** rng.integers(-12, 12), n)
        parts = []
        for v in vals:
            form = int(rng.integers(0, 4))
            parts.append(repr(float(v)) if form == 0 else ("%.3e" % v if form == 1 else
                                                            (float(v).hex() if form == 2 else "%g" % v)))
        text = "".join(p + str(rng.choice(seps)) for p in parts[:-1]) + parts[-1]
        got, ok = P.ParseZernikeCoefficients(text)
        want, wok = ref.parse_zernike(text)
        assert ok == wok and _same(got, want), text
        z = np.zeros(28)
        z[: min(n, 28)] = vals[:28]
        for sep in (",", ";"):
            assert P.FormatZernikeCoefficients(z, sep) == ref.format_zernike(z, sep)
    special = np.array([0.1, -1.2345678e-7, 1e20, 123456.5, 1234567.0, math.inf, -math.inf, -0.0, 5e-324,
                        1e-5, 0.0001, 999999.5, 9999995.0] + [0.0] * 15)
    assert P.FormatZernikeCoefficients(special) == ref.format_zernike(special)


def _small(zern_text, **kw):
    """A small GibsonLanniZernike request (hw 3, os 2, 3 planes, M 32) as both the reference's and this package's."""
    fields = dict(wavelengthNm=640.0, na=1.3, immersionIndex=1.518, sampleIndex=1.4, workingDistanceUm=170.0,
                  sampleDepthNm=300.0, pixelSizeNm=110.0, zStepNm=80.0, oversampling=2, kernelHalfWidthPx=3, nz=3,
                  pupilSamples=32, interpMode=2)
    fields.update(kw)
    z, _ = P.ParseZernikeCoefficients(zern_text)
    oreq = O.psf_request(zernike=z, **fields)
    req = P.PsfGeneratorRequest(model=P.GIBSON_LANNI_ZERNIKE, zernikeCoefficients=zern_text, **fields)
    return oreq, req


def test_disk_store_file_read_both_ways(ref, tmp_path):
    """The reference computes a kernel and stores it; LoadKernelFromDisk here reads its planes (the fingerprint and
    layout agree). Then a file written here for a request the reference has not computed is what the reference's
    ComputePsfKernelCache returns ("read from the disk cache"), planes bit for bit."""
    stamp = "%.17g" % (time.time() % 1.0)  # a coefficient text no memo holds
    oreq, req = _small("0.01,0.02,-0.03," + stamp + ",0,0,0,0,0,0,0,0,0,0,0")
    d1 = str(tmp_path / "ref")
    ok, log, planes, _ = ref.compute_psf_kernel_cache(oreq, req.zernikeCoefficients, d1, 13, 3)
    assert ok, log
    assert (tmp_path / "ref" / "psf_kernel.bin").is_file()
    P.SetPsfKernelDiskCacheDir(d1)
    c = P.LoadKernelFromDisk(req)
    assert c is not None and c.planes.shape == (3, 13, 13) and np.array_equal(c.planes, planes)
    assert c.oversampling == 2 and c.zStepNm == 80.0 and c.interpMode == 2
    other = P.PsfGeneratorRequest(**{**req.__dict__, "nz": 5})
    assert P.LoadKernelFromDisk(other) is None  # another request's fingerprint
    # written here, read by the reference
    oreq2, req2 = _small("0.04,0,0,0,0,0,0,0,0,0,0,0,0,0," + stamp, sampleDepthNm=0.0)
    fake = np.random.default_rng(9).random((3, 13, 13)).astype(np.float32)
    d2 = str(tmp_path / "ours")
    P.SetPsfKernelDiskCacheDir(d2)
    P.SaveKernelToDisk(req2, host.PsfKernelCache(fake, oversampling=2, zStepNm=80.0))
    ok, log, planes2, _ = ref.compute_psf_kernel_cache(oreq2, req2.zernikeCoefficients, d2, 13, 3)
    assert ok and any("read from the disk cache" in m for m in log), log
    assert np.array_equal(planes2, fake)
    P.SetPsfKernelDiskCacheDir("")


def test_memo_rules_with_cpu_compute(monkeypatch, tmp_path):
    calls = []

    def fake(req, ctx):
        calls.append(req.zernikeCoefficients)
        time.sleep(0.05)
        return host.PsfKernelCache(np.full((3, 5, 5), float(len(calls)), np.float32), oversampling=1,
                                   zStepNm=req.zStepNm, interpMode=req.interpMode), ""

    monkeypatch.setattr(P, "_compute_uncached", fake)
    P.SetPsfKernelDiskCacheDir("")
    P.SetPsfKernelMemoEntries(1)  # at least 2 are kept
    mk = lambda t, **kw: P.PsfGeneratorRequest(model=P.GIBSON_LANNI_ZERNIKE, zernikeCoefficients=t, nz=3,
                                               kernelHalfWidthPx=2, oversampling=1, **kw)
    tag = "%.17g" % time.time()
    a, b, c = mk(tag + ",1" + ",0" * 13), mk(tag + ",2" + ",0" * 13), mk(tag + ",3" + ",0" * 13)
    ok, ca, _ = P.ComputePsfKernelCache(a)
    log = []
    ok2, ca2, _ = P.ComputePsfKernelCache(P.PsfGeneratorRequest(**{**a.__dict__, "interpMode": 1}), log.append)
    assert ok and ok2 and len(calls) == 1 and ca2.serial == ca.serial and ca2.interpMode == 1 and ca.interpMode == 0
    assert log == ["PSF: kernel unchanged, reusing the one already computed."]
    P.ComputePsfKernelCache(b)
    P.ComputePsfKernelCache(a)       # a moves to the front
    P.ComputePsfKernelCache(c)       # b (the oldest) leaves
    assert len(calls) == 3
    P.ComputePsfKernelCache(a)
    assert len(calls) == 3
    P.ComputePsfKernelCache(b)
    assert len(calls) == 4
    # javaHome and interpMode are not part of the kernel
    assert P.SameKernel(a, P.PsfGeneratorRequest(**{**a.__dict__, "javaHome": "x", "interpMode": 3}))
    assert not P.SameKernel(a, P.PsfGeneratorRequest(**{**a.__dict__, "zStepNm": 100.0 + 1e-12}))
    # concurrent requesters of one new kernel: one computation, the others wait and take the memo entry
    d = mk(tag + ",4" + ",0" * 13)
    res = []
    th = [threading.Thread(target=lambda: res.append(P.ComputePsfKernelCache(d)[1].serial)) for _ in range(4)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    assert len(calls) == 5 and len(set(res)) == 1


def test_other_models_fail_without_a_jvm():
    ok, cache, err = P.ComputePsfKernelCache(P.PsfGeneratorRequest(model=P.RICHARDS_WOLF, nz=3))
    assert not ok and cache is None and "JVM" in err
