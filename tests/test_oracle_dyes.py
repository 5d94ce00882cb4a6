"""The oracle's dye schedules and JS-compatible math (oracle.c: orc_jsm_*, orc_label_*, orc_persistent_blinks)
pinned against the reference's own golden math lines (spec/golden/ref.txt `m`, MathCheck of
tests/parity/parity_main.cpp:41-58), fixtures the compiled reference wrote (tests/golden/dyes_ref.npz), the compiled
reference itself where it is built, and the distributions the schedules must follow."""
import json
import math
import os
import struct

import numpy as np
import pytest

import oracle as O

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
JS_PI = 3.141592653589793
LIMIT = 823549.6  # 2^19 pi/2: the end of the restated sin/cos reductions


def bits(x):
    return struct.unpack("<Q", struct.pack("<d", float(x)))[0]


def math_inputs(orc, i):
    """MathCheck's inputs of line i (tests/parity/parity_main.cpp:45-52)."""
    u1, u2 = orc.hash_unit(7, i, 0, 0), orc.hash_unit(7, i, 1, 0)
    s = (8 * JS_PI, 200.0, 1e5, 1e9)[i % 4]
    return (u1 - 0.5) * 2 * s, (u1 - 0.5) * 1500, u2 * s


def test_jsm_golden_math_lines(orc):
    g = json.load(open(os.path.join(GOLDEN, "rng_vectors.json")))
    cols = g["math_columns"]
    assert len(g["math"]) == 100
    checked = 0
    for line in g["math"]:
        x, xe, xl = math_inputs(orc, line["i"])
        want = {c: int(b, 16) for c, b in zip(cols, line["bits"])}
        assert bits(orc.jsm("exp", [xe])[0]) == want["exp"], line["i"]
        assert bits(orc.jsm("log", [xl])[0]) == want["log"], line["i"]
        if abs(x) <= LIMIT:
            assert bits(orc.jsm("sin", [x])[0]) == want["sin"], line["i"]
            assert bits(orc.jsm("cos", [x])[0]) == want["cos"], line["i"]
            checked += 1
        else:
            assert np.isnan(orc.jsm("cos", [x])[0])
    assert checked == 75


@pytest.mark.parametrize("fn", O.JSM_FUNCS)
def test_jsm_vs_reference_fixture(orc, golden, fn):
    """The reduction boundaries (n pi/2 +- ulps for n < 70, the 3 pi/4 and 2^19 pi/2 edges), exp's special cases and
    thresholds, log's subnormals and near-1 inputs: the compiled reference's bits."""
    G = golden("dyes_ref.npz")
    x, y = G["jsm_" + fn + "_x"], G["jsm_" + fn + "_y"]
    got = orc.jsm(fn, x)
    assert np.array_equal(got.view(np.uint64), y.view(np.uint64))


@pytest.mark.skipif(not os.path.exists(O.REF_PATH), reason="compiled reference not built")
def test_jsm_vs_compiled_reference_random(orc):
    R = O.Reference()
    rng = np.random.default_rng(2024)
    cases = {"sin": rng.uniform(-LIMIT, LIMIT, 20000), "cos": rng.uniform(-4, 4, 20000),
             "exp": rng.uniform(-746, 710, 20000), "log": np.exp(rng.uniform(-740, 709, 20000))}
    for fn, x in cases.items():
        assert np.array_equal(orc.jsm(fn, x).view(np.uint64), R.jsm(fn, x).view(np.uint64)), fn


def test_jsm_closed_forms(orc):
    """What the mathematics fixes: special values, identities to an ulp or two, and agreement with the platform's
    libm to one ulp (fdlibm's error bound) on ordinary arguments."""
    assert orc.jsm("exp", [0.0])[0] == 1.0 and orc.jsm("log", [1.0])[0] == 0.0
    assert orc.jsm("exp", [1.0])[0] == 2.718281828459045
    assert orc.jsm("log", [0.0])[0] == -math.inf and np.isnan(orc.jsm("log", [-1.0])[0])
    assert orc.jsm("exp", [800.0])[0] == math.inf and orc.jsm("exp", [-800.0])[0] == 0.0
    assert orc.jsm("sin", [0.0])[0] == 0.0 and orc.jsm("cos", [0.0])[0] == 1.0
    rng = np.random.default_rng(5)
    x = rng.uniform(-50, 50, 2000)
    for fn, ref in (("sin", np.sin), ("cos", np.cos), ("exp", np.exp)):
        a, b = orc.jsm(fn, x), ref(x)
        assert np.all(np.abs(a - b) <= np.spacing(np.abs(b)) + 1e-300), fn
    lx = np.exp(rng.uniform(-600, 600, 2000))
    assert np.all(np.abs(orc.jsm("log", lx) - np.log(lx)) <= np.spacing(np.abs(np.log(lx))))
    s, c = orc.jsm("sin", x), orc.jsm("cos", x)
    assert np.max(np.abs(s * s + c * c - 1)) < 4e-16
    # the log of an exact power of two is k ln2 (ln2_hi + ln2_lo)
    for k in (-1074, -1022, -3, 1, 7, 1023):
        assert abs(orc.jsm("log", [2.0 ** k])[0] - k * math.log(2)) <= abs(k) * 2e-16


LABEL = {c[0]: c for c in [
    ("dstorm_default", 0, 0, dict(), 1000.0),
    ("dstorm_multi", 0, 0, dict(activationRatePerSec=0.2, bleachProb=0.05, initialOnSec=3.0, photonCV=0.4), 300.0),
    ("dstorm_nocv", 0, 0, dict(activationRatePerSec=1.0, bleachProb=0.02, photonCV=0.0), 40.0),
    ("palm", 1, 1, dict(activationRatePerSec=0.05, bleachProb=0.3, photonCV=0.0), 500.0),
    ("palm_cv", 1, 0, dict(activationRatePerSec=0.5, bleachProb=0.001, offSec=0.2, photonCV=1.5), 8.0),
    ("palm_long", 1, 0, dict(activationRatePerSec=50.0, bleachProb=0.0, onSec=0.001, offSec=0.001), 1e9),
    ("palm_off", 1, 1, dict(activationRatePerSec=0.0), 100.0),
    ("paint", 2, 0, dict(), 100.0),
    ("widefield", 3, 0, dict(), 100.0),
]}
PERSIST = {c[0]: c for c in [
    ("win", dict(activationRatePerSec=0.3, onSec=0.2), 10.0, 20.0, None),
    ("win_nocv", dict(activationRatePerSec=2.0, photonCV=0.0), 0.0, 5.5, None),
    ("win_dense", dict(activationRatePerSec=45.0, onSec=0.01), 3.0, 4.0, None),
    ("win_empty", dict(activationRatePerSec=1.0), 5.0, 5.0, None),
    ("bins", dict(activationRatePerSec=0.7), 0.0, 0.0, (3, 9)),
    ("bins_dense", dict(activationRatePerSec=40.0), 0.0, 0.0, (0, 0)),
]}


@pytest.mark.parametrize("name", list(LABEL))
def test_label_schedule_vs_reference_fixture(orc, golden, name):
    G = golden("dyes_ref.npz")
    _, mode, pre, kw, t_max = LABEL[name]
    d = G["dyes"][:100] if name == "palm_long" else G["dyes"]
    kin = O.kinetics(**kw)
    c, b = orc.label_blinks(d, mode, kin, t_max)
    assert np.array_equal(c, G["label_" + name + "_counts"])
    assert np.array_equal(b.view(np.uint64), G["label_" + name + "_blinks"].view(np.uint64))
    has, w = orc.label_cont(d, mode, pre, kin)
    assert np.array_equal(has, G["cont_" + name + "_has"])
    assert np.array_equal((w * (has[:, None] != 0)).view(np.uint64), G["cont_" + name + "_w"].view(np.uint64))


@pytest.mark.parametrize("name", list(PERSIST))
def test_persistent_blinks_vs_reference_fixture(orc, golden, name):
    G = golden("dyes_ref.npz")
    _, kw, t0, t1, bins = PERSIST[name]
    c, b = orc.persistent_blinks(G["dyes"][:300], O.kinetics(**kw), t0, t1, bins)
    assert np.array_equal(c, G["persist_" + name + "_counts"])
    assert np.array_equal(b.view(np.uint64), G["persist_" + name + "_blinks"].view(np.uint64))


@pytest.mark.skipif(not os.path.exists(O.REF_PATH), reason="compiled reference not built")
def test_schedules_vs_compiled_reference_random_kinetics(orc):
    """Fresh addresses under a grid of kinetics, every mode: the compiled reference's blinks and windows."""
    R = O.Reference()
    rng = np.random.default_rng(99)
    d = np.stack([rng.integers(0, 2 ** 32, 300, dtype=np.uint64), rng.integers(0, 2 ** 32, 300, dtype=np.uint64),
                  rng.integers(0, 2 ** 32, 300, dtype=np.uint64)], 1).astype(np.uint32)
    for trial in range(12):
        kw = dict(activationRatePerSec=float(10 ** rng.uniform(-2, 1.5)), onSec=float(10 ** rng.uniform(-3, 0)),
                  offSec=float(10 ** rng.uniform(-2, 1)), bleachProb=float(rng.uniform(-0.1, 1.1)),
                  photonCV=float(rng.choice([0.0, rng.uniform(-0.2, 2)])),
                  initialOnSec=float(rng.choice([0.0, rng.uniform(0, 5)])))
        kin = O.kinetics(**kw)
        mode = trial % 4
        t_max = float(rng.uniform(1, 200))
        c1, b1 = orc.label_blinks(d, mode, kin, t_max)
        c2, b2 = R.label_blinks(d, mode, kin, t_max)
        assert np.array_equal(c1, c2) and np.array_equal(b1.view(np.uint64), b2.view(np.uint64)), kw
        h1, w1 = orc.label_cont(d, mode, trial % 3 == 0, kin)
        h2, w2 = R.label_cont(d, mode, trial % 3 == 0, kin)
        assert np.array_equal(h1, h2)
        assert np.array_equal((w1 * (h1[:, None] != 0)).view(np.uint64), (w2 * (h2[:, None] != 0)).view(np.uint64))
        t0 = float(rng.uniform(0, 30))
        t1 = t0 + float(rng.uniform(0, 5))
        c1, b1 = orc.persistent_blinks(d[:100], kin, t0, t1)
        c2, b2 = R.persistent_blinks(d[:100], kin, t0, t1)
        assert np.array_equal(c1, c2) and np.array_equal(b1.view(np.uint64), b2.view(np.uint64)), kw
        lo = int(rng.integers(0, 50))
        c1, b1 = orc.persistent_blinks(d[:100], kin, bins=(lo, lo + 3))
        c2, b2 = R.persistent_blinks(d[:100], kin, bins=(lo, lo + 3))
        assert np.array_equal(c1, c2) and np.array_equal(b1.view(np.uint64), b2.view(np.uint64)), kw


def addresses(n, seed):
    rng = np.random.default_rng(seed)
    return np.stack([rng.integers(0, 2 ** 32, n, dtype=np.uint64), rng.integers(0, 64, n),
                     rng.integers(0, 10 ** 6, n)], 1).astype(np.uint32)


def test_label_schedule_distributions(orc):
    """The schedule's laws: first activation Exp(rate) (P(t < T) = 1 - e^-rate T), ON times Exp(onSec), a mean-1
    log-normal brightness of CV photonCV, Geometric(bleachProb) blinks, dSTORM shifted by Exp(initialOnSec)."""
    n = 12000
    d = addresses(n, 1)
    c, b = orc.label_blinks(d, 0, O.kinetics(), 100.0)
    p = 1 - math.exp(-0.01 * 100)
    assert abs(c.mean() - p) < 5 * math.sqrt(p * (1 - p) / n)
    on = b[:, 1] - b[:, 0]
    assert abs(on.mean() - 0.05) < 5 * 0.05 / math.sqrt(len(on))
    assert abs(b[:, 2].mean() - 1) < 5 * 0.5 / math.sqrt(len(on))
    assert abs(b[:, 2].std() - 0.5) < 0.05
    # many blinks: the count is Geometric(p = 0.1) when tMax is out of reach
    c, b = orc.label_blinks(d[:4000], 1, O.kinetics(activationRatePerSec=1.0, bleachProb=0.1, photonCV=0), 1e9)
    assert abs(c.mean() - 10) < 5 * math.sqrt(90) / math.sqrt(4000)
    assert np.all(b[:, 2] == 1.0)
    # dSTORM: the initial ON window's length is the shift of every blink
    kin = O.kinetics(activationRatePerSec=0.5, initialOnSec=2.0, bleachProb=1.0)
    has, w = orc.label_cont(d[:3000], 0, 0, kin)
    assert has.all() and np.all(w[:, 2] == 2)
    assert abs(w[:, 1].mean() - 2.0) < 5 * 2.0 / math.sqrt(3000)
    c, b = orc.label_blinks(d[:3000], 0, kin, 1e9)
    assert np.all(c == 1)
    assert np.all(b[:, 0] > w[:, 1])
    assert abs((b[:, 0] - w[:, 1]).mean() - 2.0) < 5 * 2.0 / math.sqrt(3000)


def test_persistent_blinks_distributions(orc):
    """Per bin a Poisson(rate) count of blinks uniform in the bin, ON truncated at 20 onSec; the window form is the
    bins form filtered to the blinks overlapping (t0, t1)."""
    d = addresses(2000, 2)
    kin = O.kinetics(activationRatePerSec=3.0, onSec=0.3)
    c, b = orc.persistent_blinks(d, kin, bins=(0, 9))
    assert abs(c.mean() / 10 - 3.0) < 5 * math.sqrt(3.0 / 20000)
    assert np.all(np.floor(b[:, 0]) == b[:, 3])
    on = b[:, 1] - b[:, 0]
    assert on.max() <= 6.0 and abs(on.mean() - 0.3) < 0.01
    cw, bw = orc.persistent_blinks(d, kin, 4.5, 7.25)
    keep = (b[:, 0] < 7.25) & (b[:, 1] > 4.5)
    assert np.array_equal(bw[:, :3].view(np.uint64), b[keep, :3].view(np.uint64))
    # above m = 30 the count is the rounded normal approximation: mean m, variance m
    c, _ = orc.persistent_blinks(d, O.kinetics(activationRatePerSec=60.0), bins=(5, 5))
    assert abs(c.mean() - 60) < 5 * math.sqrt(60 / 2000) and abs(c.var() / 60 - 1) < 0.15


def test_block_hashes_vs_reference_fixture(orc, golden):
    """DyeH1 / NpcH1 / LaminH1 (core/src/dyes.cpp:175-222) of random blocks: Pcg4d(seed ^ salt, cx, cy, mtIndex or 0).a
    with the salts of core/src/dyes.h:33-41, as the compiled reference computes them."""
    G = golden("dyes_ref.npz")
    for seed, cx, cy, mt, dye, npc, lamin in G["h1"]:
        s = int(seed)
        assert orc.pcg4d(s ^ 0x9E3779B9, int(cx), int(cy), int(mt))[0] == dye
        assert orc.pcg4d(s ^ 0x4E504344, int(cx), int(cy), 0)[0] == npc
        assert orc.pcg4d(s ^ 0x4C414D44, int(cx), int(cy), 0)[0] == lamin
