"""Dye schedules on the GPU (iscg_label_blinks / iscg_label_cont / iscg_persistent_blinks / iscg_persistent_bins /
iscg_jsm behind host.LabelSchedule, LabelContWindows, PersistentBlinks, PersistentBlinksInBins, JsMath): bit for
bit the compiled reference's values (tests/golden/dyes_ref.npz, spec/golden/ref.txt math lines) and the oracle's."""
import json
import math
import os

import numpy as np
import pytest

import oracle as O
from insiliscope_amd import host

from test_oracle_dyes import LABEL, LIMIT, PERSIST, addresses, bits, math_inputs

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
MODES = {0: "dSTORM", 1: "PALM", 2: "DnaPaint", 3: "WideField"}


def label_of(mode, pre, kw):
    return host.Label(MODES[mode], host.Kinetics(**kw), bool(pre))


def split(counts):
    off = np.zeros(len(counts) + 1, np.int64)
    off[1:] = np.cumsum(counts)
    return off


def as_rows(rec, names):
    return np.stack([rec[n].astype(np.float64) for n in names], 1) if len(rec) else np.zeros((0, len(names)))


@pytest.mark.parametrize("fn", O.JSM_FUNCS)
def test_jsm_device_vs_reference_fixture(ctx, golden, fn):
    G = golden("dyes_ref.npz")
    x, y = G["jsm_" + fn + "_x"], G["jsm_" + fn + "_y"]
    print("jsm", fn, len(x), flush=True)
    got = host.JsMath(fn, x, ctx=ctx)
    assert np.array_equal(got.view(np.uint64), y.view(np.uint64))


def test_jsm_device_golden_math_lines(ctx, orc):
    g = json.load(open(os.path.join(GOLDEN, "rng_vectors.json")))
    cols = g["math_columns"]
    xs, xe, xl, want = [], [], [], []
    for line in g["math"]:
        x, e, lg = math_inputs(orc, line["i"])
        xs.append(x)
        xe.append(e)
        xl.append(lg)
        want.append({c: int(b, 16) for c, b in zip(cols, line["bits"])})
    s, c = host.JsMath("sin", xs, ctx=ctx), host.JsMath("cos", xs, ctx=ctx)
    e, lg = host.JsMath("exp", xe, ctx=ctx), host.JsMath("log", xl, ctx=ctx)
    for i, w in enumerate(want):
        assert bits(e[i]) == w["exp"] and bits(lg[i]) == w["log"], i
        if abs(xs[i]) <= LIMIT:
            assert bits(s[i]) == w["sin"] and bits(c[i]) == w["cos"], i
        else:
            assert np.isnan(c[i])


def test_jsm_device_vs_oracle_random(ctx, orc):
    rng = np.random.default_rng(8)
    for fn, x in {"sin": rng.uniform(-LIMIT, LIMIT, 30000), "cos": rng.uniform(-7, 7, 30000),
                  "exp": rng.uniform(-746, 710, 30000), "log": rng.uniform(0, 1, 30000)}.items():
        got = host.JsMath(fn, x, ctx=ctx)
        assert np.array_equal(got.view(np.uint64), orc.jsm(fn, x).view(np.uint64)), fn


@pytest.mark.parametrize("name", list(LABEL))
def test_label_schedule_bit_exact(ctx, golden, name):
    G = golden("dyes_ref.npz")
    _, mode, pre, kw, t_max = LABEL[name]
    d = G["dyes"][:100] if name == "palm_long" else G["dyes"]
    print("label", name, len(d), flush=True)
    off, rec = host.LabelSchedule(d, label_of(mode, pre, kw), t_max, ctx=ctx)
    assert np.array_equal(off, split(G["label_" + name + "_counts"]))
    want = G["label_" + name + "_blinks"]
    got = as_rows(rec, ("tOn", "tOff", "brightness"))
    assert np.array_equal(got.view(np.uint64), want.view(np.uint64))
    has, w = host.LabelContWindows(d, label_of(mode, pre, kw), ctx=ctx)
    assert np.array_equal(has, G["cont_" + name + "_has"] != 0)
    gw = as_rows(w, ("tOn", "tOff", "state", "aux")) * has[:, None]
    assert np.array_equal(gw.view(np.uint64), G["cont_" + name + "_w"].view(np.uint64))


@pytest.mark.parametrize("name", list(PERSIST))
def test_persistent_blinks_bit_exact(ctx, golden, name):
    G = golden("dyes_ref.npz")
    _, kw, t0, t1, bins = PERSIST[name]
    d = G["dyes"][:300]
    print("persistent", name, flush=True)
    kin = host.Kinetics(**kw)
    if bins is None:
        off, rec = host.PersistentBlinks(d, kin, t0, t1, ctx=ctx)
        got = np.concatenate([as_rows(rec, ("tOn", "tOff", "brightness")), np.zeros((len(rec), 2))], 1)
    else:
        off, rec = host.PersistentBlinksInBins(d, kin, bins[0], bins[1], ctx=ctx)
        got = as_rows(rec, ("tOn", "tOff", "brightness", "bin", "j"))
    assert np.array_equal(off, split(G["persist_" + name + "_counts"]))
    assert np.array_equal(got.view(np.uint64), G["persist_" + name + "_blinks"].view(np.uint64))


def test_capacity_and_empty(ctx, orc):
    """A capacity below the total: the total and every offset are right, the dyes whose records end within the
    capacity are written, the rest of out is untouched; no dyes: total 0."""
    d = addresses(3000, 4)
    lab = label_of(0, 0, dict(activationRatePerSec=0.3, bleachProb=0.1))
    off, rec = host.LabelSchedule(d, lab, 200.0, ctx=ctx)
    total = int(off[-1])
    assert total == len(rec) and total > 1000
    cap = total // 3
    offsets = np.zeros(len(d) + 1, np.int64)
    out = np.full(cap, -7.0, host.BLINK_DTYPE)
    tot = np.zeros(1, np.int64)
    dd = host.DyeAddresses(d[:, 0], d[:, 1].astype(np.int32), d[:, 2].astype(np.int32))
    ctx.check(ctx.L.iscg_label_blinks(ctx.h, dd.ctypes.data, len(d), 0, host.C.byref(lab.kin), None, 200.0,
                                      offsets.ctypes.data, out.ctypes.data, cap, tot.ctypes.data))
    assert tot[0] == total and np.array_equal(offsets, off)
    last = int(np.searchsorted(off, cap, side="right")) - 1  # dyes [0, last) end within cap
    n_in = int(off[last])
    assert np.array_equal(out[:n_in].view(np.uint8), rec[:n_in].view(np.uint8))
    assert np.all(out["tOn"][n_in:] == -7.0)
    off0, rec0 = host.LabelSchedule(np.zeros((0, 3), np.uint32), lab, 200.0, ctx=ctx)
    assert list(off0) == [0] and len(rec0) == 0
    # loops that would not end within a launch are refused on the host
    with pytest.raises(host.IscgError, match="time bins"):
        host.PersistentBlinksInBins(d[:4], host.Kinetics(activationRatePerSec=1.0), 0, 1 << 21, ctx=ctx)
    with pytest.raises(host.IscgError, match="activation rate"):
        host.PersistentBlinks(d[:4], host.Kinetics(activationRatePerSec=1e6), 0.0, 1.0, ctx=ctx)


def test_device_tensors_and_large_batch(ctx, orc):
    """One million dyes from device memory into device outputs: offsets and totals against the compiled reference
    (or the oracle's counts on a sample where the reference is not built), records of a sample bit-exact."""
    import torch
    n = 1 << 20
    d = addresses(n, 5)
    kw = dict(activationRatePerSec=0.2, bleachProb=0.05, initialOnSec=3.0, photonCV=0.4)
    kin = host.Kinetics(**kw)
    dd = torch.from_numpy(d.view(np.int32).reshape(-1)).cuda()
    offsets = torch.zeros(n + 1, dtype=torch.int64, device="cuda")
    total = torch.zeros(1, dtype=torch.int64, device="cuda")
    ctx.follow_torch()
    ctx.check(ctx.L.iscg_label_blinks(ctx.h, dd.data_ptr(), n, 0, host.C.byref(kin), None, 300.0, offsets.data_ptr(),
                                      None, 0, total.data_ptr()))
    torch.cuda.synchronize()
    cap = int(total.item())
    out = torch.zeros(cap * 3, dtype=torch.float64, device="cuda")
    ctx.check(ctx.L.iscg_label_blinks(ctx.h, dd.data_ptr(), n, 0, host.C.byref(kin), None, 300.0, offsets.data_ptr(),
                                      out.data_ptr(), cap, total.data_ptr()))
    torch.cuda.synchronize()
    off = offsets.cpu().numpy()
    rec = out.cpu().numpy().reshape(-1, 3)
    print("large batch total", cap, flush=True)
    okin = O.kinetics(**kw)
    if os.path.exists(O.REF_PATH):
        c, b = O.Reference().label_blinks(d, 0, okin, 300.0)
        assert np.array_equal(off, split(c))
        assert np.array_equal(rec.view(np.uint64), b.view(np.uint64))
    sample = np.random.default_rng(0).choice(n, 500, replace=False)
    c, b = orc.label_blinks(d[sample], 0, okin, 300.0)
    assert np.array_equal(off[sample + 1] - off[sample], c)
    got = np.concatenate([rec[off[i]:off[i + 1]] for i in sample])
    assert np.array_equal(got.view(np.uint64), b.view(np.uint64))


def test_block_hashes(ctx, golden):
    G = golden("dyes_ref.npz")["h1"]
    for seed in np.unique(G[:, 0])[:8]:
        rows = G[G[:, 0] == seed]
        assert np.array_equal(host.DyeH1(int(seed), rows[:, 1], rows[:, 2], rows[:, 3], ctx=ctx), rows[:, 4])
        assert np.array_equal(host.NpcH1(int(seed), rows[:, 1], rows[:, 2], ctx=ctx), rows[:, 5])
        assert np.array_equal(host.LaminH1(int(seed), rows[:, 1], rows[:, 2], ctx=ctx), rows[:, 6])


# ---- kinetics histories ------------------------------------------------------------------------------------------

from test_oracle_dyes_hist import LABEL_H, LABEL_KIN, N_PERSIST, PERSIST_H  # noqa: E402


def hist_of(t_start, kws):
    return host.KineticsHistory(t_start, [host.Kinetics(**k) for k in kws])


@pytest.mark.parametrize("name", list(LABEL_H))
def test_history_label_schedule_bit_exact(ctx, golden, name):
    """LabelSchedule under a KineticsHistory (DyeScheduleHistory / HazardWalk, core/src/dyes.cpp:336-471): blinks
    and continuous windows, the compiled reference's bits (tests/golden/dyes_hist_ref.npz)."""
    G = golden("dyes_hist_ref.npz")
    _, mode, pre, t_start, kws, t_max = LABEL_H[name]
    d = G["dyes"]
    print("history label", name, len(t_start), "segments", flush=True)
    lab = label_of(mode, pre, LABEL_KIN)
    h = hist_of(t_start, kws)
    off, rec = host.LabelSchedule(d, lab, t_max, ctx=ctx, hist=h)
    assert np.array_equal(off, split(G["label_" + name + "_counts"]))
    got = as_rows(rec, ("tOn", "tOff", "brightness"))
    assert np.array_equal(got.view(np.uint64), G["label_" + name + "_blinks"].view(np.uint64))
    has, w = host.LabelContWindows(d, lab, ctx=ctx, hist=h)
    assert np.array_equal(has, G["cont_" + name + "_has"] != 0)
    gw = as_rows(w, ("tOn", "tOff", "state", "aux")) * has[:, None]
    assert np.array_equal(gw.view(np.uint64), G["cont_" + name + "_w"].view(np.uint64))


@pytest.mark.parametrize("name", list(PERSIST_H))
def test_history_persistent_bit_exact(ctx, golden, name):
    """PersistentBlinks / PersistentBlinksInBins under a history (PersistentGen's pieces, core/src/dyes.cpp:559-626):
    the compiled reference's bits."""
    G = golden("dyes_hist_ref.npz")
    _, t_start, kws, t0, t1, bins = PERSIST_H[name]
    d = G["dyes"][:N_PERSIST]
    print("history persistent", name, flush=True)
    kin, h = host.Kinetics(**LABEL_KIN), hist_of(t_start, kws)
    if bins is None:
        off, rec = host.PersistentBlinks(d, kin, t0, t1, ctx=ctx, hist=h)
        got = np.concatenate([as_rows(rec, ("tOn", "tOff", "brightness")), np.zeros((len(rec), 2))], 1)
    else:
        off, rec = host.PersistentBlinksInBins(d, kin, bins[0], bins[1], ctx=ctx, hist=h)
        got = as_rows(rec, ("tOn", "tOff", "brightness", "bin", "j"))
    assert np.array_equal(off, split(G["persist_" + name + "_counts"]))
    assert np.array_equal(got.view(np.uint64), G["persist_" + name + "_blinks"].view(np.uint64))


def test_history_window_is_its_bins_filtered(ctx, golden):
    """The window form under a history is the bin form over the window's bins (from the lookback of the longest ON
    of any segment) keeping the blinks that overlap (t0, t1), in the same order."""
    G = golden("dyes_hist_ref.npz")
    d = G["dyes"][:N_PERSIST]
    kin = host.Kinetics(**LABEL_KIN)
    t_start = [0.0, 2.3, 5.0, 11.7]
    kws = [dict(activationRatePerSec=2.0), dict(activationRatePerSec=0.5, onSec=0.4),
           dict(activationRatePerSec=8.0, onSec=0.01), dict(activationRatePerSec=31.0, onSec=0.02, photonCV=0.0)]
    h = hist_of(t_start, kws)
    t0, t1 = 9.25, 14.5
    b0 = max(0, int(np.floor(t0 - 20 * 0.4)))
    off_w, win = host.PersistentBlinks(d, kin, t0, t1, ctx=ctx, hist=h)
    off_b, bins = host.PersistentBlinksInBins(d, kin, b0, int(np.floor(t1)), ctx=ctx, hist=h)
    assert len(win) > 100
    for i in range(len(d)):
        b = bins[off_b[i]:off_b[i + 1]]
        keep = b[(b["tOn"] < t1) & (b["tOff"] > t0)]
        w = win[off_w[i]:off_w[i + 1]]
        assert np.array_equal(as_rows(w, ("tOn", "tOff", "brightness")).view(np.uint64),
                              as_rows(keep, ("tOn", "tOff", "brightness")).view(np.uint64)), i


def test_history_one_segment_and_none(ctx, golden):
    """One segment is its kinetics from t = 0 (the label's ignored); an empty history is no history."""
    G = golden("dyes_hist_ref.npz")
    d = G["dyes"]
    k = dict(activationRatePerSec=0.7, bleachProb=0.2, initialOnSec=0.3)
    for mode in (0, 1):
        o0, r0 = host.LabelSchedule(d, label_of(mode, 0, k), 50.0, ctx=ctx)
        o1, r1 = host.LabelSchedule(d, label_of(mode, 0, LABEL_KIN), 50.0, ctx=ctx, hist=hist_of([0.0], [k]))
        o2, r2 = host.LabelSchedule(d, label_of(mode, 0, k), 50.0, ctx=ctx, hist=hist_of([], []))
        assert np.array_equal(o0, o1) and np.array_equal(r0.view(np.uint8), r1.view(np.uint8))
        assert np.array_equal(o0, o2) and np.array_equal(r0.view(np.uint8), r2.view(np.uint8))


def test_history_bad_segments_refused(ctx):
    d = addresses(10, 3)
    lab = label_of(0, 0, LABEL_KIN)
    for t_start in ([0.5, 1.0], [0.0, 2.0, 1.0], [0.0, float("nan")], [0.0, 0.0]):
        with pytest.raises(host.IscgError):
            host.LabelSchedule(d, lab, 10.0, ctx=ctx, hist=hist_of(t_start, [LABEL_KIN] * len(t_start)))
    with pytest.raises(host.IscgError):
        host.PersistentBlinks(d, host.Kinetics(), 0.0, 5.0, ctx=ctx,
                              hist=hist_of([0.0, 1.0], [LABEL_KIN, dict(activationRatePerSec=2e5)]))
