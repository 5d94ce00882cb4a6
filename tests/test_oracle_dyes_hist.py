"""Kinetics histories (KineticsHistory, core/src/dyes.h:138-149; DyeScheduleHistory, HazardWalk and PersistentGen's
pieces, core/src/dyes.cpp:323-626): the fixtures the compiled reference wrote (tests/golden/dyes_hist_ref.npz) are
pinned by what a history must leave as it was -- every blink before the first change is the single-segment
schedule's bit for bit, every time bin before the first change or after the last is the single-segment bins' (both
from the oracle's restatement of DyeSchedule / PersistentGen) -- and by the compiled reference where it is built."""
import importlib.util
import os

import numpy as np
import pytest

import oracle as O

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
needs_ref = pytest.mark.skipif(not os.path.exists(O.REF_PATH), reason="compiled reference not built")


def cases():
    spec = importlib.util.spec_from_file_location("mg", os.path.join(GOLDEN, "make_golden.py"))
    mg = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mg)
    return ({c[0]: c for c in mg.DYE_HIST_LABEL_CASES}, {c[0]: c for c in mg.DYE_HIST_PERSIST_CASES})


LABEL_H, PERSIST_H = cases()
LABEL_KIN = dict(activationRatePerSec=0.01)   # the label's own kinetics: a history of two or more segments hides it
N_PERSIST = 80


def split(counts):
    off = np.zeros(len(counts) + 1, np.int64)
    off[1:] = np.cumsum(counts)
    return off


@pytest.mark.parametrize("name", [n for n, c in LABEL_H.items() if c[1] in (0, 1)])
def test_blinks_before_the_first_change_are_the_single_segment_schedule(orc, golden, name):
    """A change at T leaves the past: each dye's blinks whose cycles end before T, and the start and brightness of
    the last blink starting before T, are DyeSchedule's under segment 0's kinetics; so are the continuous windows
    that end before T (the initial ON, the pre state). A flipped walk (a wrong segment end, a scale taken as a
    rate) or a cycle summed in another order changes these."""
    G = golden("dyes_hist_ref.npz")
    _, mode, pre, t_start, kws, t_max = LABEL_H[name]
    T = t_start[1]
    d = G["dyes"]
    c0, b0 = orc.label_blinks(d, mode, O.kinetics(**kws[0]), t_max)
    o0, oh = split(c0), split(G["label_" + name + "_counts"])
    H = G["label_" + name + "_blinks"]
    checked = 0
    for i in range(len(d)):
        P, Hi = b0[o0[i]:o0[i + 1]], H[oh[i]:oh[i + 1]]
        k = int((P[:, 0] < T).sum())
        m = max(k - 1, 0)
        assert len(Hi) >= k, (name, i)
        assert np.array_equal(Hi[:m].view(np.uint64), P[:m].view(np.uint64)), (name, i)
        if k:
            assert Hi[k - 1, 0] == P[k - 1, 0] and Hi[k - 1, 2] == P[k - 1, 2], (name, i)
        checked += k
    has0, w0 = orc.label_cont(d, mode, pre, O.kinetics(**kws[0]))
    hh, wh = G["cont_" + name + "_has"], G["cont_" + name + "_w"]
    before = (has0 != 0) & (w0[:, 1] < T)
    assert np.array_equal(hh[before], has0[before])
    assert np.array_equal(wh[before].view(np.uint64), w0[before].view(np.uint64))
    print(name, "blinks before the change", checked, "windows", int(before.sum()))
    if kws[0].get("activationRatePerSec", 0.01) > 0:
        assert checked + int(before.sum()) > 0


@pytest.mark.parametrize("name", [n for n, c in PERSIST_H.items() if c[5] is not None])
def test_bins_off_the_changes_are_the_single_segment_bins(orc, golden, name):
    """Persistent bins: a bin wholly inside one segment is PersistentBlinksInBins of that segment's kinetics, bit
    for bit; a bin a change splits keeps its own draws (piece 0, at the kinetics of its start) up to the change, then
    each later segment's piece, numbered j | i << 24, in order."""
    G = golden("dyes_hist_ref.npz")
    _, t_start, kws, _, _, (lo, hi) = PERSIST_H[name]
    d = G["dyes"][:N_PERSIST]
    oh = split(G["persist_" + name + "_counts"])
    H = G["persist_" + name + "_blinks"]
    starts = np.array(t_start)
    for b in range(lo, hi + 1):
        seg = int(np.searchsorted(starts, float(b), side="right")) - 1
        inside = starts[(starts > b) & (starts < b + 1)]
        c, P = orc.persistent_blinks(d, O.kinetics(**kws[seg]), bins=(b, b))
        op = split(c)
        for i in range(len(d)):
            Hi = H[oh[i]:oh[i + 1]]
            Hb = Hi[Hi[:, 3] == b]
            Pi = P[op[i]:op[i + 1]]
            piece = Hb[:, 4].astype(np.int64) >> 24
            assert np.all(np.diff(piece) >= 0), (name, b, i)
            if len(inside) == 0:
                assert np.array_equal(Hb.view(np.uint64), Pi.view(np.uint64)), (name, b, i)
            else:
                keep = Pi[Pi[:, 0] < inside[0]]
                assert np.array_equal(Hb[piece == 0].view(np.uint64), keep.view(np.uint64)), (name, b, i)
                for p in np.unique(piece[piece > 0]):
                    s = seg + int(p)
                    a = starts[s]
                    e = starts[s + 1] if s + 1 < len(starts) and starts[s + 1] < b + 1 else b + 1.0
                    rows = Hb[piece == p]
                    assert np.all((rows[:, 0] >= a) & (rows[:, 0] < e)), (name, b, i, p)


@needs_ref
def test_fixture_is_the_compiled_references():
    R = O.Reference()
    G = np.load(os.path.join(GOLDEN, "dyes_hist_ref.npz"))
    d = G["dyes"]
    kin = O.kinetics(**LABEL_KIN)
    for name, mode, pre, t_start, kws, t_max in LABEL_H.values():
        hist = (t_start, [O.kinetics(**k) for k in kws])
        c, b = R.label_blinks(d, mode, kin, t_max, hist=hist)
        assert np.array_equal(c, G["label_" + name + "_counts"])
        assert np.array_equal(b.view(np.uint64), G["label_" + name + "_blinks"].view(np.uint64))
    for name, t_start, kws, t0, t1, bins in PERSIST_H.values():
        hist = (t_start, [O.kinetics(**k) for k in kws])
        c, b = R.persistent_blinks(d[:N_PERSIST], kin, t0, t1, bins, hist=hist)
        assert np.array_equal(b.view(np.uint64), G["persist_" + name + "_blinks"].view(np.uint64))


@needs_ref
def test_one_segment_is_its_kinetics_from_zero():
    """A history of one segment is that segment's kinetics from t = 0 (LabelSchedule's `one`), the label's own
    kinetics ignored."""
    R = O.Reference()
    G = np.load(os.path.join(GOLDEN, "dyes_hist_ref.npz"))
    d = G["dyes"]
    k = O.kinetics(activationRatePerSec=0.7, bleachProb=0.2, initialOnSec=0.3)
    for mode in (0, 1):
        c0, b0 = R.label_blinks(d, mode, k, 50.0)
        c1, b1 = R.label_blinks(d, mode, O.kinetics(**LABEL_KIN), 50.0, hist=([0.0], [k]))
        assert np.array_equal(c0, c1) and np.array_equal(b0.view(np.uint64), b1.view(np.uint64))
    c0, b0 = R.persistent_blinks(d[:50], k, 3.0, 9.0)
    c1, b1 = R.persistent_blinks(d[:50], O.kinetics(**LABEL_KIN), 3.0, 9.0, hist=([0.0], [k]))
    assert np.array_equal(b0.view(np.uint64), b1.view(np.uint64))
