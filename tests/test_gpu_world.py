"""The device world (world.hip through iscg_world_*): the dye generators against the reference's frozen goldens
(npd / lmd lines and DyesInBlock digests of committed inputs, tests/golden/world_inputs.npz), the device jsm::cbrt /
hypot against the golden math lines, and every window query (SitesInWindow, EventsInWindow, ContinuousInWindow,
Density3dInWindow) against the compiled reference's World on the seed-1 cell field, bit for bit and in order."""
import json
import os
import struct

import numpy as np
import pytest

import oracle as O
from insiliscope_amd import host

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
HUGE = (-1e9, -1e9, 1e9, 1e9, -1e9, 1e9)
NPC_LINKS = [(0.0, 0.0, 2.0, 5.0), (10.0, 4.0, 1.0, 6.0)]
LAMIN_LINKS = [(0.0, 0.0, 0.0, 0.0), (19.3, 15.3, 0.5, 2.0)]


def bits(x):
    return struct.unpack("<Q", struct.pack("<d", float(x)))[0]


def fmt(v):
    return "%.17g" % (0.0 if v == 0 else v)


def seq_sum(col):
    s = 0.0
    for v in col:
        s += float(v)
    return s


def test_device_cbrt_hypot_vs_golden_math_lines(orc, ctx):
    """jsm::cbrt, the two- and three-argument hypot, exp and log on the device: the golden `m` lines' bits
    (MathCheck, tests/parity/parity_main.cpp:41-58; inputs from HashUnit(7, i, j, 0))."""
    g = json.load(open(os.path.join(GOLDEN, "rng_vectors.json")))
    cols = g["math_columns"]
    x, ay, ax, xe, xl, want = [], [], [], [], [], []
    for line in g["math"]:
        i = line["i"]
        u1, u2, u3 = orc.hash_unit(7, i, 0, 0), orc.hash_unit(7, i, 1, 0), orc.hash_unit(7, i, 2, 0)
        s = (8 * 3.141592653589793, 200.0, 1e5, 1e9)[i % 4]
        x.append((u1 - 0.5) * 2 * s)
        ay.append((u2 - 0.5) * s)
        ax.append((u3 - 0.5) * s)
        xe.append((u1 - 0.5) * 1500)
        xl.append(u2 * s)
        want.append({c: int(b, 16) for c, b in zip(cols, line["bits"])})
    x, ay, ax = np.array(x), np.array(ay), np.array(ax)
    cb = host.JsMath("cbrt", x, ctx)
    h2 = host.JsMath("hypot2", np.stack([ay, ax], 1), ctx)
    h3 = host.JsMath("hypot3", np.stack([ay, ax, x], 1), ctx)
    ex = host.JsMath("exp", np.array(xe), ctx)
    lg = host.JsMath("log", np.array(xl), ctx)
    for i, w in enumerate(want):
        assert bits(cb[i]) == w["cbrt"], i
        assert bits(h2[i]) == w["hypot"], i
        assert bits(h3[i]) == w["hypot3"], i
        assert bits(ex[i]) == w["exp"], i
        assert bits(lg[i]) == w["log"], i


def test_device_cbrt_hypot_special_and_random_vs_reference(ctx):
    """cbrt of zeros, subnormals, infinities, NaN and 2^20 log-uniform values of both signs; hypot of mixed scales,
    zeros, infinities and NaN: the compiled reference's jsm:: bits."""
    if not os.path.exists(O.REF_PATH):
        pytest.skip("compiled reference not built")
    R = O.Reference()
    rng = np.random.default_rng(5)
    v = np.concatenate([[0.0, -0.0, 5e-324, -5e-324, 2.2250738585072014e-308, 1e-310, np.inf, -np.inf, np.nan, 1.0,
                         8.0, -27.0, 1e300, -1e-300],
                        rng.choice([-1, 1], 1 << 20) * np.exp(rng.uniform(-700, 700, 1 << 20))])
    got = host.JsMath("cbrt", v, ctx)
    want = R.jsm_geom("cbrt", v)
    assert np.array_equal(got.view(np.uint64), want.view(np.uint64))
    a = np.concatenate([[0.0, 0.0, np.inf, 3.0, np.nan, 1e-320], rng.normal(size=1 << 16) * 10.0 **
                        rng.integers(-300, 300, 1 << 16)])
    b = np.concatenate([[0.0, -2.0, np.nan, -np.inf, 1.0, 1e-320], rng.normal(size=1 << 16)])
    c = np.concatenate([[0.0, 1.0, 2.0, 0.0, np.inf, 1e-320], rng.normal(size=1 << 16) * 1e5])
    for fn, args in (("hypot2", (a, b)), ("hypot3", (a, b, c))):
        got = host.JsMath(fn, np.stack(args, 1), ctx)
        want = R.jsm_geom(fn, *args)
        assert np.array_equal(got.view(np.uint64), want.view(np.uint64)), fn


def inputs():
    return np.load(os.path.join(GOLDEN, "world_inputs.npz"))


def one_cell_geometry(cx, cy, npcs=None, lamin=None, sites=None, reach=0.0, mts=()):
    """A geometry of one cell at the origin (packRot 0: cell-local = world) holding these elements."""
    cells = np.zeros(1, host.WORLD_CELL_DTYPE)
    cells[0] = (cx, cy, 0.0, 0.0, 0.0, 1e9)
    counts = [len(p) for p, _ in mts]
    g = host.WorldGeometry(cells=cells, mtStart=np.array([0, len(mts)], np.int32),
                           ptStart=np.concatenate([[0], np.cumsum(counts)]).astype(np.int64),
                           pts=np.concatenate([p for p, _ in mts]) if mts else np.zeros((0, 3)),
                           mtIndex=np.array([i for _, i in mts], np.int32) if mts else None)
    if npcs is not None:
        a = np.zeros(len(npcs), host.NPC_DTYPE)
        a["c"], a["n"], a["e1"], a["e2"] = npcs[:, 0:3], npcs[:, 3:6], npcs[:, 6:9], npcs[:, 9:12]
        a["cand"] = npcs[:, 12].astype(np.int32)
        g.npcs, g.npcStart = a, np.array([0, len(a)], np.int32)
    if lamin is not None:
        a = np.zeros(len(lamin), host.LAMIN_DTYPE)
        a["c"], a["n"] = lamin[:, 0:3], lamin[:, 3:6]
        a["cand"] = lamin[:, 6].astype(np.int32)
        g.lamin, g.laminStart = a, np.array([0, len(a)], np.int32)
    if sites is not None:
        g.npcSites, g.npcSiteReachUm = sites, reach
    return g


def test_envelope_dyes_vs_golden_lines(ctx):
    """NpcDyesInBlock / LaminDyesInBlock on the device (the walk's pore and lamina blocks, unrounded cell-local dyes)
    of the golden cases' first 40 pores / 200 lamina points, density 0.7, fluorescent fraction 0.9, both linkages:
    the golden npd / lmd lines (count, position sums in the reference's order, every 7th dye's pore, site, id and
    position)."""
    I = inputs()
    lines = open(os.path.join(GOLDEN, "world_golden.txt")).read().splitlines()
    checked = 0
    for ln in lines:
        t = ln.split()
        if t[0] not in ("npd", "lmd"):
            continue
        cid, v = t[1], int(t[4])
        seed, cx, cy, which = (int(x) for x in I[f"{cid}_meta"])
        rows = I[f"{cid}_rows"]
        if which == 1:
            g = one_cell_geometry(cx, cy, npcs=rows, sites=I[f"{cid}_sites"], reach=float(I[f"{cid}_reach"][0]))
        else:
            g = one_cell_geometry(cx, cy, lamin=rows)
        W = host.World(seed, g, ctx)
        links = NPC_LINKS if which == 1 else LAMIN_LINKS
        W.SetLabel(which, host.Label(density=0.7, fluorescentFraction=0.9, linkage=links[v]))
        d = W.BlockDyes(*HUGE)
        W.close()
        assert (d["structure"] == which).all()
        d = d[np.lexsort((d["k"], d["element"]))]
        got = [t[0], cid, str(cx), str(cy), str(v), str(len(d)), fmt(seq_sum(d["x"])), fmt(seq_sum(d["y"])),
               fmt(seq_sum(d["z"]))]
        assert got == t[:9], (cid, v)
        for j in range((len(t) - 9) // 16):
            r = d[7 * j]
            tok = t[9 + 16 * j: 9 + 16 * (j + 1)]
            assert tok[:4] == [str(int(r["element"])), str(int(r["k"])), str(int(r["n"])), str(int(r["id"]))], (cid, j)
            assert tok[4:7] == [fmt(r["x"]), fmt(r["y"]), fmt(r["z"])], (cid, v, j)
        checked += 1
    assert checked == 16


def test_mt_dyes_vs_reference_digests(ctx):
    """DyesInBlock on the device for every block of the committed microtubules (golden case m08, microtubules 0, 16,
    .., 112; frames built on the device): per-block counts and the sha256 of every dye's position, site and id equal
    the compiled reference's, for a nested-draw label (0.7, 0.9) and a flexible-binder linkage."""
    import hashlib
    I = inputs()
    labels = [(0.7, 0.9, (12.0, 0.0, 2.0, 5.0)), (1.0, 1.0, (10.0, 4.0, 1.0, 6.0))]
    starts = np.concatenate([[0], np.cumsum(I["mt_counts"])])
    seed, cx, cy = (int(v) for v in I["mt_meta"])
    mts = [(I["mt_pts"][starts[j]:starts[j + 1]], int(i)) for j, i in enumerate(I["mt_ids"])]
    W = host.World(seed, one_cell_geometry(cx, cy, mts=mts), ctx)
    bstart = I["mt_block_starts"]
    d_i = 0
    for li, (dens, ff, lk) in enumerate(labels):
        W.SetLabel(0, host.Label(density=dens, fluorescentFraction=ff, linkage=lk))
        d = W.BlockDyes(*HUGE)
        assert (d["structure"] == 0).all()
        # blocks in walk order: microtubule by microtubule, each one's blocks in order
        assert (np.diff(d["block"]) >= 0).all()
        for j, (_, i) in enumerate(mts):
            want_counts = I["mt_block_counts"][bstart[d_i]:bstart[d_i + 1]]
            mine = d[d["element"] == i]
            blocks = np.unique(mine["block"])
            assert len(mine) == want_counts.sum(), (li, i)
            got_counts = np.array([(mine["block"] == b).sum() for b in blocks])
            assert np.array_equal(got_counts, want_counts[want_counts > 0]), (li, i)
            a = np.stack([mine["x"], mine["y"], mine["z"]], 1).astype(np.float64)
            b = np.stack([mine["k"], mine["n"], mine["id"]], 1).astype(np.int64)
            assert hashlib.sha256(a.tobytes() + b.tobytes()).hexdigest() == I["mt_digests"][d_i], (li, i)
            if li == 0 and j == 0:
                want = I["mt_first_rows"]
                assert np.array_equal(a[: len(want)], want[:, :3]) and np.array_equal(b[: len(want)],
                                                                                    want[:, 3:6].astype(np.int64))
            d_i += 1
    W.close()


# ---- the window queries against the compiled reference's World on the seed-1 cell field -------------------------

needs_ref = pytest.mark.skipif(not os.path.exists(O.REF_PATH), reason="compiled reference not built")
FIELD = (0.0, 0.0, 25.6, 25.6)   # the widefield workload's 256 x 256 px field at 100 nm


def geometry_from(gd):
    return host.WorldGeometry.from_rows(**gd)


@pytest.fixture(scope="module")
def field(ctx):
    """The seed-1 cell field (laminDensity 20, so every structure carries dyes) around the 25.6 um field: the
    reference World and a device World over the same cells' geometry."""
    if not os.path.exists(O.REF_PATH):
        pytest.skip("compiled reference not built")
    R = O.Reference()
    w = R.world_new(1, {"laminDensity": 20})
    gd = R.world_geometry(w, *FIELD)
    print("field: cells", len(gd["cells"]), "microtubules", int(gd["mtStart"][-1]), "points", len(gd["pts"]),
          "pores", len(gd["npcRows"]), "lamina", len(gd["laminRows"]), flush=True)
    W = host.World(1, geometry_from(gd), ctx)
    yield R, w, W
    W.close()
    R.world_free(w)


LABELS = {
    # the reference's defaults: every structure DNA-PAINT at 0.7
    "defaults": {},
    # dSTORM microtubules with an initial ON phase and a flexible binder, PALM pores with a pre state and a
    # fluorescent fraction, WideField lamina
    "mixed": {0: dict(mode=0, kin=(0.5, 0.05, 1.0, 0.3, 0.5, 0.2), linkage=(12.0, 3.0, 2.0, 5.0)),
              1: dict(mode=1, kin=(0.2, 0.1, 0.5, 0.5, 0.4, 0.0), preState=1, fluorescentFraction=0.6),
              2: dict(mode=3, density=0.5)},
    # long dSTORM schedules (about 50 blinks per dye before bleaching): blocks of more than 4096 blinks
    "long": {0: dict(mode=0, kin=(5.0, 0.02, 0.05, 0.02, 0.5, 0.0)), 1: dict(mode=2, kin=(2.0, 0.1, 1.0, 1.0, 0.5,
                                                                                          0.0)),
             2: dict(mode=1, kin=(1.0, 0.05, 0.2, 0.05, 0.0, 0.0))},
}


def set_labels(R, w, W, name):
    for s in range(3):
        spec = dict(LABELS[name].get(s, {}))
        kin = spec.pop("kin", (0.01, 0.05, 1.0, 1.0, 0.5, 0.0))
        mode = spec.pop("mode", 2)
        pre = spec.pop("preState", 0)
        dens = spec.pop("density", 0.7)
        ff = spec.pop("fluorescentFraction", 1.0)
        lk = spec.pop("linkage", None)
        assert R.world_set_label(w, s, R.label_vector(dens, ff, mode, kin, pre, lk)) == 0
        W.SetLabel(s, host.Label(mode=("dSTORM", "PALM", "DnaPaint", "WideField")[mode], kin=host.Kinetics(*kin),
                                 preState=bool(pre), density=dens, fluorescentFraction=ff, linkage=lk))


def event_matrix(e):
    return np.stack([e["x"], e["y"], e["z"], e["tOn"], e["tOff"], e["brightness"], e["aux"], e["id"].astype(float),
                     e["structure"].astype(float), e["state"].astype(float)], 1)


def same_rows(got, want):
    return got.shape == want.shape and np.array_equal(got.view(np.uint64), want.view(np.uint64))


WINDOWS = [(0.0, 10.0, 3.0, 13.0, -1e9, 1e9), (-3.0, 12.0, 4.0, 18.5, 0.5, 2.25), (0.0, 0.0, 25.6, 25.6, 0.0, 3.2)]


@needs_ref
@pytest.mark.parametrize("label", ["defaults", "mixed"])
def test_sites_vs_reference(field, label):
    R, w, W = field
    set_labels(R, w, W, label)
    for win in WINDOWS:
        want = R.world_sites(w, win)
        d = W.SitesInWindow(*win)
        got = np.stack([d["x"], d["y"], d["z"], d["id"].astype(float), d["cx"].astype(float), d["cy"].astype(float),
                        d["mtIndex"].astype(float), d["k"].astype(float), d["n"].astype(float),
                        d["structure"].astype(float)], 1)
        print(label, win, "sites", len(want), np.bincount(want[:, 9].astype(int), minlength=3), flush=True)
        assert len(want) > 1000
        assert same_rows(got, want), (label, win)


@needs_ref
@pytest.mark.parametrize("label,times", [("defaults", [(0.0, 1.0), (3.5, 4.25)]),
                                         ("mixed", [(0.0, 2.0), (1.5, 3.0), (0.25, 0.26)]),
                                         ("long", [(0.0, 10.0), (4.0, 4.5)])])
def test_events_vs_reference(field, label, times):
    R, w, W = field
    set_labels(R, w, W, label)
    wins = WINDOWS[:2] if label == "long" else WINDOWS
    for win in wins:
        for t0, t1 in times:
            want = R.world_events(w, win, t0, t1)
            got = event_matrix(W.EventsInWindow(*win, t0, t1))
            print(label, win, (t0, t1), "events", len(want), flush=True)
            assert same_rows(got, want), (label, win, t0, t1)


@needs_ref
def test_long_blocks_take_the_merge_passes(field):
    """The 'long' label's blocks hold more blinks than one LDS sort takes (4096): the merge passes run, and the
    reference's order (tOn, then dye and blink order) holds across their seams."""
    R, w, W = field
    set_labels(R, w, W, "long")
    win, (t0, t1) = WINDOWS[0], (0.0, 10.0)
    got = W.EventsInWindow(*win, t0, t1)
    st = got["structure"]
    # one block's run: consecutive microtubule events with non-decreasing tOn
    runs, n = [], 1
    for i in range(1, len(got)):
        if st[i] == 0 and st[i - 1] == 0 and got["tOn"][i] >= got["tOn"][i - 1]:
            n += 1
        else:
            runs.append(n)
            n = 1
    runs.append(n)
    assert max(runs) > 4096


@needs_ref
@pytest.mark.parametrize("tmin", [0.0, 0.1, 5.0])
def test_continuous_vs_reference(field, tmin):
    R, w, W = field
    set_labels(R, w, W, "mixed")
    for win in WINDOWS:
        want = R.world_events(w, win, tmin, 0.0, continuous=True)
        got = event_matrix(W.ContinuousInWindow(*win, tmin))
        print(win, tmin, "continuous", len(want), flush=True)
        assert len(want) > 0
        assert same_rows(got, want), (win, tmin)


@needs_ref
@pytest.mark.parametrize("mask", [0xF, 0x1, 0x6])
def test_density3d_vs_reference(field, mask):
    """Density3dInWindow over the 25.6 um field at 100 x 100 x 50 nm voxels (256 x 256 x 64), per structure mask,
    and the one-plane DensityInWindow."""
    R, w, W = field
    set_labels(R, w, W, "mixed")
    win = (0.0, 0.0, 25.6, 25.6, 0.0, 3.2)
    tot, want = R.world_density3d(w, win, 256, 256, 64, mask)
    gt, got = W.Density3dInWindow(*win, 256, 256, 64, mask)
    print("density", hex(mask), tot, flush=True)
    assert gt == tot and np.array_equal(got, want)
    tot1, want1 = R.world_density3d(w, (1.0, 2.0, 20.0, 21.5, -1e9, 1e9), 190, 195, 1, 0xF)
    gt1, got1 = W.DensityInWindow(1.0, 2.0, 20.0, 21.5, -1e9, 1e9, 190, 195)
    assert gt1 == tot1 and np.array_equal(got1, want1[0])


@needs_ref
def test_queries_on_device_tensors(field):
    """A density written straight into a device tensor (the widefield path's input, resident in HBM) equals the host
    answer."""
    import torch
    R, w, W = field
    set_labels(R, w, W, "defaults")
    win = (0.0, 0.0, 25.6, 25.6, 0.0, 3.2)
    out = torch.zeros((64, 256, 256), dtype=torch.float32, device="cuda")
    tot, _ = W.Density3dInWindow(*win, 256, 256, 64, 0xF, out=out)
    torch.cuda.synchronize()
    tot_h, host_out = W.Density3dInWindow(*win, 256, 256, 64, 0xF)
    assert tot == tot_h and np.array_equal(out.cpu().numpy(), host_out)


# A kinetics history for every structure (World::SetKineticsHistory): three segments; per segment the kinetics of
# structures 0..3 (rate, on, off, bleach, cv, initial on). Segment 0 carries initial ON times, so dSTORM's walk of it
# crosses a change; segment 1 switches PALM pores and DNA-PAINT on harder; segment 2 slows everything.
HIST_T = [0.0, 0.8, 2.6]
HIST_KIN = [
    [(0.5, 0.05, 1.0, 0.3, 0.5, 1.2), (0.2, 0.1, 0.5, 0.5, 0.4, 0.0), (0.3, 0.2, 1.0, 1.0, 0.5, 0.0),
     (0.01, 0.05, 1.0, 1.0, 0.5, 0.0)],
    [(4.0, 0.02, 0.1, 0.05, 0.0, 0.0), (3.0, 0.05, 0.2, 0.1, 0.8, 0.0), (6.0, 0.03, 1.0, 1.0, 0.0, 0.0),
     (0.01, 0.05, 1.0, 1.0, 0.5, 0.0)],
    [(0.1, 0.3, 3.0, 0.5, 1.0, 2.0), (0.05, 0.4, 2.0, 0.9, 0.2, 0.0), (0.7, 0.5, 1.0, 1.0, 0.3, 0.0),
     (0.01, 0.05, 1.0, 1.0, 0.5, 0.0)],
]


@needs_ref
@pytest.mark.parametrize("label", ["mixed", "defaults", "long"])
def test_events_under_kinetics_history_vs_reference(field, label):
    """EventsInWindow and ContinuousInWindow under World::SetKineticsHistory (HistoryOf per structure: dSTORM /
    PALM through DyeScheduleHistory, DNA-PAINT through PersistentGen's pieces and the lookback of the longest ON of
    any segment) against the compiled reference's World; then cleared, the labels' own kinetics again."""
    R, w, W = field
    set_labels(R, w, W, label)
    flat = [k for seg in HIST_KIN for k in seg]
    assert R.world_set_history(w, HIST_T, [O.Kinetics(*k) for k in flat]) == 0
    W.SetKineticsHistory(HIST_T, [host.Kinetics(*k) for k in flat])
    try:
        for win in WINDOWS[:2]:
            for t0, t1 in [(0.0, 2.0), (2.2, 3.1), (0.75, 0.85)]:
                want = R.world_events(w, win, t0, t1)
                got = event_matrix(W.EventsInWindow(*win, t0, t1))
                print(label, win, (t0, t1), "history events", len(want), flush=True)
                assert same_rows(got, want), (label, win, t0, t1)
            want = R.world_events(w, win, 0.1, 0.0, continuous=True)
            got = event_matrix(W.ContinuousInWindow(*win, 0.1))
            assert same_rows(got, want), (label, win, "continuous")
    finally:
        assert R.world_set_history(w, [], []) == 0
        W.SetKineticsHistory([], [])
    want = R.world_events(w, WINDOWS[0], 0.0, 1.0)
    assert same_rows(event_matrix(W.EventsInWindow(*WINDOWS[0], 0.0, 1.0)), want)


def test_set_history_refuses_bad_input(ctx):
    W = host.World(3, one_cell_geometry(0, 0), ctx)
    good = [host.Kinetics()] * 8
    for t in ([0.5, 1.0], [0.0, 0.0], [0.0, float("inf")]):
        with pytest.raises(host.IscgError):
            W.SetKineticsHistory(t, good)
    with pytest.raises(host.IscgError):
        W.SetKineticsHistory([0.0, 1.0], [host.Kinetics(onSec=0.0)] + good[1:])
    W.SetKineticsHistory([0.0, 1.0], good)
    W.close()


# ---- the cell field through the mean field, end to end ---------------------------------------------------------

def _world_wf():
    import importlib.util
    spec = importlib.util.spec_from_file_location("mg", os.path.join(GOLDEN, "make_golden.py"))
    mg = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mg)
    return mg.WORLD_WF_WINDOW, mg.WORLD_WF_FOCI


WF_TOL = 1e-5  # relative L2 per focus image: the FP32 3D FFT path against the reference's FP32 per-focus path


def _focus_series_vs_fixture(ctx, golden, dens, fft):
    G = golden("world_wf_ref.npz")
    planes = golden("psf_ref.npz")["small_planes"]
    f0, df, nf = _world_wf()[1]
    cache = host.PsfKernelCache(planes, oversampling=4, zStepNm=50.0)
    wf = host.WidefieldStack(ctx)
    assert wf.Plan(cache, 64, 64, 64, df, 0, f0, nf, fft=fft), wf.err
    img = np.zeros(wf.image_shape, np.float32)
    assert wf.FocusSeries(dens, outImages=img), wf.err
    want = G["images"]
    errs = [float(np.sqrt(((img[f].astype(np.float64) - want[f]) ** 2).sum() / (want[f].astype(np.float64) ** 2).sum()))
            for f in range(nf)]
    print("cell-field window focus series, fft", fft, "rel L2 max %.3g" % max(errs), flush=True)
    assert max(errs) <= WF_TOL, errs


@pytest.mark.parametrize("fft", [(0, 0, 0), (128, 512, 128)])
def test_cell_field_window_focus_series_vs_reference(ctx, golden, fft):
    """The committed window of the real seed-1 cell field (tests/golden/world_wf_ref.npz: the compiled reference's
    Density3dInWindow of a 6.4 um square, 64 planes, every structure DNA-PAINT at 0.7) through the device mean field:
    the reference's WidefieldScene focus series of that density (8 foci, the small 65^2 x 9-plane PSF), per focus
    within 1e-5 relative L2, on the five-pass and the slab paths."""
    G = golden("world_wf_ref.npz")
    dens = np.zeros(int(np.prod(G["shape"])), np.float32)
    dens[G["idx"]] = G["val"]
    _focus_series_vs_fixture(ctx, golden, dens.reshape(tuple(G["shape"])), fft)


@needs_ref
def test_cell_field_density_on_device_into_the_mean_field(field, golden):
    """The same window generated on the device (World::Density3dInWindow into a device tensor) is the committed
    density bit for bit, and that tensor, left in HBM, goes straight into the focus series."""
    import torch
    R, w, W = field
    set_labels(R, w, W, "defaults")
    G = golden("world_wf_ref.npz")
    win = _world_wf()[0]
    out = torch.zeros((64, 64, 64), dtype=torch.float32, device="cuda")
    total, out = W.Density3dInWindow(*win, 64, 64, 64, 0xF, out=out)
    torch.cuda.synchronize()
    want = np.zeros(int(np.prod(G["shape"])), np.float32)
    want[G["idx"]] = G["val"]
    assert total == int(G["total"])
    assert np.array_equal(out.cpu().numpy().reshape(-1).view(np.uint32), want.view(np.uint32))
    _focus_series_vs_fixture(W.ctx, golden, out, (128, 512, 128))
