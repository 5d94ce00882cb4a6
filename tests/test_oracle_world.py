"""The dye-site generators of the cell field (core/src/dyes.cpp:127-253, BuildMtFrames core/src/microtubules.cpp:701)
pinned to the reference's frozen goldens (spec/golden/ref.txt: `npd` / `lmd` NpcDyesInBlock / LaminDyesInBlock
lines, `npc` / `lam` pore and lamina lines, `mtp` / `mtl` microtubule and lattice-site lines, made by
tests/parity/parity_main.cpp:61-230), through the compiled reference (the checker the device world is compared with)
and through the committed inputs the GPU tests run on (tests/golden/world_inputs.npz)."""
import os

import numpy as np
import pytest

import oracle as O

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
needs_ref = pytest.mark.skipif(not os.path.exists(O.REF_PATH), reason="compiled reference not built")

# the two linkages of each envelope case (parity_main.cpp:134-136, 186-188): the structure's built-in one
# (NormLinkage(LegacyLinkage)) and a flexible binder
NPC_LINKS = [(0.0, 0.0, 2.0, 5.0), (10.0, 4.0, 1.0, 6.0)]
LAMIN_LINKS = [(0.0, 0.0, 0.0, 0.0), (19.3, 15.3, 0.5, 2.0)]


def fmt(v):
    """parity_main's PrintD: %.17g, a zero printed without its sign."""
    return "%.17g" % (0.0 if v == 0 else v)


def seq_sum(col):
    s = 0.0
    for v in col:
        s += float(v)
    return s


def golden_lines():
    lines = open(os.path.join(GOLDEN, "world_golden.txt")).read().splitlines()
    params, cases = {}, {}
    for ln in lines:
        t = ln.split()
        if t[0] == "params":
            params[t[1]] = {kv.split("=")[0]: float(kv.split("=")[1]) for kv in t[2:]}
        elif t[0] in ("npcs", "lamin", "cells"):
            cases[t[1]] = (int(t[2]), t[3])
    return lines, params, cases


def inputs():
    return np.load(os.path.join(GOLDEN, "world_inputs.npz"))


def envelope_tokens(kind, cid, cx, cy, v, dyes):
    """An npd / lmd line (parity_main.cpp:165-177, 217-229) from rows pore, k, n, id, pos, site, rigid, tip."""
    t = [kind, cid, str(cx), str(cy), str(v), str(len(dyes))]
    t += [fmt(seq_sum(dyes[:, 4])), fmt(seq_sum(dyes[:, 5])), fmt(seq_sum(dyes[:, 6]))]
    for r in dyes[::7]:
        t += [str(int(r[0])), str(int(r[1])), str(int(r[2])), str(int(r[3]))] + [fmt(x) for x in r[4:16]]
    return t


@needs_ref
def test_reference_reproduces_golden_envelope_dye_lines():
    """The compiled reference's NpcDyesInBlock / LaminDyesInBlock of the first 40 pores (200 lamina points) of each
    golden case cell, density 0.7, fluorescent fraction 0.9, both linkages: the golden npd / lmd lines, token for
    token (counts, position sums, every 7th dye with its site, binder ends and tip)."""
    R = O.Reference()
    lines, params, cases = golden_lines()
    checked = 0
    for ln in lines:
        t = ln.split()
        if t[0] not in ("npd", "lmd"):
            continue
        cid, cx, cy, v = t[1], int(t[2]), int(t[3]), int(t[4])
        seed, pname = cases[cid]
        which = 1 if t[0] == "npd" else 2
        rows, sites, _ = R.case_envelope(seed, params[pname], cx, cy, which)
        ids = np.arange(min(len(rows), 40 if which == 1 else 200), dtype=np.int32)
        links = NPC_LINKS if which == 1 else LAMIN_LINKS
        dyes = R.envelope_dyes(which, seed, cx, cy, rows, ids, sites, 0.7, 0.9, links[v])
        got = envelope_tokens(t[0], cid, cx, cy, v, dyes)
        assert got[: len(t)] == t, (cid, v)
        checked += 1
    assert checked == 16


def test_committed_inputs_match_golden_envelope_lines():
    """The committed pores and lamina points (the GPU tests' inputs) are the golden npc / lam lines' samples
    (every 23rd point: candidate, piece, centre, normal, e1, e2)."""
    lines, _, _ = golden_lines()
    I = inputs()
    checked = 0
    for ln in lines:
        t = ln.split()
        if t[0] not in ("npc", "lam"):
            continue
        rows = I[f"{t[1]}_rows"]
        header, per = (18, 14) if t[0] == "npc" else (13, 8)
        vals = 12 if t[0] == "npc" else 6
        samples = (len(t) - header) // per
        assert samples >= 2
        for j in range(samples):
            tok = t[header + j * per: header + (j + 1) * per]
            r = rows[23 * j]
            assert tok[0] == str(int(r[vals])) and tok[1] == str(int(r[vals + 1])), (t[1], j)
            assert tok[2:] == [fmt(x) for x in r[:vals]], (t[1], j)
            checked += 1
    assert checked >= 20


def mtp_tokens(cid, cx, cy, i, pts):
    t = ["mtp", cid, str(cx), str(cy), str(i), str(len(pts))]
    t += [fmt(seq_sum(pts[:, 0])), fmt(seq_sum(pts[:, 1])), fmt(seq_sum(pts[:, 2]))]
    for p in pts[::37]:
        t += [fmt(x) for x in p]
    t += [fmt(x) for x in pts[-1]]
    return t


def test_committed_microtubules_match_golden_mtp_lines():
    """The committed microtubule points (m08, microtubules 0, 16, .., 112) are the golden mtp lines: count, sums,
    every 37th point and the last."""
    lines, _, _ = golden_lines()
    I = inputs()
    starts = np.concatenate([[0], np.cumsum(I["mt_counts"])])
    cx, cy = int(I["mt_meta"][1]), int(I["mt_meta"][2])
    by_i = {int(ln.split()[4]): ln.split() for ln in lines if ln.startswith("mtp m08 ")}
    for j, i in enumerate(I["mt_ids"]):
        pts = I["mt_pts"][starts[j]:starts[j + 1]]
        assert mtp_tokens("m08", cx, cy, int(i), pts) == by_i[int(i)], i


@needs_ref
def test_reference_reproduces_golden_mtl_lines():
    """BuildMtFrames + MtSiteGeometry of the compiled reference on the committed microtubules: the golden mtl lines
    (parity_main.cpp:61-80: sites of a 0.1 um window at arc 0.5 um, phase 0.3, every linker draw 0.5; att, tip, dye
    of the first 13)."""
    R = O.Reference()
    lines, _, _ = golden_lines()
    I = inputs()
    starts = np.concatenate([[0], np.cumsum(I["mt_counts"])])
    cx, cy = int(I["mt_meta"][1]), int(I["mt_meta"][2])
    by_i = {int(ln.split()[4]): ln.split() for ln in lines if ln.startswith("mtl m08 ")}
    for j, i in enumerate(I["mt_ids"]):
        pts = I["mt_pts"][starts[j]:starts[j + 1]]
        total = R.mt_frames(pts)[0][-1]
        lenUm = min(0.1, total)
        startUm = min(max(0.0, 0.5), total - lenUm)
        sites = []
        for k in range(13):
            th, off = R.mt_protofilament(0.3, k)
            sNm = off
            while sNm < lenUm * 1000 - 1e-9:
                sites.append(R.mt_site_geometry(pts, startUm + sNm * 1e-3, th, 0.5, 0.5, 0.5))
                sNm += 8.0
        t = ["mtl", "m08", str(cx), str(cy), str(int(i)), str(len(sites))]
        for g in sites[:13]:
            t += [fmt(x) for x in g]
        assert t == by_i[int(i)], i


@needs_ref
def test_mt_digests_are_the_references_dyes_in_block():
    """The committed DyesInBlock digests (the GPU test's expectation) are what the compiled reference gives."""
    import importlib.util
    spec = importlib.util.spec_from_file_location("mg", os.path.join(GOLDEN, "make_golden.py"))
    mg = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mg)
    R = O.Reference()
    I = inputs()
    starts = np.concatenate([[0], np.cumsum(I["mt_counts"])])
    seed, cx, cy = (int(v) for v in I["mt_meta"])
    d = 0
    for dens, ff, lk in mg.WORLD_MT_LABELS:
        for j, i in enumerate(I["mt_ids"]):
            pts = I["mt_pts"][starts[j]:starts[j + 1]]
            total = R.mt_frames(pts)[0][-1]
            rows = [R.dyes_in_block(seed, cx, cy, int(i), pts, b, dens, ff, lk) for b in range(int(np.ceil(total)))]
            assert mg.world_mt_digest(np.concatenate(rows)) == I["mt_digests"][d]
            d += 1


def test_frames_closed_forms():
    """BuildMtFrames' invariants on the committed microtubules through the compiled reference: T unit, U unit and
    orthogonal to T, V = T x U, cum increasing by the segment lengths (a parallel-transport frame)."""
    if not os.path.exists(O.REF_PATH):
        pytest.skip("compiled reference not built")
    R = O.Reference()
    I = inputs()
    pts = I["mt_pts"][: I["mt_counts"][0]]
    cum, T, U, V = R.mt_frames(pts)
    seg = np.linalg.norm(np.diff(pts, axis=0), axis=1)
    assert np.allclose(np.diff(cum), seg, rtol=1e-14, atol=1e-15)
    assert np.allclose(np.linalg.norm(T, axis=1), 1, atol=1e-14)
    assert np.allclose(np.linalg.norm(U, axis=1), 1, atol=1e-14)
    assert np.abs((T * U).sum(axis=1)).max() < 1e-14
    assert np.allclose(V, np.cross(T, U), atol=1e-15)


def _mg():
    import importlib.util
    spec = importlib.util.spec_from_file_location("mg", os.path.join(GOLDEN, "make_golden.py"))
    mg = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mg)
    return mg


def test_cell_field_window_images_are_its_density_convolved(orc):
    """The end-to-end fixture's reference images (tests/golden/world_wf_ref.npz) are the oracle's FP64 direct
    per-focus convolution of its density with the resampled small PSF (PlanFocus's blend and clamp), to FP32
    rounding: the fixture pairs a density with its focus series, whatever produced either."""
    mg = _mg()
    G = np.load(os.path.join(GOLDEN, "world_wf_ref.npz"))
    P = np.load(os.path.join(GOLDEN, "psf_ref.npz"))
    dens = np.zeros(int(np.prod(G["shape"])), np.float32)
    dens[G["idx"]] = G["val"]
    dens = dens.reshape(tuple(G["shape"]))
    assert int(G["total"]) == int(dens.sum())
    f0, df, nf = mg.WORLD_WF_FOCI
    R = orc.wf_radius(65, 4, 1)
    K = np.stack([orc.wf_kernel(p, 4, 1, R) for p in P["small_planes"]])
    want = np.maximum(orc.wf_direct(dens, df, 0, K, 50.0, f0 + df * np.arange(nf)), 0.0)
    for f in range(nf):
        e = np.sqrt(((G["images"][f] - want[f]) ** 2).sum() / (want[f] ** 2).sum())
        assert e <= 1e-5, (f, e)


@needs_ref
def test_cell_field_window_fixture_is_the_references():
    """The committed window is the compiled reference's World density of the real seed-1 field and its
    WidefieldScene focus series, as written."""
    mg = _mg()
    R = O.Reference()
    w = R.world_new(1, {"laminDensity": 20})
    for s in range(3):
        assert R.world_set_label(w, s, R.label_vector()) == 0
    total, dens = R.world_density3d(w, mg.WORLD_WF_WINDOW, 64, 64, 64, 0xF)
    R.world_free(w)
    G = np.load(os.path.join(GOLDEN, "world_wf_ref.npz"))
    assert total == int(G["total"])
    idx = np.flatnonzero(dens)
    assert np.array_equal(idx, G["idx"]) and np.array_equal(dens.reshape(-1)[idx], G["val"])
    P = np.load(os.path.join(GOLDEN, "psf_ref.npz"))
    f0, df, nf = mg.WORLD_WF_FOCI
    imgs = R.wf_focus_series(dens, 50.0, P["small_planes"], 4, 50.0, f0 + df * np.arange(nf))
    assert np.array_equal(imgs.view(np.uint32), G["images"].view(np.uint32))
