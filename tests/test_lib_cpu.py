"""CPU-side checks of the product: the C ABI library exports what include/iscg.h declares (no GPU calls), and the
host-side mirror of the reference's per-emitter arithmetic equals the compiled reference's output."""
import os
import re
import subprocess

import numpy as np
import pytest

import oracle as O
from insiliscope_amd import _lib, configs, host

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "iscg.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(iscg_\w+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    assert os.path.exists(_lib.LIB_PATH), "libiscg.so not built"
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True, text=True,
                         check=True).stdout
    exported = set(re.findall(r"\sT\s(iscg_\w+)", out))
    declared = declared_symbols()
    assert declared, "no declarations parsed"
    missing = [s for s in declared if s not in exported]
    assert not missing, f"declared but not exported: {missing}"


def test_ctypes_signatures_cover_header():
    assert set(declared_symbols()) == set(_lib.SIGNATURES), "SIGNATURES out of step with include/iscg.h"


def test_library_loads_without_gpu():
    L = _lib.lib()
    assert L.iscg_abi_version() == 1


def test_struct_layouts():
    import ctypes as C
    assert C.sizeof(_lib.Camera) == 56
    assert C.sizeof(_lib.PsfRequest) == C.sizeof(O.PsfRequest)
    assert host.FRAME_EMITTER_DTYPE.itemsize == 32
    for dt, st in ((host.WORLD_CELL_DTYPE, _lib.WorldCell), (host.NPC_DTYPE, _lib.Npc),
                   (host.LAMIN_DTYPE, _lib.LaminPoint), (host.WORLD_DYE_DTYPE, _lib.WorldDye),
                   (host.BLOCK_DYE_DTYPE, _lib.BlockDye), (host.WORLD_EVENT_DTYPE, _lib.WorldEvent)):
        assert dt.itemsize == C.sizeof(st), st
        for name in dt.names:
            assert dt.fields[name][1] == getattr(st, name).offset, (st, name)
    assert host.SPLAT_EMITTER_DTYPE.itemsize == 64


def test_struct_layouts_match_header(tmp_path):
    """Every ctypes struct has the size and field offsets of its C typedef (gcc compiles a probe of iscg.h)."""
    import ctypes as C
    import shutil
    import subprocess
    gcc = shutil.which("gcc")
    if gcc is None:
        pytest.skip("gcc not found")
    pairs = {"iscg_camera": _lib.Camera, "iscg_kernel_view": _lib.KernelView, "iscg_wf_desc": _lib.WfDesc,
             "iscg_wf_dep": _lib.WfDep, "iscg_wf_job": _lib.WfJob, "iscg_roi_model": _lib.RoiModel,
             "iscg_psf_request": _lib.PsfRequest, "iscg_world_cell": _lib.WorldCell, "iscg_npc": _lib.Npc,
             "iscg_lamin_point": _lib.LaminPoint, "iscg_world_geometry": _lib.WorldGeometry,
             "iscg_label": _lib.Label, "iscg_world_dye": _lib.WorldDye, "iscg_block_dye": _lib.BlockDye,
             "iscg_world_event": _lib.WorldEvent, "iscg_kinetics": _lib.Kinetics,
             "iscg_kinetics_history": _lib.KineticsHistory}
    lines = ["#include <stdio.h>", "#include <stddef.h>", '#include "iscg.h"', "int main(void) {"]
    for cname, py in pairs.items():
        lines.append(f'printf("{cname} <sizeof> %zu\\n", sizeof({cname}));')
        for fname, _ in py._fields_:
            lines.append(f'printf("{cname} {fname} %zu\\n", offsetof({cname}, {fname}));')
    lines += ["return 0;", "}"]
    src = tmp_path / "probe.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "probe"
    inc = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include")
    subprocess.run([gcc, "-std=c99", "-I", inc, str(src), "-o", str(exe)], check=True)
    got = {}
    for line in subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.splitlines():
        cname, what, val = line.split()
        got[(cname, what)] = int(val)
    for cname, py in pairs.items():
        assert got[(cname, "<sizeof>")] == C.sizeof(py), cname
        for fname, _ in py._fields_:
            assert got[(cname, fname)] == getattr(py, fname).offset, (cname, fname)


def test_host_dye_arguments():
    """The host mirrors of isc::Label and isc::KineticsHistory build their C arguments without a GPU: the label's
    fields in order, the history's segment count and arrays (n = 0 for none)."""
    lab = host.Label("PALM", host.Kinetics(0.5, 0.1), True, density=0.4)
    c = lab.to_c()
    assert (c.mode, c.preState, c.density, c.kin.activationRatePerSec, c.kin.onSec) == (1, 1, 0.4, 0.5, 0.1)
    h = host.KineticsHistory([0.0, 2.5, 6.0], [host.Kinetics(1.0), host.Kinetics(2.0), host.Kinetics(3.0)])
    assert h._c.n == 3 and h._c.tStart[1] == 2.5 and h._c.kin[2].activationRatePerSec == 3.0
    assert host.KineticsHistory([], [])._c.n == 0
    with pytest.raises(host.IscgError):
        host.KineticsHistory([0.0, 1.0], [host.Kinetics()])


def splat_fixture(golden):
    G, P = golden("splat_ref.npz"), golden("psf_ref.npz")
    ev = G["events"].view(host.BLINK_EVENT_DTYPE).reshape(-1)
    w, h = (int(x) for x in G["w_h"])
    return G, P, ev, w, h


def test_collect_frame_emitters_matches_reference(golden):
    G, P, ev, w, h = splat_fixture(golden)
    cache = host.PsfKernelCache(P["small_planes"], oversampling=4, zStepNm=50.0, interpMode=2)
    for f in range(3):
        got = host.CollectFrameEmitters(ev, f, 100.0, 1.0, psfCache=cache)
        want = G[f"emitters_{f}"].view(host.FRAME_EMITTER_DTYPE)
        assert got.tobytes() == want.tobytes()


def test_collect_gpu_emitters_matches_reference(golden):
    G, P, ev, w, h = splat_fixture(golden)
    cache = host.PsfKernelCache(P["small_planes"], oversampling=4, zStepNm=50.0, interpMode=2)
    for f in range(3):
        em = host.CollectFrameEmitters(ev, f, 100.0, 1.0, psfCache=cache)
        got = host.CollectGpuEmitters(em, cache)
        want = G[f"gpu_emitters_{f}"].view(host.SPLAT_EMITTER_DTYPE)
        assert got.tobytes() == want.tobytes()


def test_nearest_z_matches_oracle(orc):
    cache = host.PsfKernelCache(np.zeros((81, 5, 5), np.float32), oversampling=1, zStepNm=15.0)
    z = np.array([-2.0, -0.6, -0.0075, 0.0, 0.0074999, 0.0075, 0.6, 0.6075, 3.0, -0.5925])
    idx, cl = cache.NearestZIndex(z)
    for zi, i, c in zip(z, idx, cl):
        assert (i, c) == orc.nearest_z(81, 15.0, zi)


def test_synthetic_workloads_are_deterministic():
    ev1, s1 = configs.c4_events(seed=1, n_frames=5)
    ev2, s2 = configs.c4_events(seed=1, n_frames=5)
    assert ev1.tobytes() == ev2.tobytes() and np.array_equal(s1, s2)
    counts = np.diff(s1)
    assert counts.min() > 650 and counts.max() < 950  # Poisson(800)
    assert np.all((ev1["zNm"] >= -600) & (ev1["zNm"] <= 600))
    z = configs.zernike_noll(1, range(5, 16), amplitude=0.05)
    assert np.count_nonzero(z) == 11 and np.max(np.abs(z)) <= 0.05
    assert z[0] == 0 and z[4] == 0  # piston and defocus (Noll 1, 4) untouched


def test_noll_to_osa_parity_rule():
    """Noll's rule: even j -> cos (m >= 0 in OSA), odd j -> sin (m < 0); same radial order n."""
    def osa_nm(j):
        for n in range(10):
            for l in range(n + 1):
                if n * (n + 1) // 2 + l == j:
                    return n, -n + 2 * l
    def noll_n(j):
        n = 0
        while (n + 1) * (n + 2) // 2 < j:
            n += 1
        return n
    for noll, osa in configs.NOLL_TO_OSA.items():
        n, m = osa_nm(osa)
        assert n == noll_n(noll)
        if m != 0:
            assert (noll % 2 == 0) == (m > 0)


def test_host_calls_fail_loudly_without_gpu():
    """No CPU fallback: creating a context with no device raises."""
    try:
        import torch
        if torch.cuda.is_available():
            pytest.skip("a GPU is present")
    except ImportError:
        pass
    with pytest.raises(host.IscgError):
        host.Context(0)


def test_noise_setup_maps_equal_reference(golden):
    """NoiseSetup's camera maps (host code in the library: std::mt19937_64, uniform_real_distribution, Box-Muller)
    equal the compiled reference's bit for bit, for the whole sensor and a window of its rows."""
    G = golden("maps_ref.npz")
    s = host.NoiseSetup.Make(1234567, 48, 40, 100.0, 2.5, 2.0, 0.05, 1.6, 0.2)
    for got, want in zip((s.offsetMap, s.gainMap, s.readNoiseMap), G["maps_full"]):
        assert np.array_equal(got, want)
    assert s.noiseSeed == (1234567 ^ 0x9E3779B9) & 0xFFFFFFFF and s.pixelBase == 0
    w = host.NoiseSetup.Make(-77, 48, 12, 100.0, 2.5, 1.0 / 300.0, 3.0, 50.0, 1.5, sensorHeight=40, rowY0=17)
    for got, want in zip((w.offsetMap, w.gainMap, w.readNoiseMap), G["maps_window"]):
        assert np.array_equal(got, want)
    assert w.pixelBase == 17 * 48
    assert (w.gainMap >= np.float32(0.04 / 300.0) * 0.999).all() and (w.readNoiseMap >= 0).all()


def test_psf_geometry_helpers_equal_reference(golden):
    G = golden("maps_ref.npz")
    for args, want in zip(G["half_width_in"], G["half_width_out"]):
        assert host.PsfKernelHalfWidthPx(*args) == want
    for (o, p, m), want in zip(G["capped_in"], G["capped_out"]):
        assert host.CappedPsfOversampling(int(o), p, m) == want


def test_bucket_events_by_frame_vs_reference():
    """host.BucketEventsByFrame against the compiled reference's BucketEventsByFrame: events spanning frames,
    before the movie, after it, ending at 0, fractional, NaN and infinite times; identical lists in event order."""
    import oracle as O
    from insiliscope_amd import host
    rng = np.random.default_rng(11)
    n = 4000
    ev = np.zeros(n, host.BLINK_EVENT_DTYPE)
    ts = rng.uniform(-5.0, 60.0, n)
    ev["tStart"] = ts
    ev["tEnd"] = ts + rng.exponential(1.5, n)
    special = [(-1e300, 2.5), (3.0, np.inf), (np.nan, 4.0), (2.0, np.nan), (-np.inf, 0.0), (49.999, 50.0),
               (50.0, 51.0), (0.0, 0.0), (-0.5, 1e-300), (7.0, 3.0), (-2.0, 1e30)]
    for i, (a, b) in enumerate(special):
        ev["tStart"][i], ev["tEnd"][i] = a, b
    R = O.Reference()
    for nF in (0, 1, 50):
        start, idx = host.BucketEventsByFrame(ev, nF)
        counts, want = R.bucket_events(ev, nF)
        assert np.array_equal(np.diff(start), counts) and np.array_equal(idx, want), nF


def _tiff_script(ops, page_u16, page_f32, page_rows, rows):
    for op in ops:
        kind = op[0]
        if kind == "u16":
            assert page_u16(*op[1:]) == 0
        elif kind == "f32":
            assert page_f32(*op[1:]) == 0
        elif kind == "prows":
            assert page_rows(*op[1:]) == 0
        else:
            assert rows(*op[1:]) == 0


def test_tiff_writer_bytes_identical_to_reference(tmp_path):
    """The movie file (iscg_tiff_*) against the reference's TiffWriter compiled from cli/tiff_writer.h: the same
    calls (16-bit and float pages, descriptions stored in place and out of it, odd sizes that need the pad byte, a
    page written in strips: first rows, then the others), classic, forced BigTIFF and BigTIFF by expected size;
    identical files. The writer refuses rows outside a page."""
    import ctypes as C
    import oracle as O
    from insiliscope_amd import _lib, host
    R = O.Reference()
    L = _lib.lib()
    rng = np.random.default_rng(5)
    a = rng.integers(0, 65536, (23, 37), dtype=np.uint16)
    b = rng.integers(0, 65536, (8, 9), dtype=np.uint16)
    f = rng.normal(size=(5, 7)).astype(np.float32)
    s = rng.integers(0, 65536, (17, 11), dtype=np.uint16)
    def ops(u16, f32, prows, rows):
        return [("u16", u16(a), 37, 23, b""), ("u16", u16(b), 9, 8, b"abc"), ("u16", u16(b), 9, 8, b"abcdefg"),
                ("f32", f32(f), 7, 5, b"a float page with a longer description {\"k\": 1}"),
                ("prows", u16(s), 4, 11, 17, b"strips"), ("u16", u16(a), 37, 23, b"x"),
                ("rows", 4, 4, 9, u16(s[4:13])), ("rows", 4, 13, 4, u16(s[13:]))]
    for mode, (expected, big) in {"classic": (0, 0), "forced": (0, 1), "expected": (1 << 32, 0)}.items():
        keep = []
        def ptr(x):
            x = np.ascontiguousarray(x)
            keep.append(x)
            return x.ctypes.data
        pr, po = str(tmp_path / f"ref_{mode}.tif"), str(tmp_path / f"ours_{mode}.tif")
        rt = R.L.ref_tiff_open(pr.encode(), expected, big)
        _tiff_script(ops(ptr, ptr, None, None), lambda p, w, h, d: R.L.ref_tiff_page_u16(rt, p, w, h, d),
                     lambda p, w, h, d: R.L.ref_tiff_page_f32(rt, p, w, h, d),
                     lambda p, r, w, h, d: R.L.ref_tiff_page_rows_u16(rt, p, r, w, h, d),
                     lambda pg, y0, n, p: R.L.ref_tiff_rows_u16(rt, pg, y0, n, p))
        assert R.L.ref_tiff_close(rt) == 0
        t = C.c_void_p()
        assert L.iscg_tiff_open(po.encode(), expected, big, C.byref(t)) == 0
        _tiff_script(ops(ptr, ptr, None, None), lambda p, w, h, d: L.iscg_tiff_page_u16(t, p, w, h, d),
                     lambda p, w, h, d: L.iscg_tiff_page_f32(t, p, w, h, d),
                     lambda p, r, w, h, d: L.iscg_tiff_page_rows_u16(t, p, r, w, h, d),
                     lambda pg, y0, n, p: L.iscg_tiff_rows_u16(t, pg, y0, n, p))
        assert L.iscg_tiff_rows_u16(t, 9, 0, 1, ptr(s)) != 0 and b"lie outside page index 9" in L.iscg_tiff_error(t)
        L.iscg_tiff_destroy(t)
        ours, ref = open(po, "rb").read(), open(pr, "rb").read()
        assert ours == ref, mode
        assert ref[2] == (43 if big or expected else 42)
    assert host.TiffWriter.Bytes(512, 512, 2000, 2, 100) == R.L.ref_tiff_bytes(512, 512, 2000, 2, 100)
    # the host mirror: a stack of frames (the description on the first page) equals page-by-page writes
    frames = rng.integers(0, 4096, (3, 6, 10), dtype=np.uint16)
    w1 = host.TiffWriter(str(tmp_path / "stack.tif"))
    assert w1.Frames(frames, "movie") and w1.pages() == 3 and w1.Close()
    w2 = host.TiffWriter(str(tmp_path / "pages.tif"))
    assert all(w2.Page(fr, 10, 6, "movie" if i == 0 else "") for i, fr in enumerate(frames)) and w2.Close()
    assert open(tmp_path / "stack.tif", "rb").read() == open(tmp_path / "pages.tif", "rb").read()


@pytest.mark.skipif(not os.path.exists(os.path.join(ROOT, "oracle", "_ref", "libiscbridge.so")),
                    reason="libiscbridge.so not built (needs the reference checkout at build time)")
def test_engine_links_the_dropins():
    """The reference's engine built with this build's drop-ins: sim::SplatGpuHost's and sim::GpuSimulator's methods
    are defined by insiliscope_amd/integration (SplatGpuIscg.cpp, GpuSimIscg.cpp) and call libiscg's C ABI, which is
    the only thing the library leaves undefined; a short CPU movie renders through the engine."""
    import oracle as O
    path = O.BRIDGE_PATH
    defined = subprocess.run(["nm", "-DC", "--defined-only", path], capture_output=True, text=True, check=True).stdout
    for sym in ("sim::SplatGpuHost::RenderFrames", "sim::SplatGpuHost::SetKernels", "sim::MakeSplatGpuHost",
                "sim::GpuSimulator::Create", "sim::GpuSimulator::RenderFrames", "sim::RenderScopeMovie"):
        assert sym in defined, sym
    undefined = subprocess.run(["nm", "-D", "--undefined-only", path], capture_output=True, text=True,
                               check=True).stdout
    iscg = set(re.findall(r"\bU (iscg_\w+)", undefined))
    assert {"iscg_splat_render_records", "iscg_splat_set_kernels", "iscg_wf_images"} <= iscg
    adu, info = O.Bridge().movie({"size": 48, "frames": 2, "world-seed": 1, "seed": 1, "disk-cache": 0}, False,
                                 48 * 48 * 4)
    assert adu.shape == (2, 48, 48) and info["host"] == "cpu" and adu.min() > 0
