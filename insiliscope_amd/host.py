"""Host-side mirror of the reference renderer's GPU seams, over libiscg's C ABI.

Names, arguments and error behaviour follow the reference's C++ interfaces (Sim/ = adapter/inSiliScope/Simulation/):
calls return ``True``/``False`` and leave the reason in ``err`` where the reference returns ``bool`` with a
``std::string& err``; results come back as numpy arrays (host) or are written into caller tensors (device).

Array arguments may be numpy arrays (host memory, staged by the library) or torch tensors on the context's GPU
(used in place). There is no CPU fallback here: without libiscg.so every call raises.
"""
from __future__ import annotations

import ctypes as C
import dataclasses
import itertools
import sys
import threading
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from ._lib import IscgError

# ---- records -------------------------------------------------------------------------------------------------

# iscg_frame_emitter: sim::FrameEmitter (Sim/SMLMSimulation.h:158) as the splat takes it
FRAME_EMITTER_DTYPE = np.dtype([("xPx", "f8"), ("yPx", "f8"), ("photons", "f8"), ("zIndex", "i4"),
                                ("kernel", "i4")])
# iscg_splat_emitter: sim::GpuSplatEmitter (Sim/GpuSimD3D11.h:50-59), 64 bytes
SPLAT_EMITTER_DTYPE = np.dtype([("x0", "i4"), ("y0", "i4"), ("bx", "i4"), ("by", "i4"), ("plane", "i4"),
                                ("nTaps", "i4"), ("photons", "f4"), ("cutRadius2", "i4"), ("wx", "f4", 4),
                                ("wy", "f4", 4)])
# sim::BlinkEvent's hot-path fields (Sim/SMLMSimulation.h:82-105)
# iscg_blink_event: sim::BlinkEvent's image fields, 56 bytes (the device collect's input)
DEVICE_EVENT_DTYPE = np.dtype([("xUm", "f8"), ("yUm", "f8"), ("zNm", "f8"), ("tStart", "f8"), ("tEnd", "f8"),
                               ("brightness", "f8"), ("excite", "f8")])
BLINK_EVENT_DTYPE = np.dtype([("xUm", "f8"), ("yUm", "f8"), ("zNm", "f8"), ("tStart", "f8"), ("tEnd", "f8"),
                              ("brightness", "f8")])
assert FRAME_EMITTER_DTYPE.itemsize == 32 and SPLAT_EMITTER_DTYPE.itemsize == 64

INTERP = {"Nearest": 0, "Linear": 1, "Cubic": 2}


def _ptr(a):
    """(pointer, keep-alive) for a numpy array or torch tensor (None -> NULL)."""
    if a is None:
        return None, None
    if isinstance(a, np.ndarray):
        if not a.flags["C_CONTIGUOUS"]:
            a = np.ascontiguousarray(a)
        return a.ctypes.data, a
    if hasattr(a, "data_ptr"):  # torch tensor (device or host)
        if not a.is_contiguous():
            raise IscgError("tensor arguments must be contiguous")
        return a.data_ptr(), a
    raise TypeError(f"unsupported array type {type(a)!r}")


def _numel(a) -> int:
    """Element count of a numpy array or torch tensor (np.size does not count a tensor's elements)."""
    return int(a.numel()) if hasattr(a, "numel") else int(np.size(a))


def _arr(a, dtype):
    if a is None:
        return None
    if hasattr(a, "data_ptr"):
        return a
    return np.ascontiguousarray(a, dtype=dtype)


# ---- context ---------------------------------------------------------------------------------------------------

class Context:
    """One iscg_context: a device and its stream. Not re-entrant; one per thread (Sim/GpuSimD3D11.h:28-30)."""

    def __init__(self, device: int = 0):
        self.L = _lib.lib()
        h = C.c_void_p()
        if self.L.iscg_create(device, C.byref(h)) != 0:
            raise IscgError(self.L.iscg_last_error(None).decode())
        self.h = h
        self.device = device
        self._own_stream = self.L.iscg_stream(h) or 0
        self._cur_stream = self._own_stream

    def close(self):
        if getattr(self, "h", None):
            self.L.iscg_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def error(self) -> str:
        return self.L.iscg_last_error(self.h).decode()

    def check(self, rc: int, what: str = "") -> None:
        if rc != 0:
            raise IscgError(f"{what}: {self.error()}" if what else self.error())

    def synchronize(self):
        self.check(self.L.iscg_synchronize(self.h), "iscg_synchronize")

    @property
    def stream(self) -> int:
        return self.L.iscg_stream(self.h) or 0

    def set_stream(self, stream: int | None):
        """Enqueue later work on `stream` (a hipStream_t as int; None: the context's own stream)."""
        s = self._own_stream if stream is None else int(stream)
        if s != self._cur_stream:
            self.check(self.L.iscg_set_stream(self.h, s or None), "iscg_set_stream")
            self._cur_stream = s

    def follow_torch(self):
        """Once torch has initialised the GPU, enqueue on torch's current stream of this device, so tensors made or
        read by torch kernels are ordered with the library's work; the context's own stream otherwise."""
        torch = sys.modules.get("torch")
        if torch is not None and torch.cuda.is_initialized():
            self.set_stream(torch.cuda.current_stream(self.device).cuda_stream)
        else:
            self.set_stream(None)

    def inject_failure(self, count: int, message: str = "injected failure"):
        self.check(self.L.iscg_inject_failure(self.h, count, message.encode()))

    def set_noise_path(self, exact: bool):
        """exact=True: every noise pixel through the FP64 chain; False: the verified FP32 filter (same ADU)."""
        self.check(self.L.iscg_noise_set_path(self.h, 1 if exact else 0))

    def set_autotune(self, on: bool):
        """on=True (the default): the first call with a new shape key times the candidates of each tunable launch
        parameter and keeps the fastest (the results are identical for every candidate); False: fixed defaults."""
        self.check(self.L.iscg_autotune_set(self.h, 1 if on else 0), "iscg_autotune_set")

    def autotune_choices(self) -> dict:
        """The kept choices, {shape key: value}."""
        n = self.L.iscg_autotune_dump(self.h, None, 0)
        if n < 0:
            raise IscgError(self.error())
        buf = C.create_string_buffer(n + 1)
        self.L.iscg_autotune_dump(self.h, buf, n + 1)
        out = {}
        for line in buf.value.decode().splitlines():
            k, v = line.rsplit("=", 1)
            out[k] = int(v)
        return out

    def load_autotune(self, choices: dict | str) -> int:
        """Adds kept choices (a dict from autotune_choices, or its key=value text); returns how many."""
        text = choices if isinstance(choices, str) else "".join(f"{k}={v}\n" for k, v in choices.items())
        n = self.L.iscg_autotune_load(self.h, text.encode())
        if n < 0:
            raise IscgError(self.error())
        return n

    def noise_bounds(self) -> dict:
        out = (C.c_double * 3)()
        self.check(self.L.iscg_noise_bounds(self.h, out))
        return {"fp32_path": bool(out[0]), "gauss_bound": out[1], "threshold_rel_bound": out[2]}


_default: dict[int, Context] = {}


def default_context(device: int = 0) -> Context:
    if device not in _default:
        _default[device] = Context(device)
    return _default[device]


# ---- camera noise --------------------------------------------------------------------------------------------

@dataclass
class CameraNoiseParams:
    """sim::CameraNoiseParams (Sim/SMLMNoise.h:35-50), same defaults."""

    quantumEfficiency: float = 0.85
    darkCurrentElectrons: float = 0.0
    gainPhotonsPerAdu: float = 1.0
    readNoiseElectrons: float = 1.5
    emccd: bool = False
    emGain: float = 300.0
    cicElectrons: float = 0.002
    bitDepth: int = 16

    def to_c(self) -> _lib.Camera:
        return _lib.Camera(self.quantumEfficiency, self.darkCurrentElectrons, self.gainPhotonsPerAdu,
                           self.readNoiseElectrons, int(bool(self.emccd)), int(self.bitDepth), self.emGain,
                           self.cicElectrons)


def Pcg4d(a, b, c, d, ctx: Context | None = None):
    """isc::Pcg4d (core/src/rng.h:18-38) of arrays of addresses on the GPU: (n, 4) uint32 words."""
    ctx = ctx or default_context()
    arrs = [np.ascontiguousarray(np.asarray(x, np.int64) & 0xFFFFFFFF, np.uint32) for x in (a, b, c, d)]
    n = int(arrs[0].size)
    out = np.zeros((n, 4), np.uint32)
    ctx.follow_torch()
    ctx.check(ctx.L.iscg_pcg4d(ctx.h, *[x.ctypes.data for x in arrs], n, out.ctypes.data), "iscg_pcg4d")
    return out


def HashUnits(seed: int, cx, cy, k, ctx: Context | None = None):
    """isc::HashUnit (core/src/rng.h:40-44) of arrays of (cx, cy, k) on the GPU: float64 (0, 1) draws."""
    ctx = ctx or default_context()
    x = np.ascontiguousarray(np.asarray(cx, np.int64).astype(np.int32))
    y = np.ascontiguousarray(np.asarray(cy, np.int64).astype(np.int32))
    kk = np.ascontiguousarray(np.asarray(k, np.int64) & 0xFFFFFFFF, np.uint32)
    out = np.zeros(x.size)
    ctx.follow_torch()
    ctx.check(ctx.L.iscg_hash_units(ctx.h, seed & 0xFFFFFFFF, x.ctypes.data, y.ctypes.data, kk.ctypes.data, x.size,
                                    out.ctypes.data), "iscg_hash_units")
    return out


def HashStream(seed: int, cx: int, cy: int, base: int, n: int, ctx: Context | None = None):
    """The first n draws of isc::HashStream(seed, cx, cy, base) (core/src/rng.h:46-52) on the GPU."""
    ctx = ctx or default_context()
    out = np.zeros(n)
    ctx.follow_torch()
    ctx.check(ctx.L.iscg_hash_stream(ctx.h, seed & 0xFFFFFFFF, int(np.int32(cx)), int(np.int32(cy)),
                                     base & 0xFFFFFFFF, n, out.ctypes.data), "iscg_hash_stream")
    return out


@dataclass
class NoiseSetup:
    """The camera's maps and seeds for a random seed (NoiseSetup, Sim/ScopeMovie.cpp:973-996; the maps of
    Pixel*Map::Generate, Sim/SMLMNoise.cpp:89-140), made on the host by iscg_noise_setup. sensorHeight > 0: a
    window of rows rowY0 .. rowY0 + height - 1 of that sensor."""

    offsetMap: np.ndarray
    gainMap: np.ndarray
    readNoiseMap: np.ndarray
    noiseSeed: int
    pixelBase: int

    @classmethod
    def Make(cls, seed: int, width: int, height: int, offsetAdu: float, offsetStdAdu: float,
             gainPhotonsPerAdu: float, pixelGainStdFraction: float, readNoiseElectrons: float,
             pixelReadNoiseStdFraction: float, sensorHeight: int = 0, rowY0: int = 0) -> "NoiseSetup":
        L = _lib.lib()
        maps = [np.zeros((height, width), np.float32) for _ in range(3)]
        ns, pb = C.c_uint32(), C.c_uint32()
        rc = L.iscg_noise_setup(int(seed), width, height, sensorHeight, rowY0, offsetAdu, offsetStdAdu,
                                gainPhotonsPerAdu, pixelGainStdFraction, readNoiseElectrons, pixelReadNoiseStdFraction,
                                maps[0].ctypes.data, maps[1].ctypes.data, maps[2].ctypes.data, C.byref(ns),
                                C.byref(pb))
        if rc != 0:
            raise IscgError("iscg_noise_setup: bad geometry")
        return cls(maps[0], maps[1], maps[2], ns.value, pb.value)


def PsfKernelHalfWidthPx(halfWidthNm: float, pixelSizeNm: float, wavelengthNm: float, na: float) -> int:
    """PsfKernelHalfWidthPx (Sim/PsfGeneratorBridge.cpp:1000)."""
    return int(_lib.lib().iscg_psf_kernel_half_width_px(halfWidthNm, pixelSizeNm, wavelengthNm, na))


def CappedPsfOversampling(oversampling: int, pixelSizeNm: float, maxSampleNm: float) -> int:
    """CappedPsfOversampling (Sim/PsfGeneratorBridge.cpp:1011)."""
    return int(_lib.lib().iscg_capped_psf_oversampling(oversampling, pixelSizeNm, maxSampleNm))


def ApplyNoiseChainFrames(photons, cam: CameraNoiseParams, noiseSeed: int, frames, offsetMap=None, gainMap=None,
                          readNoiseMap=None, pixelBase: int = 0, out=None, ctx: Context | None = None):
    """ApplyNoiseChain (Sim/SMLMNoise.h:136) over a batch: photons (nFrames, height, width) float32 -> ADU uint16.
    frames: the frame id of each image (the noise stream). Device tensors in, device tensor out."""
    ctx = ctx or default_context()
    ph = _arr(photons, np.float32)
    n_frames, height, width = ph.shape
    fr = _arr(frames, np.uint32)
    if out is None:
        if hasattr(ph, "data_ptr") and getattr(ph, "is_cuda", False):
            import torch
            # int16 storage holds the uint16 ADU bits (view as uint16 on the host)
            out = torch.empty((n_frames, height, width), dtype=torch.int16, device=ph.device)
        else:
            out = np.zeros((n_frames, height, width), np.uint16)
    keep = [(_ptr(x)) for x in (ph, fr, _arr(offsetMap, np.float32), _arr(gainMap, np.float32),
                                _arr(readNoiseMap, np.float32), out)]
    p = [k[0] for k in keep]
    cc = cam.to_c()
    ctx.follow_torch()
    ctx.check(ctx.L.iscg_noise(ctx.h, p[0], width, height, C.byref(cc), p[2], p[3], p[4], noiseSeed & 0xFFFFFFFF,
                               p[1], n_frames, pixelBase, p[5]), "iscg_noise")
    return out


def ApplyNoiseChain(photonImage, width: int, height: int, cam: CameraNoiseParams, offsetMap=None, gainMap=None,
                    readNoiseMap=None, noiseSeed: int = 0, frame: int = 0, pixelBase: int = 0,
                    ctx: Context | None = None):
    """sim::ApplyNoiseChain (Sim/SMLMNoise.h:136) for one frame: photonImage (height*width floats) -> uint16 ADU.
    Maps whose size is not width*height are ignored, as the reference's size checks do (SMLMNoise.cpp:151-158)."""
    ph = np.ascontiguousarray(photonImage, np.float32).reshape(1, height, width)
    n = width * height

    def fit(m):
        return None if m is None or np.size(m) != n else np.ascontiguousarray(m, np.float32)

    out = ApplyNoiseChainFrames(ph, cam, noiseSeed, np.array([frame], np.uint32), fit(offsetMap), fit(gainMap),
                                fit(readNoiseMap), pixelBase, ctx=ctx)
    return out.reshape(height, width)


# ---- kernel caches and the splat -----------------------------------------------------------------------------

_serials = itertools.count(1)


@dataclass
class PsfHaloSpans:
    """sim::PsfHaloSpans (Sim/PsfGeneratorBridge.h:214-225)."""

    lo: np.ndarray          # int16 (nz, 2 camRad + 1)
    hi: np.ndarray
    radius2: np.ndarray     # int32 (nz,)
    camRad: int
    threshold: float = 0.0
    keptFraction: float = 0.0
    serial: int = field(default_factory=lambda: next(_serials))


@dataclass
class PsfKernelCache:
    """sim::PsfKernelCache (Sim/PsfGeneratorBridge.h:243-293): sum-1 planes of one kernel stack."""

    planes: object                  # (nz, size, size) float32, numpy or device tensor
    oversampling: int = 1
    zStepNm: float = 100.0
    interpMode: int = 0
    halo: PsfHaloSpans | None = None
    serial: int = field(default_factory=lambda: next(_serials))

    @property
    def nz(self) -> int:
        return int(self.planes.shape[0])

    @property
    def sizeOversampled(self) -> int:
        return int(self.planes.shape[1])

    @property
    def halfWidthOversampled(self) -> int:
        return (self.sizeOversampled - 1) // 2

    @property
    def blockSumWidth(self) -> int:
        return self.sizeOversampled + self.oversampling - 1

    @property
    def camRad(self) -> int:
        return self.halfWidthOversampled // self.oversampling

    @property
    def SplatSerial(self) -> int:
        """What a splat reads (Sim/PsfGeneratorBridge.h:273): the halo cut's serial, else the planes'."""
        return self.halo.serial if self.halo is not None else self.serial

    def NearestZIndex(self, zUm):
        """PsfKernelCache::NearestZIndex (Sim/PsfGeneratorBridge.cpp:40-60), vectorised: (index, clamped)."""
        z = np.asarray(zUm, np.float64)
        if self.nz <= 1 or self.zStepNm <= 0.0:
            return np.zeros(z.shape, np.int32), np.zeros(z.shape, bool)
        idx_f = z * 1000.0 / self.zStepNm + (self.nz - 1) / 2.0
        # std::lround: half away from zero
        idx = np.where(idx_f >= 0, np.floor(idx_f + 0.5), np.ceil(idx_f - 0.5)).astype(np.int64)
        clamped = (idx < 0) | (idx >= self.nz)
        return np.clip(idx, 0, self.nz - 1).astype(np.int32), clamped


class TiffWriter:
    """TiffWriter (cli/tiff_writer.h:24-199) over iscg_tiff_*: a little-endian baseline TIFF, one uncompressed grey
    page per frame (uint16 ADU or float32), classic or BigTIFF, the reference's bytes. Host code (no GPU needed);
    Frames() also takes a device tensor of frames."""

    @staticmethod
    def Bytes(w: int, h: int, pages: int, bytesPerSample: int, descBytes: int) -> int:
        return int(_lib.lib().iscg_tiff_bytes(w, h, pages, bytesPerSample, descBytes))

    def __init__(self, path: str, expectedBytes: int = 0, bigTiff: bool = False):
        self.L = _lib.lib()
        self.t = C.c_void_p()
        self.L.iscg_tiff_open(str(path).encode(), int(expectedBytes), int(bool(bigTiff)), C.byref(self.t))

    def _rc(self, rc: int) -> bool:
        return rc == 0

    def ok(self) -> bool:
        return bool(self.t) and not self.error()

    def error(self) -> str:
        return self.L.iscg_tiff_error(self.t).decode()

    def _info(self):
        p, b, g = C.c_uint64(), C.c_uint64(), C.c_int32()
        self.L.iscg_tiff_info(self.t, C.byref(p), C.byref(b), C.byref(g))
        return p.value, b.value, bool(g.value)

    def big(self) -> bool:
        return self._info()[2]

    def pages(self) -> int:
        return self._info()[0]

    def bytes(self) -> int:
        return self._info()[1]

    def Page(self, px, w: int, h: int, desc: str = "") -> bool:
        a = np.ascontiguousarray(px)
        if a.size != w * h:
            raise ValueError("a page holds w * h samples")
        if a.dtype == np.uint16:
            return self._rc(self.L.iscg_tiff_page_u16(self.t, a.ctypes.data, w, h, desc.encode()))
        a = np.ascontiguousarray(a, np.float32)
        return self._rc(self.L.iscg_tiff_page_f32(self.t, a.ctypes.data, w, h, desc.encode()))

    def PageRows(self, px, rows: int, w: int, h: int, desc: str = "") -> bool:
        a = np.ascontiguousarray(px, np.uint16)
        if a.size < min(rows, h) * w:
            raise ValueError("fewer samples than rows * w")
        return self._rc(self.L.iscg_tiff_page_rows_u16(self.t, a.ctypes.data, rows, w, h, desc.encode()))

    def Rows(self, page: int, y0: int, rows: int, px) -> bool:
        a = np.ascontiguousarray(px, np.uint16)
        return self._rc(self.L.iscg_tiff_rows_u16(self.t, page, y0, rows, a.ctypes.data))

    def Frames(self, frames, firstDesc: str = "") -> bool:
        """Every frame of a (n, h, w) uint16 array or device tensor as a page, the description on the first."""
        n, h, w = frames.shape
        p, keep = _ptr(frames if hasattr(frames, "data_ptr") else np.ascontiguousarray(frames, np.uint16))
        rc = self.L.iscg_tiff_frames_u16(self.t, p, n, w, h, firstDesc.encode())
        del keep
        return self._rc(rc)

    def Close(self) -> bool:
        return self._rc(self.L.iscg_tiff_close(self.t))

    def __del__(self):
        try:
            if self.t:
                self.L.iscg_tiff_destroy(self.t)
                self.t = C.c_void_p()
        except Exception:
            pass


_halo_memo: list = []   # ((serial, threshold + 10 interpMode), PsfHaloSpans), newest first, at most 8
_halo_lock = threading.Lock()


def WithHaloCut(cache: PsfKernelCache, threshold: float, ctx: Context | None = None) -> PsfKernelCache:
    """sim::WithHaloCut (Sim/PsfGeneratorBridge.cpp:1188-1243): a copy of cache whose splats leave out every camera
    pixel whose share of an emitter's photons is below threshold (the spans computed on the GPU, iscg_halo_spans);
    threshold <= 0 or a kernel under one camera pixel: the whole square. Spans are kept per (kernel serial,
    threshold, interpolation), the last 8."""
    out = dataclasses.replace(cache, halo=None)
    os_ = max(1, cache.oversampling)
    camRad = cache.halfWidthOversampled // os_
    if not threshold > 0.0 or camRad < 1 or camRad > 16000:
        return out
    key = (cache.serial, threshold + 10.0 * int(cache.interpMode))
    with _halo_lock:
        for k, spans in _halo_memo:
            if k == key:
                out.halo = spans
                return out
    ctx = ctx or default_context()
    nz, n = cache.nz, 2 * camRad + 1
    lo = np.zeros((nz, n), np.int16)
    hi = np.zeros((nz, n), np.int16)
    r2 = np.zeros(nz, np.int32)
    kept = C.c_double()
    planes, keep = _ptr(_arr(cache.planes, np.float32))
    ctx.follow_torch()
    rc = ctx.L.iscg_halo_spans(ctx.h, planes, nz, cache.sizeOversampled, os_, int(cache.interpMode), float(threshold),
                               lo.ctypes.data, hi.ctypes.data, r2.ctypes.data, C.byref(kept))
    del keep
    if rc < 0:
        raise IscgError("iscg_halo_spans: " + ctx.error())
    spans = PsfHaloSpans(lo, hi, r2, camRad=camRad, threshold=float(threshold))
    spans.keptFraction = kept.value
    spans.serial = next(_serials)
    with _halo_lock:
        _halo_memo.insert(0, (key, spans))
        del _halo_memo[8:]
    out.halo = spans
    return out


def _to_long(x):
    """static_cast<long>(std::floor(x)) as x86-64 evaluates it: values outside the long range, and NaN, give
    LONG_MIN (the conversion's "integer indefinite")."""
    f = np.floor(np.asarray(x, np.float64))
    ok = np.isfinite(f) & (f >= -2.0 ** 63) & (f < 2.0 ** 63)
    return np.where(ok, np.where(ok, f, 0.0).astype(np.int64), np.iinfo(np.int64).min)


def BucketEventsByFrame(events, nFrames: int):
    """sim::BucketEventsByFrame (Sim/SMLMSimulation.cpp:203-217): for each frame f < nFrames the indices of the
    events that overlap it (tEnd > 0, tStart < nFrames, f in [floor(tStart), floor(tEnd)] clipped to the movie), in
    event order. Returns (start, idx): frame f's list is idx[start[f]:start[f + 1]] (start int64, idx uint32)."""
    ev = np.asarray(events, dtype=BLINK_EVENT_DTYPE)
    n = max(0, int(nFrames))
    ts, te = ev["tStart"], ev["tEnd"]
    keep = (te > 0.0) & (ts < float(n))
    f0 = np.maximum(0, _to_long(ts))
    f1 = np.minimum(n - 1, _to_long(te))
    cnt = np.where(keep & (f1 >= f0), f1 - np.minimum(f0, f1) + 1, 0).astype(np.int64)
    which = np.repeat(np.arange(len(ev), dtype=np.int64), cnt)
    first = np.repeat(np.cumsum(cnt) - cnt, cnt)
    frame = np.repeat(f0, cnt) + (np.arange(len(which), dtype=np.int64) - first)
    order = np.argsort(frame, kind="stable")
    start = np.zeros(n + 1, np.int64)
    start[1:] = np.cumsum(np.bincount(frame, minlength=n)[:n]) if n else []
    return start, which[order].astype(np.uint32)


def CollectFrameEmitters(events, frameIndex: int, pixelSizeNm: float, photonsPerBlink: float,
                         driftOffsetXPx: float = 0.0, driftOffsetYPx: float = 0.0,
                         psfCache: PsfKernelCache | None = None, globalZOffsetUm: float = 0.0, kernel: int = 0):
    """sim::CollectFrameEmitters (Sim/SMLMSimulation.cpp:55-101) without an illumination field: the frame's emitters
    in event order as iscg_frame_emitter records (the same FP64 operations, vectorised)."""
    ev = np.asarray(events, dtype=BLINK_EVENT_DTYPE)
    f = float(frameIndex)
    ov = np.minimum(f + 1.0, ev["tEnd"]) - np.maximum(f, ev["tStart"])
    keep = ov > 0.0
    ev, ov = ev[keep], np.minimum(ov[keep], 1.0)
    out = np.zeros(len(ev), FRAME_EMITTER_DTYPE)
    site_x = ev["xUm"] * 1000.0 / pixelSizeNm
    site_y = ev["yUm"] * 1000.0 / pixelSizeNm
    out["xPx"] = site_x + driftOffsetXPx
    out["yPx"] = site_y + driftOffsetYPx
    ph = photonsPerBlink * ov
    b = ev["brightness"]
    out["photons"] = np.where(b != 1.0, ph * b, ph)
    if psfCache is not None:
        out["zIndex"], _ = psfCache.NearestZIndex(ev["zNm"] / 1000.0 - globalZOffsetUm)
    out["kernel"] = kernel
    return out


def _device_events(events):
    ev = np.asarray(events)
    if ev.dtype == DEVICE_EVENT_DTYPE:
        return np.ascontiguousarray(ev)
    e7 = np.zeros(len(ev), DEVICE_EVENT_DTYPE)
    for k in ("xUm", "yUm", "zNm", "tStart", "tEnd", "brightness"):
        e7[k] = ev[k]
    e7["excite"] = ev["excite"] if "excite" in (ev.dtype.names or ()) else 1.0
    return e7


def BucketEventsByFrameGpu(events, nFrames: int, ctx: Context | None = None):
    """sim::BucketEventsByFrame on the GPU (iscg_bucket_events): (start int64 [nFrames + 1], idx uint32), as
    BucketEventsByFrame returns them."""
    ctx = ctx or default_context()
    ev = _device_events(events)
    n = max(0, int(nFrames))
    start = np.zeros(n + 1, np.uint32)
    total = C.c_uint64(0)
    cap = 4 * len(ev) + 16
    idx = np.zeros(cap, np.uint32)
    ctx.follow_torch()
    rc = ctx.L.iscg_bucket_events(ctx.h, ev.ctypes.data if len(ev) else None, len(ev), n, start.ctypes.data,
                                  idx.ctypes.data, cap, C.byref(total))
    if rc != 0 and total.value > cap:
        idx = np.zeros(total.value, np.uint32)
        rc = ctx.L.iscg_bucket_events(ctx.h, ev.ctypes.data, len(ev), n, start.ctypes.data, idx.ctypes.data,
                                      total.value, C.byref(total))
    ctx.check(rc, "iscg_bucket_events")
    return start.astype(np.int64), idx[: total.value]


def CollectFrameEmittersBatch(events, bucketStart, bucketIdx, frameIndex, width: int, height: int,
                              pixelSizeNm: float, photonsPerBlink: float, drift=None,
                              psfCache: PsfKernelCache | None = None, globalZOffsetUm: float = 0.0, kernel: int = 0,
                              illumination=None, ctx: Context | None = None):
    """sim::CollectFrameEmitters (Sim/SMLMSimulation.cpp:55-101) for a batch of frames on the GPU
    (iscg_collect_frame_emitters). events: DEVICE_EVENT_DTYPE records (or BLINK_EVENT_DTYPE, excite 1); frame k
    takes events bucketIdx[bucketStart[k]:bucketStart[k+1]] (BucketEventsByFrame). drift: (nFrames, 2) pixel offsets.
    Returns (records FRAME_EMITTER_DTYPE, start offsets nFrames + 1, (z clamped, z total))."""
    ctx = ctx or default_context()
    ev = _device_events(events)
    st = np.ascontiguousarray(bucketStart, np.uint32)
    idx = np.ascontiguousarray(bucketIdx, np.uint32)
    fi = np.ascontiguousarray(frameIndex, np.int64)
    n = len(fi)
    dr = None if drift is None else np.ascontiguousarray(drift, np.float64).reshape(n, 2)
    il = None if illumination is None else np.ascontiguousarray(illumination, np.float32)
    cap = int(st[-1]) if len(st) else 0
    out = np.zeros(max(cap, 1), FRAME_EMITTER_DTYPE)
    start = np.zeros(n + 1, np.uint32)
    z = np.zeros(2, np.int64)
    nz = psfCache.nz if psfCache is not None else 0
    zs = psfCache.zStepNm if psfCache is not None else 0.0
    ctx.follow_torch()
    ctx.check(ctx.L.iscg_collect_frame_emitters(
        ctx.h, ev.ctypes.data if len(ev) else None, len(ev), st.ctypes.data, idx.ctypes.data if len(idx) else None,
        fi.ctypes.data, n, width, height, pixelSizeNm, photonsPerBlink, None if dr is None else dr.ctypes.data, nz,
        zs, globalZOffsetUm, kernel, None if il is None else il.ctypes.data, out.ctypes.data, len(out),
        start.ctypes.data, z.ctypes.data), "iscg_collect_frame_emitters")
    return out[: int(start[-1])], start, (int(z[0]), int(z[1]))


def _catmull_rom(f):
    f2 = f * f
    f3 = f2 * f
    return np.stack([-0.5 * f3 + f2 - 0.5 * f, 1.5 * f3 - 2.5 * f2 + 1.0, -1.5 * f3 + 2.0 * f2 + 0.5 * f,
                     0.5 * f3 - 0.5 * f2], axis=-1)


def CollectGpuEmitters(frameEmitters, cache: PsfKernelCache, kernel: int = 0, minPhotons: float = 0.0):
    """sim::CollectGpuEmitters (Sim/SMLMSimulation.cpp:219-256) from a frame's emitters: SplatSetup
    (Sim/PsfGeneratorBridge.cpp:1382-1422) on the host, 64-byte GpuSplatEmitter records with float weights."""
    em = np.asarray(frameEmitters, FRAME_EMITTER_DTYPE)
    em = em[~((em["photons"] <= 0.0) | (em["photons"] < minPhotons))]
    os_ = max(1, cache.oversampling)
    kc = (cache.sizeOversampled - 1) / 2.0
    x0 = np.floor(em["xPx"] + 0.5)
    y0 = np.floor(em["yPx"] + 0.5)
    tx = kc + (x0 - em["xPx"] - 0.5) * os_ + 0.5
    ty = kc + (y0 - em["yPx"] - 0.5) * os_ + 0.5
    out = np.zeros(len(em), SPLAT_EMITTER_DTYPE)
    out["x0"], out["y0"] = x0, y0
    wx = np.zeros((len(em), 4))
    wy = np.zeros((len(em), 4))
    if cache.interpMode == 0:
        out["bx"], out["by"], out["nTaps"] = np.floor(tx + 0.5), np.floor(ty + 0.5), 1
        wx[:, 0] = wy[:, 0] = 1.0
    else:
        bx, by = np.floor(tx), np.floor(ty)
        fx, fy = tx - bx, ty - by
        if cache.interpMode == 2:
            out["bx"], out["by"], out["nTaps"] = bx - 1, by - 1, 4
            wx, wy = _catmull_rom(fx), _catmull_rom(fy)
        else:
            out["bx"], out["by"], out["nTaps"] = bx, by, 2
            wx[:, 0], wx[:, 1], wy[:, 0], wy[:, 1] = 1.0 - fx, fx, 1.0 - fy, fy
    out["wx"], out["wy"] = wx.astype(np.float32), wy.astype(np.float32)
    out["plane"] = em["zIndex"] | (kernel << 16)
    out["photons"] = em["photons"].astype(np.float32)
    if cache.halo is not None:
        out["cutRadius2"] = cache.halo.radius2[np.clip(em["zIndex"], 0, cache.nz - 1)]
    else:
        out["cutRadius2"] = np.iinfo(np.int32).max
    return out


class SplatGpuHost:
    """sim::SplatGpuHost / sim::GpuSimulator (Sim/SplatGpu.h:40-77, Sim/GpuSimD3D11.h:61-105) on libiscg."""

    kMaxKernels = 4

    def __init__(self, ctx: Context | None = None):
        self.ctx = ctx or Context()
        self.err = ""
        self.width = self.height = 0

    def _call(self, rc: int) -> bool:
        if rc != 0:
            self.err = self.ctx.error()
            return False
        self.err = ""
        return True

    def SetKernels(self, kernels: list[PsfKernelCache]) -> bool:
        if len(kernels) > self.kMaxKernels:
            self.err = "at most 4 kernels"
            return False
        views = (_lib.KernelView * max(1, len(kernels)))()
        keep = []
        for i, k in enumerate(kernels):
            pl, a = _ptr(_arr(k.planes, np.float32))
            keep.append(a)
            lo = hi = None
            if k.halo is not None:
                lo, a1 = _ptr(np.ascontiguousarray(k.halo.lo, np.int16))
                hi, a2 = _ptr(np.ascontiguousarray(k.halo.hi, np.int16))
                keep += [a1, a2]
            views[i] = _lib.KernelView(pl, lo, hi, k.zStepNm, k.SplatSerial, k.nz, k.sizeOversampled, k.oversampling,
                                       k.interpMode)
        self.kernels = list(kernels)
        self.ctx.follow_torch()
        return self._call(self.ctx.L.iscg_splat_set_kernels(self.ctx.h, views, len(kernels)))

    def SetStatic(self, width: int, height: int, pixelBase: int = 0, offset=None, gain=None, readNoise=None,
                  background=None, flatBackground: float = 0.0, key: int = 0) -> bool:
        n = width * height
        # maps whose size is not width*height are ignored, as the reference's size checks do (SMLMNoise.cpp:151-158)
        fit = lambda m: None if m is None or _numel(m) != n else _arr(m, np.float32)
        ps = [_ptr(fit(m)) for m in (offset, gain, readNoise, background)]
        self.width, self.height = width, height
        self.ctx.follow_torch()
        return self._call(self.ctx.L.iscg_splat_set_static(self.ctx.h, width, height, pixelBase, ps[0][0], ps[1][0],
                                                           ps[2][0], ps[3][0], float(flatBackground), key))

    def RenderBatch(self, emitters, emStart, frames, cam: CameraNoiseParams | None, seed: int, bgScales=None,
                    extras=None, minPhotons: float = 0.0, outAdu=None, outPhotons=None, records: bool = False):
        """One call over a flat batch: emitters (iscg_frame_emitter or GpuSplatEmitter records) with per-frame
        offsets emStart (nFrames + 1). Writes outAdu / outPhotons (allocated as numpy when None and wanted)."""
        em = _arr(emitters, SPLAT_EMITTER_DTYPE if records else FRAME_EMITTER_DTYPE)
        st = _arr(emStart, np.uint32)
        n_frames = int(st.shape[0]) - 1
        fr = _arr(frames, np.uint32)
        ptrs = [_ptr(x) for x in (em, st, fr, _arr(bgScales, np.float32), _arr(extras, np.float32), outAdu,
                                  outPhotons)]
        cc = cam.to_c() if cam is not None else None
        n_em = int(em.shape[0]) if em is not None else 0
        L = self.ctx.L
        self.ctx.follow_torch()
        if records:
            rc = L.iscg_splat_render_records(self.ctx.h, ptrs[0][0], n_em, ptrs[1][0], n_frames, ptrs[2][0],
                                             ptrs[3][0], ptrs[4][0], C.byref(cc) if cc else None, seed & 0xFFFFFFFF,
                                             ptrs[5][0], ptrs[6][0])
        else:
            rc = L.iscg_splat_render_frames(self.ctx.h, ptrs[0][0], n_em, ptrs[1][0], n_frames, ptrs[2][0],
                                            ptrs[3][0], ptrs[4][0], C.byref(cc) if cc else None, seed & 0xFFFFFFFF,
                                            float(minPhotons), ptrs[5][0], ptrs[6][0])
        return self._call(rc)

    def RenderFrames(self, emitters: list, frames, extras, cam: CameraNoiseParams, seed: int, outs: list,
                     bgScales=None) -> bool:
        """SplatGpuHost::RenderFrames (Sim/SplatGpu.h:57): emitters[k] (GpuSplatEmitter records) of frame k onto
        extras[k] (or the background), noise keyed by (seed, frames[k]); outs[k] receives the ADU."""
        counts = [len(e) for e in emitters]
        start = np.zeros(len(emitters) + 1, np.uint32)
        start[1:] = np.cumsum(counts)
        flat = np.concatenate([np.asarray(e, SPLAT_EMITTER_DTYPE) for e in emitters]) if sum(counts) else \
            np.zeros(0, SPLAT_EMITTER_DTYPE)
        npx = self.width * self.height
        ext = None
        if extras is not None and any(x is not None for x in extras):
            ext = np.zeros((len(emitters), npx), np.float32)
            for k, x in enumerate(extras):
                ext[k] = 0.0 if x is None else np.asarray(x, np.float32).reshape(-1)
        adu = np.zeros((len(emitters), npx), np.uint16)
        if not self.RenderBatch(flat, start, np.asarray(frames, np.uint32), cam, seed, bgScales, ext,
                                outAdu=adu, records=True):
            return False
        for k in range(len(emitters)):
            outs[k][...] = adu[k].reshape(np.shape(outs[k]))
        return True

    def MaxBatchFrames(self) -> int:
        """Frames per call within a 1 GiB output budget (the device holds the whole batch in HBM)."""
        return max(1, (1 << 30) // max(1, 6 * self.width * self.height))


class SmlmMovieGpu:
    """An SMLM movie's frames on the device from its blink events, as the reference's SMLM loop makes them frame by
    frame on the CPU (Sim/ScopeMovie.cpp:6183 buckets a movie's events once; :7248-7259 collects and renders each
    frame, :7642-7650 applies the camera noise): the events are uploaded and bucketed by frame once
    (BucketEventsByFrame, Sim/SMLMSimulation.cpp:203-217); each Render computes the PSF stack on the GPU
    (ComputeZernikePsfPlanes + the sum-1 normalisation, from a device-held PsfLibrary), its block sums, every frame's
    emitters (CollectFrameEmitters, Sim/SMLMSimulation.cpp:55-101), and splats and noises all frames in one batch
    (RenderPhotonImage + ApplyNoiseChain). Every buffer is a torch tensor on the context's device, and a render
    waits for nothing on the host, so Capture() records it as one HIP graph that Replay() runs again.

    The splat is handed the bucket total (the (frame, event) pairs, known after bucketing) as its emitter count:
    the collect keeps at most that many, each frame's range comes from the device offsets, and the records beyond
    the kept ones are never splatted (they start zeroed, so the setup pass reads valid indices)."""

    def __init__(self, events, nFrames: int, width: int, height: int, pixelSizeNm: float, photonsPerBlink: float,
                 flatBackground: float, offset=None, ctx: Context | None = None):
        import torch
        self.ctx = ctx or default_context()
        ev7 = _device_events(events)
        self.nEvents, self.nFrames, self.width, self.height = len(ev7), int(nFrames), int(width), int(height)
        self.pixelSizeNm, self.photonsPerBlink = float(pixelSizeNm), float(photonsPerBlink)
        dev = "cuda"
        self.d_ev = torch.from_numpy(ev7.view(np.uint8).reshape(len(ev7), ev7.dtype.itemsize)).to(dev)
        self.d_bst = torch.empty(self.nFrames + 1, dtype=torch.int32, device=dev)
        cap = 4 * len(ev7) + 16
        self.d_bidx = torch.empty(cap, dtype=torch.int32, device=dev)
        self.d_fi = torch.arange(self.nFrames, dtype=torch.int64, device=dev)
        self.d_start = torch.empty(self.nFrames + 1, dtype=torch.int32, device=dev)
        self.d_frames = torch.arange(self.nFrames, dtype=torch.int32, device=dev)
        total = C.c_uint64(0)
        self.ctx.follow_torch()
        rc = self.ctx.L.iscg_bucket_events(self.ctx.h, self.d_ev.data_ptr(), self.nEvents, self.nFrames,
                                           self.d_bst.data_ptr(), self.d_bidx.data_ptr(), cap, C.byref(total))
        if rc != 0 and total.value > cap:
            self.d_bidx = torch.empty(int(total.value), dtype=torch.int32, device=dev)
            rc = self.ctx.L.iscg_bucket_events(self.ctx.h, self.d_ev.data_ptr(), self.nEvents, self.nFrames,
                                               self.d_bst.data_ptr(), self.d_bidx.data_ptr(), int(total.value),
                                               C.byref(total))
        self.ctx.check(rc, "iscg_bucket_events")
        self.nPairs = int(total.value)
        self.d_ems = torch.zeros((max(1, self.nPairs), FRAME_EMITTER_DTYPE.itemsize), dtype=torch.uint8, device=dev)
        self.splat = SplatGpuHost(self.ctx)
        if offset is not None and not hasattr(offset, "data_ptr"):
            offset = torch.from_numpy(np.ascontiguousarray(offset, np.float32)).to(dev)
        self.offset = offset
        if not self.splat.SetStatic(self.width, self.height, offset=offset, flatBackground=flatBackground):
            raise IscgError(self.splat.err)
        self.planes = None
        self.cache = None
        self.library = None
        self._req = None
        self._serial = 0
        self.graph = None

    def Render(self, psfRequest, oversampling: int, zStepNm: float, interpMode: int, cam: CameraNoiseParams | None,
               seed: int, outAdu=None, outPhotons=None) -> bool:
        """The whole movie: PSF -> block sums -> emitters -> splat + noise into outAdu / outPhotons ((nFrames, h, w)
        uint16 / float32 tensors). False with splat.err set when a step fails."""
        import torch
        if self.library is None or self._req is not psfRequest:
            self._req = psfRequest
            self.library = PsfLibrary([psfRequest], ctx=self.ctx)
            self.planes = torch.empty(self.library.shape, dtype=torch.float32, device="cuda")
        self.library.Run(self.planes, normalize=True)
        self._serial += 1
        cache = PsfKernelCache(self.planes[0], oversampling=oversampling, zStepNm=zStepNm, interpMode=interpMode)
        cache.serial = self._serial  # a new stack: its block sums are made again
        return self.RenderFrames(cache, cam, seed, outAdu, outPhotons)

    def RenderFrames(self, cache: PsfKernelCache, cam: CameraNoiseParams | None, seed: int, outAdu=None,
                     outPhotons=None) -> bool:
        """The movie on a given kernel stack: every frame's emitters collected from the events, then the splat and
        the noise of all frames (no PSF computation)."""
        self.cache = cache
        ctx = self.ctx
        ctx.follow_torch()
        ctx.check(ctx.L.iscg_collect_frame_emitters(
            ctx.h, self.d_ev.data_ptr(), self.nEvents, self.d_bst.data_ptr(), self.d_bidx.data_ptr(),
            self.d_fi.data_ptr(), self.nFrames, self.width, self.height, self.pixelSizeNm, self.photonsPerBlink, None,
            cache.nz, float(cache.zStepNm), 0.0, 0, None, self.d_ems.data_ptr(), self.d_ems.shape[0],
            self.d_start.data_ptr(), None), "iscg_collect_frame_emitters")
        return self.splat.SetKernels([cache]) and self.splat.RenderBatch(
            self.d_ems[: self.nPairs], self.d_start, self.d_frames, cam, seed, outAdu=outAdu, outPhotons=outPhotons)

    def Capture(self, psfRequest, oversampling: int, zStepNm: float, interpMode: int, cam: CameraNoiseParams | None,
                seed: int, outAdu=None, outPhotons=None):
        """Records one Render (after an eager one on the capture stream: buffers made, launch choices tuned) as a
        HIP graph (torch.cuda.CUDAGraph); Replay() then runs the whole movie again into the same outputs."""
        import torch
        stream = torch.cuda.Stream()
        stream.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(stream):
            if not self.Render(psfRequest, oversampling, zStepNm, interpMode, cam, seed, outAdu, outPhotons):
                raise IscgError(self.splat.err)
        torch.cuda.synchronize()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g, stream=stream):
            if not self.Render(psfRequest, oversampling, zStepNm, interpMode, cam, seed, outAdu, outPhotons):
                raise IscgError(self.splat.err)
        self.graph = g
        self.ctx.set_stream(None)
        return g

    def Replay(self):
        self.graph.replay()

    def Emitters(self):
        """The last Render's frame emitters (FRAME_EMITTER_DTYPE) and per-frame offsets, copied to the host."""
        st = self.d_start.cpu().numpy().astype(np.uint32)
        ems = self.d_ems[: int(st[-1])].cpu().numpy().reshape(-1).view(FRAME_EMITTER_DTYPE)
        return ems, st


# ---- PSF planes --------------------------------------------------------------------------------------------------

def psf_request(fields) -> _lib.PsfRequest:
    """An iscg_psf_request from a dict of sim::PsfGeneratorRequest's numeric fields (zernike: 28 OSA values or a
    {index: waves} dict); missing fields take the reference's defaults (Sim/PsfGeneratorBridge.h:91-170)."""
    if isinstance(fields, _lib.PsfRequest):
        return fields
    d = dict(wavelengthNm=660.0, na=1.4, immersionIndex=1.518, sampleIndex=1.518, workingDistanceUm=150.0,
             sampleDepthNm=0.0, pixelSizeNm=100.0, zStepNm=100.0, maskWaist=1.0, maskType=0, maskModes=5,
             oversampling=4, kernelHalfWidthPx=8, nz=1, pupilSamples=0, interpMode=0)
    d.update(fields)
    z = np.zeros(28)
    zern = d.pop("zernike", None)
    if isinstance(zern, dict):
        for j, v in zern.items():
            z[j] = v
    elif zern is not None:
        z[: len(zern)] = zern
    r = _lib.PsfRequest(**{k: v for k, v in d.items() if k in dict(_lib.PsfRequest._fields_)})
    for j in range(28):
        r.zernike[j] = float(z[j])
    return r


def PsfGeometry(fields):
    """(size, nz, pupilSamples) of a request (GeometryFor + ZernikePupilSamples, Sim/ZernikePsf.cpp:118-148)."""
    r = psf_request(fields)
    s, z, m = C.c_int32(), C.c_int32(), C.c_int32()
    _lib.lib().iscg_psf_geometry(C.byref(r), C.byref(s), C.byref(z), C.byref(m))
    return s.value, z.value, m.value


def PsfRequestArray(requests):
    """The C array of iscg_psf_request for a list of requests (dicts or PsfRequest): build it once and pass it to
    ComputePsfPlanesBatch when the same requests are computed again."""
    reqs = [psf_request(r) for r in requests]
    return (_lib.PsfRequest * len(reqs))(*reqs)


def ComputePsfPlanesBatch(requests, normalize: bool = False, out=None, ctx: Context | None = None):
    """ComputeZernikePsfPlanes (Sim/ZernikePsf.h:51) for many requests sharing size, nz and pupil samples, in one
    launch: (nReq, nz, size, size) float32 (sum-1 planes when normalize, as BuildZernikePsfKernelCache).
    requests: a list of requests, or a PsfRequestArray."""
    ctx = ctx or default_context()
    arr = requests if isinstance(requests, C.Array) else PsfRequestArray(requests)
    reqs = arr
    size, nz, _ = PsfGeometry(arr[0])
    if out is None:
        out = np.zeros((len(reqs), nz, size, size), np.float32)
    p, keep = _ptr(out)
    cap = int(np.prod(out.shape)) if hasattr(out, "shape") else 0
    ctx.follow_torch()
    ctx.check(ctx.L.iscg_psf_planes(ctx.h, arr, len(reqs), int(bool(normalize)), p, cap), "iscg_psf_planes")
    del keep
    return out


class PsfLibrary:
    """A fixed set of PSF requests held on the device (iscg_psf_library_*): Run() computes their planes with kernels
    only, so it can be captured into a HIP graph (torch.cuda.CUDAGraph) and replayed; Update() changes the requests
    (same count and geometry) for the next run or replay. Each Run equals ComputePsfPlanesBatch of the requests."""

    def __init__(self, requests, ctx: Context | None = None):
        self.ctx = ctx or default_context()
        self._arr = requests if isinstance(requests, C.Array) else PsfRequestArray(requests)
        h = C.c_void_p()
        self.ctx.follow_torch()
        self.ctx.check(self.ctx.L.iscg_psf_library_create(self.ctx.h, self._arr, len(self._arr), C.byref(h)),
                       "iscg_psf_library_create")
        self.h = h
        n, size, nz = C.c_int32(), C.c_int32(), C.c_int32()
        self.ctx.L.iscg_psf_library_geometry(h, C.byref(n), C.byref(size), C.byref(nz))
        self.shape = (n.value, nz.value, size.value, size.value)

    def Update(self, requests):
        self._arr = requests if isinstance(requests, C.Array) else PsfRequestArray(requests)
        self.ctx.follow_torch()
        self.ctx.check(self.ctx.L.iscg_psf_library_update(self.ctx.h, self.h, self._arr, len(self._arr)),
                       "iscg_psf_library_update")

    def Run(self, out=None, normalize: bool = False):
        if out is None:
            out = np.zeros(self.shape, np.float32)
        p, keep = _ptr(out)
        self.ctx.follow_torch()
        self.ctx.check(self.ctx.L.iscg_psf_library_run(self.ctx.h, self.h, int(bool(normalize)), p,
                                                       int(np.prod(self.shape))), "iscg_psf_library_run")
        del keep
        return out

    def close(self):
        if getattr(self, "h", None):
            self.ctx.L.iscg_psf_library_destroy(self.ctx.h, self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def ComputeZernikePsfPlanes(fields, ctx: Context | None = None):
    """sim::ComputeZernikePsfPlanes for one request: (nz, size, size) raw |E|^2 planes."""
    return ComputePsfPlanesBatch([fields], ctx=ctx)[0]


def BuildZernikePsfKernelCache(fields, ctx: Context | None = None) -> PsfKernelCache:
    """sim::BuildZernikePsfKernelCache (Sim/ZernikePsf.cpp:382): the sum-1 planes as a kernel cache."""
    r = psf_request(fields)
    planes = ComputePsfPlanesBatch([r], normalize=True, ctx=ctx)[0]
    return PsfKernelCache(planes, oversampling=max(1, r.oversampling), zStepNm=r.zStepNm, interpMode=r.interpMode)


# ---- mean field --------------------------------------------------------------------------------------------------

class BinnedBlinkRenderer:
    """sim::BinnedBlinkRenderer (Sim/BinnedBlinks.h:37-62): the binned blink regime on the GPU. Setup makes the
    per-plane kernel spectra once; Render / RenderFrames add the emitters' photon images (FFT-convolved per PSF
    plane on a grid of u cells per pixel, max(0, .), binned) to the frames."""

    def __init__(self, ctx: Context | None = None):
        self.ctx = ctx or default_context()
        self.err = ""
        self.size = None

    def _call(self, rc: int) -> bool:
        self.err = "" if rc == 0 else self.ctx.error()
        return rc == 0

    def Setup(self, kernel: PsfKernelCache | None, sigmaPx: float, width: int, height: int, upscale: int) -> bool:
        """BinnedBlinkRenderer::Setup (Sim/BinnedBlinks.cpp:94): kernel None (or without planes) = the in-focus
        Gaussian of sigmaPx."""
        view, keep = None, None
        r2 = -1
        if kernel is not None and kernel.planes is not None:
            pl, keep = _ptr(_arr(kernel.planes, np.float32))
            view = _lib.KernelView(pl, None, None, kernel.zStepNm, kernel.serial, kernel.nz, kernel.sizeOversampled,
                                   kernel.oversampling, kernel.interpMode)
            if kernel.halo is not None and len(kernel.halo.radius2):
                r2 = int(np.max(kernel.halo.radius2))
        self.ctx.follow_torch()
        ok = self._call(self.ctx.L.iscg_binned_setup(self.ctx.h, C.byref(view) if view is not None else None, r2,
                                                      float(sigmaPx), width, height, upscale))
        del keep
        self.size = (width, height) if ok else None
        return ok

    def Valid(self) -> bool:
        return self.size is not None

    def _geometry(self):
        u, R = C.c_int32(0), C.c_int32(0)
        if self.ctx.L.iscg_binned_geometry(self.ctx.h, C.byref(u), C.byref(R)) != 0:
            return 1, 1
        return u.value, R.value

    def Upscale(self) -> int:
        return self._geometry()[0]

    def RadiusCells(self) -> int:
        return self._geometry()[1]

    def RenderFrames(self, frames, imgs) -> bool:
        """frames: per frame a FRAME_EMITTER_DTYPE array (event order); imgs: (nF, H, W) float32 numpy or device
        tensor, added to."""
        em = np.concatenate([np.asarray(f, FRAME_EMITTER_DTYPE) for f in frames]) if len(frames) else \
            np.zeros(0, FRAME_EMITTER_DTYPE)
        start = np.zeros(len(frames) + 1, np.int32)
        start[1:] = np.cumsum([len(f) for f in frames])
        em = np.ascontiguousarray(em)
        out, keep = _ptr(imgs)
        self.ctx.follow_torch()
        rc = self.ctx.L.iscg_binned_render(self.ctx.h, em.ctypes.data if em.size else None, start.ctypes.data,
                                           len(frames), out)
        del keep
        return self._call(rc)

    def Render(self, emitters, img, alone=None) -> bool:
        """BinnedBlinkRenderer::Render (Sim/BinnedBlinks.cpp:165): one frame; img (H x W, float32 numpy, added to
        in place) and alone (optional, the same values added)."""
        if not isinstance(img, np.ndarray) or img.dtype != np.float32 or not img.flags["C_CONTIGUOUS"]:
            raise TypeError("img: a C-contiguous float32 numpy array")
        if alone is None:
            return self.RenderFrames([emitters], img.reshape(1, *img.shape))
        add = np.zeros_like(img)  # 0 + (float) sum: the very values the reference adds to img and to alone
        if not self.RenderFrames([emitters], add.reshape(1, *add.shape)):
            return False
        img += add
        alone += add
        return True


class WidefieldStack:
    """A focus series of a dye density as one 3D FFT convolution on the GPU: the per-focus work of
    sim::WidefieldScene (FocusSeries, Sim/WidefieldRender.h:588) for foci spaced by the dye-plane thickness, with
    the PSF as a KernelWidefieldPsf of a PsfKernelCache (Sim/WidefieldRender.h:219-236)."""

    def __init__(self, ctx: Context | None = None):
        self.ctx = ctx or Context()
        self.err = ""
        self.desc = None

    def _call(self, rc: int) -> bool:
        self.err = "" if rc == 0 else self.ctx.error()
        return rc == 0

    def Plan(self, cache: PsfKernelCache, nx: int, ny: int, nzD: int, dzUm: float, k0: int, focus0Um: float,
             nF: int, fov=(0, 0, 0, 0), fft=(0, 0, 0), kernelRadius: int = 0) -> bool:
        d = _lib.WfDesc(dzUm=dzUm, focus0Um=focus0Um, k0=k0, nx=nx, ny=ny, nzD=nzD, nF=nF, fovX0=fov[0],
                        fovY0=fov[1], cw=fov[2], ch=fov[3], NX=fft[0], NY=fft[1], NZ=fft[2], upscale=1)
        pl, keep = _ptr(_arr(cache.planes, np.float32))
        self.ctx.follow_torch()
        ok = self._call(self.ctx.L.iscg_wf_conv3d_plan(self.ctx.h, C.byref(d), pl, cache.nz, cache.sizeOversampled,
                                                       cache.oversampling, cache.zStepNm, kernelRadius))
        del keep
        if ok:
            cw = fov[2] or nx - fov[0]
            ch = fov[3] or ny - fov[1]
            self.desc = (nx, ny, nzD, nF, cw, ch)
        return ok

    @property
    def image_shape(self):
        nx, ny, nzD, nF, cw, ch = self.desc
        return (nF, ch, cw)

    def FocusSeries(self, density, cam: CameraNoiseParams | None = None, offset=None, background: float = 0.0,
                    seed: int = 0, frameBase: int = 0, pixelBase: int = 0, outImages=None, outAdu=None) -> bool:
        """density (nzD, ny, nx) -> outImages (nF, ch, cw) float32 frames and/or outAdu uint16 (camera noise)."""
        ptrs = [_ptr(x) for x in (_arr(density, np.float32), _arr(offset, np.float32), outImages, outAdu)]
        cc = cam.to_c() if cam is not None else None
        self.ctx.follow_torch()
        return self._call(self.ctx.L.iscg_wf_conv3d(self.ctx.h, ptrs[0][0], C.byref(cc) if cc else None,
                                                    ptrs[1][0], float(background), seed & 0xFFFFFFFF, frameBase,
                                                    pixelBase, ptrs[2][0], ptrs[3][0]))


@dataclass
class WidefieldGpuJob:
    """sim::WidefieldGpuJob (Sim/WidefieldRender.h:306-340): the GPU's share of one focus.

    kernels: [(p, spec)] with spec the (NY, NX/2 + 1) complex64 half spectrum of kernel plane p;
    planes: [(key, cells, values)] sparse dye planes (cell = y * nx + x) the host does not hold yet;
    channels: per channel a list of deps (key, p0, p1, w0, w1, two)."""

    NX: int = 0
    NY: int = 0
    nx: int = 0
    ny: int = 0
    fovX0: int = 0
    fovY0: int = 0
    cw: int = 0
    ch: int = 0
    fracX: float = 0.0
    fracY: float = 0.0
    geometry: int = 0
    kernels: list = field(default_factory=list)
    planes: list = field(default_factory=list)
    channels: list = field(default_factory=list)
    hasPersistent: bool = False
    kMaxGpuFft = 2048


class WidefieldGpuHost:
    """A WidefieldAccelerator (Sim/WidefieldRender.h:342-361) on libiscg: plane spectra resident in HBM by key per
    geometry, kernel spectra per (geometry, p); Images runs the per-focus MAC, phase ramp and cropped inverse on the
    GPU. Methods return bool with the message in .err, as the reference's do."""

    def __init__(self, ctx: Context | None = None):
        self.ctx = ctx or Context()
        self.err = ""

    def HasPlane(self, geometry: int, key: int) -> bool:
        rc = self.ctx.L.iscg_wf_has_plane(self.ctx.h, geometry, key)
        if rc < 0:
            raise IscgError(self.ctx.error())
        return rc == 1

    def Fits(self, geometry: int, NX: int, NY: int, planes: int) -> bool:
        rc = self.ctx.L.iscg_wf_fits(self.ctx.h, geometry, NX, NY, planes)
        if rc < 0:
            raise IscgError(self.ctx.error())
        return rc == 1

    def SetBudget(self, nbytes: int):
        self.ctx.check(self.ctx.L.iscg_wf_set_budget(self.ctx.h, int(nbytes)))

    def Images(self, job: WidefieldGpuJob, out: list | None = None, outArray=None) -> bool:
        """One (ch, cw) float32 image per job channel, appended to `out` (numpy), or written to outArray
        (nChannels, ch, cw) (numpy, or a device tensor left on the device)."""
        keep = []
        kp = np.array([k[0] for k in job.kernels], np.int32)
        specs = []
        for _, sp in job.kernels:
            a = sp if hasattr(sp, "data_ptr") else np.ascontiguousarray(sp, np.complex64)
            p, k = _ptr(a)
            specs.append(p)
            keep.append(k)
        spec_ptrs = (C.c_void_p * max(1, len(specs)))(*specs)
        keys = np.array([p[0] for p in job.planes], np.uint64)
        cells = [np.ascontiguousarray(p[1], np.uint32).ravel() for p in job.planes]
        vals = [np.ascontiguousarray(p[2], np.float32).ravel() for p in job.planes]
        starts = np.zeros(len(job.planes) + 1, np.uint32)
        if job.planes:
            starts[1:] = np.cumsum([len(c) for c in cells])
        cat_c = np.concatenate(cells) if cells else np.zeros(0, np.uint32)
        cat_v = np.concatenate(vals) if vals else np.zeros(0, np.float32)
        deps = [d for ch in job.channels for d in ch]
        dep_arr = (_lib.WfDep * max(1, len(deps)))()
        for i, (key, p0, p1, w0, w1, two) in enumerate(deps):
            dep_arr[i] = _lib.WfDep(int(key), int(p0), int(p1), float(w0), float(w1), int(bool(two)), 0)
        cstart = np.zeros(len(job.channels) + 1, np.uint32)
        if job.channels:
            cstart[1:] = np.cumsum([len(ch) for ch in job.channels])
        nC = len(job.channels)
        if outArray is None:
            outArray = np.zeros((nC, job.ch, job.cw), np.float32)
        po, ko = _ptr(outArray)
        J = _lib.WfJob(NX=job.NX, NY=job.NY, nx=job.nx, ny=job.ny, fovX0=job.fovX0, fovY0=job.fovY0, cw=job.cw,
                       ch=job.ch, fracX=job.fracX, fracY=job.fracY, geometry=job.geometry,
                       nKernels=len(job.kernels), nPlanes=len(job.planes), nChannels=nC,
                       hasPersistent=int(bool(job.hasPersistent)), kernelP=kp.ctypes.data,
                       kernelSpec=C.cast(spec_ptrs, C.c_void_p), planeKey=keys.ctypes.data,
                       planeStart=starts.ctypes.data, cells=cat_c.ctypes.data, values=cat_v.ctypes.data,
                       channelStart=cstart.ctypes.data, deps=C.cast(dep_arr, C.c_void_p))
        self.ctx.follow_torch()
        rc = self.ctx.L.iscg_wf_images(self.ctx.h, C.byref(J), po)
        self.err = "" if rc == 0 else self.ctx.error()
        if rc != 0:
            return False
        if out is not None:
            out.clear()
            if isinstance(outArray, np.ndarray):
                out.extend(outArray[c] for c in range(nC))
            else:
                out.extend(outArray[c] for c in range(nC))
        del keep, ko
        return True


# ---- drifting samples on the mean field (Sim/Drift.h:118-157, Sim/WidefieldRender.h:292-300,880-911) ------------

kDriftFocusStepNm = 10.0


@dataclass
class DriftFocusGrid:
    """sim::DriftFocusGrid (Sim/Drift.h:136-144, Sim/Drift.cpp:145-166): foci every kDriftFocusStepNm over the
    trajectory's z range; a frame's image interpolates its two neighbours."""

    k0: int = 0
    n: int = 1

    @staticmethod
    def For(zLoNm: float, zHiNm: float) -> "DriftFocusGrid":
        lo = int(np.floor(zLoNm / kDriftFocusStepNm))
        hi = int(np.ceil(zHiNm / kDriftFocusStepNm))
        return DriftFocusGrid(lo, hi - lo + 1)

    def DzNm(self, k: int) -> float:
        return (self.k0 + k) * kDriftFocusStepNm

    def Weights(self, dzNm: float):
        """(k, w): image(dz) = (1 - w) image(k) + w image(k + 1)."""
        if self.n <= 1:
            return 0, 0.0
        t = dzNm / kDriftFocusStepNm - float(self.k0)
        k = min(self.n - 2, max(0, int(np.floor(t))))
        return k, min(1.0, max(0.0, t - k))


def DriftRange(driftNm):
    """sim::DriftRange (Sim/Drift.cpp:124-137): the trajectory's bounds, (xLo, xHi, yLo, yHi, zLo, zHi) nm, each
    including 0."""
    d = np.asarray(driftNm, np.float64).reshape(-1, 3)
    lo = np.minimum(d.min(axis=0), 0.0) if len(d) else np.zeros(3)
    hi = np.maximum(d.max(axis=0), 0.0) if len(d) else np.zeros(3)
    return lo[0], hi[0], lo[1], hi[1], lo[2], hi[2]


def WidefieldShiftedFrames(desc: dict, spectra, frames, a=None, cam=None, ctx: Context | None = None):
    """RenderShiftedImages (Sim/WidefieldRender.cpp:2905-2935) for a batch of frames (iscg_wf_shifted_frames).
    desc: NX, NY, fovX0, fovY0, cw, ch, upscale, fracX, fracY, hasPersistent; spectra (nImages, nChannels, NY,
    NX/2 + 1) complex64 (numpy) or the same as a float32 device tensor; frames: (i0, i1, w, dxCells, dyCells) per
    frame; a: (nFrames, nBleach) coefficients or None; cam (nFrames, ch/up, cw/up) float32, added to (zeros when
    None). Returns cam."""
    ctx = ctx or default_context()
    if isinstance(spectra, np.ndarray):
        sp = np.ascontiguousarray(spectra, np.complex64)
        nImages, nCh = sp.shape[0], sp.shape[1]
        sp = sp.view(np.float32)
    else:
        sp = spectra
        nImages, nCh = int(desc["nImages"]), int(desc["nChannels"])
    up = max(1, int(desc.get("upscale", 1)))
    fr = (_lib.WfShiftFrame * max(1, len(frames)))()
    for k, f in enumerate(frames):
        fr[k] = _lib.WfShiftFrame(int(f[0]), int(f[1]), float(f[2]), float(f[3]), float(f[4]))
    if cam is None:
        cam = np.zeros((len(frames), desc["ch"] // up, desc["cw"] // up), np.float32)
    d = _lib.WfShiftDesc(desc["NX"], desc["NY"], desc["fovX0"], desc["fovY0"], desc["cw"], desc["ch"], up, nImages,
                         nCh, int(desc.get("hasPersistent", 1)), float(desc.get("fracX", 0.0)),
                         float(desc.get("fracY", 0.0)))
    aa = None if a is None else np.ascontiguousarray(a, np.float64)
    ps, k1 = _ptr(sp)
    pa, k2 = _ptr(aa)
    pc, k3 = _ptr(cam)
    ctx.follow_torch()
    ctx.check(ctx.L.iscg_wf_shifted_frames(ctx.h, C.byref(d), ps, fr, pa, len(frames), pc), "iscg_wf_shifted_frames")
    return cam


class WidefieldDriftFrames:
    """sim::WidefieldDriftFrames (Sim/WidefieldRender.h:880-911, .cpp:2937-2977) on the GPU: the focus images on the
    drift's focus grid around each base focus (their kept spectra, SetSeries), each frame interpolated in z and
    shifted in xy by iscg_wf_shifted_frames, many frames per call."""

    def __init__(self, ctx: Context | None = None):
        self.ctx = ctx or default_context()
        self.grid = DriftFocusGrid()
        self.base = []
        self.desc = None
        self.spectra = None

    def Begin(self, bounds, baseFocusWorldUm):
        """bounds: DriftRange(trajectory); baseFocusWorldUm: the undrifted focal planes."""
        self.grid = DriftFocusGrid.For(bounds[4], bounds[5])
        self.base = list(baseFocusWorldUm)
        self.spectra = None

    def Foci(self):
        """The foci of the series, base-major (WidefieldDriftFrames::Refresh: the sample sits dz higher, so the
        focal plane is dz lower in it)."""
        return [f - self.grid.DzNm(k) / 1000.0 for f in self.base for k in range(self.grid.n)]

    def SetSeries(self, desc: dict, spectra, pitchNm: float):
        """The series' kept spectra (len(Foci()) images, in Foci() order) and the grid pitch."""
        self.desc, self.spectra, self.pitchNm = dict(desc), spectra, float(pitchNm)

    def Render(self, base: int, driftNm, a=None, cam=None):
        """The frames of base focus `base` at each drift (x, y, z nm), added to cam."""
        frames = []
        for dx, dy, dz in np.asarray(driftNm, np.float64).reshape(-1, 3):
            k, w = self.grid.Weights(dz)
            i = base * self.grid.n + k
            two = w != 0.0 and k + 1 < self.grid.n
            frames.append((i, i + 1 if two else -1, w if two else 0.0, dx / self.pitchNm, dy / self.pitchNm))
        return WidefieldShiftedFrames(self.desc, self.spectra, frames, a, cam, ctx=self.ctx)


def WidefieldImagesRender(channels, hasPersistent: bool, upscale: int, a, base=None, cam: CameraNoiseParams | None = None,
                          offset=None, seed: int = 0, frames=None, pixelBase: int = 0, outPhotons=None, outAdu=None,
                          ctx: Context | None = None):
    """WidefieldImages::Render (Sim/WidefieldRender.cpp:547-580) of cell images channels (nCh, ch, cw) -- channel 0
    persistent when hasPersistent -- for each row of bleach coefficients a (nFrames, nCh - hasPersistent), onto base
    (nFrames, ch/u, cw/u) or zeros; with cam, the camera noise of the frames (frame ids `frames`). Returns
    (photons or None, adu or None); device tensors in give device tensors out."""
    ctx = ctx or default_context()
    chans = _arr(channels, np.float32)
    nC, chh, cw = (int(x) for x in chans.shape)
    aa = _arr(a, np.float64)
    n_frames = int(aa.shape[0]) if aa is not None and len(aa.shape) == 2 else 1
    u = max(1, int(upscale))
    shape = (n_frames, chh // u, cw // u)
    dev = hasattr(chans, "is_cuda") and chans.is_cuda
    if outPhotons is None and cam is None:
        if dev:
            import torch
            outPhotons = torch.empty(shape, dtype=torch.float32, device=chans.device)
        else:
            outPhotons = np.zeros(shape, np.float32)
    if cam is not None and outAdu is None:
        if dev:
            import torch
            outAdu = torch.empty(shape, dtype=torch.int16, device=chans.device)
        else:
            outAdu = np.zeros(shape, np.uint16)
    fr = _arr(frames if frames is not None else np.arange(n_frames), np.uint32) if cam is not None else None
    ptrs = [_ptr(x) for x in (chans, aa, _arr(base, np.float32), _arr(offset, np.float32), fr, outPhotons, outAdu)]
    cc = cam.to_c() if cam is not None else None
    ctx.follow_torch()
    ctx.check(ctx.L.iscg_wf_render(ctx.h, ptrs[0][0], nC, int(bool(hasPersistent)), cw, chh, u, ptrs[1][0], n_frames,
                                   ptrs[2][0], C.byref(cc) if cc else None, ptrs[3][0], seed & 0xFFFFFFFF,
                                   ptrs[4][0], pixelBase, ptrs[5][0], ptrs[6][0]), "iscg_wf_render")
    return outPhotons, outAdu


# ---- differentiable ROI model -------------------------------------------------------------------------------------

ROI_PARAMS_DTYPE = np.dtype([("x", "f8"), ("y", "f8"), ("z", "f8"), ("photons", "f8"), ("background", "f8")])


def roi_model(**kw) -> _lib.RoiModel:
    """iscg_roi_model with the C5 optics as defaults (670 nm, NA 1.45, ni 1.518, ns 1.33, 100 nm pixels, M 64,
    15 x 15 pixel ROIs, 4 x 4 sub-samples per pixel)."""
    d = dict(wavelengthNm=670.0, na=1.45, immersionIndex=1.518, sampleIndex=1.33, sampleDepthNm=0.0,
             pixelSizeNm=100.0, pupilSamples=64, roiSize=15, oversampling=4, reserved=0)
    d.update(kw)
    return _lib.RoiModel(**d)


def _zern(z):
    a = np.zeros(28)
    z = np.asarray(z, np.float64).reshape(-1)
    a[: len(z)] = z
    return (C.c_double * 28)(*a)


def RoiForward(model, zernike, params, out=None, ctx: Context | None = None):
    """Expected images mu (nRoi, R, R) float32 of emitter ROIs (iscg_roi_forward)."""
    ctx = ctx or default_context()
    prm = _arr(params, ROI_PARAMS_DTYPE)
    n = int(prm.shape[0])
    if out is None:
        out = np.zeros((n, model.roiSize, model.roiSize), np.float32)
    pp, k1 = _ptr(prm)
    po, k2 = _ptr(out)
    ctx.follow_torch()
    ctx.check(ctx.L.iscg_roi_forward(ctx.h, C.byref(model), _zern(zernike), pp, n, po), "iscg_roi_forward")
    return out


def RoiBackward(model, zernike, params, data, gradParams=None, ctx: Context | None = None):
    """Poisson NLL of data and its gradient (iscg_roi_backward): (nll, d/d(x,y,z,N,b) (nRoi, 5), d/d zernike (28))."""
    ctx = ctx or default_context()
    prm = _arr(params, ROI_PARAMS_DTYPE)
    n = int(prm.shape[0])
    d = _arr(data, np.float32)
    if gradParams is None:
        gradParams = np.zeros((n, 5), np.float64)
    gz = (C.c_double * 28)()
    nll = C.c_double()
    pp, k1 = _ptr(prm)
    pd, k2 = _ptr(d)
    pg, k3 = _ptr(gradParams)
    ctx.follow_torch()
    ctx.check(ctx.L.iscg_roi_backward(ctx.h, C.byref(model), _zern(zernike), pp, n, pd, pg, gz, C.byref(nll)),
              "iscg_roi_backward")
    return nll.value, gradParams, np.array(gz[:])


def PhaseRetrieval(model, params, data, modes, zernike0=None, iterations: int = 200, ctx: Context | None = None):
    """Zernike phase retrieval: the coefficients of the OSA slots `modes` (waves) that minimise the Poisson NLL of
    the data (e.g. a bead z-stack: one ROI per plane with its known defocus), by L-BFGS on the adjoint gradient.
    Returns (zernike[28], final NLL, iterations used)."""
    from scipy.optimize import minimize
    ctx = ctx or default_context()
    z0 = np.zeros(28) if zernike0 is None else np.array(zernike0, np.float64)
    modes = list(modes)
    data = np.ascontiguousarray(data, np.float32)

    def f(v):
        z = z0.copy()
        z[modes] = v
        nll, _, gz = RoiBackward(model, z, params, data, ctx=ctx)
        return nll, gz[modes]

    res = minimize(f, z0[modes], jac=True, method="L-BFGS-B", options={"maxiter": iterations, "gtol": 1e-10,
                                                                        "ftol": 1e-15})
    z = z0.copy()
    z[modes] = res.x
    return z, float(res.fun), int(res.nit)


def SetDefaultBlinkSplatGpu(make):
    """Installs a host factory, as sim::SetDefaultBlinkSplatGpu (Sim/ScopeMovie.h:378) does for the viewer."""
    global _default_splat_factory
    _default_splat_factory = make


_default_splat_factory = None


def DefaultBlinkSplatGpu():
    return _default_splat_factory() if _default_splat_factory else SplatGpuHost()


__all__ = ["Context", "CameraNoiseParams", "ApplyNoiseChain", "ApplyNoiseChainFrames", "PsfKernelCache",
           "PsfHaloSpans", "SplatGpuHost", "WidefieldStack", "ComputePsfPlanesBatch", "ComputeZernikePsfPlanes",
           "BuildZernikePsfKernelCache", "PsfGeometry", "psf_request", "ROI_PARAMS_DTYPE", "roi_model", "RoiForward",
           "RoiBackward", "PhaseRetrieval", "CollectFrameEmitters", "CollectGpuEmitters", "FRAME_EMITTER_DTYPE",
           "SPLAT_EMITTER_DTYPE", "BLINK_EVENT_DTYPE", "DEVICE_EVENT_DTYPE", "INTERP", "IscgError"]


# ---- dye schedules (core/src/dyes.h:129-261) ---------------------------------------------------------------------

# iscg_dye_address (h1 of the structure block, site k, n), iscg_blink = isc::Blink, iscg_bin_blink = isc::BinBlink,
# iscg_cont_window = isc::ContWindow (core/src/dyes.h:190-200)
DYE_ADDRESS_DTYPE = np.dtype([("h1", "u4"), ("k", "i4"), ("n", "i4")])
BLINK_DTYPE = np.dtype([("tOn", "f8"), ("tOff", "f8"), ("brightness", "f8")])
BIN_BLINK_DTYPE = np.dtype([("tOn", "f8"), ("tOff", "f8"), ("brightness", "f8"), ("bin", "u4"), ("j", "u4")])
CONT_WINDOW_DTYPE = np.dtype({"names": ["tOn", "tOff", "state", "aux"], "formats": ["f8", "f8", "u1", "f8"],
                              "offsets": [0, 8, 16, 24], "itemsize": 32})
assert DYE_ADDRESS_DTYPE.itemsize == 12 and BLINK_DTYPE.itemsize == 24 and BIN_BLINK_DTYPE.itemsize == 32

# isc::LabelMode (core/src/dyes.h:155); isc::EventState of a continuous window (:159)
LABEL_MODES = {"dSTORM": 0, "PALM": 1, "DnaPaint": 2, "WideField": 3}
STATE_PRE, STATE_INITIAL_ON, STATE_ALWAYS_ON = 1, 2, 3


def Kinetics(activationRatePerSec: float = 0.01, onSec: float = 0.05, offSec: float = 1.0, bleachProb: float = 1.0,
             photonCV: float = 0.5, initialOnSec: float = 0.0) -> _lib.Kinetics:
    """isc::Kinetics with the reference's defaults (core/src/dyes.h:129-136)."""
    return _lib.Kinetics(activationRatePerSec, onSec, offSec, bleachProb, photonCV, initialOnSec)


class KineticsHistory:
    """isc::KineticsHistory (core/src/dyes.h:138-149): segment i of constant kinetics kin[i] from tStart[i] (tStart[0]
    = 0, then increasing) to tStart[i+1], the last for ever. The schedule calls take it where the reference's take
    theirs (LabelSchedule, PersistentBlinks, PersistentBlinksInBins: hist=...)."""

    def __init__(self, tStart, kin):
        self.tStart = np.ascontiguousarray(tStart, np.float64)
        kin = list(kin)
        if len(kin) != self.tStart.size:
            raise IscgError("KineticsHistory: one set of kinetics per segment")
        self.kin = (_lib.Kinetics * max(1, len(kin)))(*kin)
        self._c = _lib.KineticsHistory(len(kin), 0, self.tStart.ctypes.data_as(C.POINTER(C.c_double)), self.kin)

    def ref(self):
        return C.byref(self._c)


def _hist(hist):
    return None if hist is None else hist.ref()


@dataclass
class Label:
    """The fields of isc::Label (core/src/dyes.h:166-179) the schedules and the world read: mode, kinetics, PALM pre
    state, the labelled density and fluorescent fraction, and the linkage (binderNm, binderFlexNm, linkerMinNm,
    linkerMaxNm; None: the structure's built-in one, LegacyLinkage). Defaults are the reference's."""

    mode: str = "DnaPaint"
    kin: _lib.Kinetics = field(default_factory=Kinetics)
    preState: bool = False
    density: float = 0.7
    fluorescentFraction: float = 1.0
    linkage: tuple | None = None

    def to_c(self) -> _lib.Label:
        lk = self.linkage if self.linkage is not None else (0.0, 0.0, 0.0, 0.0)
        return _lib.Label(float(self.density), float(self.fluorescentFraction), LABEL_MODES[self.mode],
                          int(self.preState), self.kin, int(self.linkage is not None), 0, *(float(v) for v in lk))


# core/src/dyes.h:33-41
DYE_SALT, NPC_DYE_SALT, LAMIN_DYE_SALT = 0x9E3779B9, 0x4E504344, 0x4C414D44


def DyeH1(seed: int, cx, cy, mtIndex, ctx: Context | None = None):
    """isc::DyeH1 (core/src/dyes.cpp:175-178) of arrays of (cx, cy, mtIndex): Pcg4d(seed ^ DYE_SALT, cx, cy,
    mtIndex).a, on the GPU."""
    cx = np.atleast_1d(np.asarray(cx, np.int64))
    return Pcg4d(np.full(cx.size, (seed ^ DYE_SALT) & 0xFFFFFFFF), cx, cy, mtIndex, ctx)[:, 0]


def NpcH1(seed: int, cx, cy, ctx: Context | None = None):
    """isc::NpcH1 (core/src/dyes.cpp:180-183): Pcg4d(seed ^ NPC_DYE_SALT, cx, cy, 0).a."""
    cx = np.atleast_1d(np.asarray(cx, np.int64))
    return Pcg4d(np.full(cx.size, (seed ^ NPC_DYE_SALT) & 0xFFFFFFFF), cx, cy, np.zeros(cx.size), ctx)[:, 0]


def LaminH1(seed: int, cx, cy, ctx: Context | None = None):
    """isc::LaminH1 (core/src/dyes.cpp:219-222): Pcg4d(seed ^ LAMIN_DYE_SALT, cx, cy, 0).a."""
    cx = np.atleast_1d(np.asarray(cx, np.int64))
    return Pcg4d(np.full(cx.size, (seed ^ LAMIN_DYE_SALT) & 0xFFFFFFFF), cx, cy, np.zeros(cx.size), ctx)[:, 0]


def DyeAddresses(h1, k, n):
    """An iscg_dye_address array from per-dye h1 (DyeH1 of the block, core/src/dyes.cpp:175), k and n."""
    h1 = np.asarray(h1, np.int64)
    out = np.zeros(h1.size, DYE_ADDRESS_DTYPE)
    out["h1"] = (h1 & 0xFFFFFFFF).astype(np.uint32)
    out["k"] = np.asarray(k, np.int64).astype(np.int32)
    out["n"] = np.asarray(n, np.int64).astype(np.int32)
    return out


def _dyes(dyes):
    if hasattr(dyes, "data_ptr"):
        if dyes.numel() * dyes.element_size() % 12:
            raise IscgError("dye addresses: 12 bytes per dye")
        return dyes, dyes.numel() * dyes.element_size() // 12
    d = np.ascontiguousarray(dyes)
    if d.dtype != DYE_ADDRESS_DTYPE:
        d = np.ascontiguousarray(np.asarray(dyes, np.int64).reshape(-1, 3).astype(np.uint32)).view(DYE_ADDRESS_DTYPE)
        d = d.reshape(-1)
    return d, d.size


def _records(ctx, fn, args, n, dtype, capacity):
    """The two-step capacity call shared by the schedule entry points: (offsets n + 1, records)."""
    offsets = np.zeros(n + 1, np.int64)
    total = np.zeros(1, np.int64)
    cap = int(capacity) if capacity is not None else max(64, 4 * n)
    for _ in range(2):
        out = np.zeros(cap, dtype)
        ctx.check(fn(ctx.h, *args, offsets.ctypes.data, out.ctypes.data if cap else None, cap, total.ctypes.data),
                  fn.__name__)
        if total[0] <= cap:
            return offsets, out[: total[0]]
        cap = int(total[0])
    raise IscgError("dye schedule: record count changed between calls")


def LabelSchedule(dyes, label: Label, tMax: float = float("inf"), ctx: Context | None = None, capacity=None,
                  hist: KineticsHistory | None = None):
    """isc::LabelSchedule's blinks (core/src/dyes.cpp:440-471) for a batch of dyes of one label on the GPU: (offsets,
    blinks) with dye i's blinks at blinks[offsets[i]:offsets[i+1]] (BLINK_DTYPE), in time order, bit-identical to
    the reference's; under hist, the label's kinetics per segment of the dye clock (DyeScheduleHistory)."""
    ctx = ctx or default_context()
    d, n = _dyes(dyes)
    p, keep = _ptr(d)
    ctx.follow_torch()
    return _records(ctx, ctx.L.iscg_label_blinks, (p, n, LABEL_MODES[label.mode], C.byref(label.kin), _hist(hist),
                                                   float(tMax)), n, BLINK_DTYPE, capacity)


def LabelContWindows(dyes, label: Label, ctx: Context | None = None, hist: KineticsHistory | None = None):
    """isc::LabelSchedule's continuous windows (core/src/dyes.cpp:447-477): (has, windows) per dye, CONT_WINDOW_DTYPE
    (at most one per dye)."""
    ctx = ctx or default_context()
    d, n = _dyes(dyes)
    p, keep = _ptr(d)
    has = np.zeros(n, np.uint8)
    out = np.zeros(n, CONT_WINDOW_DTYPE)
    ctx.follow_torch()
    ctx.check(ctx.L.iscg_label_cont(ctx.h, p, n, LABEL_MODES[label.mode], int(label.preState), C.byref(label.kin),
                                    _hist(hist), out.ctypes.data, has.ctypes.data), "iscg_label_cont")
    return has.astype(bool), out


def PersistentBlinks(dyes, kin: _lib.Kinetics, t0: float, t1: float, ctx: Context | None = None, capacity=None,
                     hist: KineticsHistory | None = None):
    """isc::PersistentBlinks (core/src/dyes.cpp:636-646) of a batch of persistent (DNA-PAINT) dyes: the blinks
    overlapping (t0, t1), as (offsets, blinks)."""
    ctx = ctx or default_context()
    d, n = _dyes(dyes)
    p, keep = _ptr(d)
    ctx.follow_torch()
    return _records(ctx, ctx.L.iscg_persistent_blinks, (p, n, C.byref(kin), _hist(hist), float(t0), float(t1)), n,
                    BLINK_DTYPE, capacity)


def PersistentBlinksInBins(dyes, kin: _lib.Kinetics, binLo: int, binHi: int, ctx: Context | None = None,
                           capacity=None, hist: KineticsHistory | None = None):
    """isc::PersistentBlinksInBins (core/src/dyes.cpp:648-655): every blink of bins [binLo, binHi] with its bin and
    index (BIN_BLINK_DTYPE), as (offsets, blinks)."""
    ctx = ctx or default_context()
    d, n = _dyes(dyes)
    p, keep = _ptr(d)
    ctx.follow_torch()
    return _records(ctx, ctx.L.iscg_persistent_bins, (p, n, C.byref(kin), _hist(hist), int(binLo), int(binHi)), n,
                    BIN_BLINK_DTYPE, capacity)


def JsMath(fn: str, x, ctx: Context | None = None):
    """The device's jsm::sin / cos / exp / log / cbrt (core/src/jsmath.h) of each element of x (float64), or V8's
    Math.hypot ("hypot2" / "hypot3") of each row of x (n x 2 or n x 3)."""
    ctx = ctx or default_context()
    x = np.ascontiguousarray(x, np.float64)
    fns = ("sin", "cos", "exp", "log", "cbrt", "hypot2", "hypot3")
    k = fns.index(fn)
    n = x.shape[0] if k >= 5 else x.size
    out = np.empty(n) if k >= 5 else np.empty_like(x)
    ctx.follow_torch()
    ctx.check(ctx.L.iscg_jsm(ctx.h, k, x.ctypes.data, n, out.ctypes.data), "iscg_jsm")
    return out


# ---- the cell field's dye blocks and window queries ----------------------------------------------------------------

# isc::World's structures (core/src/world.h:54-58) and the rows of its queries (iscg_world_dye, iscg_world_event)
STRUCTURE_MT, STRUCTURE_NPC, STRUCTURE_LAMIN, STRUCTURE_ORIGAMI = 0, 1, 2, 3
WORLD_CELL_DTYPE = np.dtype({"names": ["cx", "cy", "x", "y", "packRot", "rOuter"],
                             "formats": [np.int32, np.int32, np.float64, np.float64, np.float64, np.float64],
                             "offsets": [0, 4, 8, 16, 24, 32], "itemsize": 40})
NPC_DTYPE = np.dtype({"names": ["c", "n", "e1", "e2", "cand"],
                      "formats": [(np.float64, 3), (np.float64, 3), (np.float64, 3), (np.float64, 3), np.int32],
                      "offsets": [0, 24, 48, 72, 96], "itemsize": 104})
LAMIN_DTYPE = np.dtype({"names": ["c", "n", "cand"], "formats": [(np.float64, 3), (np.float64, 3), np.int32],
                        "offsets": [0, 24, 48], "itemsize": 56})
WORLD_DYE_DTYPE = np.dtype({"names": ["x", "y", "z", "id", "cx", "cy", "mtIndex", "k", "n", "structure"],
                            "formats": [np.float64] * 3 + [np.uint32] + [np.int32] * 6,
                            "offsets": [0, 8, 16, 24, 28, 32, 36, 40, 44, 48], "itemsize": 56})
BLOCK_DYE_DTYPE = np.dtype({"names": ["x", "y", "z", "id", "block", "structure", "element", "k", "n"],
                            "formats": [np.float64] * 3 + [np.uint32] + [np.int32] * 5,
                            "offsets": [0, 8, 16, 24, 28, 32, 36, 40, 44], "itemsize": 56})
WORLD_EVENT_DTYPE = np.dtype({"names": ["x", "y", "z", "tOn", "tOff", "brightness", "aux", "id", "structure", "state"],
                              "formats": [np.float64] * 7 + [np.uint32, np.uint8, np.uint8],
                              "offsets": [0, 8, 16, 24, 32, 40, 48, 56, 60, 61], "itemsize": 64})


@dataclass
class WorldGeometry:
    """A cell field's cells as isc::CellAssets holds them (core/src/world.h:58-90), flattened: cells in the walk's
    order (WORLD_CELL_DTYPE), cell c's microtubules [mtStart[c], mtStart[c+1]) with microtubule m's centreline
    pts[ptStart[m]:ptStart[m+1]] (cell-local um) and mtIndex[m] its index in the cell (None: its position), the
    pores (NPC_DTYPE) and lamina points (LAMIN_DTYPE) of cell c at [npcStart[c], npcStart[c+1]) and [laminStart[c],
    laminStart[c+1]) in candidate order, NpcAttachmentPoints (64 x 3 nm) and NpcSiteReachUm."""

    cells: np.ndarray
    mtStart: np.ndarray
    ptStart: np.ndarray
    pts: np.ndarray
    mtIndex: np.ndarray | None = None
    npcStart: np.ndarray | None = None
    npcs: np.ndarray | None = None
    laminStart: np.ndarray | None = None
    lamin: np.ndarray | None = None
    npcSites: np.ndarray = field(default_factory=lambda: np.zeros(192))
    npcSiteReachUm: float = 0.0

    @classmethod
    def from_rows(cls, cells, mtStart, ptStart, pts, npcStart=None, npcRows=None, laminStart=None, laminRows=None,
                  npcSites=None, npcSiteReachUm=0.0, mtIndex=None):
        """From row arrays: cells with fields cx, cy, x, y, packRot, rOuter; pores as rows c, n, e1, e2, cand (13
        values); lamina points as rows c, n, cand (7 values)."""
        c = np.zeros(len(cells), WORLD_CELL_DTYPE)
        for k in WORLD_CELL_DTYPE.names:
            c[k] = cells[k]
        g = cls(cells=c, mtStart=np.asarray(mtStart, np.int32), ptStart=np.asarray(ptStart, np.int64),
                pts=np.asarray(pts, np.float64).reshape(-1, 3), mtIndex=mtIndex)
        if npcRows is not None and len(npcRows):
            r = np.asarray(npcRows, np.float64)
            a = np.zeros(len(r), NPC_DTYPE)
            a["c"], a["n"], a["e1"], a["e2"] = r[:, 0:3], r[:, 3:6], r[:, 6:9], r[:, 9:12]
            a["cand"] = r[:, 12].astype(np.int32)
            g.npcs, g.npcStart = a, np.asarray(npcStart, np.int32)
        if laminRows is not None and len(laminRows):
            r = np.asarray(laminRows, np.float64)
            a = np.zeros(len(r), LAMIN_DTYPE)
            a["c"], a["n"], a["cand"] = r[:, 0:3], r[:, 3:6], r[:, 6].astype(np.int32)
            g.lamin, g.laminStart = a, np.asarray(laminStart, np.int32)
        if npcSites is not None:
            g.npcSites = np.asarray(npcSites, np.float64)
        g.npcSiteReachUm = float(npcSiteReachUm)
        return g


class World:
    """isc::World's dye blocks and window queries (core/src/world.h:152-300) over a cell field's geometry kept on the
    GPU: SitesInWindow, EventsInWindow, ContinuousInWindow, Density3dInWindow and DensityInWindow answer as the
    reference's do, bit for bit and in its order, for one set of kinetics or one kinetics history per structure."""

    def __init__(self, seed: int, geometry: WorldGeometry, ctx: Context | None = None):
        self.ctx = ctx or default_context()
        g = geometry
        keep = []

        def arr(a, dtype):
            if a is None:
                return None
            a = np.ascontiguousarray(a, dtype)
            keep.append(a)
            return a.ctypes.data

        cells = np.ascontiguousarray(g.cells, WORLD_CELL_DTYPE)
        keep.append(cells)
        c = _lib.WorldGeometry()
        c.nCells = cells.size
        c.cells = cells.ctypes.data if cells.size else None
        c.mtStart = arr(g.mtStart, np.int32)
        c.ptStart = arr(g.ptStart, np.int64)
        c.mtIndex = arr(g.mtIndex, np.int32)
        c.pts = arr(np.asarray(g.pts, np.float64).reshape(-1), np.float64)
        c.npcStart = arr(g.npcStart, np.int32)
        c.npcs = arr(g.npcs, NPC_DTYPE) if g.npcs is not None and len(g.npcs) else None
        c.laminStart = arr(g.laminStart, np.int32)
        c.lamin = arr(g.lamin, LAMIN_DTYPE) if g.lamin is not None and len(g.lamin) else None
        c.npcSites[:] = [float(v) for v in np.asarray(g.npcSites, np.float64).reshape(-1)[:192]]
        c.npcSiteReachUm = float(g.npcSiteReachUm)
        h = C.c_void_p()
        self.ctx.follow_torch()
        self.ctx.check(self.ctx.L.iscg_world_create(self.ctx.h, int(seed) & 0xFFFFFFFF, C.byref(c), C.byref(h)),
                       "iscg_world_create")
        self.w = h.value
        self._cap = {}

    def close(self):
        if getattr(self, "w", None):
            self.ctx.L.iscg_world_destroy(self.w)
            self.w = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def SetLabel(self, structure: int, label: Label) -> None:
        """World::SetLabel (core/src/world.cpp:1080)."""
        lab = label.to_c()
        self.ctx.check(self.ctx.L.iscg_world_set_label(self.w, int(structure), C.byref(lab)), "iscg_world_set_label")

    def SetKineticsHistory(self, tStart, kin) -> None:
        """World::SetKineticsHistory (core/src/world.cpp:933): segment starts tStart (0, then increasing) and
        kin[i * 4 + s], structure s's kinetics in segment i; empty clears it."""
        t = np.ascontiguousarray(tStart, np.float64)
        k = (_lib.Kinetics * max(1, len(kin)))(*kin)
        if len(kin) != 4 * t.size:
            raise IscgError("SetKineticsHistory: 4 kinetics per segment")
        self.ctx.check(self.ctx.L.iscg_world_set_history(self.w, t.ctypes.data, C.addressof(k), int(t.size)),
                       "iscg_world_set_history")

    def _rows(self, fn, key, dtype, args):
        win = np.ascontiguousarray(args[0], np.float64)
        total = np.zeros(1, np.int64)
        cap = self._cap.get(key, 1 << 16)
        self.ctx.follow_torch()
        for _ in range(2):
            out = np.zeros(cap, dtype)
            self.ctx.check(fn(self.w, win.ctypes.data, *args[1:], out.ctypes.data, cap, total.ctypes.data),
                           fn.__name__)
            if total[0] <= cap:
                self._cap[key] = max(cap, int(total[0]))
                return out[: total[0]]
            cap = int(total[0])
        raise IscgError(f"{fn.__name__}: row count changed between calls")

    def BlockDyes(self, x0, y0, x1, y1, zMin, zMax):
        """DyesInBlock / NpcDyesInBlock / LaminDyesInBlock (core/src/dyes.cpp:127-253) of every block the window's
        walk takes: BLOCK_DYE_DTYPE rows, cell-local and unrounded, blocks in walk order."""
        return self._rows(self.ctx.L.iscg_world_block_dyes, "blocks", BLOCK_DYE_DTYPE,
                          ((x0, y0, x1, y1, zMin, zMax),))

    def SitesInWindow(self, x0, y0, x1, y1, zMin, zMax):
        """World::SitesInWindow (core/src/world.cpp:1212): WORLD_DYE_DTYPE rows."""
        return self._rows(self.ctx.L.iscg_world_sites, "sites", WORLD_DYE_DTYPE, ((x0, y0, x1, y1, zMin, zMax),))

    def EventsInWindow(self, x0, y0, x1, y1, zMin, zMax, t0, t1):
        """World::EventsInWindow (core/src/world.cpp:1219): WORLD_EVENT_DTYPE rows."""
        return self._rows(self.ctx.L.iscg_world_events, "events", WORLD_EVENT_DTYPE,
                          ((x0, y0, x1, y1, zMin, zMax), float(t0), float(t1)))

    def ContinuousInWindow(self, x0, y0, x1, y1, zMin, zMax, tMin):
        """World::ContinuousInWindow (core/src/world.cpp:1279): WORLD_EVENT_DTYPE rows."""
        return self._rows(self.ctx.L.iscg_world_continuous, "cont", WORLD_EVENT_DTYPE,
                          ((x0, y0, x1, y1, zMin, zMax), float(tMin)))

    def Density3dInWindow(self, x0, y0, x1, y1, zMin, zMax, nx: int, ny: int, nz: int, structureMask: int = 0xF,
                          out=None):
        """World::Density3dInWindow (core/src/world.cpp:1519): (total, counts[nz, ny, nx] float32). out may be a
        device tensor of nx * ny * nz float32."""
        win = np.array([x0, y0, x1, y1, zMin, zMax], np.float64)
        if out is None:
            out = np.zeros((nz, ny, nx), np.float32)
        if _numel(out) != nx * ny * nz:
            raise IscgError("Density3dInWindow: out must hold nx * ny * nz values")
        p, keep = _ptr(out)
        total = np.zeros(1, np.int64)
        self.ctx.follow_torch()
        self.ctx.check(self.ctx.L.iscg_world_density3d(self.w, win.ctypes.data, int(nx), int(ny), int(nz),
                                                       int(structureMask), p, total.ctypes.data),
                       "iscg_world_density3d")
        return int(total[0]), out

    def DensityInWindow(self, x0, y0, x1, y1, zMin, zMax, nx: int, ny: int, out=None):
        """World::DensityInWindow (core/src/world.cpp:1510): every structure, one plane."""
        total, out = self.Density3dInWindow(x0, y0, x1, y1, zMin, zMax, nx, ny, 1, 0xF, out)
        return total, out.reshape(ny, nx) if isinstance(out, np.ndarray) else out
