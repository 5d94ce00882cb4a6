"""The PSF kernel request and its caches, as the reference renderer keeps them around the PSF seam.

- `PsfGeneratorRequest` (Sim/PsfGeneratorBridge.h:91-160): every field of the reference's request, the Zernike
  coefficients as the reference's text (15 or 28 positional OSA values, Sim/SMLMZernike.h).
- `ParseZernikeCoefficients` / `FormatZernikeCoefficients` (Sim/SMLMZernike.cpp:16-86): the text <-> 28 values.
- `ComputePsfKernelCache` (Sim/PsfGeneratorBridge.cpp:895-969): the memo of the last kernels (SameKernel, most
  recent first, at least 2 entries, a second requester of a kernel being computed waits for it), the opt-in disk
  store of the last kernel (psf_kernel.bin: the reference's file layout and FNV-1a fingerprint, :684-833), and on a
  miss the kernel computed on the GPU (iscg_psf_planes, sum-1 planes).

A kernel cache made here is a `host.PsfKernelCache` (planes as numpy float32), the type the splat host takes.
"""
from __future__ import annotations

import dataclasses
import math
import os
import re
import struct
import threading
from dataclasses import dataclass
from fractions import Fraction

import numpy as np

from . import host

# PsfModelKind (Sim/PsfGeneratorBridge.h:42-54)
GAUSSIAN, RICHARDS_WOLF, GIBSON_LANNI, GIBSON_LANNI_ZERNIKE = 0, 1, 2, 3

# ---- Zernike coefficient text (Sim/SMLMZernike.cpp:16-86) -----------------------------------------------------------

_DBL_MIN = 2.2250738585072014e-308
_NUM = re.compile(
    r"[ \t\n\v\f\r]*([+-]?)("
    r"(?P<inf>(?i:infinity|inf))|"
    r"(?P<nan>(?i:nan)(?:\([0-9A-Za-z_]*\))?)|"
    r"(?P<hex>0[xX](?:[0-9a-fA-F]+(?:\.[0-9a-fA-F]*)?|\.[0-9a-fA-F]+)(?:[pP][+-]?[0-9]+)?)|"
    r"(?P<dec>(?:[0-9]+(?:\.[0-9]*)?|\.[0-9]+)(?:[eE][+-]?[0-9]+)?))")


def _stod(token: str):
    """std::stod of a token: the longest prefix strtod converts (leading whitespace skipped), or None where stod
    throws (nothing converted, or ERANGE: overflow, or an inexact result below the smallest normal double)."""
    m = _NUM.match(token)
    if not m:
        return None
    sign = -1.0 if m.group(1) == "-" else 1.0
    if m.group("inf"):
        return sign * math.inf
    if m.group("nan"):
        return math.copysign(math.nan, sign)
    if m.group("hex"):
        h = m.group("hex")
        if "p" not in h.lower():
            h += "p0"
        v = float.fromhex(h)
        mant, exp = re.split(r"[pP]", h)
        digits = mant[2:].replace(".", "")
        frac = len(mant[2:].partition(".")[2])
        exact = Fraction(int(digits, 16)) * Fraction(2) ** (int(exp) - 4 * frac)
    else:
        d = m.group("dec")
        v = float(d)
        exact = Fraction(d)
    if math.isinf(v):
        return None
    # glibc's strtod reports ERANGE for a tiny result (subnormal, or zero from a non-zero value) that is inexact
    if abs(v) < _DBL_MIN and exact != 0 and Fraction(v) != exact:
        return None
    return sign * v


def ParseZernikeCoefficients(text: str):
    """sim::ParseZernikeCoefficients (Sim/SMLMZernike.cpp:28-86): (28 float64 values, ok). Commas, semicolons,
    spaces and tabs separate values; two separators with no value between are an error; 15 values fill slots 0-14,
    28 all slots; anything else, or a value std::stod rejects, gives (zeros, False)."""
    zero = np.zeros(28)
    tokens, token, pending = [], "", False
    for ch in text:
        sep = ch in ",;"
        if sep or ch in " \t":
            if token:
                tokens.append(token)
                token = ""
                pending = False
            if sep:
                if pending:
                    return zero, False
                pending = True
            continue
        token += ch
    if token:
        tokens.append(token)
    if len(tokens) not in (15, 28):
        return zero, False
    out = np.zeros(28)
    for i, t in enumerate(tokens):
        v = _stod(t)
        if v is None:
            return zero, False
        out[i] = v
    return out, True


def FormatZernikeCoefficients(coeffs, separator: str = ",") -> str:
    """sim::FormatZernikeCoefficients (Sim/SMLMZernike.cpp:16-26): each value as an ostream prints a double (6
    significant digits, %g), joined by separator."""
    return separator.join("%g" % float(v) for v in coeffs)


# ---- the request ---------------------------------------------------------------------------------------------------

@dataclass
class PsfGeneratorRequest:
    """sim::PsfGeneratorRequest (Sim/PsfGeneratorBridge.h:91-160) with the reference's defaults."""

    model: int = RICHARDS_WOLF
    wavelengthNm: float = 660.0
    na: float = 1.4
    immersionIndex: float = 1.518
    sampleIndex: float = 1.518
    workingDistanceUm: float = 150.0
    sampleDepthNm: float = 0.0
    pixelSizeNm: float = 100.0
    zernikeCoefficients: str = "0,0,0,0,0,0,0,0,0,0,0,0,0,0,0"
    maskType: int = 0
    maskModes: int = 5
    maskWaist: float = 1.0
    oversampling: int = 4
    kernelHalfWidthPx: int = 8
    nz: int = 1
    zStepNm: float = 100.0
    pupilSamples: int = 0
    javaHome: str = ""
    interpMode: int = 0

    def psf_request(self):
        """The iscg_psf_request of the GPU seam (the coefficient text parsed as ComputeZernikePsfPlanes does: a text
        that does not parse gives no aberration, Sim/ZernikePsf.cpp:174-176)."""
        z, _ = ParseZernikeCoefficients(self.zernikeCoefficients)
        return host.psf_request(dict(
            wavelengthNm=self.wavelengthNm, na=self.na, immersionIndex=self.immersionIndex,
            sampleIndex=self.sampleIndex, workingDistanceUm=self.workingDistanceUm, sampleDepthNm=self.sampleDepthNm,
            pixelSizeNm=self.pixelSizeNm, zStepNm=self.zStepNm, maskWaist=self.maskWaist, maskType=self.maskType,
            maskModes=self.maskModes, oversampling=self.oversampling, kernelHalfWidthPx=self.kernelHalfWidthPx,
            nz=self.nz, pupilSamples=self.pupilSamples, interpMode=self.interpMode, zernike=z))


def SameKernel(a: PsfGeneratorRequest, b: PsfGeneratorRequest) -> bool:
    """SameKernel / PsfRequestsSameKernel (Sim/PsfGeneratorBridge.cpp:672-683): every field the planes depend on
    (all but interpMode and javaHome), compared exactly."""
    keys = ("model", "wavelengthNm", "na", "immersionIndex", "sampleIndex", "workingDistanceUm", "sampleDepthNm",
            "pixelSizeNm", "zernikeCoefficients", "maskType", "maskModes", "maskWaist", "oversampling",
            "kernelHalfWidthPx", "nz", "zStepNm", "pupilSamples")
    return all(getattr(a, k) == getattr(b, k) for k in keys)


PsfRequestsSameKernel = SameKernel

# ---- the disk store (Sim/PsfGeneratorBridge.cpp:684-833) -----------------------------------------------------------

PSF_FILE_MAGIC = b"ISCPSF01"
PSF_KERNEL_CODE_VERSION = "2026-10-07"   # the reference's kernel code version: its files are taken as ours
PSF_DISK_MAX_BYTES = 512.0 * 1024 * 1024
_FNV_OFFSET, _FNV_PRIME, _MASK64 = 14695981039346656037, 1099511628211, (1 << 64) - 1

_disk_lock = threading.Lock()
_disk_dir = ""


def _fnv_mix(h: int, data: bytes) -> int:
    for b in data:
        h = ((h ^ b) * _FNV_PRIME) & _MASK64
    return h


def KernelFingerprint(r: PsfGeneratorRequest) -> int:
    """KernelFingerprint (Sim/PsfGeneratorBridge.cpp:704-717): FNV-1a 64 over the code version, 7 int32 (model,
    maskType, maskModes, oversampling, kernelHalfWidthPx, nz, pupilSamples), 9 doubles (wavelength, NA, ni, ns,
    working distance, depth, pixel, mask waist, z step) and the coefficient text's bytes."""
    h = _fnv_mix(_FNV_OFFSET, PSF_KERNEL_CODE_VERSION.encode())
    h = _fnv_mix(h, struct.pack("<7i", r.model, r.maskType, r.maskModes, r.oversampling, r.kernelHalfWidthPx, r.nz,
                                r.pupilSamples))
    h = _fnv_mix(h, struct.pack("<9d", r.wavelengthNm, r.na, r.immersionIndex, r.sampleIndex, r.workingDistanceUm,
                                r.sampleDepthNm, r.pixelSizeNm, r.maskWaist, r.zStepNm))
    return _fnv_mix(h, r.zernikeCoefficients.encode())


def SetPsfKernelDiskCacheDir(d: str) -> None:
    """SetPsfKernelDiskCacheDir (Sim/PsfGeneratorBridge.h:359): psf_kernel.bin in d ("" = no disk store)."""
    global _disk_dir
    with _disk_lock:
        _disk_dir = d


def PsfKernelDiskCacheDir() -> str:
    with _disk_lock:
        return _disk_dir


def _disk_path() -> str:
    d = PsfKernelDiskCacheDir()
    return os.path.join(d, "psf_kernel.bin") if d else ""


def LoadKernelFromDisk(req: PsfGeneratorRequest):
    """LoadKernelFromDisk (Sim/PsfGeneratorBridge.cpp:727-776): the stored kernel when its fingerprint is the
    request's and its geometry is sane, else None."""
    path = _disk_path()
    if not path or not os.path.isfile(path):
        return None
    try:
        with open(path, "rb") as f:
            head = f.read(8 + 8 + 20 + 8)
            if len(head) != 44 or head[:8] != PSF_FILE_MAGIC:
                return None
            fp, = struct.unpack("<Q", head[8:16])
            os_, half, size, nz, bw = struct.unpack("<5i", head[16:36])
            zstep, = struct.unpack("<d", head[36:44])
            if fp != KernelFingerprint(req):
                return None
            if (os_ < 1 or os_ > 64 or half < 0 or size != 2 * half + 1 or size > 8192 or nz < 1 or nz > 4096 or
                    bw != size + os_ - 1):
                return None
            n = nz * size * size
            planes = np.frombuffer(f.read(4 * n), dtype="<f4")
            if planes.size != n:
                return None
    except OSError:
        return None
    return host.PsfKernelCache(planes.reshape(nz, size, size).astype(np.float32), oversampling=os_, zStepNm=zstep,
                               interpMode=req.interpMode)


def SaveKernelToDisk(req: PsfGeneratorRequest, c: host.PsfKernelCache) -> None:
    """SaveKernelToDisk (Sim/PsfGeneratorBridge.cpp:778-813): the whole file rewritten through a temporary (planes
    up to 512 MB)."""
    path = _disk_path()
    planes = np.asarray(c.planes)
    if not path or planes.size == 0:
        return
    if float(c.nz) * c.sizeOversampled * c.sizeOversampled * 4 > PSF_DISK_MAX_BYTES:
        return
    os.makedirs(os.path.dirname(path), exist_ok=True)
    tmp = path + ".tmp"
    try:
        with open(tmp, "wb") as f:
            f.write(PSF_FILE_MAGIC)
            f.write(struct.pack("<Q", KernelFingerprint(req)))
            f.write(struct.pack("<5i", c.oversampling, c.halfWidthOversampled, c.sizeOversampled, c.nz,
                                c.blockSumWidth))
            f.write(struct.pack("<d", c.zStepNm))
            f.write(np.ascontiguousarray(planes, dtype="<f4").tobytes())
        os.replace(tmp, path)
    except OSError:
        try:
            os.remove(tmp)
        except OSError:
            pass


# ---- the memo (Sim/PsfGeneratorBridge.cpp:815-969) ------------------------------------------------------------------

class _KernelMemo:
    def __init__(self):
        self.cv = threading.Condition()
        self.entries: list[tuple[PsfGeneratorRequest, host.PsfKernelCache]] = []   # most recent first
        self.in_flight: list[PsfGeneratorRequest] = []
        self.keep = 2


_memo = _KernelMemo()


def SetPsfKernelMemoEntries(n: int) -> None:
    """SetPsfKernelMemoEntries (Sim/PsfGeneratorBridge.h:363): the memo keeps max(2, n) kernels."""
    with _memo.cv:
        _memo.keep = max(2, int(n))


def _compute_uncached(req: PsfGeneratorRequest, ctx):
    """ComputePsfKernelCacheUncached (Sim/PsfGeneratorBridge.cpp:835-871): GibsonLanniZernike on the GPU (the
    sum-1 planes of BuildZernikePsfKernelCache); the JVM models are not available here."""
    if req.model != GIBSON_LANNI_ZERNIKE:
        return None, ("The RichardsWolf/GibsonLanni models need the embedded PSFGenerator JVM bridge, only "
                      "implemented for Windows.")
    r = req.psf_request()
    if not (r.na > 0.0 and r.wavelengthNm > 0.0 and r.immersionIndex > 0.0 and r.sampleIndex > 0.0 and
            r.pixelSizeNm > 0.0 and r.zStepNm > 0.0):
        return None, "GibsonLanniZernike PSF: NA, wavelength, refractive indices, pixel size and z step must be > 0."
    planes = host.ComputePsfPlanesBatch([r], normalize=True, ctx=ctx)[0]
    return host.PsfKernelCache(planes, oversampling=max(1, r.oversampling), zStepNm=r.zStepNm,
                               interpMode=r.interpMode), ""


def ComputePsfKernelCache(req: PsfGeneratorRequest, logCallback=None, ctx=None):
    """sim::ComputePsfKernelCache (Sim/PsfGeneratorBridge.cpp:895-969): (ok, cache, error). A request with the same
    kernel as a memo entry returns that entry's planes (same serial) with the request's interpMode and moves it to
    the front; while another thread computes the same kernel, waits for it; else reads the disk store, or computes
    (on the GPU) and stores, then puts the kernel at the front of the memo (max(2, entries) kept)."""
    m = _memo
    with m.cv:
        while True:
            for i, (r, c) in enumerate(m.entries):
                if SameKernel(r, req):
                    m.entries.insert(0, m.entries.pop(i))
                    if logCallback:
                        logCallback("PSF: kernel unchanged, reusing the one already computed.")
                    return True, dataclasses.replace(c, interpMode=req.interpMode), ""
            if not any(SameKernel(r, req) for r in m.in_flight):
                break
            m.cv.wait()
        m.in_flight.append(req)
    try:
        cache = LoadKernelFromDisk(req)
        if cache is not None:
            if logCallback:
                logCallback("PSF: kernel read from the disk cache (" + PsfKernelDiskCacheDir() + ").")
        else:
            if req.model == GIBSON_LANNI_ZERNIKE and logCallback:
                logCallback("PSF: computing the GibsonLanniZernike kernel (GPU, %d px square, %d Z planes)..." % (
                    2 * max(1, req.kernelHalfWidthPx) * max(1, req.oversampling) + 1, max(3, req.nz)))
            cache, err = _compute_uncached(req, ctx)
            if cache is None:
                return False, None, err
            SaveKernelToDisk(req, cache)
        with m.cv:
            m.entries.insert(0, (dataclasses.replace(req), cache))
            del m.entries[m.keep:]
        return True, cache, ""
    finally:
        with m.cv:
            for i, r in enumerate(m.in_flight):
                if SameKernel(r, req):
                    del m.in_flight[i]
                    break
            m.cv.notify_all()
