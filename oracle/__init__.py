"""ctypes bindings of the CPU checkers.

TEST INFRASTRUCTURE ONLY. ``liboracle.so`` is the FP64 C restatement of the reference's image-formation path
(``oracle.c``); ``_ref/libiscref.so`` is the reference's own CPU code compiled from its sources (``ref_shim.cpp``),
present only where it was built. Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline`` leg
import this package; the product library never does.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "build", "liboracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libiscref.so")
REFERENCE_ROOT = "/root/reference"


def build(ref: bool = True) -> None:
    """Compile liboracle.so, and libiscref.so and libiscbridge.so (which links libiscg.so: build it first) when the
    reference checkout is present."""
    subprocess.run(["make", "-C", HERE, "oracle"], check=True, stdout=subprocess.DEVNULL)
    if ref and os.path.isdir(REFERENCE_ROOT):
        subprocess.run(["make", "-C", HERE, "-j8", "ref", "bridge", f"REF={REFERENCE_ROOT}"], check=True,
                       stdout=subprocess.DEVNULL)


class Camera(C.Structure):
    """orc_camera / sim::CameraNoiseParams (Sim/SMLMNoise.h:35-50)."""

    _fields_ = [("quantumEfficiency", C.c_double), ("darkCurrentElectrons", C.c_double),
                ("gainPhotonsPerAdu", C.c_double), ("readNoiseElectrons", C.c_double),
                ("emccd", C.c_int32), ("bitDepth", C.c_int32), ("emGain", C.c_double),
                ("cicElectrons", C.c_double)]


class PsfRequest(C.Structure):
    """orc_psf_request: the numeric fields of sim::PsfGeneratorRequest (Sim/PsfGeneratorBridge.h:91-170)."""

    _fields_ = [("wavelengthNm", C.c_double), ("na", C.c_double), ("immersionIndex", C.c_double),
                ("sampleIndex", C.c_double), ("workingDistanceUm", C.c_double), ("sampleDepthNm", C.c_double),
                ("pixelSizeNm", C.c_double), ("zStepNm", C.c_double), ("maskWaist", C.c_double),
                ("zernike", C.c_double * 28), ("maskType", C.c_int32), ("maskModes", C.c_int32),
                ("oversampling", C.c_int32), ("kernelHalfWidthPx", C.c_int32), ("nz", C.c_int32),
                ("pupilSamples", C.c_int32), ("interpMode", C.c_int32), ("reserved", C.c_int32)]


class RoiModel(C.Structure):
    _fields_ = [("wavelengthNm", C.c_double), ("na", C.c_double), ("immersionIndex", C.c_double),
                ("sampleIndex", C.c_double), ("sampleDepthNm", C.c_double), ("pixelSizeNm", C.c_double),
                ("pupilSamples", C.c_int32), ("roiSize", C.c_int32), ("oversampling", C.c_int32),
                ("reserved", C.c_int32)]


ROI_PARAMS_DTYPE = np.dtype([("x", "f8"), ("y", "f8"), ("z", "f8"), ("photons", "f8"), ("background", "f8")])


class Emitter(C.Structure):
    _fields_ = [("xPx", C.c_double), ("yPx", C.c_double), ("photons", C.c_double),
                ("zIndex", C.c_int32), ("kernel", C.c_int32)]


class SplatSetupT(C.Structure):
    _fields_ = [("x0", C.c_int32), ("y0", C.c_int32), ("bx", C.c_int32), ("by", C.c_int32),
                ("nTaps", C.c_int32), ("pad", C.c_int32), ("wx", C.c_double * 4), ("wy", C.c_double * 4)]


EVENT_DTYPE = np.dtype([("xUm", "f8"), ("yUm", "f8"), ("zNm", "f8"), ("tStart", "f8"), ("tEnd", "f8"),
                        ("brightness", "f8")])
EMITTER_DTYPE = np.dtype([("xPx", "f8"), ("yPx", "f8"), ("photons", "f8"), ("zIndex", "i4"), ("kernel", "i4")])
# sim::GpuSplatEmitter (Sim/GpuSimD3D11.h:50-59)
GPU_EMITTER_DTYPE = np.dtype([("x0", "i4"), ("y0", "i4"), ("bx", "i4"), ("by", "i4"), ("plane", "i4"),
                              ("nTaps", "i4"), ("photons", "f4"), ("cutRadius2", "i4"), ("wx", "f4", 4),
                              ("wy", "f4", 4)])


class Kinetics(C.Structure):
    """core/src/dyes.h:129-136."""
    _fields_ = [(n, C.c_double) for n in ("activationRatePerSec", "onSec", "offSec", "bleachProb", "photonCV",
                                          "initialOnSec")]


def kinetics(**kw) -> Kinetics:
    """Kinetics with the reference's defaults (core/src/dyes.h:130-135), overridden by kw."""
    k = Kinetics(0.01, 0.05, 1.0, 1.0, 0.5, 0.0)
    for a, v in kw.items():
        setattr(k, a, float(v))
    return k


JSM_FUNCS = ("sin", "cos", "exp", "log")


def _p(a, ct):
    return None if a is None else a.ctypes.data_as(C.POINTER(ct))


def camera(**kw) -> Camera:
    c = Camera(quantumEfficiency=1.0, darkCurrentElectrons=0.0, gainPhotonsPerAdu=1.0, readNoiseElectrons=1.5,
               emccd=0, bitDepth=16, emGain=300.0, cicElectrons=0.002)
    for k, v in kw.items():
        setattr(c, k, v)
    return c


def psf_request(**kw) -> PsfRequest:
    r = PsfRequest(wavelengthNm=660.0, na=1.4, immersionIndex=1.518, sampleIndex=1.518, workingDistanceUm=150.0,
                   sampleDepthNm=0.0, pixelSizeNm=100.0, zStepNm=100.0, maskWaist=1.0, maskType=0, maskModes=5,
                   oversampling=4, kernelHalfWidthPx=8, nz=1, pupilSamples=0, interpMode=0, reserved=0)
    for k, v in kw.items():
        if k == "zernike":
            z = np.zeros(28)
            if isinstance(v, dict):
                for j, a in v.items():
                    z[j] = a
            else:
                z[: len(v)] = v
            for j in range(28):
                r.zernike[j] = float(z[j])
        else:
            setattr(r, k, v)
    return r


class Oracle:
    """The FP64 restatement (oracle.c)."""

    def __init__(self, path: str = LIB_PATH):
        if not os.path.exists(path):
            build(ref=False)
        L = C.CDLL(path)
        self.L = L
        D, I32, U32, I64 = C.c_double, C.c_int32, C.c_uint32, C.c_int64
        PD, PF, PI16 = C.POINTER(C.c_double), C.POINTER(C.c_float), C.POINTER(C.c_int16)
        L.orc_pcg4d.argtypes = [C.POINTER(C.c_uint32)]
        L.orc_hash_unit.argtypes = [U32, I32, I32, U32]
        L.orc_hash_unit.restype = D
        L.orc_hash_stream.argtypes = [U32, I32, I32, U32, I32, PD]
        L.orc_counter_uniforms.argtypes = [U32, U32, U32, I32, PD]
        L.orc_counter_gauss.argtypes = [U32, U32, U32, I32, PD]
        L.orc_counter_poisson.argtypes = [U32, U32, U32, D]
        L.orc_counter_poisson.restype = D
        L.orc_counter_gamma.argtypes = [U32, U32, U32, D]
        L.orc_counter_gamma.restype = D
        L.orc_noise_chain.argtypes = [PF, C.POINTER(C.c_uint16), U32, U32, C.POINTER(Camera), PF, PF, PF, U32, U32,
                                      U32]
        L.orc_psf_geometry.argtypes = [C.POINTER(PsfRequest), C.POINTER(I32), C.POINTER(I32), C.POINTER(I32)]
        L.orc_psf_planes.argtypes = [C.POINTER(PsfRequest), PF]
        L.orc_psf_planes.restype = I32
        L.orc_psf_fields.argtypes = [C.POINTER(PsfRequest), PD]
        L.orc_psf_fields.restype = I32
        L.orc_psf_normalize.argtypes = [PF, I32, I32]
        L.orc_block_sums.argtypes = [PF, I32, I32, PF]
        L.orc_nearest_z.argtypes = [I32, D, D, C.POINTER(I32)]
        L.orc_nearest_z.restype = I32
        L.orc_splat_setup.argtypes = [I32, I32, D, D, I32, C.POINTER(SplatSetupT)]
        L.orc_render_frame.argtypes = [PF, U32, U32, D, PF, D, C.c_void_p, I32, PF, I32, I32, I32, I32, I32, PI16,
                                       PI16, D]
        L.orc_fft_shift_tile.argtypes = [PF, I32, D, D, PF]
        L.orc_render_frame_fft.argtypes = [PF, U32, U32, D, PF, D, C.c_void_p, I32, PF, I32, I32, I32, PI16, PI16,
                                           D]
        L.orc_wf_images_render.argtypes = [PF, I32, I32, U32, U32, I32, PD, I32, PF]
        PI32 = C.POINTER(C.c_int32)
        L.orc_binned_geometry.argtypes = [I32, I32, I32, I32, D, I32, PI32, PI32]
        L.orc_binned_gauss.argtypes = [D, I32, I32, PF]
        L.orc_binned_render.argtypes = [PF, I32, I32, I32, I32, I32, PD, I32, PF]
        L.orc_wf_radius.argtypes = [I32, I32, I32]
        L.orc_wf_radius.restype = I32
        L.orc_wf_kernel.argtypes = [PF, I32, I32, I32, I32, PF]
        L.orc_wf_direct.argtypes = [PF, I32, I32, I32, D, I64, PF, I32, I32, D, PD, I32, PD]
        L.orc_wf_plane_weight.argtypes = [D, D, I32, D, C.POINTER(I32), PD]
        L.orc_roi_forward.argtypes = [C.POINTER(RoiModel), PD, C.c_void_p, I32, PD]
        L.orc_roi_backward.argtypes = [C.POINTER(RoiModel), PD, C.c_void_p, I32, PD, PD, PD]
        L.orc_roi_backward.restype = D
        for f in JSM_FUNCS:
            getattr(L, "orc_jsm_" + f).argtypes = [D]
            getattr(L, "orc_jsm_" + f).restype = D
        PK = C.POINTER(Kinetics)
        L.orc_label_blinks.argtypes = [U32, I32, I32, I32, PK, D, PD, I32]
        L.orc_label_blinks.restype = I32
        L.orc_label_cont.argtypes = [U32, I32, I32, I32, I32, PK, PD]
        L.orc_label_cont.restype = I32
        L.orc_persistent_blinks.argtypes = [U32, I32, I32, PK, I32, D, D, I64, I64, PD, I32]
        L.orc_persistent_blinks.restype = I32

    # -- RNG --
    def pcg4d(self, a, b, c, d):
        v = (C.c_uint32 * 4)(a & 0xFFFFFFFF, b & 0xFFFFFFFF, c & 0xFFFFFFFF, d & 0xFFFFFFFF)
        self.L.orc_pcg4d(v)
        return tuple(v)

    def hash_unit(self, seed, cx, cy, k):
        return self.L.orc_hash_unit(seed & 0xFFFFFFFF, _i32(cx), _i32(cy), k & 0xFFFFFFFF)

    def hash_stream(self, seed, cx, cy, base, n):
        out = np.zeros(n)
        self.L.orc_hash_stream(seed & 0xFFFFFFFF, _i32(cx), _i32(cy), base & 0xFFFFFFFF, n, _p(out, C.c_double))
        return out

    def counter_uniforms(self, seed, frame, pixel, n):
        out = np.zeros(n)
        self.L.orc_counter_uniforms(seed, frame, pixel, n, _p(out, C.c_double))
        return out

    def counter_gauss(self, seed, frame, pixel, n):
        out = np.zeros(n)
        self.L.orc_counter_gauss(seed, frame, pixel, n, _p(out, C.c_double))
        return out

    def counter_poisson(self, seed, frame, pixel, l):
        return self.L.orc_counter_poisson(seed, frame, pixel, l)

    def counter_gamma(self, seed, frame, pixel, k):
        return self.L.orc_counter_gamma(seed, frame, pixel, k)

    # -- noise --
    def noise(self, photons, cam: Camera, seed, frame, offset=None, gain=None, read=None, pixel_base=0):
        photons = np.ascontiguousarray(photons, dtype=np.float32)
        h, w = photons.shape
        out = np.zeros((h, w), np.uint16)
        f = lambda a: None if a is None else np.ascontiguousarray(a, np.float32)
        offset, gain, read = f(offset), f(gain), f(read)
        self.L.orc_noise_chain(_p(photons, C.c_float), _p(out, C.c_uint16), w, h, C.byref(cam),
                               _p(offset, C.c_float), _p(gain, C.c_float), _p(read, C.c_float), seed, frame,
                               pixel_base)
        return out

    # -- PSF --
    def psf_geometry(self, req):
        s, nz, m = C.c_int32(), C.c_int32(), C.c_int32()
        self.L.orc_psf_geometry(C.byref(req), C.byref(s), C.byref(nz), C.byref(m))
        return s.value, nz.value, m.value

    def psf_planes(self, req, normalize=False):
        size, nz, _ = self.psf_geometry(req)
        out = np.zeros((nz, size, size), np.float32)
        if self.L.orc_psf_planes(C.byref(req), _p(out, C.c_float)) != 0:
            raise ValueError("bad PSF request")
        if normalize:
            self.L.orc_psf_normalize(_p(out, C.c_float), nz, size)
        return out

    def psf_fields(self, req):
        size, nz, _ = self.psf_geometry(req)
        out = np.zeros((nz, size, size, 2), np.float64)
        if self.L.orc_psf_fields(C.byref(req), _p(out, C.c_double)) != 0:
            raise ValueError("bad PSF request")
        return out[..., 0] + 1j * out[..., 1]

    def normalize(self, planes):
        planes = np.ascontiguousarray(planes, np.float32).copy()
        self.L.orc_psf_normalize(_p(planes, C.c_float), planes.shape[0], planes.shape[1])
        return planes

    def block_sums(self, kernel, os_):
        kernel = np.ascontiguousarray(kernel, np.float32)
        n = kernel.shape[0]
        out = np.zeros((n + os_ - 1, n + os_ - 1), np.float32)
        self.L.orc_block_sums(_p(kernel, C.c_float), n, os_, _p(out, C.c_float))
        return out

    def nearest_z(self, nz, z_step_nm, z_um):
        c = C.c_int32()
        return self.L.orc_nearest_z(nz, z_step_nm, z_um, C.byref(c)), bool(c.value)

    def splat_setup(self, size, os_, x, y, interp):
        r = SplatSetupT()
        self.L.orc_splat_setup(size, os_, x, y, interp, C.byref(r))
        return r

    def render_frame(self, w, h, bg, emitters, block_sums, size, os_, interp, bg_map=None, bg_scale=1.0,
                     span_lo=None, span_hi=None, min_photons=0.0):
        emitters = np.ascontiguousarray(emitters, dtype=EMITTER_DTYPE)
        bs = np.ascontiguousarray(block_sums, np.float32)
        img = np.zeros((h, w), np.float32)
        bg_map = None if bg_map is None else np.ascontiguousarray(bg_map, np.float32)
        lo = None if span_lo is None else np.ascontiguousarray(span_lo, np.int16)
        hi = None if span_hi is None else np.ascontiguousarray(span_hi, np.int16)
        self.L.orc_render_frame(_p(img, C.c_float), w, h, bg, _p(bg_map, C.c_float), bg_scale,
                                emitters.ctypes.data if len(emitters) else None, len(emitters),
                                _p(bs, C.c_float), bs.shape[0], bs.shape[1], os_, size, interp,
                                _p(lo, C.c_int16), _p(hi, C.c_int16), min_photons)
        return img

    # -- mean field --
    def fft_shift_tile(self, kernel, shift_x, shift_y):
        k = np.ascontiguousarray(kernel, np.float32)
        out = np.zeros_like(k)
        self.L.orc_fft_shift_tile(_p(k, C.c_float), k.shape[0], shift_x, shift_y, _p(out, C.c_float))
        return out

    def render_frame_fft(self, w, h, bg, emitters, planes, os_, span_lo=None, span_hi=None, min_photons=0.0):
        emitters = np.ascontiguousarray(emitters, dtype=EMITTER_DTYPE)
        planes = np.ascontiguousarray(planes, np.float32)
        img = np.zeros((h, w), np.float32)
        lo = None if span_lo is None else np.ascontiguousarray(span_lo, np.int16)
        hi = None if span_hi is None else np.ascontiguousarray(span_hi, np.int16)
        self.L.orc_render_frame_fft(_p(img, C.c_float), w, h, bg, None, 1.0,
                                    emitters.ctypes.data if len(emitters) else None, len(emitters),
                                    _p(planes, C.c_float), planes.shape[0], os_, planes.shape[1],
                                    _p(lo, C.c_int16), _p(hi, C.c_int16), min_photons)
        return img

    def wf_images_render(self, channels, has_persistent, upscale, a, cam=None):
        ch_ = np.ascontiguousarray(channels, np.float32)
        nC, chh, cw = ch_.shape
        a = np.ascontiguousarray(a, np.float64)
        if cam is None:
            cam = np.zeros((chh // max(1, upscale), cw // max(1, upscale)), np.float32)
        cam = np.ascontiguousarray(cam, np.float32).copy()
        self.L.orc_wf_images_render(_p(ch_, C.c_float), nC, int(has_persistent), cw, chh, upscale, _p(a, C.c_double),
                                    len(a), _p(cam, C.c_float))
        return cam

    def wf_radius(self, size, os_, upscale):
        return self.L.orc_wf_radius(size, os_, upscale)

    def binned_geometry(self, planes, os_, upscale, sigma_px=0.0, halo_r2=-1):
        """(u, R) of BinnedBlinkRenderer::Setup (Sim/BinnedBlinks.cpp:94-129); planes None: the Gaussian."""
        u, R = C.c_int32(0), C.c_int32(0)
        size = 0 if planes is None else np.asarray(planes).shape[-1]
        self.L.orc_binned_geometry(int(planes is not None), size, os_, upscale, sigma_px, halo_r2, C.byref(u),
                                   C.byref(R))
        return u.value, R.value

    def binned_kernels(self, planes, os_, upscale, R, sigma_px=0.0):
        """The per-plane grid kernels nz x (2R+1)^2: KernelWidefieldPsf::Kernel, or the Gaussian (:146-160)."""
        if planes is None:
            u, _ = self.binned_geometry(None, 1, upscale, sigma_px)
            out = np.zeros((1, 2 * R + 1, 2 * R + 1), np.float32)
            self.L.orc_binned_gauss(sigma_px, u, R, _p(out, C.c_float))
            return out
        return np.stack([self.wf_kernel(p, os_, upscale, R) for p in planes])

    def binned_render(self, kernels, u, R, W, H, emitters, img=None):
        """BinnedBlinkRenderer::Render (Sim/BinnedBlinks.cpp:165-246) of one frame's emitters (n x 4: xPx, yPx,
        photons, zIndex) added to img (H x W), as a direct FP64 convolution of the per-plane deposits."""
        k = np.ascontiguousarray(kernels, np.float32)
        em = np.ascontiguousarray(emitters, np.float64).reshape(-1, 4)
        img = np.zeros((H, W), np.float32) if img is None else np.ascontiguousarray(img, np.float32).copy()
        self.L.orc_binned_render(_p(k, C.c_float), k.shape[0], u, R, W, H, _p(em, C.c_double), em.shape[0],
                                 _p(img, C.c_float))
        return img

    def wf_kernel(self, plane, os_, upscale, R):
        plane = np.ascontiguousarray(plane, np.float32)
        out = np.zeros((2 * R + 1, 2 * R + 1), np.float32)
        self.L.orc_wf_kernel(_p(plane, C.c_float), plane.shape[0], os_, upscale, R, _p(out, C.c_float))
        return out

    def wf_plane_weight(self, centre_um, focus_um, nzP, z_step_nm):
        """PlanFocus's blend of one dye plane at one focus: (p0, fr): plane p0 weighs 1 - fr, plane p0 + 1 fr."""
        p0, fr = C.c_int32(), C.c_double()
        self.L.orc_wf_plane_weight(centre_um, focus_um, nzP, z_step_nm, C.byref(p0), C.byref(fr))
        return p0.value, fr.value

    def wf_direct(self, density, dz_um, k0, kernels, z_step_nm, foci_um):
        density = np.ascontiguousarray(density, np.float32)
        kernels = np.ascontiguousarray(kernels, np.float32)
        foci = np.ascontiguousarray(foci_um, np.float64)
        nzD, ny, nx = density.shape
        nzP, D, _ = kernels.shape
        out = np.zeros((len(foci), ny, nx))
        self.L.orc_wf_direct(_p(density, C.c_float), nx, ny, nzD, dz_um, k0, _p(kernels, C.c_float), nzP,
                             (D - 1) // 2, z_step_nm, _p(foci, C.c_double), len(foci), _p(out, C.c_double))
        return out


    # -- ROI model and adjoint --
    def roi_forward(self, model, zernike, params):
        params = np.ascontiguousarray(params, ROI_PARAMS_DTYPE)
        z = np.ascontiguousarray(zernike, np.float64)
        R = model.roiSize
        mu = np.zeros((len(params), R, R))
        self.L.orc_roi_forward(C.byref(model), _p(z, C.c_double), params.ctypes.data, len(params),
                               _p(mu, C.c_double))
        return mu

    def roi_backward(self, model, zernike, params, data):
        params = np.ascontiguousarray(params, ROI_PARAMS_DTYPE)
        z = np.ascontiguousarray(zernike, np.float64)
        d = np.ascontiguousarray(data, np.float64)
        gp = np.zeros((len(params), 5))
        gz = np.zeros(28)
        nll = self.L.orc_roi_backward(C.byref(model), _p(z, C.c_double), params.ctypes.data, len(params),
                                      _p(d, C.c_double), _p(gp, C.c_double), _p(gz, C.c_double))
        return nll, gp, gz


    # -- JS-compatible math and dye schedules --
    def jsm(self, fn: str, x):
        """jsm::sin / cos / exp / log (fdlibm as V8 runs it) of each element of x."""
        g = getattr(self.L, "orc_jsm_" + fn)
        x = np.asarray(x, np.float64)
        return np.array([g(float(v)) for v in x.ravel()], np.float64).reshape(x.shape)

    def label_blinks(self, dyes, mode, kin, t_max):
        """LabelSchedule's blinks per dye (dyes: n x 3 uint32 h1, k, n). Returns (counts, blinks total x 3)."""
        dyes = np.asarray(dyes, np.uint32).reshape(-1, 3)
        buf = np.zeros((1024, 3))
        counts = np.zeros(len(dyes), np.int32)
        out = []
        for i, (h1, k, n) in enumerate(dyes):
            c = self.L.orc_label_blinks(int(h1), int(np.int32(k)), int(np.int32(n)), mode, C.byref(kin), t_max,
                                        _p(buf, C.c_double), buf.shape[0])
            if c > buf.shape[0]:
                buf = np.zeros((c, 3))
                c = self.L.orc_label_blinks(int(h1), int(np.int32(k)), int(np.int32(n)), mode, C.byref(kin), t_max,
                                            _p(buf, C.c_double), buf.shape[0])
            counts[i] = c
            out.append(buf[:c].copy())
        return counts, (np.concatenate(out) if out else np.zeros((0, 3)))

    def label_cont(self, dyes, mode, pre_state, kin):
        """LabelSchedule's continuous window per dye: (has n, windows n x 4 tOn, tOff, state, aux)."""
        dyes = np.asarray(dyes, np.uint32).reshape(-1, 3)
        has = np.zeros(len(dyes), np.int32)
        w = np.zeros((len(dyes), 4))
        one = np.zeros(4)
        for i, (h1, k, n) in enumerate(dyes):
            has[i] = self.L.orc_label_cont(int(h1), int(np.int32(k)), int(np.int32(n)), mode, int(pre_state),
                                           C.byref(kin), _p(one, C.c_double))
            if has[i]:
                w[i] = one
        return has, w

    def persistent_blinks(self, dyes, kin, t0=0.0, t1=0.0, bins=None):
        """PersistentBlinks over the window (t0, t1), or PersistentBlinksInBins over bins = (lo, hi): (counts,
        blinks total x 5 tOn, tOff, brightness, bin, j; bin and j 0 for the window form)."""
        dyes = np.asarray(dyes, np.uint32).reshape(-1, 3)
        buf = np.zeros((4096, 5))
        counts = np.zeros(len(dyes), np.int32)
        out = []
        lo, hi = bins if bins is not None else (0, 0)
        for i, (h1, k, n) in enumerate(dyes):
            args = (int(h1), int(np.int32(k)), int(np.int32(n)), C.byref(kin), int(bins is not None), t0, t1, lo, hi)
            c = self.L.orc_persistent_blinks(*args, _p(buf, C.c_double), buf.shape[0])
            if c > buf.shape[0]:
                buf = np.zeros((c, 5))
                c = self.L.orc_persistent_blinks(*args, _p(buf, C.c_double), buf.shape[0])
            counts[i] = c
            b = buf[:c].copy()
            if bins is None:
                b[:, 3:] = 0
            out.append(b)
        return counts, (np.concatenate(out) if out else np.zeros((0, 5)))


def roi_model(**kw) -> RoiModel:
    m = RoiModel(wavelengthNm=670.0, na=1.45, immersionIndex=1.518, sampleIndex=1.33, sampleDepthNm=0.0,
                 pixelSizeNm=100.0, pupilSamples=64, roiSize=15, oversampling=4, reserved=0)
    for k, v in kw.items():
        setattr(m, k, v)
    return m


class Reference:
    """The reference's own CPU functions (ref_shim.cpp over the compiled reference sources)."""

    def __init__(self, path: str = REF_PATH):
        if not os.path.exists(path):
            raise FileNotFoundError(path)
        L = C.CDLL(path)
        self.L = L
        D, I32, U32, I64 = C.c_double, C.c_int32, C.c_uint32, C.c_int64
        PF, PI16, PI32, PD = C.POINTER(C.c_float), C.POINTER(C.c_int16), C.POINTER(C.c_int32), C.POINTER(C.c_double)
        L.ref_psf_planes.argtypes = [C.POINTER(PsfRequest), PF]
        L.ref_psf_planes.restype = I32
        L.ref_zernike_pupil_samples.argtypes = [C.POINTER(PsfRequest)]
        L.ref_zernike_pupil_samples.restype = I32
        L.ref_psf_kernel_cache.argtypes = [C.POINTER(PsfRequest), PF, PF]
        L.ref_psf_kernel_cache.restype = I32
        L.ref_block_sums.argtypes = [PF, I32, I32, PF]
        L.ref_noise_chain.argtypes = [PF, C.POINTER(C.c_uint16), U32, U32, C.POINTER(Camera), PF, PF, PF, U32, U32,
                                      U32, I32]
        L.ref_collect_emitters.argtypes = [C.c_void_p, I32, I64, U32, U32, D, D, I32, D, D, C.c_void_p, I32]
        L.ref_collect_emitters.restype = I32
        L.ref_render_frame.argtypes = [PF, U32, U32, C.c_void_p, I32, I64, D, D, D, PF, I32, I32, I32, D, I32, D, D,
                                       D, I32]
        L.ref_render_frame.restype = I32
        L.ref_halo_spans.argtypes = [PF, I32, I32, I32, I32, D, PI16, PI16, PI32]
        L.ref_halo_spans.restype = I32
        L.ref_collect_gpu_emitters.argtypes = [C.c_void_p, I32, I64, U32, U32, D, D, PF, I32, I32, I32, D, I32, D,
                                               C.c_void_p, I32]
        L.ref_collect_gpu_emitters.restype = I32
        L.ref_wf_kernel.argtypes = [PF, I32, I32, I32, D, I32, I32, I32, PF]
        L.ref_wf_kernel.restype = I32
        L.ref_render_noise_frames.argtypes = [C.c_void_p, I32, PI32, I32, U32, U32, D, D, PF, I32, I32, I32, D, I32,
                                              C.POINTER(Camera), PF, U32, C.POINTER(C.c_uint16)]
        L.ref_render_noise_frames.restype = I32
        L.ref_psf_library.argtypes = [C.POINTER(PsfRequest), I32, PF]
        L.ref_psf_library.restype = I32
        L.ref_wf_focus_series.argtypes = [PF, I32, I32, I32, D, PF, I32, I32, I32, D, D, C.POINTER(D), I32, PF]
        L.ref_wf_focus_series.restype = I32
        L.ref_wf_capture_jobs.argtypes = [PF, I32, I32, I32, D, PF, I32, I32, I32, D, D, C.POINTER(D), I32, D, D,
                                          C.c_void_p, C.c_size_t, C.POINTER(C.c_size_t), PF]
        L.ref_wf_capture_jobs.restype = I32
        L.ref_wf_drift_frames.argtypes = [PF, I32, I32, I32, D, PF, I32, I32, I32, D, D, D, PD, I32, PF, PF,
                                          C.c_size_t, PI32, PD, PD]
        L.ref_wf_drift_frames.restype = I32
        L.ref_noise_setup.argtypes = [C.c_int64, U32, U32, U32, U32, D, D, D, D, D, D, PF, PF, PF]
        L.ref_psf_half_width_px.argtypes = [D, D, D, D]
        L.ref_psf_half_width_px.restype = I32
        L.ref_capped_psf_oversampling.argtypes = [I32, D, D]
        L.ref_capped_psf_oversampling.restype = I32
        L.ref_collect_emitters_full.argtypes = [C.c_void_p, I32, I64, U32, U32, D, D, D, D, I32, D, D, PF, C.c_void_p,
                                                I32, C.POINTER(C.c_int64)]
        L.ref_collect_emitters_full.restype = I32
        L.ref_bucket_events.argtypes = [C.c_void_p, I32, I64, C.POINTER(C.c_int64), C.POINTER(C.c_uint32), I64]
        L.ref_bucket_events.restype = I64
        L.ref_tiff_open.argtypes = [C.c_char_p, C.c_uint64, I32]
        L.ref_tiff_open.restype = C.c_void_p
        for nm in ("ref_tiff_page_u16", "ref_tiff_page_f32"):
            getattr(L, nm).argtypes = [C.c_void_p, C.c_void_p, U32, U32, C.c_char_p]
            getattr(L, nm).restype = I32
        L.ref_tiff_page_rows_u16.argtypes = [C.c_void_p, C.c_void_p, U32, U32, U32, C.c_char_p]
        L.ref_tiff_page_rows_u16.restype = I32
        L.ref_tiff_rows_u16.argtypes = [C.c_void_p, C.c_uint64, U32, U32, C.c_void_p]
        L.ref_tiff_rows_u16.restype = I32
        L.ref_tiff_close.argtypes = [C.c_void_p]
        L.ref_tiff_close.restype = I32
        L.ref_tiff_bytes.argtypes = [C.c_uint64] * 5
        L.ref_tiff_bytes.restype = C.c_uint64
        L.ref_parse_zernike.argtypes = [C.c_char_p, PD]
        L.ref_parse_zernike.restype = I32
        L.ref_format_zernike.argtypes = [PD, C.c_char, C.c_char_p, I32]
        L.ref_format_zernike.restype = I32
        L.ref_compute_psf_kernel_cache.argtypes = [C.POINTER(PsfRequest), C.c_char_p, C.c_char_p, I32, C.c_char_p,
                                                   I32, PF, C.POINTER(C.c_uint64)]
        L.ref_compute_psf_kernel_cache.restype = I32
        L.ref_wf_images_render.argtypes = [PF, I32, I32, U32, U32, I32, PD, I32, PF]
        L.ref_wf_images_render.restype = I32
        L.ref_binned_render.argtypes = [C.c_void_p, I32, I32, I32, D, D, I32, I32, I32, PD, PI32, I32, PF, PI32]
        L.ref_binned_render.restype = I32
        L.ref_jsm.argtypes = [I32, PD, I64, PD]
        L.ref_dye_h1.argtypes = [I32, U32, I32, I32, I32]
        L.ref_dye_h1.restype = U32
        PK = C.POINTER(Kinetics)
        L.ref_label_blinks.argtypes = [C.c_void_p, I32, I32, PK, D, PI32, PD, I64, I32, PD, PK]
        L.ref_label_blinks.restype = I64
        L.ref_label_cont.argtypes = [C.c_void_p, I32, I32, I32, PK, PI32, PD, I32, PD, PK]
        L.ref_persistent_blinks.argtypes = [C.c_void_p, I32, PK, I32, D, D, I64, I64, PI32, PD, I64, I32, PD, PK]
        L.ref_persistent_blinks.restype = I64

    def psf_planes(self, req, size, nz):
        out = np.zeros((nz, size, size), np.float32)
        if self.L.ref_psf_planes(C.byref(req), _p(out, C.c_float)) != 0:
            raise ValueError("bad PSF request")
        return out

    def pupil_samples(self, req):
        return int(self.L.ref_zernike_pupil_samples(C.byref(req)))

    def psf_kernel_cache(self, req, size, nz, os_):
        bw = size + os_ - 1
        planes = np.zeros((nz, size, size), np.float32)
        bs = np.zeros((nz, bw, bw), np.float32)
        if self.L.ref_psf_kernel_cache(C.byref(req), _p(planes, C.c_float), _p(bs, C.c_float)) != 0:
            raise ValueError("bad PSF request")
        return planes, bs

    def block_sums(self, kernel, os_):
        kernel = np.ascontiguousarray(kernel, np.float32)
        n = kernel.shape[0]
        out = np.zeros((n + os_ - 1, n + os_ - 1), np.float32)
        self.L.ref_block_sums(_p(kernel, C.c_float), n, os_, _p(out, C.c_float))
        return out

    def noise(self, photons, cam, seed, frame, offset=None, gain=None, read=None, pixel_base=0, parallel=False):
        photons = np.ascontiguousarray(photons, dtype=np.float32)
        h, w = photons.shape
        out = np.zeros((h, w), np.uint16)
        f = lambda a: None if a is None else np.ascontiguousarray(a, np.float32)
        offset, gain, read = f(offset), f(gain), f(read)
        self.L.ref_noise_chain(_p(photons, C.c_float), _p(out, C.c_uint16), w, h, C.byref(cam),
                               _p(offset, C.c_float), _p(gain, C.c_float), _p(read, C.c_float), seed, frame,
                               pixel_base, int(parallel))
        return out

    def collect_emitters(self, events, frame, w, h, pixel_nm, photons_per_blink, nz, z_step_nm, global_z_um=0.0):
        events = np.ascontiguousarray(events, dtype=EVENT_DTYPE)
        out = np.zeros(max(1, len(events)), dtype=EMITTER_DTYPE)
        n = self.L.ref_collect_emitters(events.ctypes.data, len(events), frame, w, h, pixel_nm, photons_per_blink,
                                        nz, z_step_nm, global_z_um, out.ctypes.data, len(out))
        return out[:n]

    def render_frame(self, w, h, events, frame, pixel_nm, photons_per_blink, bg, planes, os_, z_step_nm, interp,
                     global_z_um=0.0, halo=0.0, min_photons=0.0, parallel=False):
        events = np.ascontiguousarray(events, dtype=EVENT_DTYPE)
        planes = np.ascontiguousarray(planes, np.float32)
        img = np.zeros((h, w), np.float32)
        self.L.ref_render_frame(_p(img, C.c_float), w, h, events.ctypes.data, len(events), frame, pixel_nm,
                                photons_per_blink, bg, _p(planes, C.c_float), planes.shape[0], planes.shape[1], os_,
                                z_step_nm, interp, global_z_um, halo, min_photons, int(parallel))
        return img

    def halo_spans(self, planes, os_, interp, threshold):
        planes = np.ascontiguousarray(planes, np.float32)
        nz, size = planes.shape[0], planes.shape[1]
        cam_rad = ((size - 1) // 2) // os_
        n = 2 * cam_rad + 1
        lo = np.zeros((nz, n), np.int16)
        hi = np.zeros((nz, n), np.int16)
        r2 = np.zeros(nz, np.int32)
        r = self.L.ref_halo_spans(_p(planes, C.c_float), nz, size, os_, interp, threshold, _p(lo, C.c_int16),
                                  _p(hi, C.c_int16), _p(r2, C.c_int32))
        if r < 0:
            raise ValueError("no halo cut")
        return lo, hi, r2

    def collect_gpu_emitters(self, events, frame, w, h, pixel_nm, photons_per_blink, planes, os_, z_step_nm, interp,
                             global_z_um=0.0):
        events = np.ascontiguousarray(events, dtype=EVENT_DTYPE)
        planes = np.ascontiguousarray(planes, np.float32)
        out = np.zeros(max(1, len(events)), dtype=GPU_EMITTER_DTYPE)
        n = self.L.ref_collect_gpu_emitters(events.ctypes.data, len(events), frame, w, h, pixel_nm,
                                            photons_per_blink, _p(planes, C.c_float), planes.shape[0],
                                            planes.shape[1], os_, z_step_nm, interp, global_z_um, out.ctypes.data,
                                            len(out))
        return out[:n]

    def wf_radius(self, planes, os_, upscale):
        planes = np.ascontiguousarray(planes, np.float32)
        return self.L.ref_wf_kernel(_p(planes, C.c_float), planes.shape[0], planes.shape[1], os_, 100.0, upscale, 0,
                                    0, None)

    def wf_kernel(self, planes, os_, upscale, p, R):
        planes = np.ascontiguousarray(planes, np.float32)
        out = np.zeros((2 * R + 1, 2 * R + 1), np.float32)
        self.L.ref_wf_kernel(_p(planes, C.c_float), planes.shape[0], planes.shape[1], os_, 100.0, upscale, p, R,
                             _p(out, C.c_float))
        return out

    def render_noise_frames(self, events, ev_start, w, h, pixel_nm, bg, planes, os_, z_step_nm, interp, cam, offset,
                            seed):
        events = np.ascontiguousarray(events, dtype=EVENT_DTYPE)
        ev_start = np.ascontiguousarray(ev_start, np.int32)
        planes = np.ascontiguousarray(planes, np.float32)
        n_frames = len(ev_start) - 1
        out = np.zeros((n_frames, h, w), np.uint16)
        offset = None if offset is None else np.ascontiguousarray(offset, np.float32)
        self.L.ref_render_noise_frames(events.ctypes.data, len(events), _p(ev_start, C.c_int32), n_frames, w, h,
                                       pixel_nm, bg, _p(planes, C.c_float), planes.shape[0], planes.shape[1], os_,
                                       z_step_nm, interp, C.byref(cam), _p(offset, C.c_float), seed,
                                       _p(out, C.c_uint16))
        return out

    def wf_focus_series(self, density, dz_nm, planes, os_, z_step_nm, foci_um, kernel_cap_um=12.8):
        """WidefieldScene::UpdateFromGrid + FocusSeries on a persistent population, flat illumination over the FOV,
        unit dose and efficiency: (nF, ny, nx) images."""
        density = np.ascontiguousarray(density, np.float32)
        planes = np.ascontiguousarray(planes, np.float32)
        foci = np.ascontiguousarray(foci_um, np.float64)
        nz, ny, nx = density.shape
        out = np.zeros((len(foci), ny, nx), np.float32)
        r = self.L.ref_wf_focus_series(_p(density, C.c_float), nx, ny, nz, dz_nm, _p(planes, C.c_float),
                                       planes.shape[0], planes.shape[1], os_, z_step_nm, kernel_cap_um,
                                       _p(foci, C.c_double), len(foci), _p(out, C.c_float))
        if r < 0:
            raise RuntimeError("WidefieldScene failed")
        return out

    def wf_drift_frames(self, density, dz_nm, planes, os_, z_step_nm, base_focus_um, drift_nm, kernel_cap_um=12.8):
        """WidefieldDriftFrames on a persistent population: (frames (nF, ny, nx), spectra (nImages, NY, NX/2+1)
        complex64, geometry dict)."""
        density = np.ascontiguousarray(density, np.float32)
        planes = np.ascontiguousarray(planes, np.float32)
        drift = np.ascontiguousarray(drift_nm, np.float64).reshape(-1, 3)
        nz, ny, nx = density.shape
        cam = np.zeros((len(drift), ny, nx), np.float32)
        cap = 1 << 26
        spec = np.zeros(cap, np.float32)
        geo = np.zeros(9, np.int32)
        frac = np.zeros(2)
        pitch = C.c_double()
        r = self.L.ref_wf_drift_frames(_p(density, C.c_float), nx, ny, nz, dz_nm, _p(planes, C.c_float),
                                       planes.shape[0], planes.shape[1], os_, z_step_nm, kernel_cap_um, base_focus_um,
                                       _p(drift, C.c_double), len(drift), _p(cam, C.c_float), _p(spec, C.c_float),
                                       cap, _p(geo, C.c_int32), _p(frac, C.c_double), C.byref(pitch))
        if r != 0:
            raise RuntimeError(f"WidefieldDriftFrames failed ({r})")
        NX, NY, n = int(geo[0]), int(geo[1]), int(geo[7])
        W = NX // 2 + 1
        spectra = spec[: 2 * n * NY * W].view(np.complex64).reshape(n, NY, W).copy()
        g = dict(NX=NX, NY=NY, fovX0=int(geo[2]), fovY0=int(geo[3]), cw=int(geo[4]), ch=int(geo[5]),
                 upscale=int(geo[6]), nImages=n, k0=int(geo[8]), fracX=float(frac[0]), fracY=float(frac[1]),
                 pitchNm=pitch.value)
        return cam, spectra, g

    def wf_capture_jobs(self, density, dz_nm, planes, os_, z_step_nm, foci_um, dx_um=0.0, dy_um=0.0,
                        kernel_cap_um=12.8):
        """The WidefieldGpuJobs a GPU-mode WidefieldScene sends its accelerator over a focus series (planes resident
        after the first job), and that scene's own CPU images: (jobs, (nF, ny, nx) images)."""
        density = np.ascontiguousarray(density, np.float32)
        planes = np.ascontiguousarray(planes, np.float32)
        foci = np.ascontiguousarray(foci_um, np.float64)
        nz, ny, nx = density.shape
        out = np.zeros((len(foci), ny, nx), np.float32)
        cap = 1 << 28
        blob = np.zeros(cap, np.uint8)
        used = C.c_size_t(0)
        r = self.L.ref_wf_capture_jobs(_p(density, C.c_float), nx, ny, nz, dz_nm, _p(planes, C.c_float),
                                       planes.shape[0], planes.shape[1], os_, z_step_nm, kernel_cap_um,
                                       _p(foci, C.c_double), len(foci), dx_um, dy_um, blob.ctypes.data, cap,
                                       C.byref(used), _p(out, C.c_float))
        if r < 0:
            raise RuntimeError(f"WidefieldScene job capture failed ({r})")
        return parse_wf_jobs(blob[:used.value].tobytes(), r), out

    def wf_images_render(self, channels, has_persistent, upscale, a, cam=None):
        """WidefieldImages::Render of channel images [nCh][ch][cw] with bleach coefficients a, added to cam."""
        ch_ = np.ascontiguousarray(channels, np.float32)
        nC, chh, cw = ch_.shape
        a = np.ascontiguousarray(a, np.float64)
        if cam is None:
            cam = np.zeros((chh // max(1, upscale), cw // max(1, upscale)), np.float32)
        cam = np.ascontiguousarray(cam, np.float32).copy()
        self.L.ref_wf_images_render(_p(ch_, C.c_float), nC, int(has_persistent), cw, chh, upscale,
                                    _p(a, C.c_double), len(a), _p(cam, C.c_float))
        return cam

    def binned_render(self, planes, os_, upscale, W, H, frames, sigma_px=0.0, halo_threshold=0.0, img=None):
        """BinnedBlinkRenderer Setup + Render per frame (frames: a list of n x 4 emitter arrays xPx, yPx, photons,
        zIndex); planes None: the Gaussian of sigma_px. Returns (images nF x H x W added to img, (u, R))."""
        em = np.ascontiguousarray(np.concatenate([np.asarray(f, np.float64).reshape(-1, 4) for f in frames])
                                  if frames else np.zeros((0, 4)), np.float64)
        start = np.zeros(len(frames) + 1, np.int32)
        start[1:] = np.cumsum([np.asarray(f).reshape(-1, 4).shape[0] for f in frames])
        out = (np.zeros((len(frames), H, W), np.float32) if img is None
               else np.ascontiguousarray(img, np.float32).copy())
        uR = np.zeros(2, np.int32)
        pl = None
        nz = size = 0
        if planes is not None:
            pl = np.ascontiguousarray(planes, np.float32)
            nz, size = pl.shape[0], pl.shape[1]
        rc = self.L.ref_binned_render(None if pl is None else pl.ctypes.data, nz, size, os_, halo_threshold,
                                      sigma_px, W, H, upscale, _p(em, C.c_double), _p(start, C.c_int32), len(frames),
                                      _p(out, C.c_float), _p(uR, C.c_int32))
        if rc != 0:
            raise RuntimeError("BinnedBlinkRenderer::Setup failed")
        return out, (int(uR[0]), int(uR[1]))

    def dye_h1(self, which, seed, cx, cy, mt=0):
        """DyeH1 (which 0), NpcH1 (1), LaminH1 (2) of one block."""
        return int(self.L.ref_dye_h1(which, seed & 0xFFFFFFFF, int(np.int32(cx)), int(np.int32(cy)), int(mt)))

    def jsm(self, fn: str, x):
        """The compiled reference's jsm:: function (fdlibm as V8 runs it) of each element of x."""
        x = np.ascontiguousarray(x, np.float64)
        out = np.empty_like(x)
        self.L.ref_jsm(JSM_FUNCS.index(fn), _p(x, C.c_double), x.size, _p(out, C.c_double))
        return out

    @staticmethod
    def _hist(hist):
        """(n, tStart, kinetics array) of a history (tStart, [Kinetics]) or None."""
        if hist is None:
            return 0, None, None
        t = np.ascontiguousarray(hist[0], np.float64)
        k = (Kinetics * len(hist[1]))(*hist[1])
        return len(hist[1]), t, k

    def label_blinks(self, dyes, mode, kin, t_max, hist=None):
        """LabelSchedule (core/src/dyes.cpp:440) blinks per dye, under hist = (tStart, [Kinetics]) if given:
        (counts, blinks total x 3)."""
        d = np.ascontiguousarray(dyes, np.uint32).reshape(-1, 3)
        counts = np.zeros(len(d), np.int32)
        n, t, k = self._hist(hist)
        h = (n, None if t is None else _p(t, C.c_double), k)
        total = self.L.ref_label_blinks(d.ctypes.data, len(d), mode, C.byref(kin), t_max, _p(counts, C.c_int32),
                                        None, 0, *h)
        out = np.zeros((max(total, 1), 3))
        self.L.ref_label_blinks(d.ctypes.data, len(d), mode, C.byref(kin), t_max, _p(counts, C.c_int32),
                                _p(out, C.c_double), total, *h)
        return counts, out[:total]

    def label_cont(self, dyes, mode, pre_state, kin, hist=None):
        d = np.ascontiguousarray(dyes, np.uint32).reshape(-1, 3)
        has = np.zeros(len(d), np.int32)
        w = np.zeros((len(d), 4))
        n, t, k = self._hist(hist)
        self.L.ref_label_cont(d.ctypes.data, len(d), mode, int(pre_state), C.byref(kin), _p(has, C.c_int32),
                              _p(w, C.c_double), n, None if t is None else _p(t, C.c_double), k)
        return has, w

    def persistent_blinks(self, dyes, kin, t0=0.0, t1=0.0, bins=None, hist=None):
        """PersistentBlinks (window) or PersistentBlinksInBins (bins = (lo, hi)), under hist if given: (counts,
        blinks total x 5)."""
        d = np.ascontiguousarray(dyes, np.uint32).reshape(-1, 3)
        counts = np.zeros(len(d), np.int32)
        lo, hi = bins if bins is not None else (0, 0)
        args = (d.ctypes.data, len(d), C.byref(kin), int(bins is not None), t0, t1, lo, hi, _p(counts, C.c_int32))
        n, t, k = self._hist(hist)
        h = (n, None if t is None else _p(t, C.c_double), k)
        total = self.L.ref_persistent_blinks(*args, None, 0, *h)
        out = np.zeros((max(total, 1), 5))
        self.L.ref_persistent_blinks(*args, _p(out, C.c_double), total, *h)
        return counts, out[:total]

    def noise_setup(self, seed, w, h, offset, offset_std, gain, gain_std, read, read_std, sensor_h=0, row_y0=0):
        maps = [np.zeros((h, w), np.float32) for _ in range(3)]
        self.L.ref_noise_setup(seed, w, h, sensor_h, row_y0, offset, offset_std, gain, gain_std, read, read_std,
                               *[_p(m, C.c_float) for m in maps])
        return maps

    def psf_half_width_px(self, half_nm, pixel_nm, wavelength_nm, na):
        return self.L.ref_psf_half_width_px(half_nm, pixel_nm, wavelength_nm, na)

    def capped_psf_oversampling(self, os_, pixel_nm, max_sample_nm):
        return self.L.ref_capped_psf_oversampling(os_, pixel_nm, max_sample_nm)

    def collect_emitters_full(self, ev7, frame, w, h, pixel_nm, photons_per_blink, drift=(0.0, 0.0), nz=0,
                              z_step_nm=0.0, global_z_um=0.0, illum=None):
        """CollectFrameEmitters with excite, drift and an illumination field: (records (x, y, photons, zIndex,
        kernel), (z clamped, z total))."""
        ev7 = np.ascontiguousarray(ev7, np.float64).reshape(-1, 7)
        out = np.zeros(max(1, len(ev7)), EMITTER_DTYPE)
        z = np.zeros(2, np.int64)
        il = None if illum is None else np.ascontiguousarray(illum, np.float32)
        n = self.L.ref_collect_emitters_full(ev7.ctypes.data, len(ev7), frame, w, h, pixel_nm, photons_per_blink,
                                             drift[0], drift[1], nz, z_step_nm, global_z_um,
                                             None if il is None else _p(il, C.c_float), out.ctypes.data, len(out),
                                             z.ctypes.data_as(C.POINTER(C.c_int64)))
        return out[:n], (int(z[0]), int(z[1]))

    def bucket_events(self, events, n_frames):
        """BucketEventsByFrame: (counts per frame, the concatenated index lists)."""
        events = np.ascontiguousarray(events, dtype=EVENT_DTYPE)
        counts = np.zeros(max(0, n_frames), np.int64)
        cap = 4 * len(events) + 16
        while True:
            idx = np.zeros(cap, np.uint32)
            n = self.L.ref_bucket_events(events.ctypes.data, len(events), n_frames, _p(counts, C.c_int64),
                                         _p(idx, C.c_uint32), cap)
            if n >= 0:
                return counts, idx[:n]
            cap *= 4

    def parse_zernike(self, text: str):
        """ParseZernikeCoefficients: (28 values, ok)."""
        out = np.zeros(28)
        ok = self.L.ref_parse_zernike(text.encode("latin-1"), _p(out, C.c_double))
        return out, bool(ok)

    def format_zernike(self, values, separator=","):
        v = np.ascontiguousarray(values, np.float64)
        buf = C.create_string_buffer(4096)
        self.L.ref_format_zernike(_p(v, C.c_double), separator.encode(), buf, 4096)
        return buf.value.decode()

    def compute_psf_kernel_cache(self, req, zernike_text, disk_dir, size, nz, memo_entries=0):
        """ComputePsfKernelCache with the disk store in disk_dir: (ok, log lines, sum-1 planes, serial)."""
        log = C.create_string_buffer(1 << 14)
        planes = np.zeros((nz, size, size), np.float32)
        serial = C.c_uint64()
        rc = self.L.ref_compute_psf_kernel_cache(C.byref(req), None if zernike_text is None else
                                                 zernike_text.encode(), disk_dir.encode(), memo_entries, log,
                                                 1 << 14, _p(planes, C.c_float), C.byref(serial))
        return rc == 0, log.value.decode().splitlines(), planes, serial.value

    def psf_library(self, reqs, out=None):
        arr = (PsfRequest * len(reqs))(*reqs)
        return self.L.ref_psf_library(arr, len(reqs), _p(out, C.c_float))

    # ---- the world (core/src/world.{h,cpp}) --------------------------------------------------------------------

    def _world_sigs(self):
        L = self.L
        if getattr(self, "_world_ready", False):
            return L
        D, I32, U32, I64, P = C.c_double, C.c_int32, C.c_uint32, C.c_int64, C.c_void_p
        PD, PI32, PF = C.POINTER(C.c_double), C.POINTER(C.c_int32), C.POINTER(C.c_float)
        L.ref_world_new.argtypes = [U32, C.POINTER(C.c_char_p), PD, I32]
        L.ref_world_new.restype = P
        L.ref_world_free.argtypes = [P]
        L.ref_world_set_label.argtypes = [P, I32, PD]
        L.ref_world_set_label.restype = I32
        L.ref_world_set_history.argtypes = [P, PD, C.POINTER(Kinetics), I32]
        L.ref_world_set_history.restype = I32
        L.ref_world_get_label.argtypes = [P, I32, PD]
        L.ref_world_cells.argtypes = [P, D, D, D, D, PD, I32]
        L.ref_world_cells.restype = I32
        L.ref_world_cell_mts.argtypes = [P, I32, I32, PD, I64, PI32, I32, PI32]
        L.ref_world_cell_mts.restype = I64
        L.ref_mt_frames.argtypes = [PD, I32, PD, PD, PD, PD]
        L.ref_world_cell_envelope.argtypes = [P, I32, I32, I32, PD, I32]
        L.ref_world_cell_envelope.restype = I32
        L.ref_world_npc_sites.argtypes = [P, PD, PD]
        L.ref_world_sites.argtypes = [P, D, D, D, D, D, D, PD, I64]
        L.ref_world_sites.restype = I64
        L.ref_world_events.argtypes = [P, I32, D, D, D, D, D, D, D, D, PD, I64]
        L.ref_world_events.restype = I64
        L.ref_world_density3d.argtypes = [P, D, D, D, D, D, D, I32, I32, I32, U32, PF]
        L.ref_world_density3d.restype = I64
        L.ref_dyes_in_block.argtypes = [U32, I32, I32, I32, PD, I32, I32, D, D, PD, PD, I32]
        L.ref_dyes_in_block.restype = I32
        L.ref_jsm_geom.argtypes = [I32, PD, PD, PD, I64, PD]
        L.ref_case_mts.argtypes = [U32, C.POINTER(C.c_char_p), PD, I32, I32, I32, PD, I64, PI32, I32, PI32]
        L.ref_case_mts.restype = I64
        L.ref_case_envelope.argtypes = [U32, C.POINTER(C.c_char_p), PD, I32, I32, I32, I32, PD, I32, PD, PD]
        L.ref_case_envelope.restype = I32
        L.ref_envelope_dyes.argtypes = [I32, U32, I32, I32, PD, I32, PI32, I32, PD, D, D, PD, PD, I32]
        L.ref_envelope_dyes.restype = I32
        L.ref_mt_site_geometry.argtypes = [PD, I32, D, D, D, D, D, PD]
        L.ref_mt_protofilament.argtypes = [D, I32, PD, PD]
        L.ref_mt_density.argtypes = [U32, PD, I32, PI32, C.POINTER(C.c_int64), PD, D, D, PD, PD, I32, I32, I32, PF]
        L.ref_mt_density.restype = I64
        self._world_ready = True
        return L

    @staticmethod
    def _params(params):
        params = params or {}
        names = (C.c_char_p * max(1, len(params)))(*[k.encode() for k in params])
        vals = np.array([float(v) for v in params.values()] or [0.0])
        return names, vals, len(params)

    def case_mts(self, seed, params, cx, cy):
        """BuildMicrotubulesForCell of the raw candidate cell (cx, cy): (points (n, 3), per-microtubule counts)."""
        L = self._world_sigs()
        names, vals, n = self._params(params)
        nm = C.c_int32()
        tot = L.ref_case_mts(seed & 0xFFFFFFFF, names, _p(vals, C.c_double), n, cx, cy, None, 0, None, 0,
                             C.byref(nm))
        if tot < 0:
            raise ValueError("no cell there, or an unknown parameter")
        xyz = np.zeros((max(tot, 1), 3))
        lens = np.zeros(max(nm.value, 1), np.int32)
        L.ref_case_mts(seed & 0xFFFFFFFF, names, _p(vals, C.c_double), n, cx, cy, _p(xyz, C.c_double), tot,
                       _p(lens, C.c_int32), nm.value, C.byref(nm))
        return xyz[:tot], lens[: nm.value]

    def case_envelope(self, seed, params, cx, cy, which):
        """BuildNpcsForCell (which 1: rows c, n, e1, e2, cand, piece) or BuildLaminForCell (2: c, n, cand, piece) of
        the raw candidate cell, with NpcAttachmentPoints and NpcSiteReachUm: (rows, sites, reach)."""
        L = self._world_sigs()
        names, vals, n = self._params(params)
        width = 14 if which == 1 else 8
        sites = np.zeros(192)
        reach = C.c_double()
        cnt = L.ref_case_envelope(seed & 0xFFFFFFFF, names, _p(vals, C.c_double), n, cx, cy, which, None, 0,
                                  _p(sites, C.c_double), C.byref(reach))
        if cnt < 0:
            raise ValueError("no cell there, or an unknown parameter")
        rows = np.zeros((max(cnt, 1), width))
        L.ref_case_envelope(seed & 0xFFFFFFFF, names, _p(vals, C.c_double), n, cx, cy, which, _p(rows, C.c_double),
                            cnt, _p(sites, C.c_double), C.byref(reach))
        return rows[:cnt], sites, reach.value

    def envelope_dyes(self, which, seed, cx, cy, rows, ids, sites, density, ff, linkage):
        """NpcDyesInBlock (which 1) or LaminDyesInBlock (2): rows pore, k, n, id, pos, site, rigid, tip."""
        L = self._world_sigs()
        rows = np.ascontiguousarray(rows, np.float64)
        ids = np.ascontiguousarray(ids, np.int32)
        sites = np.ascontiguousarray(sites, np.float64)
        lk = np.ascontiguousarray(linkage, np.float64)
        cap = max(1, len(ids) * (64 if which == 1 else 1))
        out = np.zeros((cap, 16))
        n = L.ref_envelope_dyes(which, seed & 0xFFFFFFFF, cx, cy, _p(rows, C.c_double), len(rows),
                                _p(ids, C.c_int32), len(ids), _p(sites, C.c_double), density, ff,
                                _p(lk, C.c_double), _p(out, C.c_double), cap)
        return out[:n]

    def mt_site_geometry(self, pts, S, theta, r1, r2, r3):
        """MtSiteGeometry: (att, tip, dye) as 9 doubles."""
        pts = np.ascontiguousarray(pts, np.float64)
        out = np.zeros(9)
        self._world_sigs().ref_mt_site_geometry(_p(pts, C.c_double), len(pts), S, theta, r1, r2, r3,
                                                _p(out, C.c_double))
        return out

    def mt_density(self, seed, cells, mtStart, ptStart, pts, density, ff, linkage, win, nx, ny, nz):
        """GenerateDyes + Density3dInWindow's binning over given microtubules in one thread (cells: rows cx, cy, x,
        y, packRot): (dyes counted, counts)."""
        L = self._world_sigs()
        cells = np.ascontiguousarray(cells, np.float64)
        mtStart = np.ascontiguousarray(mtStart, np.int32)
        ptStart = np.ascontiguousarray(ptStart, np.int64)
        pts = np.ascontiguousarray(pts, np.float64)
        lk = np.ascontiguousarray(linkage, np.float64)
        w = np.ascontiguousarray(win, np.float64)
        out = np.zeros((nz, ny, nx), np.float32)
        n = L.ref_mt_density(seed & 0xFFFFFFFF, _p(cells, C.c_double), len(cells), _p(mtStart, C.c_int32),
                             ptStart.ctypes.data_as(C.POINTER(C.c_int64)), _p(pts, C.c_double), density, ff,
                             _p(lk, C.c_double), _p(w, C.c_double), nx, ny, nz, _p(out, C.c_float))
        return n, out

    def mt_protofilament(self, phase, k):
        th, off = C.c_double(), C.c_double()
        self._world_sigs().ref_mt_protofilament(phase, k, C.byref(th), C.byref(off))
        return th.value, off.value

    def world_new(self, seed: int, params: dict | None = None):
        """A reference World (core/src/world.h:152) of seed with the default params and these overrides."""
        L = self._world_sigs()
        params = params or {}
        names = (C.c_char_p * max(1, len(params)))(*[k.encode() for k in params])
        vals = np.array([float(v) for v in params.values()] or [0.0])
        w = L.ref_world_new(seed & 0xFFFFFFFF, names, _p(vals, C.c_double), len(params))
        if not w:
            raise ValueError(f"unknown world parameter in {params}")
        return w

    def world_free(self, w):
        self._world_sigs().ref_world_free(w)

    @staticmethod
    def label_vector(density=0.7, fluorescentFraction=1.0, mode=2, kin=(0.01, 0.05, 1.0, 1.0, 0.5, 0.0),
                     preState=0, linkage=None):
        """A label in the C ABI's ISC_LABEL layout (core/include/insiliscope/insiliscope.h:218-278)."""
        lk = linkage if linkage is not None else (np.nan,) * 4
        return np.array([density, fluorescentFraction, mode, *kin, preState, 0, 90, 0, 0, 0, 0, *lk], np.float64)

    def world_set_label(self, w, structure: int, vec) -> int:
        vec = np.ascontiguousarray(vec, np.float64)
        return self._world_sigs().ref_world_set_label(w, structure, _p(vec, C.c_double))

    def world_set_history(self, w, tStart, kin) -> int:
        """World::SetKineticsHistory (core/src/world.cpp:933): kin[i * 4 + s] per segment i and structure s."""
        t = np.ascontiguousarray(tStart, np.float64)
        k = (Kinetics * max(1, len(kin)))(*kin)
        return self._world_sigs().ref_world_set_history(w, _p(t, C.c_double), k, len(t))

    def world_get_label(self, w, structure: int):
        out = np.zeros(14)
        self._world_sigs().ref_world_get_label(w, structure, _p(out, C.c_double))
        return out

    def world_cells(self, w, x0, y0, x1, y1):
        """World::CellsInRect: rows cx, cy, x, y, packRot, rOuter in its order."""
        L = self._world_sigs()
        out = np.zeros((512, 6))
        n = L.ref_world_cells(w, x0, y0, x1, y1, _p(out, C.c_double), 512)
        if n > 512:
            out = np.zeros((n, 6))
            L.ref_world_cells(w, x0, y0, x1, y1, _p(out, C.c_double), n)
        return out[:n]

    def world_cell_mts(self, w, cx, cy):
        """The microtubules of a cell (World::Assets): (points (n, 3), per-microtubule point counts)."""
        L = self._world_sigs()
        nm = C.c_int32()
        tot = L.ref_world_cell_mts(w, cx, cy, None, 0, None, 0, C.byref(nm))
        xyz = np.zeros((max(tot, 1), 3))
        lens = np.zeros(max(nm.value, 1), np.int32)
        L.ref_world_cell_mts(w, cx, cy, _p(xyz, C.c_double), tot, _p(lens, C.c_int32), nm.value, C.byref(nm))
        return xyz[:tot], lens[: nm.value]

    def world_cell_envelope(self, w, which: int, cx, cy):
        """The pores (which 1: rows c, n, e1, e2, cand) or lamina points (2: c, n, cand) of a cell."""
        L = self._world_sigs()
        width = 13 if which == 1 else 7
        n = L.ref_world_cell_envelope(w, which, cx, cy, None, 0)
        out = np.zeros((max(n, 1), width))
        if n > 0:
            L.ref_world_cell_envelope(w, which, cx, cy, _p(out, C.c_double), n)
        return out[: max(n, 0)]

    def world_npc_sites(self, w):
        sites = np.zeros(192)
        r = C.c_double()
        self._world_sigs().ref_world_npc_sites(w, _p(sites, C.c_double), C.byref(r))
        return sites, r.value

    def world_geometry(self, w, x0, y0, x1, y1, npcs=True, lamin=True):
        """The geometry of the cells CellsInRect returns for the rect, as the fields of the product's
        WorldGeometry (a dict): cells, microtubule ranges and points, pores and lamina points."""
        rows = self.world_cells(w, x0, y0, x1, y1)
        cells = np.zeros(len(rows), [("cx", "<i4"), ("cy", "<i4"), ("x", "<f8"), ("y", "<f8"), ("packRot", "<f8"),
                                     ("rOuter", "<f8")])
        mtStart, ptStart, pts = [0], [0], []
        npcStart, laminStart, npcRows, laminRows = [0], [0], [], []
        for i, r in enumerate(rows):
            cells[i] = (int(r[0]), int(r[1]), r[2], r[3], r[4], r[5])
            xyz, lens = self.world_cell_mts(w, int(r[0]), int(r[1]))
            pts.append(xyz)
            for ln in lens:
                ptStart.append(ptStart[-1] + int(ln))
            mtStart.append(mtStart[-1] + len(lens))
            pr = self.world_cell_envelope(w, 1, int(r[0]), int(r[1])) if npcs else np.zeros((0, 13))
            lr = self.world_cell_envelope(w, 2, int(r[0]), int(r[1])) if lamin else np.zeros((0, 7))
            npcRows.append(pr)
            laminRows.append(lr)
            npcStart.append(npcStart[-1] + len(pr))
            laminStart.append(laminStart[-1] + len(lr))
        sites, reach = self.world_npc_sites(w)
        npcRows = np.concatenate(npcRows) if npcRows else np.zeros((0, 13))
        laminRows = np.concatenate(laminRows) if laminRows else np.zeros((0, 7))
        return dict(cells=cells, mtStart=np.array(mtStart, np.int32), ptStart=np.array(ptStart, np.int64),
                    pts=np.concatenate(pts) if pts else np.zeros((0, 3)), npcStart=np.array(npcStart, np.int32),
                    npcRows=npcRows, laminStart=np.array(laminStart, np.int32), laminRows=laminRows,
                    npcSites=sites, npcSiteReachUm=reach)

    def world_sites(self, w, win):
        L = self._world_sigs()
        n = L.ref_world_sites(w, *win, None, 0)
        out = np.zeros((max(n, 1), 10))
        L.ref_world_sites(w, *win, _p(out, C.c_double), n)
        return out[:n]

    def world_events(self, w, win, t0, t1, continuous=False):
        """World::EventsInWindow (continuous False; t0, t1) or ContinuousInWindow (tMin = t0): rows x, y, z, tOn,
        tOff, brightness, aux, id, structure, state."""
        L = self._world_sigs()
        which = 1 if continuous else 0
        n = L.ref_world_events(w, which, *win, t0, t1, None, 0)
        out = np.zeros((max(n, 1), 10))
        L.ref_world_events(w, which, *win, t0, t1, _p(out, C.c_double), n)
        return out[:n]

    def world_density3d(self, w, win, nx, ny, nz, mask=0xF):
        out = np.zeros((nz, ny, nx), np.float32)
        tot = self._world_sigs().ref_world_density3d(w, *win, nx, ny, nz, mask, _p(out, C.c_float))
        return tot, out

    def mt_frames(self, pts):
        """BuildMtFrames (core/src/microtubules.cpp:701): (cum, T, U, V)."""
        pts = np.ascontiguousarray(pts, np.float64)
        n = len(pts)
        cum = np.zeros(n)
        T, U, V = (np.zeros((max(n - 1, 1), 3)) for _ in range(3))
        self._world_sigs().ref_mt_frames(_p(pts, C.c_double), n, _p(cum, C.c_double), _p(T, C.c_double),
                                         _p(U, C.c_double), _p(V, C.c_double))
        return cum, T[: max(n - 1, 0)], U[: max(n - 1, 0)], V[: max(n - 1, 0)]

    def dyes_in_block(self, seed, cx, cy, mt_index, pts, block, density, ff, linkage=(12.0, 0.0, 2.0, 5.0)):
        """DyesInBlock (core/src/dyes.cpp:127): rows x, y, z, k, n, id, theta (cell-local)."""
        L = self._world_sigs()
        pts = np.ascontiguousarray(pts, np.float64)
        lk = np.array(linkage, np.float64)
        out = np.zeros((4096, 7))
        n = L.ref_dyes_in_block(seed & 0xFFFFFFFF, cx, cy, mt_index, _p(pts, C.c_double), len(pts), block, density,
                                ff, _p(lk, C.c_double), _p(out, C.c_double), 4096)
        return out[:n]

    def jsm_geom(self, fn: str, x, y=None, z=None):
        """jsm::cbrt (fn "cbrt"), 2-argument hypot ("hypot2") or 3-argument hypot ("hypot3") elementwise."""
        x = np.ascontiguousarray(x, np.float64)
        y = np.ascontiguousarray(x if y is None else y, np.float64)
        z = np.ascontiguousarray(x if z is None else z, np.float64)
        out = np.zeros_like(x)
        self._world_sigs().ref_jsm_geom(("cbrt", "hypot2", "hypot3").index(fn), _p(x, C.c_double),
                                        _p(y, C.c_double), _p(z, C.c_double), x.size, _p(out, C.c_double))
        return out


# ---- WidefieldGpuJob (Sim/WidefieldRender.h:306-340): parsing captured jobs and the CPU statement of the hosts --

BRIDGE_PATH = os.path.join(HERE, "_ref", "libiscbridge.so")


class Bridge:
    """The reference's engine with this build's link-time drop-ins (bridge_check.cpp over libiscg.so)."""

    def __init__(self, path: str = BRIDGE_PATH):
        if not os.path.exists(path):
            raise FileNotFoundError(path)
        L = C.CDLL(path)
        self.L = L
        D, I32, U32, I64 = C.c_double, C.c_int32, C.c_uint32, C.c_int64
        PD, PF, PU16 = C.POINTER(D), C.POINTER(C.c_float), C.POINTER(C.c_uint16)
        L.bridge_error.restype = C.c_char_p
        L.bridge_splat_frames.argtypes = [I32, PD, I32, I32, U32, U32, D, D, D, PF, I32, I32, I32, D, I32, PD, D, U32,
                                          PU16, PU16]
        L.bridge_splat_frames.restype = I32
        L.bridge_movie.argtypes = [C.POINTER(C.c_char_p), PD, I32, I32, PU16, C.c_uint64, C.POINTER(I64),
                                   C.c_char_p, I32]
        L.bridge_movie.restype = I64
        L.bridge_wf_scene.argtypes = [PF, I32, I32, I32, D, PF, I32, I32, I32, D, D, PD, I32, D, D, PF, PF,
                                      C.POINTER(I32)]
        L.bridge_wf_scene.restype = I32

    def error(self) -> str:
        return self.L.bridge_error().decode()

    def splat_frames(self, via, events, n_frames, w, h, pixel_nm, per_blink, bg, planes, os_, z_step_nm, interp,
                     cam8, offset, seed):
        ev = np.ascontiguousarray(np.stack([events[k] for k in EVENT_DTYPE.names], axis=1), np.float64)
        planes = np.ascontiguousarray(planes, np.float32)
        cam = np.ascontiguousarray(cam8, np.float64)
        gpu = np.zeros((n_frames, h, w), np.uint16)
        cpu = np.zeros((n_frames, h, w), np.uint16)
        rc = self.L.bridge_splat_frames(via, _p(ev, C.c_double), len(ev), n_frames, w, h, pixel_nm, per_blink, bg,
                                        _p(planes, C.c_float), planes.shape[0], planes.shape[1], os_, z_step_nm,
                                        interp, _p(cam, C.c_double), offset, seed, _p(gpu, C.c_uint16),
                                        _p(cpu, C.c_uint16))
        if rc != 0:
            raise RuntimeError(self.error())
        return gpu, cpu

    def movie(self, options: dict, gpu: bool, max_pixels: int):
        names = (C.c_char_p * len(options))(*[k.encode() for k in options])
        vals = np.ascontiguousarray(list(options.values()), np.float64)
        out = np.zeros(max_pixels, np.uint16)
        info = (C.c_int64 * 4)()
        text = C.create_string_buffer(256)
        n = self.L.bridge_movie(names, _p(vals, C.c_double), len(options), int(gpu), _p(out, C.c_uint16), max_pixels,
                                info, text, 256)
        if n < 0:
            raise RuntimeError(self.error())
        w, h = int(info[0]), int(info[1])
        return out[: n * w * h].reshape(n, h, w), {"frames": int(info[2]), "blinks": int(info[3]),
                                                   "host": text.value.decode()}

    def wf_scene(self, density, dz_nm, planes, os_, z_step_nm, foci_um, kernel_cap_um=12.8, dx_um=0.0, dy_um=0.0):
        density = np.ascontiguousarray(density, np.float32)
        planes = np.ascontiguousarray(planes, np.float32)
        foci = np.ascontiguousarray(foci_um, np.float64)
        nz, ny, nx = density.shape
        gpu = np.zeros((len(foci), ny, nx), np.float32)
        cpu = np.zeros_like(gpu)
        used = C.c_int32(0)
        rc = self.L.bridge_wf_scene(_p(density, C.c_float), nx, ny, nz, dz_nm, _p(planes, C.c_float),
                                    planes.shape[0], planes.shape[1], os_, z_step_nm, kernel_cap_um,
                                    _p(foci, C.c_double), len(foci), dx_um, dy_um, _p(gpu, C.c_float),
                                    _p(cpu, C.c_float), C.byref(used))
        if rc != 0:
            raise RuntimeError(self.error())
        return gpu, cpu, bool(used.value)


def parse_wf_jobs(blob: bytes, n: int) -> list:
    """Jobs in ref_wf_capture_jobs's byte layout -> namespaces with WidefieldGpuJob's fields (kernels [(p, spec)],
    planes [(key, cells, values)], channels [[(key, p0, p1, w0, w1, two)]])."""
    from types import SimpleNamespace
    jobs, o = [], 0

    def take(fmt_dtype, count=1):
        nonlocal o
        a = np.frombuffer(blob, dtype=fmt_dtype, count=count, offset=o)
        o += a.nbytes
        return a

    for _ in range(n):
        NX, NY, nx, ny, fovX0, fovY0, cw, ch = (int(v) for v in take("<u4", 8))
        fracX, fracY = (float(v) for v in take("<f8", 2))
        geometry = int(take("<u8")[0])
        nK, nP, nC, hasP = (int(v) for v in take("<i4", 4))
        W = NX // 2 + 1
        kernels = []
        for _k in range(nK):
            p = int(take("<i4")[0])
            spec = take("<c8", W * NY).reshape(NY, W).copy()
            kernels.append((p, spec))
        planes = []
        for _p_ in range(nP):
            key = int(take("<u8")[0])
            cnt = int(take("<u4")[0])
            cells = take("<u4", cnt).copy()
            vals = take("<f4", cnt).copy()
            planes.append((key, cells, vals))
        channels = []
        for _c in range(nC):
            nd = int(take("<u4")[0])
            deps = []
            for _d in range(nd):
                key = int(take("<u8")[0])
                p0, p1 = (int(v) for v in take("<i4", 2))
                w0, w1 = (float(v) for v in take("<f4", 2))
                two = bool(take("<i4")[0])
                deps.append((key, p0, p1, w0, w1, two))
            channels.append(deps)
        jobs.append(SimpleNamespace(NX=NX, NY=NY, nx=nx, ny=ny, fovX0=fovX0, fovY0=fovY0, cw=cw, ch=ch,
                                    fracX=fracX, fracY=fracY, geometry=geometry, kernels=kernels, planes=planes,
                                    channels=channels, hasPersistent=bool(hasP)))
    if o != len(blob):
        raise ValueError("trailing bytes in the job blob")
    return jobs


def pack_wf_jobs(jobs, prefix: str, out: dict) -> None:
    """Jobs -> flat arrays under `prefix` in an npz dict (kernel spectra stored once per p: fixed in a geometry)."""
    out[prefix + "n"] = np.array(len(jobs))
    specs = {}
    for i, j in enumerate(jobs):
        q = f"{prefix}{i}_"
        out[q + "u32"] = np.array([j.NX, j.NY, j.nx, j.ny, j.fovX0, j.fovY0, j.cw, j.ch], np.uint32)
        out[q + "f64"] = np.array([j.fracX, j.fracY])
        out[q + "u64"] = np.array([j.geometry], np.uint64)
        out[q + "hasPersistent"] = np.array(int(j.hasPersistent))
        out[q + "kp"] = np.array([k[0] for k in j.kernels], np.int32)
        for p, spec in j.kernels:
            if p in specs and not np.array_equal(specs[p], spec):
                raise ValueError("kernel spectrum changed within a geometry")
            specs[p] = spec
        out[q + "keys"] = np.array([pl[0] for pl in j.planes], np.uint64)
        out[q + "starts"] = np.concatenate([[0], np.cumsum([len(pl[1]) for pl in j.planes])]).astype(np.uint32)
        out[q + "cells"] = np.concatenate([pl[1] for pl in j.planes] or [np.zeros(0, np.uint32)]).astype(np.uint32)
        out[q + "values"] = np.concatenate([pl[2] for pl in j.planes] or [np.zeros(0, np.float32)]).astype(np.float32)
        deps = [d for ch in j.channels for d in ch]
        out[q + "dep_keys"] = np.array([d[0] for d in deps], np.uint64)
        out[q + "dep_p"] = np.array([[d[1], d[2], int(d[5])] for d in deps], np.int32).reshape(-1, 3)
        out[q + "dep_w"] = np.array([[d[3], d[4]] for d in deps], np.float32).reshape(-1, 2)
        out[q + "chan"] = np.concatenate([[0], np.cumsum([len(ch) for ch in j.channels])]).astype(np.uint32)
    for p, spec in specs.items():
        out[f"{prefix}kernel{p}"] = spec


def unpack_wf_jobs(npz, prefix: str) -> list:
    from types import SimpleNamespace
    jobs = []
    for i in range(int(npz[prefix + "n"])):
        q = f"{prefix}{i}_"
        NX, NY, nx, ny, fovX0, fovY0, cw, ch = (int(v) for v in npz[q + "u32"])
        fracX, fracY = (float(v) for v in npz[q + "f64"])
        kernels = [(int(p), npz[f"{prefix}kernel{int(p)}"]) for p in npz[q + "kp"]]
        st, cells, vals = npz[q + "starts"], npz[q + "cells"], npz[q + "values"]
        planes = [(int(k), cells[st[a]:st[a + 1]], vals[st[a]:st[a + 1]]) for a, k in enumerate(npz[q + "keys"])]
        dk, dp, dw, cs = npz[q + "dep_keys"], npz[q + "dep_p"], npz[q + "dep_w"], npz[q + "chan"]
        deps = [(int(dk[t]), int(dp[t, 0]), int(dp[t, 1]), float(dw[t, 0]), float(dw[t, 1]), bool(dp[t, 2]))
                for t in range(len(dk))]
        channels = [deps[cs[c]:cs[c + 1]] for c in range(len(cs) - 1)]
        jobs.append(SimpleNamespace(NX=NX, NY=NY, nx=nx, ny=ny, fovX0=fovX0, fovY0=fovY0, cw=cw, ch=ch,
                                    fracX=fracX, fracY=fracY, geometry=int(npz[q + "u64"][0]), kernels=kernels,
                                    planes=planes, channels=channels,
                                    hasPersistent=bool(npz[q + "hasPersistent"])))
    return jobs


class JobReferenceHost:
    """The GPU hosts' per-focus algorithm (restated from JobReferenceHost, cli/widefield_check.cpp:816-902) in FP64:
    plane spectra kept by key within one geometry; per channel S = sum_d A (w0 K0 + w1 K1); the phase ramp of
    ApplyShiftRamp (Sim/DriftShift.cpp:24-48), cosine only at the Nyquist bins; the normalised inverse, cropped."""

    def __init__(self):
        self.geometry = None
        self.shape = None
        self.planes = {}
        self.jobPlanes = 0

    def HasPlane(self, geometry, key) -> bool:
        return geometry == self.geometry and key in self.planes

    def Images(self, job) -> list:
        NX, NY = job.NX, job.NY
        if job.geometry != self.geometry or self.shape != (NX, NY):
            self.planes = {}
            self.geometry = job.geometry
            self.shape = (NX, NY)
        for key, cells, values in job.planes:
            img = np.zeros((NY, NX))
            cells = np.asarray(cells, np.int64)
            img[cells // job.nx, cells % job.nx] = np.asarray(values, np.float64)
            self.planes[key] = np.fft.rfft2(img)
            self.jobPlanes += 1
        K = {p: np.asarray(spec, np.complex128) for p, spec in job.kernels}
        W = NX // 2 + 1
        kx = np.arange(W)
        ky = np.arange(NY)
        kys = np.where(ky <= NY // 2, ky, ky - NY).astype(np.float64)
        ax = 2 * np.pi * kx * job.fracX / NX
        ay = 2 * np.pi * kys * job.fracY / NY
        px = np.where(kx == NX // 2, np.cos(ax) + 0j, np.exp(1j * ax))
        py = np.where(ky == NY // 2, np.cos(ay) + 0j, np.exp(1j * ay))
        out = []
        for deps in job.channels:
            acc = np.zeros((NY, W), np.complex128)
            for key, p0, p1, w0, w1, two in deps:
                if key not in self.planes or p0 not in K or (two and p1 not in K):
                    raise KeyError("job refers to a plane or kernel it does not carry")
                k = w0 * K[p0]
                if two:
                    k = k + w1 * K[p1]
                acc += self.planes[key] * k
            acc *= py[:, None] * px[None, :]
            img = np.fft.irfft2(acc, s=(NY, NX))
            out.append(img[job.fovY0:job.fovY0 + job.ch, job.fovX0:job.fovX0 + job.cw])
        return out


def _i32(v):
    v &= 0xFFFFFFFF
    return v - (1 << 32) if v >= (1 << 31) else v


def load_oracle() -> Oracle:
    return Oracle()


def load_reference() -> Reference | None:
    try:
        return Reference()
    except (FileNotFoundError, OSError):
        return None
