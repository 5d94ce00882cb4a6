"""ctypes binding of libiscg.so (include/iscg.h). Loading fails loudly when the library is missing: there is no
CPU fallback behind this package."""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libiscg.so")


class IscgError(RuntimeError):
    """A failed libiscg call (the message is iscg_last_error's)."""


class Camera(C.Structure):
    """iscg_camera = sim::CameraNoiseParams (Sim/SMLMNoise.h:35-50)."""

    _fields_ = [("quantumEfficiency", C.c_double), ("darkCurrentElectrons", C.c_double),
                ("gainPhotonsPerAdu", C.c_double), ("readNoiseElectrons", C.c_double),
                ("emccd", C.c_int32), ("bitDepth", C.c_int32), ("emGain", C.c_double),
                ("cicElectrons", C.c_double)]


class KernelView(C.Structure):
    """iscg_kernel_view: one kernel stack for the splat."""

    _fields_ = [("planes", C.c_void_p), ("spanLo", C.c_void_p), ("spanHi", C.c_void_p), ("zStepNm", C.c_double),
                ("serial", C.c_uint64), ("nz", C.c_int32), ("size", C.c_int32), ("oversampling", C.c_int32),
                ("interpMode", C.c_int32)]


class WfDesc(C.Structure):
    """iscg_wf_desc: a dye grid and its focus series."""

    _fields_ = [("dzUm", C.c_double), ("focus0Um", C.c_double), ("k0", C.c_int64), ("nx", C.c_uint32),
                ("ny", C.c_uint32), ("nzD", C.c_uint32), ("nF", C.c_uint32), ("fovX0", C.c_uint32),
                ("fovY0", C.c_uint32), ("cw", C.c_uint32), ("ch", C.c_uint32), ("NX", C.c_uint32),
                ("NY", C.c_uint32), ("NZ", C.c_uint32), ("upscale", C.c_int32)]


class WfDep(C.Structure):
    """iscg_wf_dep: WidefieldGpuJob::Dep (Sim/WidefieldRender.h:325-331)."""

    _fields_ = [("key", C.c_uint64), ("p0", C.c_int32), ("p1", C.c_int32), ("w0", C.c_float), ("w1", C.c_float),
                ("two", C.c_int32), ("reserved", C.c_int32)]


class WfJob(C.Structure):
    """iscg_wf_job: WidefieldGpuJob (Sim/WidefieldRender.h:306-340) flattened."""

    _fields_ = [("NX", C.c_uint32), ("NY", C.c_uint32), ("nx", C.c_uint32), ("ny", C.c_uint32),
                ("fovX0", C.c_uint32), ("fovY0", C.c_uint32), ("cw", C.c_uint32), ("ch", C.c_uint32),
                ("fracX", C.c_double), ("fracY", C.c_double), ("geometry", C.c_uint64), ("nKernels", C.c_int32),
                ("nPlanes", C.c_int32), ("nChannels", C.c_int32), ("hasPersistent", C.c_int32),
                ("kernelP", C.c_void_p), ("kernelSpec", C.c_void_p), ("planeKey", C.c_void_p),
                ("planeStart", C.c_void_p), ("cells", C.c_void_p), ("values", C.c_void_p),
                ("channelStart", C.c_void_p), ("deps", C.c_void_p)]


class WfShiftDesc(C.Structure):
    """iscg_wf_shift_desc: the kept spectra of a drift's focus images (Sim/WidefieldRender.h:281-290)."""

    _fields_ = [("NX", C.c_uint32), ("NY", C.c_uint32), ("fovX0", C.c_uint32), ("fovY0", C.c_uint32),
                ("cw", C.c_uint32), ("ch", C.c_uint32), ("upscale", C.c_int32), ("nImages", C.c_int32),
                ("nChannels", C.c_int32), ("hasPersistent", C.c_int32), ("fracX", C.c_double), ("fracY", C.c_double)]


class WfShiftFrame(C.Structure):
    """iscg_wf_shift_frame: one drift frame (WidefieldDriftFrames::Render, Sim/WidefieldRender.cpp:2964-2977)."""

    _fields_ = [("i0", C.c_int32), ("i1", C.c_int32), ("w", C.c_double), ("dxCells", C.c_double),
                ("dyCells", C.c_double)]


class RoiModel(C.Structure):
    """iscg_roi_model."""

    _fields_ = [("wavelengthNm", C.c_double), ("na", C.c_double), ("immersionIndex", C.c_double),
                ("sampleIndex", C.c_double), ("sampleDepthNm", C.c_double), ("pixelSizeNm", C.c_double),
                ("pupilSamples", C.c_int32), ("roiSize", C.c_int32), ("oversampling", C.c_int32),
                ("reserved", C.c_int32)]


class PsfRequest(C.Structure):
    """iscg_psf_request: the numeric fields of sim::PsfGeneratorRequest (Sim/PsfGeneratorBridge.h:91-170)."""

    _fields_ = [("wavelengthNm", C.c_double), ("na", C.c_double), ("immersionIndex", C.c_double),
                ("sampleIndex", C.c_double), ("workingDistanceUm", C.c_double), ("sampleDepthNm", C.c_double),
                ("pixelSizeNm", C.c_double), ("zStepNm", C.c_double), ("maskWaist", C.c_double),
                ("zernike", C.c_double * 28), ("maskType", C.c_int32), ("maskModes", C.c_int32),
                ("oversampling", C.c_int32), ("kernelHalfWidthPx", C.c_int32), ("nz", C.c_int32),
                ("pupilSamples", C.c_int32), ("interpMode", C.c_int32), ("reserved", C.c_int32)]


class Kinetics(C.Structure):
    """iscg_kinetics = isc::Kinetics (core/src/dyes.h:129-136)."""
    _fields_ = [(n, C.c_double) for n in ("activationRatePerSec", "onSec", "offSec", "bleachProb", "photonCV",
                                          "initialOnSec")]


class KineticsHistory(C.Structure):
    _fields_ = [("n", C.c_int32), ("reserved", C.c_int32), ("tStart", C.POINTER(C.c_double)),
                ("kin", C.POINTER(Kinetics))]


class WorldCell(C.Structure):
    """iscg_world_cell: the fields of isc::Cell the block walk reads (core/src/cells.h:48-75)."""
    _fields_ = [("cx", C.c_int32), ("cy", C.c_int32), ("x", C.c_double), ("y", C.c_double),
                ("packRot", C.c_double), ("rOuter", C.c_double)]


class Npc(C.Structure):
    """iscg_npc = isc::Npc (core/src/npcs.h:36-40) without its surface piece."""
    _fields_ = [("c", C.c_double * 3), ("n", C.c_double * 3), ("e1", C.c_double * 3), ("e2", C.c_double * 3),
                ("cand", C.c_int32), ("reserved", C.c_int32)]


class LaminPoint(C.Structure):
    """iscg_lamin_point = isc::LaminPoint (core/src/npcs.h:44-48) without its surface piece."""
    _fields_ = [("c", C.c_double * 3), ("n", C.c_double * 3), ("cand", C.c_int32), ("reserved", C.c_int32)]


class WorldGeometry(C.Structure):
    """iscg_world_geometry: the cells' assets (isc::CellAssets, core/src/world.h:58-90) as flat host arrays."""
    _fields_ = [("nCells", C.c_int32), ("cells", C.c_void_p), ("mtStart", C.c_void_p), ("ptStart", C.c_void_p),
                ("mtIndex", C.c_void_p), ("pts", C.c_void_p), ("npcStart", C.c_void_p), ("npcs", C.c_void_p),
                ("laminStart", C.c_void_p), ("lamin", C.c_void_p), ("npcSites", C.c_double * 192),
                ("npcSiteReachUm", C.c_double)]


class Label(C.Structure):
    """iscg_label = isc::Label's emission and linkage (core/src/dyes.h:163-176)."""
    _fields_ = [("density", C.c_double), ("fluorescentFraction", C.c_double), ("mode", C.c_int32),
                ("preState", C.c_int32), ("kin", Kinetics), ("hasLinkage", C.c_int32), ("reserved", C.c_int32),
                ("binderNm", C.c_double), ("binderFlexNm", C.c_double), ("linkerMinNm", C.c_double),
                ("linkerMaxNm", C.c_double)]


class WorldDye(C.Structure):
    """iscg_world_dye = isc::WorldDye (core/src/world.h:92-99)."""
    _fields_ = [("x", C.c_double), ("y", C.c_double), ("z", C.c_double), ("id", C.c_uint32), ("cx", C.c_int32),
                ("cy", C.c_int32), ("mtIndex", C.c_int32), ("k", C.c_int32), ("n", C.c_int32),
                ("structure", C.c_int32), ("reserved", C.c_int32)]


class BlockDye(C.Structure):
    """iscg_block_dye: a generator's dye (isc::Dye / isc::NpcDye, core/src/dyes.h:263-300) with its walk block."""
    _fields_ = [("x", C.c_double), ("y", C.c_double), ("z", C.c_double), ("id", C.c_uint32), ("block", C.c_int32),
                ("structure", C.c_int32), ("element", C.c_int32), ("k", C.c_int32), ("n", C.c_int32),
                ("reserved", C.c_int32)]


class WorldEvent(C.Structure):
    """iscg_world_event = isc::WorldEvent without its cache key (core/src/world.h:103-111)."""
    _fields_ = [("x", C.c_double), ("y", C.c_double), ("z", C.c_double), ("tOn", C.c_double), ("tOff", C.c_double),
                ("brightness", C.c_double), ("aux", C.c_double), ("id", C.c_uint32), ("structure", C.c_uint8),
                ("state", C.c_uint8), ("reserved", C.c_uint8 * 2)]


_lib = None

I32, U32, U64, D, P = C.c_int32, C.c_uint32, C.c_uint64, C.c_double, C.c_void_p

# name -> (restype, argtypes); every symbol include/iscg.h declares
SIGNATURES = {
    "iscg_abi_version": (I32, []),
    "iscg_create": (I32, [I32, C.POINTER(P)]),
    "iscg_destroy": (None, [P]),
    "iscg_last_error": (C.c_char_p, [P]),
    "iscg_synchronize": (I32, [P]),
    "iscg_stream": (P, [P]),
    "iscg_inject_failure": (I32, [P, I32, C.c_char_p]),
    "iscg_profile_enable": (I32, [P, I32]),
    "iscg_profile_read": (I32, [P, C.c_char_p, C.POINTER(C.c_int64), C.POINTER(C.c_double), I32, I32]),
    "iscg_set_stream": (I32, [P, P]),
    "iscg_psf_library_create": (I32, [P, P, I32, C.POINTER(P)]),
    "iscg_psf_library_update": (I32, [P, P, P, I32]),
    "iscg_psf_library_geometry": (I32, [P, C.POINTER(I32), C.POINTER(I32), C.POINTER(I32)]),
    "iscg_psf_library_run": (I32, [P, P, I32, P, C.c_size_t]),
    "iscg_psf_library_destroy": (I32, [P, P]),
    "iscg_halo_spans": (I32, [P, P, I32, I32, I32, I32, D, P, P, P, C.POINTER(D)]),
    "iscg_tiff_bytes": (C.c_uint64, [C.c_uint64] * 5),
    "iscg_tiff_open": (I32, [C.c_char_p, C.c_uint64, I32, C.POINTER(P)]),
    "iscg_tiff_page_u16": (I32, [P, P, U32, U32, C.c_char_p]),
    "iscg_tiff_page_f32": (I32, [P, P, U32, U32, C.c_char_p]),
    "iscg_tiff_page_rows_u16": (I32, [P, P, U32, U32, U32, C.c_char_p]),
    "iscg_tiff_rows_u16": (I32, [P, C.c_uint64, U32, U32, P]),
    "iscg_tiff_frames_u16": (I32, [P, P, U32, U32, U32, C.c_char_p]),
    "iscg_tiff_info": (I32, [P, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64), C.POINTER(I32)]),
    "iscg_tiff_error": (C.c_char_p, [P]),
    "iscg_tiff_close": (I32, [P]),
    "iscg_tiff_destroy": (None, [P]),
    "iscg_collect_frame_emitters": (I32, [P, P, U32, P, P, P, U32, U32, U32, D, D, P, I32, D, D, I32, P, P, U64, P,
                                          P]),
    "iscg_bucket_events": (I32, [P, P, U32, U32, P, P, U64, C.POINTER(U64)]),
    "iscg_autotune_set": (I32, [P, I32]),
    "iscg_autotune_dump": (I32, [P, C.c_char_p, I32]),
    "iscg_autotune_load": (I32, [P, C.c_char_p]),
    "iscg_binned_setup": (I32, [P, P, I32, D, U32, U32, I32]),
    "iscg_binned_geometry": (I32, [P, P, P]),
    "iscg_binned_render": (I32, [P, P, P, I32, P]),
    "iscg_pcg4d": (I32, [P, P, P, P, P, U32, P]),
    "iscg_hash_units": (I32, [P, U32, P, P, P, U32, P]),
    "iscg_hash_stream": (I32, [P, U32, I32, I32, U32, U32, P]),
    "iscg_noise_setup": (I32, [C.c_int64, U32, U32, U32, U32, D, D, D, D, D, D, P, P, P, P, P]),
    "iscg_psf_kernel_half_width_px": (I32, [D, D, D, D]),
    "iscg_capped_psf_oversampling": (I32, [I32, D, D]),
    "iscg_wf_has_plane": (I32, [P, C.c_uint64, C.c_uint64]),
    "iscg_wf_fits": (I32, [P, C.c_uint64, U32, U32, C.c_uint64]),
    "iscg_wf_images": (I32, [P, C.POINTER(WfJob), P]),
    "iscg_wf_set_budget": (I32, [P, C.c_uint64]),
    "iscg_wf_shifted_frames": (I32, [P, C.POINTER(WfShiftDesc), P, P, P, U32, P]),
    "iscg_wf_render": (I32, [P, P, I32, I32, U32, U32, I32, P, U32, P, C.POINTER(Camera), P, U32, P, U32, P, P]),
    "iscg_noise_set_path": (I32, [P, I32]),
    "iscg_noise_bounds": (I32, [P, P]),
    "iscg_noise": (I32, [P, P, U32, U32, C.POINTER(Camera), P, P, P, U32, P, U32, U32, P]),
    "iscg_noise_noalloc": (I32, [P, P, U32, U32, C.POINTER(Camera), P, P, P, U32, P, U32, U32, P]),
    "iscg_splat_set_kernels": (I32, [P, C.POINTER(KernelView), I32]),
    "iscg_splat_set_static": (I32, [P, U32, U32, U32, P, P, P, P, D, U64]),
    "iscg_splat_render_frames": (I32, [P, P, U32, P, U32, P, P, P, C.POINTER(Camera), U32, D, P, P]),
    "iscg_splat_render_records": (I32, [P, P, U32, P, U32, P, P, P, C.POINTER(Camera), U32, P, P]),
    "iscg_psf_geometry": (I32, [C.POINTER(PsfRequest), C.POINTER(I32), C.POINTER(I32), C.POINTER(I32)]),
    "iscg_psf_planes": (I32, [P, C.POINTER(PsfRequest), I32, I32, P, C.c_size_t]),
    "iscg_roi_forward": (I32, [P, C.POINTER(RoiModel), C.POINTER(D), P, I32, P]),
    "iscg_roi_backward": (I32, [P, C.POINTER(RoiModel), C.POINTER(D), P, I32, P, P, C.POINTER(D), C.POINTER(D)]),
    "iscg_wf_conv3d_plan": (I32, [P, C.POINTER(WfDesc), P, I32, I32, I32, D, I32]),
    "iscg_wf_conv3d": (I32, [P, P, C.POINTER(Camera), P, D, U32, U32, U32, P, P]),
    "iscg_label_blinks": (I32, [P, P, I32, I32, C.POINTER(Kinetics), C.POINTER(KineticsHistory), D, P, P, C.c_int64,
                                P]),
    "iscg_label_cont": (I32, [P, P, I32, I32, I32, C.POINTER(Kinetics), C.POINTER(KineticsHistory), P, P]),
    "iscg_persistent_blinks": (I32, [P, P, I32, C.POINTER(Kinetics), C.POINTER(KineticsHistory), D, D, P, P,
                                     C.c_int64, P]),
    "iscg_persistent_bins": (I32, [P, P, I32, C.POINTER(Kinetics), C.POINTER(KineticsHistory), C.c_int64, C.c_int64,
                                   P, P, C.c_int64, P]),
    "iscg_jsm": (I32, [P, I32, P, C.c_int64, P]),
    "iscg_world_create": (I32, [P, U32, C.POINTER(WorldGeometry), C.POINTER(P)]),
    "iscg_world_destroy": (I32, [P]),
    "iscg_world_set_label": (I32, [P, I32, C.POINTER(Label)]),
    "iscg_world_set_history": (I32, [P, P, P, I32]),
    "iscg_world_block_dyes": (I32, [P, P, P, C.c_int64, P]),
    "iscg_world_sites": (I32, [P, P, P, C.c_int64, P]),
    "iscg_world_events": (I32, [P, P, D, D, P, C.c_int64, P]),
    "iscg_world_continuous": (I32, [P, P, D, P, C.c_int64, P]),
    "iscg_world_density3d": (I32, [P, P, I32, I32, I32, U32, P, P]),
}


def library_path() -> str:
    return LIB_PATH


def _share_hip_runtime():
    """Loads torch's HIP runtime first when torch is installed. torch's wheel carries its own libamdhip64 /
    libhsa-runtime64 with the same SONAMEs as ROCm's; the loader then resolves libiscg.so's dependencies to the copies
    already loaded, so torch tensors and libiscg share one runtime (two runtimes in one process cannot both open
    the GPU)."""
    try:
        import torch  # noqa: F401
    except ImportError:
        pass


def lib():
    """The loaded library (raises if libiscg.so was not built)."""
    global _lib
    if _lib is None:
        _share_hip_runtime()
        if not os.path.exists(LIB_PATH):
            raise IscgError(f"{LIB_PATH} is missing: build it first (python -m insiliscope_amd.build)")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            if hasattr(L, name):
                fn = getattr(L, name)
                fn.restype = res
                fn.argtypes = args
        _lib = L
    return _lib
