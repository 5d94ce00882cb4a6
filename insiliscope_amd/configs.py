"""The benchmark workloads, restated in the reference's conventions, with seeded synthetic inputs.

Conventions (the reference's, Sim/ = adapter/inSiliScope/Simulation/):
  * Zernike amplitudes are given per Noll index and written unchanged, in waves, into the OSA/ANSI slot the
    reference uses (unnormalised R_n^m, Sim/ZernikePsf.cpp:35-78).
  * "32x32" / "64x64" kernels are the reference's odd grids 33x33 / 65x65 (2 hw os + 1, Sim/ZernikePsf.cpp:118-130).
  * ti0 (the 170 um coverslip) cancels out of the phase (Sim/ZernikePsf.cpp:238,244); ns < NA cuts the pupil at ns.
  * "cubic spline" splats are Catmull-Rom on block sums (Sim/PsfGeneratorBridge.cpp:1250); z is nearest plane.
  * EMCCD conversion gain is 1 e-/ADU / EM gain; an explicit constant offset map carries offset 100.
  * Every random draw comes from pcg4d (core/src/rng.h:18): draw (seed, a, b, c) -> four uniforms
    ((x >> 9) + 0.5) 2^-23, normals by Box-Muller of the first two, Poisson by CounterPoisson's rule
    (Sim/SMLMCounterRng.h:104-119). Uses of the four words are fixed per quantity (the `_U*` tags below).
"""
from __future__ import annotations

import numpy as np

from .host import BLINK_EVENT_DTYPE, CameraNoiseParams

# Noll j -> OSA/ANSI j (Noll's parity rule: even j cos, odd j sin)
NOLL_TO_OSA = {1: 0, 2: 2, 3: 1, 4: 4, 5: 3, 6: 5, 7: 7, 8: 8, 9: 6, 10: 9, 11: 12, 12: 13, 13: 11, 14: 14, 15: 10,
               16: 18, 17: 17, 18: 19, 19: 16, 20: 20, 21: 15, 22: 24, 23: 23, 24: 25, 25: 22, 26: 26, 27: 21,
               28: 27}

_M = np.uint32(1664525)
_A = np.uint32(1013904223)


def pcg4d(a, b, c, d):
    """Vectorised pcg4d (core/src/rng.h:18-35) on uint32 arrays."""
    with np.errstate(over="ignore"):
        a = np.asarray(a, np.uint32) * _M + _A
        b = np.asarray(b, np.uint32) * _M + _A
        c = np.asarray(c, np.uint32) * _M + _A
        d = np.asarray(d, np.uint32) * _M + _A
        a = a + b * d
        b = b + c * a
        c = c + a * b
        d = d + b * c
        a ^= a >> np.uint32(16)
        b ^= b >> np.uint32(16)
        c ^= c >> np.uint32(16)
        d ^= d >> np.uint32(16)
        a = a + b * d
        b = b + c * a
        c = c + a * b
        d = d + b * c
    return a, b, c, d


def unit(x):
    """((x >> 9) + 0.5) 2^-23, exact in FP64 (core/src/rng.h:40)."""
    return ((np.asarray(x, np.uint32) >> np.uint32(9)).astype(np.float64) + 0.5) * 1.1920928955078125e-7


def uniforms(seed, a, b, c):
    """Four (0,1) uniforms per address (seed, a, b, c), shape (..., 4)."""
    a, b, c = np.broadcast_arrays(np.asarray(a, np.int64), np.asarray(b, np.int64), np.asarray(c, np.int64))
    w = pcg4d(np.full(a.shape, seed, np.uint32), (a & 0xFFFFFFFF).astype(np.uint32),
              (b & 0xFFFFFFFF).astype(np.uint32), (c & 0xFFFFFFFF).astype(np.uint32))
    return np.stack([unit(x) for x in w], axis=-1)


def normals(u):
    """Box-Muller of u[..., 0], u[..., 1] (radius first, angle second, Sim/SMLMCounterRng.h:95-99)."""
    return np.sqrt(-2.0 * np.log(u[..., 0])) * np.cos(6.283185307179586 * u[..., 1])


def poisson_counts(seed, tag, n, mean):
    """n Poisson(mean) counts, one address per draw: exact (Knuth) up to 60, rounded normal above (the
    CounterPoisson rule)."""
    idx = np.arange(n, dtype=np.int64)
    if mean > 60.0:
        g = normals(uniforms(seed, tag, idx, 0))
        return np.maximum(0.0, np.floor(mean + np.sqrt(mean) * g + 0.5)).astype(np.int64)
    out = np.zeros(n, np.int64)
    L = np.exp(-mean)
    for i in range(n):
        p, k, ctr = 1.0, 0, 0
        while True:
            u = uniforms(seed, tag, i, ctr)
            ctr += 1
            done = False
            for j in range(4):
                k += 1
                p *= u[j]
                if not (p > L and k < 1000):
                    done = True
                    break
            if done:
                break
        out[i] = k - 1
    return out


# quantity tags (the second pcg4d word's high bits keep the streams apart)
_U_ZERN = 0x5A45524E
_U_COUNT = 0x434F554E
_U_EMIT = 0x454D4954
_U_DENS = 0x44454E53


def zernike_noll(seed, nolls, amplitude=None, sigma=None, set_index=0):
    """An OSA 28-vector with the given Noll modes drawn uniform in [-amplitude, amplitude] or N(0, sigma^2)."""
    z = np.zeros(28)
    for j in nolls:
        u = uniforms(seed, _U_ZERN, set_index, j)
        v = (2.0 * u[0] - 1.0) * amplitude if amplitude is not None else sigma * normals(u)
        z[NOLL_TO_OSA[j]] = float(v)
    return z


# ---- PSF requests ----------------------------------------------------------------------------------------------

def psf_request_fields(**kw):
    """A PsfGeneratorRequest's numeric fields with the reference's defaults (Sim/PsfGeneratorBridge.h:91-170)."""
    r = dict(wavelengthNm=660.0, na=1.4, immersionIndex=1.518, sampleIndex=1.518, workingDistanceUm=150.0,
             sampleDepthNm=0.0, pixelSizeNm=100.0, zStepNm=100.0, maskWaist=1.0, zernike=np.zeros(28), maskType=0,
             maskModes=5, oversampling=4, kernelHalfWidthPx=8, nz=1, pupilSamples=0, interpMode=0)
    r.update(kw)
    return r


def c1_psf(seed=1, oversampling=1, kernel_half_width=16):
    """C1: NA 1.45, ni 1.518, ns 1.33, 670 nm, ti0 170 um, M 128, 41 planes at 50 nm, Noll 5-15 uniform +-0.05."""
    return psf_request_fields(wavelengthNm=670.0, na=1.45, immersionIndex=1.518, sampleIndex=1.33,
                              workingDistanceUm=170.0, sampleDepthNm=0.0, pixelSizeNm=100.0 if oversampling == 1
                              else 100.0, zStepNm=50.0, oversampling=oversampling,
                              kernelHalfWidthPx=kernel_half_width, nz=41, pupilSamples=128, interpMode=2,
                              zernike=zernike_noll(seed, range(5, 16), amplitude=0.05))


def c3_requests(seed=1, n_lambda=8, n_sets=64):
    """C3: 8 wavelengths 520-700 nm x 64 Zernike sets (Noll 4-21, N(0, 0.04^2)), 65x65 at hw 32 os 1, 41 planes at
    50 nm, M 128, NA 1.45, ni = ns = 1.518."""
    reqs = []
    for li in range(n_lambda):
        lam = 520.0 + 180.0 * li / max(1, n_lambda - 1)
        for s in range(n_sets):
            reqs.append(psf_request_fields(wavelengthNm=lam, na=1.45, immersionIndex=1.518, sampleIndex=1.518,
                                           workingDistanceUm=170.0, pixelSizeNm=100.0, zStepNm=50.0, oversampling=1,
                                           kernelHalfWidthPx=32, nz=41, pupilSamples=128, interpMode=0,
                                           zernike=zernike_noll(seed, range(4, 22), sigma=0.04, set_index=s)))
    return reqs


def c4_psf(seed=1):
    """C4 splat table: hw 8, os 4 (65x65), 81 planes at 15 nm, Catmull-Rom; NA 1.45, ni 1.518, ns 1.33, 670 nm."""
    return psf_request_fields(wavelengthNm=670.0, na=1.45, immersionIndex=1.518, sampleIndex=1.33,
                              workingDistanceUm=170.0, pixelSizeNm=100.0, zStepNm=15.0, oversampling=4,
                              kernelHalfWidthPx=8, nz=81, pupilSamples=128, interpMode=2,
                              zernike=zernike_noll(seed, range(5, 16), amplitude=0.05))


# ---- cameras ---------------------------------------------------------------------------------------------------

def scmos_c1():
    return CameraNoiseParams(quantumEfficiency=1.0, darkCurrentElectrons=0.0, gainPhotonsPerAdu=2.0,
                             readNoiseElectrons=1.5, emccd=False)


def emccd_c4():
    return CameraNoiseParams(quantumEfficiency=1.0, darkCurrentElectrons=0.0, gainPhotonsPerAdu=1.0 / 300.0,
                             readNoiseElectrons=50.0, emccd=True, emGain=300.0, cicElectrons=0.002, bitDepth=16)


# ---- blink events ----------------------------------------------------------------------------------------------

def blink_events(seed, n_frames, width, height, pixel_nm, mean_per_frame, z_half_nm, photons_median,
                 sigma_log=0.5, first_frame=0):
    """Per frame Poisson(mean) blinks, x,y uniform over the FOV, z uniform +-z_half_nm, photons log-normal (median,
    sigma_log), each blink exactly on [f, f+1) (tStart = f, tEnd = f + 1, brightness = photons, perBlink 1).
    Returns (events in frame order, per-frame offsets)."""
    counts = poisson_counts(seed, _U_COUNT, n_frames + first_frame, mean_per_frame)[first_frame:]
    start = np.zeros(n_frames + 1, np.int64)
    start[1:] = np.cumsum(counts)
    n = int(start[-1])
    ev = np.zeros(n, BLINK_EVENT_DTYPE)
    frame = np.repeat(np.arange(first_frame, first_frame + n_frames, dtype=np.int64), counts)
    idx = np.arange(n, dtype=np.int64) - np.repeat(start[:-1], counts)
    u = uniforms(seed, _U_EMIT + frame, idx, 0)
    v = uniforms(seed, _U_EMIT + frame, idx, 1)
    ev["xUm"] = u[:, 0] * width * pixel_nm / 1000.0
    ev["yUm"] = u[:, 1] * height * pixel_nm / 1000.0
    ev["zNm"] = (2.0 * u[:, 2] - 1.0) * z_half_nm
    ev["tStart"] = frame
    ev["tEnd"] = frame + 1.0
    ev["brightness"] = photons_median * np.exp(sigma_log * normals(v))
    return ev, start


def c1_events(seed=1, n_frames=50):
    """C1: 128x128 px at 100 nm, 50 frames, ~200 blinks/frame, z +-1 um, photons median 1500 (sigma_log 0.5)."""
    return blink_events(seed, n_frames, 128, 128, 100.0, 200.0, 1000.0, 1500.0)


def c4_events(seed=1, n_frames=2000, first_frame=0):
    """C4: 512x512 px at 100 nm, Poisson(800) blinks/frame, z +-600 nm, photons median 3000 (sigma_log 0.5)."""
    return blink_events(seed, n_frames, 512, 512, 100.0, 800.0, 600.0, 3000.0, first_frame=first_frame)


# ---- ROI adjoint (C5) -------------------------------------------------------------------------------------------

_U_ROI = 0x524F4949


def c5_rois(seed=1, n=4096, roi=15):
    """C5: n ROIs of roi x roi pixels; x, y uniform +-0.5 px about the centre, z uniform +-500 nm, photons
    2000 exp(0.5 N(0,1)), background uniform [5, 50]; shared Zernike Noll 4-24 from N(0, 0.03^2) (OSA slots).
    Returns (zernike[28], params) with params in ROI_PARAMS order (x, y, z, photons, background)."""
    z = zernike_noll(seed, range(4, 25), sigma=0.03)
    idx = np.arange(n, dtype=np.int64)
    u = uniforms(seed, _U_ROI, idx, 0)
    v = uniforms(seed, _U_ROI, idx, 1)
    c = (roi - 1) / 2.0 + 0.5
    params = np.stack([c + u[:, 0] - 0.5, c + u[:, 1] - 0.5, (2.0 * u[:, 2] - 1.0) * 500.0,
                       2000.0 * np.exp(0.5 * normals(v)), 5.0 + 45.0 * u[:, 3]], axis=1)
    return z, params


# ---- mean-field density ----------------------------------------------------------------------------------------

def c2_density(seed=1, nx=256, ny=256, nz=64, occupancy=0.03):
    """C2: dye counts on a nz x ny x nx grid (~3 % of voxels occupied, counts 1-3; two populations summed with
    weights 1 and 0.5), from pcg4d."""
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    u = uniforms(seed, _U_DENS + z, y, x)
    on = u[..., 0] < occupancy
    count = 1.0 + np.floor(u[..., 1] * 3.0)
    second = u[..., 2] < 0.3
    w = np.where(second, 0.5, 1.0)
    return np.where(on, count * w, 0.0).astype(np.float32)


def world_field(seed=1, n_cells=5, mts_per_cell=440, pores_per_cell=1260, field_um=25.6):
    """A synthetic cell field the size of the seed-1 field over a 25.6 um window (5 cells, about 2200 microtubules
    of 70-200 points at 0.12 um steps, 6300 pores on ellipsoidal nuclei), as WorldGeometry fields: the shapes the
    dye-site generators take, drawn from a seeded generator (the cells' own growth rules are not restated)."""
    rng = np.random.default_rng(seed)
    centres = np.array([(4.0, 5.0), (14.0, 20.0), (22.0, 6.0), (8.0, 18.0), (20.0, 14.0)])[:n_cells]
    centres = centres * field_um / 25.6
    cells = np.zeros(n_cells, [("cx", "<i4"), ("cy", "<i4"), ("x", "<f8"), ("y", "<f8"), ("packRot", "<f8"),
                               ("rOuter", "<f8")])
    mt_start, pt_start, pts = [0], [0], []
    npc_start, npcs = [0], []
    for c in range(n_cells):
        cells[c] = (c - 2, c % 2, centres[c, 0], centres[c, 1], rng.uniform(-0.6, 0.6), 15.0)
        for _ in range(mts_per_cell):
            n = int(rng.integers(70, 200))
            a0 = rng.uniform(0, 2 * np.pi)
            ang = a0 + np.cumsum(rng.normal(0, 0.08, n))
            r0 = 4.5
            step = 0.12
            x = r0 * np.cos(a0) + np.concatenate([[0], np.cumsum(step * np.cos(ang[1:]))])
            y = r0 * np.sin(a0) + np.concatenate([[0], np.cumsum(step * np.sin(ang[1:]))])
            z = np.clip(1.8 - np.linspace(0, 1.4, n) + np.cumsum(rng.normal(0, 0.01, n)), 0.2, 3.0)
            pts.append(np.stack([x, y, z], 1))
            pt_start.append(pt_start[-1] + n)
        mt_start.append(mt_start[-1] + mts_per_cell)
        # pores on the nucleus ellipsoid (semi-axes 5.5, 4.5, 1.3 um about (0, 0, 1.7)), outward normals
        u = rng.uniform(-1, 1, pores_per_cell)
        phi = rng.uniform(0, 2 * np.pi, pores_per_cell)
        s = np.sqrt(1 - u * u)
        ax = np.array([5.5, 4.5, 1.3])
        cpos = np.stack([ax[0] * s * np.cos(phi), ax[1] * s * np.sin(phi), 1.7 + ax[2] * u], 1)
        nrm = (cpos - [0, 0, 1.7]) / ax ** 2
        nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
        ref = np.where(np.abs(nrm[:, 2:3]) < 0.9, [[0.0, 0.0, 1.0]], [[1.0, 0.0, 0.0]])
        e1 = np.cross(nrm, ref)
        e1 /= np.linalg.norm(e1, axis=1, keepdims=True)
        e2 = np.cross(nrm, e1)
        rows = np.zeros((pores_per_cell, 13))
        rows[:, 0:3], rows[:, 3:6], rows[:, 6:9], rows[:, 9:12] = cpos, nrm, e1, e2
        rows[:, 12] = 2 * np.arange(pores_per_cell)
        npcs.append(rows)
        npc_start.append(npc_start[-1] + pores_per_cell)
    # 64 attachment sites: 2 rings at +-25 nm along the axis, 8 corners at 45 degrees, 4-point arcs of 20 nm spread
    k = np.arange(64)
    ring, corner, arc = k // 32, (k % 32) // 4, k % 4
    ang = corner * np.pi / 4 + (arc - 1.5) * np.radians(45) / 2 * 0.2
    sites = np.stack([53 * np.cos(ang), 53 * np.sin(ang), np.where(ring == 0, -25.0, 25.0)], 1)
    return dict(cells=cells, mtStart=np.array(mt_start, np.int32), ptStart=np.array(pt_start, np.int64),
                pts=np.concatenate(pts), npcStart=np.array(npc_start, np.int32), npcRows=np.concatenate(npcs),
                laminStart=np.zeros(n_cells + 1, np.int32), laminRows=np.zeros((0, 7)),
                npcSites=sites.reshape(-1), npcSiteReachUm=float(np.hypot(63.0, 25.0) * 1e-3))
