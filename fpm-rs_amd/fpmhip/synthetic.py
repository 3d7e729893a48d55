"""Synthetic inputs of the measured workloads (the reference has no generator
for them): smoothed Gaussian random objects, Noll Zernike pupils, and a
Poisson camera that yields exact uint16 counts. Seeded numpy generators;
nothing here is on the reconstruction path.
"""
from __future__ import annotations

import math

import numpy as np

from .model import Optics, PlanarLedArray


def base_optics(wavelength=532e-9) -> Optics:
    """532 nm, NA 0.10, 4x, 6.5 um pixels (src/simulation/presets.rs:71-82)."""
    return Optics(wavelength_vacuum_m=wavelength, objective_na=0.10, magnification=4.0, camera_pixel_size=6.5e-6)


def planar_array(k: int) -> PlanarLedArray:
    """k x k lattice, 4 mm pitch, centred reference, 90 mm below the sample."""
    return PlanarLedArray(shape=(k, k), pitch_m=(4e-3, 4e-3), reference_index=((k - 1) / 2.0, (k - 1) / 2.0),
                          translation_m=(0.0, 0.0, -90e-3))


def gaussian_smooth_periodic(field: np.ndarray, sigma: float) -> np.ndarray:
    """Periodic separable Gaussian smoothing, kernel truncated at 4 sigma and
    normalised to unit sum (a convex combination, so ranges are preserved)."""
    radius = int(math.ceil(4.0 * sigma))
    x = np.arange(-radius, radius + 1, dtype=np.float64)
    k = np.exp(-0.5 * (x / sigma) ** 2)
    k /= k.sum()
    out = np.zeros_like(field)
    for i, w in zip(range(-radius, radius + 1), k):
        out += w * np.roll(field, i, axis=0)
    out2 = np.zeros_like(out)
    for i, w in zip(range(-radius, radius + 1), k):
        out2 += w * np.roll(out, i, axis=1)
    return out2


def uniform_object(shape, seed, amp=(0.5, 1.0), phase=(-math.pi / 2, math.pi / 2), sigma=1.0):
    """Amplitude U[a,b], phase U[c,d], i.i.d., then periodic Gaussian smoothing."""
    rng = np.random.default_rng(seed)
    a = rng.uniform(amp[0], amp[1], size=shape)
    p = rng.uniform(phase[0], phase[1], size=shape)
    a = gaussian_smooth_periodic(a, sigma)
    p = gaussian_smooth_periodic(p, sigma)
    return a * np.exp(1j * p)


def random_field_object(shape, seed, amp=(0.3, 1.0), phase=(-math.pi, math.pi), sigma=2.0):
    """Two independent N(0,1) fields smoothed with sigma, min-max mapped to the
    amplitude and phase ranges."""
    rng = np.random.default_rng(seed)
    fa = gaussian_smooth_periodic(rng.standard_normal(shape), sigma)
    fp = gaussian_smooth_periodic(rng.standard_normal(shape), sigma)

    def remap(f, lo, hi):
        return lo + (hi - lo) * (f - f.min()) / (f.max() - f.min())

    return remap(fa, *amp) * np.exp(1j * remap(fp, *phase))


def noll_to_nm(j: int):
    """Noll index -> (n, m) (R. J. Noll, JOSA 66, 207 (1976))."""
    n = 0
    while (n + 1) * (n + 2) // 2 < j:
        n += 1
    k = j - n * (n + 1) // 2 - 1  # 0-based position within order n
    ms = []
    for m in range(n % 2, n + 1, 2):
        if m == 0:
            ms.append(0)
        else:
            ms.extend([m, m])
    m = ms[k]
    if m != 0:
        m = m if j % 2 == 0 else -m
    return n, m


def zernike(j: int, rho: np.ndarray, theta: np.ndarray) -> np.ndarray:
    """Orthonormal Noll Zernike Z_j on the unit disk."""
    n, m = noll_to_nm(j)
    am = abs(m)
    r = np.zeros_like(rho)
    for s in range((n - am) // 2 + 1):
        c = (-1) ** s * math.factorial(n - s) / (
            math.factorial(s) * math.factorial((n + am) // 2 - s) * math.factorial((n - am) // 2 - s))
        r += c * rho ** (n - 2 * s)
    if m == 0:
        return math.sqrt(n + 1) * r
    norm = math.sqrt(2 * (n + 1))
    return norm * r * (np.cos(am * theta) if m > 0 else np.sin(am * theta))


def zernike_pupil(model, rms_rad: float, seed: int, noll=range(4, 16)) -> np.ndarray:
    """support * exp(i phi), phi = sum_j c_j Z_j with c ~ N(0,1) rescaled so
    that RMS(phi over the support) = rms_rad; rho = |k|/k_cut on the model's
    centred low-resolution grid."""
    h, w = model.image_shape
    support = model.support.astype(bool)
    rows = np.arange(h) - h // 2
    cols = np.arange(w) - w // 2
    yy, xx = np.meshgrid(rows, cols, indexing="ij")
    # k_cut in grid pixels: the largest supported radius bounds it from below;
    # the circular support is |k| <= k_cut, sampled on the pixel grid.
    rad = np.hypot(yy, xx)
    kcut = rad[support].max() if support.any() else 1.0
    rho = np.clip(rad / kcut, 0.0, 1.0)
    theta = np.arctan2(yy, xx)
    rng = np.random.default_rng(seed)
    coeffs = rng.standard_normal(len(list(noll)))
    phi = np.zeros((h, w))
    for c, j in zip(coeffs, noll):
        phi += c * zernike(j, rho, theta)
    rms = math.sqrt(np.mean(phi[support] ** 2)) if support.any() else 1.0
    if rms > 0:
        phi *= rms_rad / rms
    return np.where(support, np.exp(1j * phi), 0.0).astype(np.complex128)


def object_spectrum(obj: np.ndarray) -> np.ndarray:
    """Centred spectrum with fpm-rs's forward normalization (x 1/M;
    src/simulation/simulator.rs:209-226)."""
    return np.fft.fftshift(np.fft.fft2(obj) / obj.size)


def poisson_counts(intensity: np.ndarray, peak_photons: float, seed: int):
    """CameraModel{photons_per_pixel = peak / max I, gain 1, offset 0, shot
    noise, 16 bits, quantize} (src/simulation/camera.rs:249-318): returns
    (uint16 counts, photons_per_pixel). Counts are clamped to 65535."""
    ppp = peak_photons / float(intensity.max())
    rng = np.random.default_rng(seed)
    counts = rng.poisson(np.maximum(intensity, 0.0) * ppp)
    counts = np.clip(counts, 0, 65535).astype(np.uint16)
    return counts, ppp
