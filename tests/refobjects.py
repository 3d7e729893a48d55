"""Closed-form test objects and fixtures restated from the reference's tests
(src/simulation/synthetic_object.rs, src/simulation/presets.rs,
tests/common/mod.rs). Inputs only; no algorithm of the path lives here."""
import math
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (os.path.join(ROOT, "fpm-rs_amd"), os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)

from fpmhip.model import ImagePlaneModel, Optics, PlanarLedArray  # noqa: E402


def mixed_test_pattern(shape):
    """SyntheticObject::mixed_test_pattern (synthetic_object.rs:275-298)."""
    h, w = shape
    cy, cx = (h - 1) / 2.0, (w - 1) / 2.0
    scale = float(min(h, w))
    out = np.empty(shape, np.complex128)
    for r in range(h):
        for c in range(w):
            y = r - cy
            x = c - cx
            radius = math.hypot(x, y)
            amp = 0.45 if ((abs(x) < 0.12 * scale and abs(y) < 0.35 * scale) or radius < 0.12 * scale) else 1.0
            ph = 0.8 * math.exp(-radius * radius / (0.08 * scale * scale)) + 0.25 * math.sin(2 * math.pi * x / scale)
            out[r, c] = complex(amp * math.cos(ph), amp * math.sin(ph))
    return out


def resolution_target(shape):
    """SyntheticObject::resolution_target (synthetic_object.rs:202-226)."""
    h, w = shape
    v = np.ones(shape, np.complex128)
    groups = [2, 3, 4, 6, 8]
    for g, period in enumerate(groups):
        top = g * h // len(groups)
        bottom = (g + 1) * h // len(groups)
        for r in range(top, bottom):
            for c in range(w):
                dark = (c // period) % 2 == 0 if g % 2 == 0 else ((r - top) // period) % 2 == 0
                if dark:
                    v[r, c] = 0.2
    return v


def phase_disk(shape, radius_pixels, phase_shift):
    """SyntheticObject::phase_disk (synthetic_object.rs:154-176)."""
    h, w = shape
    cy, cx = (h - 1) / 2.0, (w - 1) / 2.0
    out = np.empty(shape, np.complex128)
    for r in range(h):
        for c in range(w):
            ph = phase_shift if math.hypot(r - cy, c - cx) <= radius_pixels else 0.0
            out[r, c] = complex(math.cos(ph), math.sin(ph))
    return out


def direct_model():
    """tests/common/mod.rs:49-81: 8x8 -> 16x16, unit pupil, all-ones support,
    five sources at integer shifts (col,row) (-3,0),(0,-3),(0,0),(0,3),(3,0)."""
    shifts = [(-3, 0), (0, -3), (0, 0), (0, 3), (3, 0)]
    crop = np.array([[4 + r, 4 + c] for c, r in shifts], np.uint64)
    kv = np.array([[c, r] for c, r in shifts], np.float64)
    return ImagePlaneModel((8, 8), (16, 16), crop, np.zeros((5, 2)), np.ones((8, 8), np.complex128),
                           np.ones((8, 8), np.uint8), k_vectors=kv)


def base_optics():
    """presets.rs:71-82."""
    return Optics(532e-9, 0.10, 4.0, 6.5e-6, 1.0, 1.0)


def base_array():
    """presets.rs:84-97: 3x3 LEDs, 4 mm pitch, reference (1,1), 90 mm below."""
    return PlanarLedArray((3, 3), (4e-3, 4e-3), (1.0, 1.0), (0.0, 0.0, -90e-3))


def spectrum_of(obj):
    """Simulator object_spectrum (simulator.rs:209-226) through the oracle FFT."""
    import oracle
    return oracle.shift(oracle.fft2(obj, "forward"))
