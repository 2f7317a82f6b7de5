"""Analysis/synthesis windows and per-frequency sigma locations.

Restates the reference parameter helpers (python/helpers.py:8-109) for Python 3.10:
the twelve window shapes, the per-residue normalisation of the spatial window,
the separable 3-D window and the frequency-location functions used to expand a
sigma curve into the [T][16][9] sigma array.
"""

from __future__ import annotations

import itertools
import math
from typing import Callable, Dict, List, Sequence, Tuple

__all__ = ["get_location", "get_sigma", "get_window", "get_window_value", "normalize", "i0_series"]

# Cosine-sum window coefficients by window id (python/helpers.py:8-25).
COSINE_WINDOWS: Dict[int, Tuple[float, ...]] = {
    0: (0.5, -0.5),  # Hanning
    1: (0.53836, -0.46164),  # Hamming
    2: (0.42, -0.5, 0.08),  # Blackman
    3: (0.35875, -0.48829, 0.14128, -0.01168),  # 4-term Blackman-Harris
    5: (
        0.27105140069342415,
        -0.433297939234486060,
        0.218122999543110620,
        -0.065925446388030898,
        0.010811742098372268,
        -7.7658482522509342e-4,
        1.3887217350903198e-5,
    ),  # 7-term Blackman-Harris
    6: (0.2810639, -0.5208972, 0.1980399),  # flat top
    10: (0.355768, -0.487396, 0.144232, -0.012604),  # Nuttall
    11: (0.3635819, -0.4891775, 0.1365995, -0.0106411),  # Blackman-Nuttall
}

WINDOW_NAMES = {
    0: "hanning", 1: "hamming", 2: "blackman", 3: "blackman-harris-4", 4: "kaiser-bessel",
    5: "blackman-harris-7", 6: "flat-top", 7: "rectangular", 8: "bartlett", 9: "bartlett-hann",
    10: "nuttall", 11: "blackman-nuttall",
}


def i0_series(p: float) -> float:
    """Modified Bessel I0 by its power series, truncated as the reference does.

    Follows python/helpers.py:28-40: the series sum_k ((p/2)^k / k!)^2 stops after the
    term with k = 14 or as soon as a term's root (p/2)^k/k! drops to 1e-8 or below.
    """
    half = p / 2.0
    numer = 1.0
    fact = 1.0
    total = 1.0
    k = 1
    while True:
        numer *= half
        fact *= k
        term_root = numer / fact
        total += term_root * term_root
        k += 1
        if k >= 15 or term_root <= 1e-8:
            return total


def normalize(window: Sequence[float], size: int, step: int) -> List[float]:
    """Scale a window so that the squares of every residue class mod `step` sum to 1.

    Follows python/helpers.py:45-47 (norm per residue q = index mod step).
    """
    norms = []
    for rem in range(step):
        norms.append(math.sqrt(sum(window[h] ** 2 for h in range(rem, size, step))))
    return [w / norms[i % step] for i, w in enumerate(window)]


def get_window_value(location: float, size: int, mode: int, beta: float) -> float:
    """Value of window `mode` at `location` in [0, size] (python/helpers.py:52-69)."""
    phase = math.pi * location / size
    coeffs = COSINE_WINDOWS.get(mode)
    if coeffs is not None:
        return sum(c * math.cos(2 * k * phase) for k, c in enumerate(coeffs))
    if mode == 4:  # Kaiser-Bessel
        v = 2 * location / size - 1
        return i0_series(math.pi * beta * math.sqrt(1 - v * v)) / i0_series(math.pi * beta)
    if mode == 7:  # rectangular
        return 1.0
    if mode == 8:  # Bartlett
        return 1.0 - 2.0 * abs(location - size / 2) / size
    if mode == 9:  # Bartlett-Hann, kept without abs() on the linear term as the reference has it
        return 0.62 - 0.48 * (location / size - 0.5) - 0.38 * math.cos(2 * phase)
    raise ValueError(f"unknown window: {mode}")


def get_window(
    radius: int,
    block_size: int,
    block_step: int,
    spatial_window_mode: int,
    spatial_beta: float,
    temporal_window_mode: int,
    temporal_beta: float,
) -> List[float]:
    """Flattened [T][B][B] analysis window (python/helpers.py:74-94).

    w[t][y][x] = wt[t] * ws[y] * ws[x] / (sqrt(T) * B), with the spatial window ws
    sampled at i + 0.5 and normalised per residue class mod block_step.
    """
    t_len = 2 * radius + 1
    wt = [get_window_value(i + 0.5, t_len, temporal_window_mode, temporal_beta) for i in range(t_len)]
    ws = normalize(
        [get_window_value(i + 0.5, block_size, spatial_window_mode, spatial_beta) for i in range(block_size)],
        block_size,
        block_step,
    )
    scale = 1.0 / (math.sqrt(t_len) * block_size)
    plane = [a * b * scale for a, b in itertools.product(ws, ws)]
    return [t * s for t, s in itertools.product(wt, plane)]


def get_location(position: float, length: int) -> float:
    """Normalised frequency location of bin `position` of `length` (python/helpers.py:99-103)."""
    if length == 1:
        return 0.0
    half = length // 2
    return (length - position if position > half else position) / half


def get_sigma(position: float, length: int, func: Callable[[float], float]) -> float:
    """Sigma factor of one axis (python/helpers.py:106-109)."""
    return 1.0 if length == 1 else func(get_location(position, length))
