"""DFTTest argument translation: user arguments -> plugin-level filter parameters.

Restates the reference front end (python/dfttest2.py:24-360, 424-452) without
VapourSynth: `dfttest_params()` takes DFTTest()'s argument surface and returns the
arrays the plugin's DFTTest function receives (window, sigma, sigma2, pmin, pmax,
filter_type, radius, block_size, block_step, zero_mean, window_freq).
"""

from __future__ import annotations

import itertools
import math
from dataclasses import dataclass, field
from typing import Callable, List, Optional, Sequence, Union

import numpy as np

from .helpers import get_location, get_sigma, get_window

__all__ = ["FilterParams", "dfttest_params", "dfttest2_params", "rdft", "flatten", "to_func"]

LocationData = Union[Sequence[Sequence[float]], Sequence[float]]


@dataclass
class FilterParams:
    """Plugin-level parameters, in the layouts the reference plugin's DFTTest takes.

    window: T*B*B doubles, [t][y][x] (hiprtc_source/source.cpp:1025-1046 arg "window").
    sigma: 1 double (scalar) or T*B*(B/2+1) doubles, [t][ky][kx] (arg "sigma").
    window_freq: T*B*(B/2+1)*2 doubles, [t][ky][kx][re,im], or None without zero_mean.
    """

    radius: int
    block_size: int
    block_step: int
    filter_type: int
    zero_mean: bool
    window: np.ndarray
    sigma: np.ndarray
    sigma2: float
    pmin: float
    pmax: float
    window_freq: Optional[np.ndarray]
    planes: Optional[List[int]] = None
    extra: dict = field(default_factory=dict)

    @property
    def t_len(self) -> int:
        return 2 * self.radius + 1

    def sigma_array(self) -> np.ndarray:
        """Sigma expanded to [T][B][B/2+1] float64 (the CPU backend's expansion, python/dfttest2.py:364)."""
        shape = (self.t_len, self.block_size, self.block_size // 2 + 1)
        if self.sigma.size == 1:
            return np.full(shape, float(self.sigma[0]))
        return np.asarray(self.sigma, dtype=np.float64).reshape(shape)


def rdft(data: Sequence[float], shape: Sequence[int]) -> np.ndarray:
    """FP64 real DFT of `data` viewed as `shape`; returns interleaved [..., n/2+1, (re, im)].

    Same result as the plugin's RDFT (cpu_source/source.cpp:269-340): an unnormalised
    forward transform (exponent -2*pi*i), real along the last axis, complex along the
    others, output layout [d0][d1][d2/2+1] complex as consecutive (re, im) doubles.
    """
    shape = tuple(int(s) for s in shape)
    if len(shape) not in (1, 2, 3):
        raise ValueError('"shape" must be an array of ints with 1, 2 or 3 values')
    arr = np.asarray(data, dtype=np.float64)
    if arr.size != int(np.prod(shape)):
        raise ValueError("cannot reshape array")
    spec = np.fft.rfftn(arr.reshape(shape), axes=tuple(range(len(shape))))
    out = np.empty(spec.shape + (2,), dtype=np.float64)
    out[..., 0] = spec.real
    out[..., 1] = spec.imag
    return out.reshape(-1)


def flatten(data: Optional[LocationData]) -> Optional[List[float]]:
    """Flatten [(freq, sigma), ...] or [freq, sigma, ...] (python/dfttest2.py:424-434)."""
    if not data:
        return None
    if isinstance(data[0], (int, float)):
        return [float(v) for v in data]  # type: ignore[arg-type]
    return [float(v) for v in itertools.chain.from_iterable(data)]  # type: ignore[arg-type]


def to_func(data: Optional[Sequence[float]], norm: Callable[[float], float], sigma: float) -> Callable[[float], float]:
    """Piecewise-linear sigma(freq) over norm(sigma) values (python/dfttest2.py:437-452).

    Points are sorted by frequency; a query x takes the first segment whose right end
    satisfies x <= x1, and raises when x is beyond the last point.
    """
    if data is None:
        return lambda _x: norm(sigma)

    points = sorted(zip(data[::2], data[1::2]), key=lambda p: p[0])

    def func(x: float) -> float:
        for (x0, y0), (x1, y1) in zip(points, points[1:]):
            if x <= x1:
                weight = (x - x0) / (x1 - x0)
                return (1 - weight) * norm(y0) + weight * norm(y1)
        raise ValueError(f"Value {x} out of range")

    return func


def _filter_type(ftype: int, f0beta: float) -> int:
    # python/dfttest2.py:263-271
    if ftype == 0:
        if abs(f0beta - 1) < 0.00005:
            return 0
        if abs(f0beta - 0.5) < 0.0005:
            return 6
        return 5
    return ftype


def dfttest2_params(
    ftype: int = 0,
    sigma: Union[float, Callable[[float], float], Sequence[Callable[[float], float]]] = 8.0,
    sigma2: float = 8.0,
    pmin: float = 0.0,
    pmax: float = 500.0,
    sbsize: int = 16,
    sosize: int = 12,
    tbsize: int = 3,
    swin: int = 0,
    twin: int = 7,
    sbeta: float = 2.5,
    tbeta: float = 2.5,
    zmean: bool = True,
    f0beta: float = 1.0,
    ssystem: int = 0,
    planes: Optional[Union[int, Sequence[int]]] = None,
) -> FilterParams:
    """Translate DFTTest2()'s arguments to plugin parameters (python/dfttest2.py:240-360)."""
    filter_type = _filter_type(ftype, f0beta)
    radius = (tbsize - 1) // 2
    block_size = sbsize
    block_step = sbsize - sosize

    # python/dfttest2.py:283-287 holds the fused backends to radius 0..3 and block_size 16;
    # the plan (dft2_plan_create) takes those on the fused kernel and the rest, as the
    # reference's FFT-library backends do, on the generic block path, and checks its limits.
    if radius < 0 or block_size < 1 or block_step < 1 or block_step > block_size:
        raise ValueError("invalid tbsize, sbsize or sosize")

    t_len = 2 * radius + 1
    y_len = block_size
    x_len = block_size // 2 + 1

    if not isinstance(sigma, Sequence) and not callable(sigma):
        sigma_norm: Union[float, List[float]] = float(sigma)
    else:
        funcs = [sigma] if callable(sigma) else list(sigma)[:3]
        funcs += [funcs[-1]] * (3 - len(funcs))
        fx, fy, ft = funcs
        if ssystem == 0:
            st = [get_sigma(t, t_len, ft) for t in range(t_len)]
            sy = [get_sigma(y, y_len, fy) for y in range(y_len)]
            sx = [get_sigma(x, x_len, fx) for x in range(x_len)]
            sigma_norm = [a * b * c for a, b, c in itertools.product(st, sy, sx)]
        else:
            lt = [get_location(t, t_len) ** 2 for t in range(t_len)]
            ly = [get_location(y, y_len) ** 2 for y in range(y_len)]
            lx = [get_location(x, x_len) ** 2 for x in range(x_len)]
            inv_ndim = 1.0 / (3 if radius > 0 else 2)
            sigma_norm = [ft(math.sqrt((a + b + c) * inv_ndim)) for a, b, c in itertools.product(lt, ly, lx)]

    window = get_window(
        radius=radius,
        block_size=block_size,
        block_step=block_step,
        spatial_window_mode=swin,
        spatial_beta=sbeta,
        temporal_window_mode=twin,
        temporal_beta=tbeta,
    )
    wscale = math.fsum(w * w for w in window)

    if ftype < 2:
        if isinstance(sigma_norm, list):
            sigma_norm = [s * wscale for s in sigma_norm]
        else:
            sigma_norm *= wscale
        sigma2 *= wscale

    if filter_type == 5:
        pmin = f0beta  # used as the exponent, unscaled
    else:
        pmin *= wscale
    pmax *= wscale

    shape = (block_size, block_size) if radius == 0 else (t_len, block_size, block_size)
    window_freq = rdft([w * 255 for w in window], shape)

    if isinstance(planes, int):
        planes = [planes]

    return FilterParams(
        radius=radius,
        block_size=block_size,
        block_step=block_step,
        filter_type=filter_type,
        zero_mean=bool(zmean),
        window=np.asarray(window, dtype=np.float64),
        sigma=np.atleast_1d(np.asarray(sigma_norm, dtype=np.float64)),
        sigma2=float(sigma2),
        pmin=float(pmin),
        pmax=float(pmax),
        window_freq=window_freq if zmean else None,
        planes=list(planes) if planes is not None else None,
        extra={"wscale": wscale},
    )


def dfttest_params(
    ftype: int = 0,
    sigma: float = 8.0,
    sigma2: float = 8.0,
    pmin: float = 0.0,
    pmax: float = 500.0,
    sbsize: int = 16,
    smode: int = 1,
    sosize: int = 12,
    tbsize: int = 3,
    swin: int = 0,
    twin: int = 7,
    sbeta: float = 2.5,
    tbeta: float = 2.5,
    zmean: bool = True,
    f0beta: float = 1.0,
    nlocation: Optional[Sequence[int]] = None,
    alpha: Optional[float] = None,
    slocation: Optional[LocationData] = None,
    ssx: Optional[LocationData] = None,
    ssy: Optional[LocationData] = None,
    sst: Optional[LocationData] = None,
    ssystem: int = 0,
    planes: Optional[Union[int, Sequence[int]]] = None,
) -> FilterParams:
    """Validate DFTTest()'s arguments and build plugin parameters (python/dfttest2.py:167-237)."""
    if ftype < 0 or ftype > 4:
        raise ValueError("ftype must be 0, 1, 2, 3, or 4")
    if sbsize < 1:
        raise ValueError("sbsize must be greater than or equal to 1")
    if smode != 1:
        raise ValueError('"smode" must be 1')
    if sosize > sbsize // 2 and (sbsize % (sbsize - sosize) != 0):
        raise ValueError("spatial overlap greater than 50% requires that sbsize-sosize is a divisor of sbsize")
    if tbsize < 1:
        raise ValueError('"tbsize" must be at least 1')
    if not 0 <= swin <= 11:
        raise ValueError("swin must be between 0 and 11 (inclusive)")
    if not 0 <= twin <= 11:
        raise ValueError("twin must be between 0 and 11 (inclusive)")
    if nlocation is not None:
        raise ValueError('"nlocation" must be None')
    for name, loc in (("slocation", slocation), ("ssx", ssx), ("ssy", ssy), ("sst", sst)):
        if loc and len(loc) % 2 != 0:
            raise ValueError(f"number of elements in {name} must be a multiple of 2")
    if ssystem not in (0, 1):
        raise ValueError("ssystem must be 0 or 1")

    def norm(x: float) -> float:
        if slocation is not None and ssystem == 1:
            return x
        if tbsize == 1:
            return math.sqrt(x)
        return math.cbrt(x) if hasattr(math, "cbrt") else math.copysign(abs(x) ** (1.0 / 3.0), x)

    if slocation is not None:
        sigma_arg: Union[float, List[Callable[[float], float]]] = [to_func(flatten(slocation), norm, sigma)] * 3
    elif any(ss is not None for ss in (ssx, ssy, sst)):
        sigma_arg = [to_func(flatten(ss), norm, sigma) for ss in (ssx, ssy, sst)]
    else:
        sigma_arg = sigma

    return dfttest2_params(
        ftype=ftype,
        sigma=sigma_arg,
        sigma2=sigma2,
        pmin=pmin,
        pmax=pmax,
        sbsize=sbsize,
        sosize=sosize,
        tbsize=tbsize,
        swin=swin,
        twin=twin,
        sbeta=sbeta,
        tbeta=tbeta,
        zmean=zmean,
        f0beta=f0beta,
        ssystem=ssystem,
        planes=planes,
    )
