"""Synthetic workloads of many ranges: the batched list and the serving stream.

The reference runs one range per process (main.cpp:313-405); these lists are
the multi-range workloads the batched engine and the request loop are
measured on. Both are fixed by their seeds (numpy's PCG64 default_rng), so
the same lists are produced everywhere.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

PLO, PHI = 1e9, 1e14


def batch_ranges(seed: int = 1234, n: int = 256, wmin: int = 100_000, wmax: int = 10_000_000):
    """n disjoint ranges [start, start + width): starts log-uniform in
    [1e9, 1e14], widths uniform in [wmin, wmax] rounded to even. Ranges are
    taken in order of start; one that would overlap its predecessor starts
    at the predecessor's end instead (same width), which keeps them disjoint.
    """
    rng = np.random.default_rng(seed)
    starts = np.exp(rng.uniform(math.log(PLO), math.log(PHI), size=n))
    widths = rng.uniform(wmin, wmax, size=n)
    raw = [(int(round(s)), 2 * int(round(w / 2))) for s, w in zip(starts, widths)]
    raw.sort()
    out = []
    end = 0
    for s, w in raw:
        s = max(s, end)
        out.append((s, s + w))
        end = s + w
    return out


@dataclass(frozen=True)
class Request:
    rid: int
    arrival: float  # seconds after the stream starts
    pmin: int
    pmax: int


def serving_requests(seed: int = 99, n: int = 2000, rate: float = 50.0, wmin: float = 1e4, wmax: float = 1e7):
    """n range workunits with Poisson arrivals (exponential gaps of mean
    1/rate s), widths log-uniform in [wmin, wmax] rounded to even, starts
    log-uniform in [1e9, 1e14]."""
    rng = np.random.default_rng(seed)
    gaps = rng.exponential(1.0 / rate, size=n)
    starts = np.exp(rng.uniform(math.log(PLO), math.log(PHI), size=n))
    widths = np.exp(rng.uniform(math.log(wmin), math.log(wmax), size=n))
    t = np.cumsum(gaps)
    return [Request(i, float(t[i]), int(round(starts[i])), int(round(starts[i])) + 2 * int(round(widths[i] / 2)))
            for i in range(n)]


def li(x: float) -> float:
    """Logarithmic integral Li(x) = li(x) - li(2) by Ramanujan's series."""
    g = 0.5772156649015329
    lx = math.log(x)
    s, term, k = 0.0, 1.0, 1
    inner = 0.0
    while True:
        term *= lx / k
        if (k - 1) % 2 == 0:
            inner += 1.0 / (2 * ((k - 1) // 2) + 1)
        add = term * ((-1) ** (k - 1)) / (2 ** (k - 1)) * inner
        s += add
        if abs(add) < 1e-17 * abs(s) and k > 10:
            break
        k += 1
    return g + math.log(lx) + math.sqrt(x) * s - 1.045163780117492784


def pi_approx(x: float) -> float:
    """pi(x) ~ Li(x) - Li(sqrt x)/2; relative error ~1e-7 at 1e12-1e13."""
    if x < 2:
        return 0.0
    return li(x) - 0.5 * li(math.sqrt(x))
