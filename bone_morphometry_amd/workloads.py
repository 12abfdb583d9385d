"""The benchmark and test workloads: seeded synthetic volumes with their ROIs.

plain_small   128^3, sigma 3, BV/TV 0.25, full-cube ROI, 10 um, seed 1
plain_512     512^3, sigma 4, BV/TV 0.15, z-cylinder ROI of diameter 460
              centred at (255.5, 255.5), 10.5 um, seed 2
large         1024 x 1024 x 1536, sigma (5, 5, 9), BV/TV 0.12, ellipsoid ROI
              with semi-axes (500, 480, 740) about the centre, 8 um, seed 3
batched       96 volumes, axes uniform in 64..384, sigma U[2, 6], BV/TV
              U[0.08, 0.35], cylinder ROI of diameter U[0.70, 0.95]*min(W, H),
              spacing U[6, 20] um, seed 4 (per-volume parameters from a numpy
              Generator seeded with 4)
"""
import dataclasses

import numpy as np

from . import synth


@dataclasses.dataclass
class Workload:
    name: str
    width: int
    height: int
    depth: int
    bone: np.ndarray  # u8 0/255, z-major
    roi: np.ndarray   # u8 0/1
    spacing: tuple

    @property
    def n(self):
        return self.width * self.height * self.depth


def plain_small(threads=0):
    w = h = d = 128
    bone = synth.volume(w, h, d, 3.0, 0.25, 1, threads)
    return Workload("plain_small_128", w, h, d, bone, synth.roi_full(w, h, d), (10.0,) * 3)


def plain_512(threads=0):
    w = h = d = 512
    bone = synth.volume(w, h, d, 4.0, 0.15, 2, threads)
    roi = synth.roi_cylinder(w, h, d, 460.0, 255.5, 255.5)
    return Workload("plain_512", w, h, d, bone, roi, (10.5,) * 3)


def large(threads=0):
    w, h, d = 1024, 1024, 1536
    bone = synth.volume(w, h, d, (5.0, 5.0, 9.0), 0.12, 3, threads)
    roi = synth.roi_ellipsoid(w, h, d, (500.0, 480.0, 740.0))
    return Workload("large_1024x1024x1536", w, h, d, bone, roi, (8.0,) * 3)


def batched_params(count=96, seed=4, lo=64, hi=384):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        w, h, d = (int(v) for v in rng.integers(lo, hi + 1, size=3))
        sigma = float(rng.uniform(2.0, 6.0))
        bvtv = float(rng.uniform(0.08, 0.35))
        frac = float(rng.uniform(0.70, 0.95))
        sp = float(rng.uniform(6.0, 20.0))
        sub_seed = int(rng.integers(1, 2**62))
        out.append(dict(width=w, height=h, depth=d, sigma=sigma, bv_tv=bvtv,
                        diameter=frac * min(w, h), spacing=sp, seed=sub_seed))
    return out


def make_volume(p, threads=0):
    w, h, d = p["width"], p["height"], p["depth"]
    bone = synth.volume(w, h, d, p["sigma"], p["bv_tv"], p["seed"], threads)
    roi = synth.roi_cylinder(w, h, d, p["diameter"])
    return Workload(f"vol_{w}x{h}x{d}", w, h, d, bone, roi, (p["spacing"],) * 3)


def batched(count=96, seed=4, threads=0, lo=64, hi=384):
    return [make_volume(p, threads) for p in batched_params(count, seed, lo, hi)]


@dataclasses.dataclass
class Request:
    width: int
    height: int
    depth: int
    bone: np.ndarray
    roi: np.ndarray
    spacing: tuple

    @property
    def n(self):
        return self.width * self.height * self.depth


def serving_arrivals(rate=150.0, duration=60.0, seed=5):
    """Poisson arrival times (seconds) at `rate` per second over `duration`."""
    rng = np.random.default_rng(seed)
    gaps = rng.exponential(1.0 / rate, size=int(rate * duration * 1.5) + 16)
    arrivals = np.cumsum(gaps)
    return arrivals[arrivals < duration]


def serving_pool_params(count, seed, median_side=192, sigma_ln=0.7, lo=64, hi=512,
                        jitter=0.2):
    """Parameters of `count` serving requests, each its own generated volume:
    voxel count lognormal (median median_side^3, log-sigma sigma_ln, clipped
    to [lo^3, hi^3]) with near-cubic axes (each jittered by U[-jitter,
    jitter], clipped to [lo, hi]); sigma U[2, 6], BV/TV U[0.08, 0.35], a
    centred z-cylinder ROI of diameter U[0.70, 0.95]*min(w, h), spacing
    U[6, 20] um, and a generator seed."""
    rng = np.random.default_rng(seed)
    mu = np.log(float(median_side) ** 3)
    out = []
    for _ in range(count):
        nvox = float(np.clip(np.exp(rng.normal(mu, sigma_ln)), float(lo) ** 3, float(hi) ** 3))
        side = nvox ** (1.0 / 3.0)
        w, h, d = (int(np.clip(round(side * (1.0 + rng.uniform(-jitter, jitter))), lo, hi))
                   for _ in range(3))
        sigma = float(rng.uniform(2.0, 6.0))
        bvtv = float(rng.uniform(0.08, 0.35))
        frac = float(rng.uniform(0.70, 0.95))
        sp = float(rng.uniform(6.0, 20.0))
        out.append(dict(width=w, height=h, depth=d, sigma=sigma, bv_tv=bvtv,
                        diameter=frac * min(w, h), spacing=sp,
                        seed=int(rng.integers(1, 2**62))))
    return out


def serving_pool(count, seed, threads=0, **kw):
    """The pool of serving_pool_params(count, seed, **kw) as Request volumes."""
    reqs = []
    for p in serving_pool_params(count, seed, **kw):
        wl = make_volume(p, threads)
        reqs.append(Request(wl.width, wl.height, wl.depth, wl.bone, wl.roi, wl.spacing))
    return reqs


def serving_stream(rate=150.0, duration=60.0, seed=5, pool=256, median_side=192,
                   sigma_ln=0.7, lo=64, hi=512, jitter=0.2, bases=4, threads=0):
    """The serving workload: Poisson arrival times at `rate` per second over
    `duration` seconds, and a pool of request volumes whose voxel counts are
    lognormal (median median_side^3, log-sigma sigma_ln, clipped to
    [lo^3, hi^3]) with near-cubic axes (each axis jittered by U[-jitter,
    jitter], clipped to [lo, hi]).  Arrival i carries pool[i % pool].  Each
    pool volume is a crop, at a random offset, of one of `bases` generated
    hi^3 volumes (sigma U[2, 6], BV/TV U[0.08, 0.35]), with a centred
    z-cylinder ROI of diameter U[0.70, 0.95]*min(w, h) and spacing U[6, 20] um.
    Returns (arrival times in seconds, pool of Request)."""
    rng = np.random.default_rng(seed)
    gaps = rng.exponential(1.0 / rate, size=int(rate * duration * 1.5) + 16)
    arrivals = np.cumsum(gaps)
    arrivals = arrivals[arrivals < duration]
    base = []
    for _ in range(bases):
        sigma = float(rng.uniform(2.0, 6.0))
        bvtv = float(rng.uniform(0.08, 0.35))
        v = synth.volume(hi, hi, hi, sigma, bvtv, int(rng.integers(1, 2**62)), threads)
        base.append(v.reshape(hi, hi, hi))
    reqs = []
    mu = np.log(float(median_side) ** 3)
    for _ in range(pool):
        nvox = float(np.clip(np.exp(rng.normal(mu, sigma_ln)), float(lo) ** 3, float(hi) ** 3))
        side = nvox ** (1.0 / 3.0)
        w, h, d = (int(np.clip(round(side * (1.0 + rng.uniform(-jitter, jitter))), lo, hi))
                   for _ in range(3))
        b = base[int(rng.integers(0, bases))]
        z0, y0, x0 = (int(rng.integers(0, hi - s + 1)) for s in (d, h, w))
        bone = np.ascontiguousarray(b[z0:z0 + d, y0:y0 + h, x0:x0 + w]).reshape(-1)
        roi = synth.roi_cylinder(w, h, d, float(rng.uniform(0.70, 0.95)) * min(w, h))
        sp = float(rng.uniform(6.0, 20.0))
        reqs.append(Request(w, h, d, bone, roi, (sp, sp, sp)))
    return arrivals, reqs
