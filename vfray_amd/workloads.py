"""Seeded synthetic workloads (meshes, transfer functions, views, request
streams) of the benchmark configurations.

  plain_cpu   16^3 grid x 6 = 24,576 tets, ortho view along a seeded
              direction (baked into the vertices, as the reference renders
              along -z), 512x512, REF compositing.
  plain_gpu   64^3 grid x 6 = 1,572,864 tets, one perspective view (60 deg,
              eye at distance 2.5), 1024x1024, front to back, alpha 0.99.
  batch64     128^3 grid x 6 = 12,582,912 tets, 64 views on the r = 2.5
              sphere, sides from {256, 384, 512, 768, 1024}.
  concave     96^3 grid with a graded radial map and 12 voids, 32 views 768^2.
  serve       8 meshes 32^3..160^3, 4 transfer functions each, Poisson
              request stream (see vfray_amd.serve).

All randomness comes from splitmix64 streams (the same generator the C++
mesh generator uses), so every machine builds the same inputs.
"""
from __future__ import annotations

import math

import numpy as np

from . import perspective_view, synth_grid, synth_rotate, synth_tf

SEED_BASE = 0x5EED0000
ROLE_VIEWS, ROLE_REQUESTS = 6, 7
SIDES_BATCH = (256, 384, 512, 768, 1024)
SIDES_SERVE = (256, 384, 512, 640, 768, 896, 1024)
MASK = (1 << 64) - 1


class Rng:
    """splitmix64 stream, seeded as the C++ generator seeds its roles."""

    def __init__(self, seed: int, role: int):
        self.s = (seed ^ ((0xD1B54A32D192ED03 * (role + 1)) & MASK)) & MASK

    def next(self) -> int:
        self.s = (self.s + 0x9E3779B97F4A7C15) & MASK
        z = self.s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK
        return z ^ (z >> 31)

    def uni(self) -> float:
        return (self.next() >> 11) * (1.0 / 9007199254740992.0)

    def direction(self) -> np.ndarray:
        z = 2.0 * self.uni() - 1.0
        phi = 2.0 * math.pi * self.uni()
        r = math.sqrt(max(0.0, 1.0 - z * z))
        return np.array([r * math.cos(phi), r * math.sin(phi), z])

    def choice(self, seq):
        return seq[min(len(seq) - 1, int(self.uni() * len(seq)))]


def seed_of(config_index: int) -> int:
    return SEED_BASE + config_index


def plain_cpu(n: int = 16, side: int = 512):
    """(xyzs rotated so the seeded direction maps to -z, tets, tf, side)."""
    seed = seed_of(0)
    x, t = synth_grid(n, seed, 0)
    d = Rng(seed, ROLE_VIEWS).direction()
    x = synth_rotate(x, d.astype(np.float32))
    return x, t, synth_tf(seed), side


def sphere_views(seed: int, count: int, sides, tf_id: int, radius: float = 2.5, fov: float = 60.0):
    r = Rng(seed, ROLE_VIEWS)
    views = []
    for _ in range(count):
        d = r.direction()
        side = r.choice(sides) if isinstance(sides, (tuple, list)) else sides
        views.append(perspective_view(side, side, tf_id, eye=radius * d, fov_y_deg=fov))
    return views


def plain_gpu(n: int = 64):
    seed = seed_of(1)
    x, t = synth_grid(n, seed, 0)
    return x, t, synth_tf(seed), seed


def batch64(n: int = 128):
    seed = seed_of(2)
    x, t = synth_grid(n, seed, 0)
    return x, t, synth_tf(seed), seed


def concave(n: int = 96):
    seed = seed_of(3)
    x, t = synth_grid(n, seed, 1)
    return x, t, synth_tf(seed), seed


SERVE_SIZES = (32, 48, 64, 80, 96, 112, 128, 160)


def serve_meshes(sizes=SERVE_SIZES):
    """The 8 resident meshes of the serving workload and 4 TFs each."""
    seed = seed_of(4)
    out = []
    for k, n in enumerate(sizes):
        x, t = synth_grid(n, seed + 0x100 * (k + 1), 0)
        tfs = [synth_tf(seed + 0x100 * (k + 1) + j) for j in range(4)]
        out.append((x, t, tfs))
    return out


def serve_requests(n: int, n_meshes: int, tf_ids, sides=SIDES_SERVE, radius: float = 2.5):
    """n requests: uniform mesh, one of its 4 TFs, eye on the r = 2.5 sphere,
    side from `sides`.  tf_ids[m] lists mesh m's transfer-function ids."""
    from . import Request
    r = Rng(seed_of(4), ROLE_REQUESTS)
    reqs = []
    for _ in range(n):
        m = min(n_meshes - 1, int(r.uni() * n_meshes))
        tf = tf_ids[m][min(3, int(r.uni() * 4))]
        d = r.direction()
        side = r.choice(sides)
        v = perspective_view(side, side, tf, eye=radius * d, fov_y_deg=60.0)
        reqs.append(Request(m, 0, v))
    return reqs


def poisson_offsets(n: int, rate_per_s: float) -> "np.ndarray":
    """Arrival offsets (s) of a Poisson stream at `rate_per_s` (seeded)."""
    r = Rng(seed_of(4), ROLE_REQUESTS + 1)
    t, out = 0.0, []
    for _ in range(n):
        t += -math.log(max(1e-300, 1.0 - r.uni())) / rate_per_s
        out.append(t)
    return np.array(out)
