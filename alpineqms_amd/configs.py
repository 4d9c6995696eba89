"""Seeded generators of the benchmark designs, written in the reference's own
input grammar (inputs.txt + material-ID geometry, as ``master.py:16-77``
produces them).

Sizes are stated in nodes (N); ``inputs.txt`` holds cells = N - 1.

* ``config1``: 32^3 nodes, 1 mm, one centred 12^3-cell block (mur 100,
  sigma 5.8e7) in a mur 1 / sigma 1e-3 background, B0 = 0.1 T along z,
  frame velocity -1 m/s along x (the block moves +x), 50 steps of 1e-5 s.
* ``config2``: 64^3 nodes, 0.5 mm, 8 random boxes, B0 0.05 T random
  direction, 100 steps of 5e-6 s, v = 0.
* ``config3``: 128^3 nodes, 0.25 mm, 24 random boxes plus a 40x40x8-cell
  slab (sigma 3.5e7, mur 1), 200 steps of 2e-6 s, frame velocity 5 m/s.
* ``config4_designs``: 64 designs with per-axis node counts from
  {24, 32, ..., 96}, 4-12 boxes each, |B0| in [0.01, 0.2] T, 50 steps.
* ``config5_stream``: Poisson request stream with N = 8*round(exp(U(ln 32,
  ln 160))/8) and U{20..200} steps.

``ramp=True`` turns the constant frame velocity v into the reference's
accelerating trajectory v0 = 0, a = v/(Nt*dt), n_a = Nt
(``QMS_3D.f90:134-151, 2331-2350``): with a constant velocity every step of
the reference reproduces the steady state, so the ramp is the variant whose
steps carry real work.
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import List, Optional

import numpy as np

from .qms_io import Inputs, Material

SIG_MAX = 6e7


@dataclass
class Design:
    name: str
    inputs: Inputs
    geom: np.ndarray          # (cz, cy, cx) float64 material IDs

    @property
    def nodes(self):
        return self.inputs.nodes

    @property
    def dof(self) -> int:
        nx, ny, nz = self.nodes
        return 4 * nx * ny * nz - ny * nz - nx * nz - nx * ny


def _inputs(cells, h, dt, steps, B0, v_axis_x=0.0, ramp=False, mats=None, save_all="yes", tt=1,
            slice_axis="z", slice_loc=None) -> Inputs:
    if ramp and v_axis_x != 0.0:
        trajx = (0.0, v_axis_x / (steps * dt), steps)
    else:
        trajx = (v_axis_x, 0.0, 0)
    hh = tuple(float(a) for a in h) if isinstance(h, (tuple, list)) else (h, h, h)
    return Inputs(cells=tuple(cells), h=hh, dt=dt, steps=steps, tt=tt, slice_axis=slice_axis,
                  slice_loc=slice_loc or (cells[2] // 2 + 1), save_all=save_all, B0=tuple(B0),
                  G=((0.0, 0.0, 0.0), (0.0, 0.0, 0.0), (0.0, 0.0, 0.0)),
                  traj=(trajx, (0.0, 0.0, 0), (0.0, 0.0, 0)), materials=mats or [], geom_name="geom.bin")


def _random_boxes(rng, cells, nbox, first_id=2, margin=3):
    cx, cy, cz = cells
    geom = np.ones((cz, cy, cx))
    mats = [Material(1, (1.0, 1.0, 1.0), (1e-3, 1e-3, 1e-3))]
    for b in range(nbox):
        e = rng.integers(6, 21, size=3)
        lim = np.array([cx, cy, cz]) - 2 * margin
        e = np.minimum(e, np.maximum(lim, 1))
        lo = [int(rng.integers(margin, [cx, cy, cz][a] - margin - e[a] + 1)) for a in range(3)]
        mur = float(10 ** rng.uniform(0.0, 3.0))
        sig = float(10 ** rng.uniform(5.0, np.log10(SIG_MAX)))
        mid = first_id + b
        mats.append(Material(mid, (mur, mur, mur), (sig, sig, sig)))
        geom[lo[2]:lo[2] + e[2], lo[1]:lo[1] + e[1], lo[0]:lo[0] + e[0]] = mid
    return geom, mats


def _unit(rng):
    u = rng.normal(size=3)
    return u / np.linalg.norm(u)


def config1(N: int = 32, steps: int = 50, ramp: bool = False) -> Design:
    c = N - 1
    geom = np.ones((c, c, c))
    b = 12 if N >= 20 else max(2, N // 3)
    lo = (c - b) // 2
    geom[lo:lo + b, lo:lo + b, lo:lo + b] = 2.0
    mats = [Material(1, (1.0, 1.0, 1.0), (1e-3, 1e-3, 1e-3)),
            Material(2, (100.0, 100.0, 100.0), (5.8e7, 5.8e7, 5.8e7))]
    inp = _inputs((c, c, c), 1e-3, 1e-5, steps, (0.0, 0.0, 0.1), v_axis_x=-1.0, ramp=ramp, mats=mats)
    return Design(f"config1_N{N}" + ("_ramp" if ramp else ""), inp, geom)


def config2(N: int = 64, steps: int = 100, seed: int = 1002, nbox: int = 8, ramp_v: float = 0.0) -> Design:
    rng = np.random.default_rng(seed)
    c = N - 1
    geom, mats = _random_boxes(rng, (c, c, c), nbox)
    B0 = 0.05 * _unit(rng)
    inp = _inputs((c, c, c), 5e-4, 5e-6, steps, B0, v_axis_x=ramp_v, ramp=ramp_v != 0.0, mats=mats)
    return Design(f"config2_N{N}" + (f"_ramp{ramp_v:g}" if ramp_v else ""), inp, geom)


def config3(N: int = 128, steps: int = 200, seed: int = 1003, ramp: bool = True, v: float = 5.0) -> Design:
    rng = np.random.default_rng(seed)
    c = N - 1
    geom, mats = _random_boxes(rng, (c, c, c), 24)
    B0 = 0.05 * _unit(rng)
    sid = len(mats) + 1
    mats.append(Material(sid, (1.0, 1.0, 1.0), (3.5e7, 3.5e7, 3.5e7)))
    sx, sy, sz = min(40, c - 6), min(40, c - 6), min(8, c - 6)
    x0, y0, z0 = (c - sx) // 2, (c - sy) // 2, (c - sz) // 2
    geom[z0:z0 + sz, y0:y0 + sy, x0:x0 + sx] = sid
    inp = _inputs((c, c, c), 2.5e-4 * 128 / N if N != 128 else 2.5e-4, 2e-6, steps, B0, v_axis_x=v, ramp=ramp,
                  mats=mats)
    return Design(f"config3_N{N}" + ("_ramp" if ramp else ""), inp, geom)


def config4_designs(count: int = 64, steps: int = 50, sizes=None, ramp_v: float = 0.0) -> List[Design]:
    """Config 4's designs.  ramp_v != 0: the dynamic variant, a frame velocity
    ramped from 0 to ramp_v m/s along x over the run (the reference's
    trajectory grammar, QMS_3D.f90:134-151, :2331-2350), so that every step
    carries a solve; 0 is the literal config (every step after the steady
    state reproduces it)."""
    sizes = sizes or list(range(24, 97, 8))
    out = []
    for d in range(count):
        rng = np.random.default_rng(4000 + d)
        nodes = [int(rng.choice(sizes)) for _ in range(3)]
        cells = tuple(n - 1 for n in nodes)
        nbox = int(rng.integers(4, 13))
        geom, mats = _random_boxes(rng, cells, nbox)
        B0 = float(rng.uniform(0.01, 0.2)) * _unit(rng)
        inp = _inputs(cells, 5e-4, 5e-6, steps, B0, v_axis_x=ramp_v, ramp=ramp_v != 0.0, mats=mats)
        out.append(Design(f"config4_d{d}" + (f"_ramp{ramp_v:g}" if ramp_v else ""), inp, geom))
    return out


@dataclass
class Request:
    arrival: float
    design: Design


def config5_stream(count: int, rate: float, seed: int = 5000, nmin: int = 32, nmax: int = 160,
                   smin: int = 20, smax: int = 200, ramp_v: float = 0.0) -> List[Request]:
    """Config 5's request stream (Poisson arrivals at `rate` per second).
    ramp_v != 0: the dynamic variant, each request's frame velocity ramped
    from 0 to ramp_v m/s along x over its steps (as config4_designs)."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for r in range(count):
        t += float(rng.exponential(1.0 / rate))
        N = int(8 * round(np.exp(rng.uniform(np.log(nmin), np.log(nmax))) / 8))
        N = min(max(N, nmin), nmax)
        steps = int(rng.integers(smin, smax + 1))
        c = N - 1
        geom, mats = _random_boxes(rng, (c, c, c), int(rng.integers(4, 13)))
        B0 = 0.05 * _unit(rng)
        inp = _inputs((c, c, c), 5e-4, 5e-6, steps, B0, v_axis_x=ramp_v, ramp=ramp_v != 0.0, mats=mats)
        out.append(Request(t, Design(f"config5_r{r}_N{N}" + (f"_ramp{ramp_v:g}" if ramp_v else ""), inp, geom)))
    return out


def small_box_design(N: int, seed: int, nbox: int = 3, h=5e-4, dt: float = 5e-6, steps: int = 4,
                     v: float = 0.0, ramp: bool = True, gradient: bool = False, cells=None,
                     aniso: bool = False) -> Design:
    """Small designs for parity tests: random boxes, optional frame velocity
    ramp and an optional traceless symmetric background gradient.  `cells`
    (x, y, z) overrides the cubic N-1 cells per axis; `h` may be a per-axis
    (dx, dy, dz); `aniso` draws each box's mu_r and sigma per axis."""
    rng = np.random.default_rng(seed)
    cz = tuple(cells) if cells is not None else (N - 1,) * 3
    geom = np.ones((cz[2], cz[1], cz[0]))
    mats = [Material(1, (1.0, 1.0, 1.0), (1e-3, 1e-3, 1e-3))]
    for b in range(nbox):
        if cells is None:
            e = rng.integers(2, max(3, cz[0] // 3), size=3)
        else:
            e = [int(rng.integers(2, max(3, cz[a] // 3))) for a in range(3)]
        lo = [int(rng.integers(3, max(4, cz[a] - 3 - e[a]))) for a in range(3)]
        mur = float(10 ** rng.uniform(0, 3))
        sig = float(10 ** rng.uniform(5, np.log10(SIG_MAX)))
        if aniso:
            murs = tuple(float(10 ** rng.uniform(0, 3)) for _ in range(3))
            sigs = tuple(float(10 ** rng.uniform(5, np.log10(SIG_MAX))) for _ in range(3))
            mats.append(Material(b + 2, murs, sigs))
        else:
            mats.append(Material(b + 2, (mur, mur, mur), (sig, sig, sig)))
        geom[lo[2]:lo[2] + e[2], lo[1]:lo[1] + e[1], lo[0]:lo[0] + e[0]] = b + 2
    B0 = 0.05 * _unit(rng)
    inp = _inputs(cz, h, dt, steps, B0, v_axis_x=v, ramp=ramp, mats=mats)
    if gradient:
        g = rng.normal(size=(3, 3)) * 0.5
        g = (g + g.T) / 2
        g -= np.eye(3) * np.trace(g) / 3
        inp.G = tuple(tuple(float(x) for x in row) for row in g)
    return Design(f"small_{'x'.join(str(c + 1) for c in cz)}_s{seed}", inp, geom)
