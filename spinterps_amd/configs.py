"""Seeded synthetic workloads for the interpolation path.

The sizes follow the five workload definitions of this project:
  small   : 100 stations, 100 km x 100 km, 24 steps, 50 x 50 grid at 2 km
  daily   : 1000 stations, 500 km x 500 km, 365 steps, 500 x 500 grid at 1 km
  hourly  : 2000 stations, 500 km x 500 km, 8760 steps, 1000 x 1000 grid at
            500 m, one variogram per step
  stress  : ill-conditioned Gaussian steps with close station pairs
Coordinates are metres with a UTM-like origin (350000, 5300000).  Cell
centres follow the reference grid convention: x ascending, y descending,
row-major ravel (interp/prepare.py:193-234).

Per-step generators take (seed, step) so that any subset of steps can be
built without the others, which is how the large workloads are sampled.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

ORIGIN_X = 350000.0
ORIGIN_Y = 5300000.0


@dataclass
class Workload:
    name: str
    stn_x: np.ndarray          # [N]
    stn_y: np.ndarray          # [N]
    cell_x: np.ndarray         # [C]
    cell_y: np.ndarray         # [C]
    grid_shape: tuple          # (ny, nx)
    n_steps: int
    seed: int
    stn_drift: np.ndarray | None = None    # [N x m]
    cell_drift: np.ndarray | None = None   # [m x C]
    idw_exp: float = 2.0
    meta: dict = field(default_factory=dict)

    @property
    def n_stns(self):
        return self.stn_x.size

    @property
    def n_cells(self):
        return self.cell_x.size


def grid_cells(x0, y_top, cell, nx, ny):
    """Cell centres: x ascending, y descending, row-major (prepare.py:193)."""
    xs = x0 + cell / 2 + cell * np.arange(nx)
    ys = y_top - cell / 2 - cell * np.arange(ny)
    gx, gy = np.meshgrid(xs, ys)
    return gx.ravel().copy(), gy.ravel().copy()


def vg_string(parts):
    """Join (sill, name, range) into a reference-style variogram string."""
    return ' + '.join(f'{s!r} {n}({r!r})' for (s, n, r) in parts)


# ---------------------------------------------------------------------------
# small: 100 stations, 24 steps, IDW p=2 + OK '0.1 Nug + 0.9 Sph(25 km)'
# ---------------------------------------------------------------------------

def small(seed=20240601, n_cells_side=50, cell=2000.0):
    rng = np.random.default_rng(seed)
    N, T = 100, 24
    sx = ORIGIN_X + rng.uniform(0, 1e5, N)
    sy = ORIGIN_Y + rng.uniform(0, 1e5, N)
    vals = rng.gamma(0.8, 5.0, (T, N))
    nz = int(round(0.3 * N))
    for t in range(T):
        vals[t, rng.choice(N, nz, replace=False)] = 0.0
    cx, cy = grid_cells(ORIGIN_X, ORIGIN_Y + 1e5, cell, n_cells_side,
                        n_cells_side)
    w = Workload('small', sx, sy, cx, cy, (n_cells_side, n_cells_side), T,
                 seed)
    vgs = []
    for t in range(T):
        v = float(np.var(vals[t]))
        vgs.append(vg_string([(0.1 * v, 'Nug', 0.0),
                              (0.9 * v, 'Sph', 25000.0)]))
    w.meta['vals'] = vals
    w.meta['vgs'] = vgs
    return w


def small_steps(w: Workload, t0=0, t1=None):
    t1 = w.n_steps if t1 is None else t1
    return w.meta['vals'][t0:t1].copy(), list(w.meta['vgs'][t0:t1])


# ---------------------------------------------------------------------------
# daily: 1000 stations, 365 steps, 0-20 % masked per step, DEM drift
# ---------------------------------------------------------------------------

def dem(x, y, centres, heights, sigmas):
    """Sum of 2-D Gaussians, evaluated in a loop to bound memory."""
    out = np.zeros(x.shape, dtype=np.float64)
    for (cx, cy), h, s in zip(centres, heights, sigmas):
        out += h * np.exp(-(((x - cx) ** 2) + ((y - cy) ** 2)) / (2 * s * s))
    return out


def daily(seed=20240602, n_side=500, cell=1000.0, n_stns=1000,
          n_steps=365, extent=5e5):
    rng = np.random.default_rng(seed)
    sx = ORIGIN_X + rng.uniform(0, extent, n_stns)
    sy = ORIGIN_Y + rng.uniform(0, extent, n_stns)
    cx, cy = grid_cells(ORIGIN_X, ORIGIN_Y + extent, cell, n_side, n_side)
    centres = np.stack([ORIGIN_X + rng.uniform(0, extent, 20),
                        ORIGIN_Y + rng.uniform(0, extent, 20)], axis=1)
    heights = rng.uniform(0, 2500, 20)
    sigmas = rng.uniform(10e3, 80e3, 20)
    w = Workload('daily', sx, sy, cx, cy, (n_side, n_side), n_steps, seed)
    w.stn_drift = dem(sx, sy, centres, heights, sigmas)[:, None].copy()
    w.cell_drift = dem(cx, cy, centres, heights, sigmas)[None, :].copy()
    w.meta['vg'] = '0.05 Nug(0.0) + 0.95 Exp(40000.0)'
    return w


def daily_steps(w: Workload, t0=0, t1=None):
    """Values [T x N] (Gamma(0.8, 5), 30 % zeros, NaN mask of a uniform
    0-20 % of stations) and the shared variogram per step."""
    t1 = w.n_steps if t1 is None else t1
    N = w.n_stns
    vals = np.empty((t1 - t0, N))
    for k, t in enumerate(range(t0, t1)):
        r = np.random.default_rng([w.seed, 1, t])
        v = r.gamma(0.8, 5.0, N)
        v[r.choice(N, int(round(0.3 * N)), replace=False)] = 0.0
        nmask = int(round(r.uniform(0, 0.2) * N))
        v[r.choice(N, nmask, replace=False)] = np.nan
        vals[k] = v
    return vals, [w.meta['vg']] * (t1 - t0)


# ---------------------------------------------------------------------------
# hourly: 2000 stations, 8760 steps, one variogram per step
# ---------------------------------------------------------------------------

def hourly(seed=20240603, n_side=1000, cell=500.0, n_stns=2000,
           n_steps=8760, extent=5e5):
    rng = np.random.default_rng(seed)
    sx = ORIGIN_X + rng.uniform(0, extent, n_stns)
    sy = ORIGIN_Y + rng.uniform(0, extent, n_stns)
    cx, cy = grid_cells(ORIGIN_X, ORIGIN_Y + extent, cell, n_side, n_side)
    return Workload('hourly', sx, sy, cx, cy, (n_side, n_side), n_steps, seed)


def hourly_steps(w: Workload, t0=0, t1=None, models=('Sph', 'Exp', 'Gau'),
                 range_lo=5e3, range_hi=8e4, min_active=200):
    """Values lognormal(0, 1) on a random subset of U{min_active..N}
    stations (rest NaN); per step a variogram string
    '<f v> Nug(0.0) + <(1-f) v> <M>(<r>)' with v the step variance."""
    t1 = w.n_steps if t1 is None else t1
    N = w.n_stns
    vals = np.full((t1 - t0, N), np.nan)
    vgs = []
    for k, t in enumerate(range(t0, t1)):
        r = np.random.default_rng([w.seed, 2, t])
        n_act = int(r.integers(min(min_active, N), N + 1))
        act = r.choice(N, n_act, replace=False)
        vals[k, act] = r.lognormal(0.0, 1.0, n_act)
        m = models[int(r.integers(0, len(models)))]
        rg = float(r.uniform(range_lo, range_hi))
        f = float(r.uniform(0.0, 0.3))
        v = float(np.var(vals[k, act]))
        vgs.append(vg_string([(f * v, 'Nug', 0.0), ((1 - f) * v, m, rg)]))
    return vals, vgs


# ---------------------------------------------------------------------------
# stress: Gaussian model, nugget 1e-6, 10 % of stations in close pairs
# ---------------------------------------------------------------------------

def stress_stations(w: Workload, seed=20240604):
    """Replace 10 % of the stations by partners 1-10 m from another one."""
    rng = np.random.default_rng(seed)
    N = w.n_stns
    k = N // 10
    moved = rng.choice(N, k, replace=False)
    rest = np.setdiff1d(np.arange(N), moved)
    anchors = rng.choice(rest, k, replace=False)
    ang = rng.uniform(0, 2 * np.pi, k)
    dist = rng.uniform(1.0, 10.0, k)
    sx = w.stn_x.copy()
    sy = w.stn_y.copy()
    sx[moved] = sx[anchors] + dist * np.cos(ang)
    sy[moved] = sy[anchors] + dist * np.sin(ang)
    return sx, sy


def stress_steps(w: Workload, t0=0, t1=500, range_lo=5e3, range_hi=8e4):
    N = w.n_stns
    vals = np.empty((t1 - t0, N))
    vgs = []
    for k, t in enumerate(range(t0, t1)):
        r = np.random.default_rng([w.seed, 4, t])
        v = r.gamma(0.8, 5.0, N)
        v[r.choice(N, int(round(0.3 * N)), replace=False)] = 0.0
        vals[k] = v
        rg = float(r.uniform(range_lo, range_hi))
        vgs.append(vg_string([(1e-6, 'Nug', 0.0), (0.999999, 'Gau', rg)]))
    return vals, vgs
