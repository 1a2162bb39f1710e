"""Synthetic initial conditions of the benchmark workloads (seeded, NumPy float64).

* ``tgv_primitives``: inviscid Taylor-Green vortex on [0, 2pi)^3
  (examples/examples_3D/01_tgv/tgv.json, pressure offset for Mach 0.1).
* ``four_quadrant_cases``: independent 2-D four-quadrant Riemann problems on [0,1]^2 with
  seeded grid sizes and quadrant states.
* ``kelvin_helmholtz_primitives``: periodic double shear layer with a seeded multi-mode
  velocity perturbation.
"""
from __future__ import annotations

from typing import List, NamedTuple, Sequence

import numpy as np


def tgv_primitives(n: int, gamma: float = 1.4, mach: float = 0.1) -> np.ndarray:
    """rho=1, u=sin x cos y cos z, v=-cos x sin y cos z, w=0,
    p = 1/(gamma Ma^2) + (1/16)(cos 2x + cos 2y)(cos 2z + 2) on cell centers of [0,2pi)^3."""
    L = 2 * np.pi
    d = L / n
    x = np.linspace(d / 2, L - d / 2, n)
    sx, cx, c2 = np.sin(x), np.cos(x), np.cos(2 * x)
    W = np.empty((5, n, n, n))
    W[0] = 1.0
    W[1] = sx[:, None, None] * cx[None, :, None] * cx[None, None, :]
    W[2] = -cx[:, None, None] * sx[None, :, None] * cx[None, None, :]
    W[3] = 0.0
    W[4] = 1.0 / gamma / mach ** 2 + 1.0 / 16.0 * ((c2[:, None, None] + c2[None, :, None]) * (c2[None, None, :] + 2))
    return W


class QuadrantCase(NamedTuple):
    nx: int
    ny: int
    states: np.ndarray  # (4, 4): quadrant (NE, NW, SW, SE) x (rho, u, v, p)


def four_quadrant_cases(ncases: int = 64, seed: int = 0,
                        sizes: Sequence[int] = (128, 192, 256, 384, 512)) -> List[QuadrantCase]:
    """Nx, Ny drawn independently and uniformly from `sizes`; per quadrant rho ~ U(0.1, 1.5),
    p ~ U(0.03, 1.5), u, v ~ U(-1, 1)."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(ncases):
        nx, ny = (int(v) for v in rng.choice(sizes, 2))
        st = np.empty((4, 4))
        st[:, 0] = rng.uniform(0.1, 1.5, 4)
        st[:, 1] = rng.uniform(-1.0, 1.0, 4)
        st[:, 2] = rng.uniform(-1.0, 1.0, 4)
        st[:, 3] = rng.uniform(0.03, 1.5, 4)
        out.append(QuadrantCase(nx, ny, st))
    return out


def quadrant_primitives(case: QuadrantCase) -> np.ndarray:
    """(5, nx, ny, 1) primitives on [0,1]^2 split at x = 0.5, y = 0.5."""
    nx, ny = case.nx, case.ny
    x = (np.arange(nx) + 0.5) / nx
    y = (np.arange(ny) + 0.5) / ny
    X, Y = np.meshgrid(x, y, indexing="ij")
    east, north = X > 0.5, Y > 0.5
    q = np.where(north & east, 0, np.where(north & ~east, 1, np.where(~north & ~east, 2, 3)))
    st = case.states
    W = np.zeros((5, nx, ny, 1))
    W[0, ..., 0] = st[q, 0]
    W[1, ..., 0] = st[q, 1]
    W[2, ..., 0] = st[q, 2]
    W[4, ..., 0] = st[q, 3]
    return W


def kelvin_helmholtz_primitives(n: int = 256, seed: int = 0, amplitude: float = 1e-2, modes: int = 8,
                                thickness: float = 0.02, du: float = 1.0, density_ratio: float = 2.0,
                                p0: float = 2.5) -> np.ndarray:
    """Periodic [0,1]^2: tanh shear layers at y = 0.25 and 0.75 (thickness 0.02, velocity jump du,
    density ratio 2, p = 2.5) plus a seeded velocity perturbation on modes k = 1..8."""
    rng = np.random.default_rng(seed)
    x = (np.arange(n) + 0.5) / n
    X, Y = np.meshgrid(x, x, indexing="ij")
    prof = 0.5 * (np.tanh((Y - 0.25) / thickness) - np.tanh((Y - 0.75) / thickness))  # 1 in the band
    rho = 1.0 + (density_ratio - 1.0) * prof
    u = du * (prof - 0.5)
    v = np.zeros_like(X)
    ph = rng.uniform(0, 2 * np.pi, (2, modes))
    for k in range(1, modes + 1):
        env = np.exp(-((Y - 0.25) / (4 * thickness)) ** 2) + np.exp(-((Y - 0.75) / (4 * thickness)) ** 2)
        v += amplitude / modes * np.sin(2 * np.pi * k * X + ph[0, k - 1]) * env
        u += amplitude / modes * np.cos(2 * np.pi * k * X + ph[1, k - 1]) * env
    W = np.zeros((5, n, n, 1))
    W[0, ..., 0] = rho
    W[1, ..., 0] = u
    W[2, ..., 0] = v
    W[4, ..., 0] = p0
    return W


def fourier_velocity_direction(cells, seed: int = 0, modes: int = 8, amplitude: float = 1e-2) -> np.ndarray:
    """A seeded (u, v) field on the periodic unit-spaced grid of `cells` made of the Fourier modes
    k = 1..modes along every active axis (random phases, amplitude/modes each); density, w and
    pressure components are zero.  The direction of the finite-difference checks of the adjoint
    workloads: it lies in the span of their initial perturbations."""
    rng = np.random.default_rng(seed)
    cells = tuple(int(c) for c in cells)
    axes = [(np.arange(n) + 0.5) / n for n in cells]
    X = np.meshgrid(*axes, indexing="ij")
    d = np.zeros((5,) + cells)
    for comp in (1, 2):
        for a, n in enumerate(cells):
            if n == 1:
                continue
            ph = rng.uniform(0, 2 * np.pi, modes)
            for k in range(1, modes + 1):
                d[comp] += amplitude / modes * np.cos(2 * np.pi * k * X[a] + ph[k - 1])
    return d
