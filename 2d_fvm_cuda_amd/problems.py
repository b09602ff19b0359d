"""Initial-data library (host side, numpy).

Every builder returns the reference's boundary layout: a C-contiguous float64
array of shape [ny][nx][5] (row 0 = north edge, column 0 = west edge,
primitive (rho, p, u, v, 0) per cell, the outer two layers being ghost cells).

Reference builders restated here:
  test2D.cu:4-18    initialize_2Mach     (N x 4N)
  test2D.cu:20-31   initialize_jet       (N x 2N)
  test2D.cu:33-44   initialize_jet800    (N x 3N)
  test2D.cu:46-80   initialize_4shocks   (N x N, split 4N/5 columns, N/5 rows)
  test2D.cu:82-116  initialize_4cds      (N x N, split N/2)
  test2D.cu:118-133 initialize_Sedov1D   (N x N)
  wrappers.cuh:80-108 RGRP_2D_CUDA::setRPInitialData (split Nx/2, Ny/2)
Seeded recipes for the batched and adjoint workloads are project-defined
(they have no counterpart in the reference) and use splitmix64 so that the
C++ and Python sides can regenerate identical data.
"""
from __future__ import annotations

import math

import numpy as np

MASK64 = (1 << 64) - 1


def _state(rho, p, u, v):
    return np.array([rho, p, u, v, 0.0], dtype=np.float64)


def initialize_4shocks(N: int) -> np.ndarray:
    """test2D.cu:46-80: rows < N/5 north (NW | NE), columns < 4N/5 west."""
    U1 = _state(1.5, 1.5, 0.0, 0.0)
    U2 = _state(0.5323, 0.3, 1.206, 0.0)
    U3 = _state(0.138, 0.029, 1.206, 1.206)
    U4 = _state(0.5323, 0.3, 0.0, 1.206)
    U = np.empty((N, N, 5), dtype=np.float64)
    r, c = N // 5, 4 * N // 5
    U[:r, :c] = U2
    U[:r, c:] = U1
    U[r:, :c] = U3
    U[r:, c:] = U4
    return U


def set_rp_initial_data(U1, U2, U3, U4, Nx: int, Ny: int) -> np.ndarray:
    """wrappers.cuh:80-108: quadrant split at (Nx/2, Ny/2); U1 = NE, U2 = NW,
    U3 = SW, U4 = SE (row 0 is north)."""
    U = np.empty((Ny, Nx, 5), dtype=np.float64)
    cx, cy = Nx // 2, Ny // 2
    U[:cy, :cx] = np.asarray(U2, np.float64)
    U[cy:, :cx] = np.asarray(U3, np.float64)
    U[:cy, cx:] = np.asarray(U1, np.float64)
    U[cy:, cx:] = np.asarray(U4, np.float64)
    return U


def initialize_2Mach(N: int) -> np.ndarray:
    """test2D.cu:4-18: N rows x 4N columns, post-shock wedge on the west."""
    U = np.empty((N, 4 * N, 5), dtype=np.float64)
    U[:] = _state(1.4, 1.0, 0.0, 0.0)
    post = _state(8.0, 116.5, 4.125 * math.sqrt(3.0), -4.125)
    for i in range(N - 1, -1, -1):
        temp = float(N - i + 3) / math.sqrt(3.0)
        l = N // 6 + int(temp)
        U[i, :l] = post
    return U


def initialize_jet(N: int) -> np.ndarray:
    """test2D.cu:20-31: N rows x 2N columns, jet inflow on column 0."""
    U = np.empty((N, 2 * N, 5), dtype=np.float64)
    U[:] = _state(0.5, 0.4127, 0.0, 0.0)
    for i in range(N):
        if (i >= 2 * N // 5 + 1) and (i <= 3 * N // 5):
            U[i, 0] = _state(5.0, 0.4127, 800.0, 0.0)
    return U


def initialize_jet800(N: int) -> np.ndarray:
    """test2D.cu:33-44: N rows x 3N columns."""
    U = np.empty((N, 3 * N, 5), dtype=np.float64)
    U[:] = _state(0.14, 1.0, 0.0, 0.0)
    for i in range(N):
        if i > 9 * N // 10:
            U[i, 0] = _state(1.4, 1.0, 800.0, 0.0)
    return U


def initialize_4cds(N: int) -> np.ndarray:
    """test2D.cu:82-116: four contact discontinuities, split at N/2."""
    U1 = _state(1.0, 1.0, 0.75, -0.5)
    U2 = _state(2.0, 1.0, 0.75, 0.5)
    U3 = _state(1.0, 1.0, -0.75, 0.5)
    U4 = _state(3.0, 1.0, -0.75, -0.5)
    U = np.empty((N, N, 5), dtype=np.float64)
    h = N // 2
    U[:h, :h] = U2
    U[:h, h:] = U1
    U[h:, :h] = U3
    U[h:, h:] = U4
    return U


def initialize_Sedov1D(N: int) -> np.ndarray:
    """test2D.cu:118-133: every row is the same 1-D Sedov profile."""
    E0 = 3200000.0
    domain_size = 4.0
    dx = 2 * domain_size / N
    row = np.empty((N, 5), dtype=np.float64)
    row[:] = _state(1.0, 0.4 * 1e-12, 0.0, 0.0)
    row[N // 2] = _state(1.0, 0.4 * E0 / dx, 0.0, 0.0)
    row[N // 2 - 1] = _state(1.0, 0.4 * E0 / dx, 0.0, 0.0)
    return np.ascontiguousarray(np.broadcast_to(row, (N, N, 5)))


# ---------------------------------------------------------------- seeded recipes

class SplitMix64:
    """splitmix64 stream; ``uniform()`` returns 53-bit doubles in [0, 1)."""

    def __init__(self, seed: int):
        self.s = seed & MASK64

    def next(self) -> int:
        self.s = (self.s + 0x9E3779B97F4A7C15) & MASK64
        z = self.s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
        return z ^ (z >> 31)

    def uniform(self, lo: float = 0.0, hi: float = 1.0) -> float:
        return lo + (hi - lo) * ((self.next() >> 11) * (1.0 / 9007199254740992.0))

    def randint(self, lo: int, hi: int) -> int:
        """Uniform integer in [lo, hi] (inclusive)."""
        return lo + int(self.next() % (hi - lo + 1))


def random_four_quadrant(rng: SplitMix64, nx: int, ny: int) -> np.ndarray:
    """Four-quadrant Riemann data with the split point ~U[0.3,0.7]^2 and
    rho, p ~ U[0.1, 2], u, v ~ U[-1, 1] per quadrant (batched workload)."""
    sx = rng.uniform(0.3, 0.7)
    sy = rng.uniform(0.3, 0.7)
    states = []
    for _ in range(4):
        rho = rng.uniform(0.1, 2.0)
        p = rng.uniform(0.1, 2.0)
        u = rng.uniform(-1.0, 1.0)
        v = rng.uniform(-1.0, 1.0)
        states.append(_state(rho, p, u, v))
    col = int(math.floor(sx * nx))
    row = int(math.floor((1.0 - sy) * ny))  # rows above the split are north
    U = np.empty((ny, nx, 5), dtype=np.float64)
    U[:row, col:] = states[0]   # NE
    U[:row, :col] = states[1]   # NW
    U[row:, :col] = states[2]   # SW
    U[row:, col:] = states[3]   # SE
    return U


def batched_instances(count: int = 64, seed: int = 2024, lo: int = 128, hi: int = 1024):
    """The batched workload: ``count`` meshes with sides ~U{lo..hi} drawn
    independently (non-square allowed), each with its own four-quadrant data."""
    rng = SplitMix64(seed)
    out = []
    for _ in range(count):
        nx = rng.randint(lo, hi)
        ny = rng.randint(lo, hi)
        out.append(random_four_quadrant(rng, nx, ny))
    return out


def fourier_perturbed(nx: int, ny: int, seed: int, modes: int = 16, kmax: int = 8,
                      amp: float = 0.05, disc_radius: float = 0.15, disc_factor: float = 2.0,
                      u0: float = 0.1, v0: float = 0.1) -> np.ndarray:
    """Adjoint workload: rho = 1 + sum of ``modes`` seeded Fourier modes,
    p = 1 + its own independent modes (k_x, k_y ~ U{1..kmax}, a ~ U[0, amp],
    phase ~ U[0, 2 pi)), p multiplied by ``disc_factor`` inside the centred
    disc, u = u0, v = v0.  Cell centres x = (i + 1/2)/nx, y = (j + 1/2)/ny on
    [0,1]^2 with y measured from the north edge."""
    rng = SplitMix64(seed)
    x = (np.arange(nx) + 0.5) / nx
    y = (np.arange(ny) + 0.5) / ny
    X, Y = np.meshgrid(x, y)

    def field():
        f = np.ones((ny, nx))
        for _ in range(modes):
            kx = rng.randint(1, kmax)
            ky = rng.randint(1, kmax)
            a = rng.uniform(0.0, amp)
            ph = rng.uniform(0.0, 2.0 * math.pi)
            f += a * np.sin(2.0 * math.pi * (kx * X + ky * Y) + ph)
        return f

    rho = field()
    p = field()
    disc = (X - 0.5) ** 2 + (Y - 0.5) ** 2 < disc_radius ** 2
    p = np.where(disc, disc_factor * p, p)
    U = np.zeros((ny, nx, 5), dtype=np.float64)
    U[:, :, 0] = rho
    U[:, :, 1] = p
    U[:, :, 2] = u0
    U[:, :, 3] = v0
    return U
