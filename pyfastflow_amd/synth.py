"""Seeded synthetic DEMs for the benchmark workloads and tests.

The reference's scripts build terrains with third-party noise generators or
rasters (examples/exemple_raw_1.py:26-35); these are replaced by a spectral
fractal with Gaussian pits, generated in float64 on the host and cast to
float32 once, from numpy's PCG64 generator.

- spectral fractal: amplitude |k|^(-beta/2) (zero at k=0), uniform random
  phases, real part of the inverse FFT, rescaled linearly to [0, relief];
- tilt: z += slope * dx * (ny-1-row), so the south row is lowest;
- Gaussian pits: centre uniform, depth U[5,50] m, radius R ~ U[5,40] cells,
  z -= depth * exp(-rho^2 / (2 (R/2)^2)) for rho <= 2R (wrapped east-west
  when `periodic_ew`).
"""

from __future__ import annotations

import numpy as np


def spectral_fractal(ny: int, nx: int, beta: float, relief: float, rng: np.random.Generator) -> np.ndarray:
    ky = np.fft.fftfreq(ny)[:, None]
    kx = np.fft.fftfreq(nx)[None, :]
    k = np.sqrt(kx * kx + ky * ky)
    amp = np.zeros_like(k)
    nz = k > 0
    amp[nz] = k[nz] ** (-beta / 2.0)
    phase = rng.uniform(0.0, 2.0 * np.pi, size=(ny, nx))
    z = np.real(np.fft.ifft2(amp * np.exp(1j * phase)))
    z -= z.min()
    zmax = z.max()
    if zmax > 0:
        z *= relief / zmax
    return z


def add_pits(z: np.ndarray, n_pits: int, rng: np.random.Generator, periodic_ew: bool = False,
             depth=(5.0, 50.0), radius=(5.0, 40.0)) -> np.ndarray:
    ny, nx = z.shape
    for _ in range(int(n_pits)):
        cy = rng.uniform(0, ny)
        cx = rng.uniform(0, nx)
        d = rng.uniform(*depth)
        R = rng.uniform(*radius)
        s = R / 2.0
        r0, r1 = int(np.floor(cy - 2 * R)), int(np.ceil(cy + 2 * R)) + 1
        c0, c1 = int(np.floor(cx - 2 * R)), int(np.ceil(cx + 2 * R)) + 1
        rows = np.arange(max(r0, 0), min(r1, ny))
        cols = np.arange(c0, c1)
        if not periodic_ew:
            cols = cols[(cols >= 0) & (cols < nx)]
        if rows.size == 0 or cols.size == 0:
            continue
        dy = rows[:, None] + 0.5 - cy
        dxx = cols[None, :] + 0.5 - cx
        rho2 = dy * dy + dxx * dxx
        w = np.where(rho2 <= 4 * R * R, d * np.exp(-rho2 / (2 * s * s)), 0.0)
        cc = np.mod(cols, nx) if periodic_ew else cols
        z[np.ix_(rows, cc)] -= w
    return z


def dem(ny: int, nx: int, beta: float, relief: float, seed: int, n_pits: int = 0, tilt: float = 0.0, dx: float = 1.0,
        periodic_ew: bool = False) -> np.ndarray:
    """(ny, nx) float32 DEM."""
    rng = np.random.Generator(np.random.PCG64(seed))
    z = spectral_fractal(ny, nx, beta, relief, rng)
    if tilt:
        z += tilt * dx * (ny - 1 - np.arange(ny, dtype=np.float64))[:, None]
    if n_pits:
        add_pits(z, n_pits, rng, periodic_ew=periodic_ew)
    return z.astype(np.float32)


# ---- the benchmark workloads ------------------------------------------------

def plain_small(n: int = 512, seed: int = 101):
    """512^2, dx=1 m, beta=2.0, 0-500 m relief, 0.01 south tilt, open edges."""
    return dict(z=dem(n, n, 2.0, 500.0, seed, tilt=0.01, dx=1.0), nx=n, ny=n, dx=1.0, boundary_mode="normal")


def depressions(n: int = 4096, seed: int = 202, n_pits: int = 2000):
    """4096^2, dx=10 m, beta=2.0, 0-1000 m relief, 2000 Gaussian pits,
    periodic east-west; precipitation 1e-5 m/s."""
    return dict(z=dem(n, n, 2.0, 1000.0, seed, n_pits=n_pits, dx=10.0, periodic_ew=True), nx=n, ny=n, dx=10.0,
                boundary_mode="periodic_EW", precipitation=1e-5)


def graphflood_plain(n: int = 2048, seed: int = 303, n_pits: int = 500):
    """2048^2, dx=10 m, beta=2.2, 0-300 m relief, 500 pits, open edges;
    Manning n=0.033, rainfall 1e-5 m/s."""
    return dict(z=dem(n, n, 2.2, 300.0, seed, n_pits=n_pits, dx=10.0), nx=n, ny=n, dx=10.0, boundary_mode="normal",
                precipitation=1e-5, manning=0.033)


def ragged_batch(n_tiles: int = 256, seed: int = 404, lo: int = 128, hi: int = 1024):
    """Square tiles with sides log-uniform in [lo, hi]; per tile beta ~
    U[1.8,2.4], relief ~ U[100,1500] m, pits proportional to area."""
    rng = np.random.Generator(np.random.PCG64(seed))
    tiles = []
    for t in range(n_tiles):
        u = rng.uniform()
        s = int(round(lo * (hi / lo) ** u))
        beta = rng.uniform(1.8, 2.4)
        relief = rng.uniform(100.0, 1500.0)
        pits = int(round(2000 * s * s / 4096.0 ** 2))
        tseed = int(rng.integers(0, 2**31 - 1))
        tiles.append(dict(z=dem(s, s, beta, relief, tseed, n_pits=pits, dx=10.0), nx=s, ny=s, dx=10.0,
                          boundary_mode="normal"))
    return tiles
