"""Deterministic synthetic scenes for the FLIP / semi-Lagrangian workloads.

Particles are placed by a counter-based generator (Philox4x32-10, Salmon et al.,
"Parallel random numbers: as easy as 1, 2, 3", SC'11), keyed by the seed and
counted by (cell linear index, sub-sample), so any consumer (the HIP library, the
CPU checker, the C++ classes) is handed bit-identical float32 particle arrays.

Layouts follow the reference: particles are AoS ``(x, y, vx, vy)``
(include/MarkerParticleSet.h:50-56); grids are row-major ``i + j * nx``
(include/Grid.h:27-31), i.e. numpy arrays of shape ``(ny, nx)``.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

_M0 = np.uint64(0xD2511F53)
_M1 = np.uint64(0xCD9E8D57)
_W0 = np.uint32(0x9E3779B9)
_W1 = np.uint32(0xBB67AE85)
_MASK32 = np.uint64(0xFFFFFFFF)


def philox4x32_10(counter: np.ndarray, key: tuple[int, int]) -> np.ndarray:
    """Philox4x32 with 10 rounds. ``counter``: uint32 array (..., 4)."""
    c = [counter[..., q].astype(np.uint64) for q in range(4)]
    k0 = np.uint32(key[0] & 0xFFFFFFFF)
    k1 = np.uint32(key[1] & 0xFFFFFFFF)
    for _ in range(10):
        p0 = _M0 * c[0]
        p1 = _M1 * c[2]
        hi0, lo0 = p0 >> np.uint64(32), p0 & _MASK32
        hi1, lo1 = p1 >> np.uint64(32), p1 & _MASK32
        c = [hi1 ^ c[1] ^ np.uint64(k0), lo1, hi0 ^ c[3] ^ np.uint64(k1), lo0]
        k0 = np.uint32((int(k0) + int(_W0)) & 0xFFFFFFFF)
        k1 = np.uint32((int(k1) + int(_W1)) & 0xFFFFFFFF)
    return np.stack([x.astype(np.uint32) for x in c], axis=-1)


def uniform24(words: np.ndarray) -> np.ndarray:
    """Map uint32 words to float64 uniforms in [0, 1) with 24 random bits."""
    return (words >> np.uint32(8)).astype(np.float64) * (1.0 / (1 << 24))


def jittered_particles(cells_i: np.ndarray, cells_j: np.ndarray, nx: int, dx: float,
                       dy: float, seed: int, per_axis: int = 2) -> np.ndarray:
    """``per_axis**2`` stratified, uniformly jittered particles per listed cell.

    Order: cells as listed (callers list j outer, i inner), then sub-row b, then
    sub-column a, like the reference emitter's y-outer / x-inner lattice
    (src/FluidDomain.cpp:39-46).  x = (i + (a + U) / per_axis) * dx.
    """
    n_sub = per_axis * per_axis
    n = cells_i.size * n_sub
    lin = (cells_i.astype(np.int64) + cells_j.astype(np.int64) * nx)
    lin = np.repeat(lin, n_sub)
    sub = np.tile(np.arange(n_sub, dtype=np.int64), cells_i.size)
    ctr = np.zeros((n, 4), dtype=np.uint32)
    ctr[:, 0] = (lin & 0xFFFFFFFF).astype(np.uint32)
    ctr[:, 1] = sub.astype(np.uint32)
    ctr[:, 2] = (lin >> 32).astype(np.uint32)
    words = philox4x32_10(ctr, (seed, 0x5EED))
    ux, uy = uniform24(words[:, 0]), uniform24(words[:, 1])
    b = (sub // per_axis).astype(np.float64)
    a = (sub % per_axis).astype(np.float64)
    ci = np.repeat(cells_i.astype(np.float64), n_sub)
    cj = np.repeat(cells_j.astype(np.float64), n_sub)
    out = np.zeros((n, 4), dtype=np.float32)
    out[:, 0] = ((ci + (a + ux) / per_axis) * dx).astype(np.float32)
    out[:, 1] = ((cj + (b + uy) / per_axis) * dy).astype(np.float32)
    return out


@dataclass
class Scene:
    """A synthetic initial state in the reference's units."""
    nx: int
    ny: int
    length_x: float
    length_y: float
    density: float
    pic_ratio: float
    dt: float
    particles: np.ndarray            # (P, 4) float32 AoS
    u: np.ndarray | None = None      # (ny, nx) float32 front u, None = zeros
    v: np.ndarray | None = None
    solid_mask: np.ndarray | None = None  # (ny, nx) uint8, interior solids
    phi: np.ndarray | None = None    # (ny, nx) float32 level set (grid-only runs)
    name: str = ""


def dam_break(n: int, seed: int = 0, fx: float = 0.4, fy: float = 0.6,
              dt: float = 0.01, pic_ratio: float = 0.05) -> Scene:
    """Square ``n x n`` dam break, unit cells, 1-cell solid border.

    The fluid block starts at cell (1, 1) and spans ceil(fx*(n-2)) x ceil(fy*(n-2))
    cells (51 x 76 at n = 128, 409 x 614 at n = 1024) with 4 jittered particles per
    cell.  density = dt as in the reference's examples
    (examples/simple.cpp:22,55; FluidRendererGL.cpp:96,133); pic_ratio 0.05 is
    the 0.95-FLIP blend.
    """
    bw = math.ceil(fx * (n - 2) - 1e-9)
    bh = math.ceil(fy * (n - 2) - 1e-9)
    jj, ii = np.meshgrid(np.arange(1, 1 + bh), np.arange(1, 1 + bw), indexing="ij")
    parts = jittered_particles(ii.ravel(), jj.ravel(), n, 1.0, 1.0, seed)
    return Scene(n, n, float(n), float(n), dt, pic_ratio, dt, parts,
                 name=f"dam_break_{n}")


def dam_break_batch(count: int, lo: int = 64, hi: int = 512, seed: int = 0,
                    dt: float = 0.01, pic_ratio: float = 0.05) -> list[Scene]:
    """``count`` non-square dam breaks of independent random sizes ``lo..hi`` (unit
    cells, 1-cell solid border), for the packed batch of scenes.

    Scene s draws from its own generator (seeded by ``(seed, s)``): nx != ny; 1-4
    axis-aligned SOLID rectangles covering at most 15 % of the cells (redrawn until
    every liquid region touches air, so no pressure system is singular); a fluid
    rectangle at cell (1, 1) covering 20-60 % of the interior, minus the solids, with
    4 jittered particles per cell (particle seed ``seed * 100003 + s``).  All
    scenes share dt, density (= dt) and pic_ratio, as one batch must.
    """
    from scipy import ndimage
    out = []
    for s in range(count):
        rng = np.random.default_rng([seed, s])
        nx = ny = 0
        while nx == ny:
            nx, ny = (int(v) for v in rng.integers(lo, hi + 1, size=2))
        f = rng.uniform(0.2, 0.6)
        fx = rng.uniform(f, 1.0)
        bw = max(1, math.ceil(fx * (nx - 2) - 1e-9))
        bh = max(1, math.ceil(f / fx * (ny - 2) - 1e-9))
        fluid = np.zeros((ny, nx), bool)
        fluid[1:1 + bh, 1:1 + bw] = True
        while True:
            mask = np.zeros((ny, nx), bool)
            for _ in range(int(rng.integers(1, 5))):
                w = int(rng.integers(2, max(3, (nx - 2) // 4) + 1))
                h = int(rng.integers(2, max(3, (ny - 2) // 4) + 1))
                x0 = int(rng.integers(1, nx - 1 - w + 1))
                y0 = int(rng.integers(1, ny - 1 - h + 1))
                trial = mask.copy()
                trial[y0:y0 + h, x0:x0 + w] = True
                if trial.sum() <= 0.15 * nx * ny:
                    mask = trial
            liquid = fluid & ~mask
            air = np.zeros((ny, nx), bool)
            air[1:-1, 1:-1] = True
            air &= ~mask & ~liquid
            comp, n = ndimage.label(liquid)
            touch = ndimage.binary_dilation(air) & liquid
            if mask.any() and liquid.any() and len(np.unique(comp[touch])) == n:
                break
        jj, ii = np.nonzero(liquid)          # j outer, i inner
        parts = jittered_particles(ii, jj, nx, 1.0, 1.0, seed * 100003 + s)
        out.append(Scene(nx, ny, float(nx), float(ny), dt, pic_ratio, dt, parts,
                         solid_mask=mask.astype(np.uint8), name=f"dam_break_batch_{s}"))
    return out


def interior_mask(nx: int, ny: int) -> np.ndarray:
    m = np.zeros((ny, nx), dtype=bool)
    m[1:-1, 1:-1] = True
    return m


def fourier_streamfunction_velocity(n: int, seed: int = 0, modes: int = 64,
                                    kmin: float = 1.0, kmax: float = 32.0):
    """Discretely divergence-free MAC velocity from a node-centred streamfunction.

    psi at node (i, j) (world (i, j), unit cells) = sum_m A_m sin(2 pi k_m.x / L
    + phi_m), |k_m| in [kmin, kmax], A_m ~ |k_m|^(-5/3).  u(i, j) = psi(i, j+1) -
    psi(i, j) and v(i, j) = -(psi(i+1, j) - psi(i, j)) so that the discrete
    divergence of every cell vanishes identically.  psi is multiplied by the
    window sin(pi (x-1)/(n-2)) sin(pi (y-1)/(n-2)), which vanishes on the wall
    nodes x, y in {1, n-1}, so wall faces carry no flux.  Returns float32 (u, v).
    """
    rng = np.random.default_rng(seed)
    ang = rng.uniform(0, 2 * np.pi, modes)
    mag = np.exp(rng.uniform(np.log(kmin), np.log(kmax), modes))
    kx, ky = mag * np.cos(ang), mag * np.sin(ang)
    amp = mag ** (-5.0 / 3.0)
    ph = rng.uniform(0, 2 * np.pi, modes)
    L = float(n)
    y, x = np.meshgrid(np.arange(n + 1, dtype=np.float64),
                       np.arange(n + 1, dtype=np.float64), indexing="ij")
    psi = np.zeros_like(x)
    for m in range(modes):
        psi += amp[m] * np.sin(2 * np.pi * (kx[m] * x + ky[m] * y) / L + ph[m])
    wx = np.sin(np.pi * np.clip((x - 1.0) / (n - 2), 0.0, 1.0))
    wy = np.sin(np.pi * np.clip((y - 1.0) / (n - 2), 0.0, 1.0))
    psi *= wx * wy
    psi[:2, :] = 0
    psi[:, :2] = 0
    psi[n - 1:, :] = 0
    psi[:, n - 1:] = 0
    u = (psi[1:n + 1, :n] - psi[:n, :n])
    v = -(psi[:n, 1:n + 1] - psi[:n, :n])
    return u.astype(np.float32), v.astype(np.float32)


def grid_only(n: int, seed: int = 0) -> Scene:
    """Fully fluid grid-only scene: phi = -1 everywhere (border SOLID via labels),
    a divergence-free Fourier velocity and dt = 0.5 / max|u|."""
    u, v = fourier_streamfunction_velocity(n, seed)
    vmax = float(max(np.abs(u).max(), np.abs(v).max()))
    dt = 0.5 / vmax if vmax > 0 else 0.01
    phi = -np.ones((n, n), dtype=np.float32)
    return Scene(n, n, float(n), float(n), dt, 0.0, dt,
                 np.zeros((0, 4), np.float32), u=u, v=v, phi=phi,
                 name=f"grid_only_{n}")
