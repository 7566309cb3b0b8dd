"""Seeded synthetic Lyman-alpha skewer data.

A Gaussian random field with covariance
``var_s * exp(-d_perp^2/(2 l_perp^2) - dz^2/(2 l_para^2))`` is drawn by real
FFT on a periodic grid, sampled by periodic trilinear interpolation along
skewers at uniform-random (x, y) with cell-centred pixels in z, and observed
with per-pixel Gaussian noise.  The output is an array of the reference's
``Pixel`` records ``{x, y, z, n, d}`` (lib/pixel.h:16-22), the layout
``read_pixel_data`` reads (lib/util.cc:11-23, README.md:32-36).

The named workloads below are the five benchmark configurations (sizes,
boxes, noise models); their seeds and unspecified details (map grids for the
larger cases, the FFT cell size) are fixed here.
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np

try:  # scipy's FFT is multi-threaded; numpy's is the fallback.
    import scipy.fft as _fft

    def _rfftn(a):
        return _fft.rfftn(a, workers=-1)

    def _irfftn(a, s):
        return _fft.irfftn(a, s=s, workers=-1)
except ImportError:  # pragma: no cover
    def _rfftn(a):
        return np.fft.rfftn(a)

    def _irfftn(a, s):
        return np.fft.irfftn(a, s=s)


@dataclasses.dataclass
class Workload:
    """One synthetic reconstruction problem."""
    name: str
    box: tuple            # (Lx, Ly, Lz) in Mpc/h
    n_skewers: int
    n_per_skewer: int
    var_s: float = 0.05
    l_perp: float = 2.5
    l_para: float = 2.1
    noise: str = "uniform"       # "uniform" U[0.1,0.3] | "lognormal"
    grid_cell: float = 0.5       # FFT cell size (Mpc/h), <= l/4
    map_shape: tuple = (16, 16, 64)
    tol: float = 1e-6
    max_iter: int = 1000
    step_r: int = 50
    seed: int = 1

    @property
    def n(self) -> int:
        return self.n_skewers * self.n_per_skewer


def workload(name: str) -> Workload:
    """The five benchmark configurations (plus tiny ones for tests)."""
    if name == "c1":   # 8,192 px, the CPU-runnable case
        return Workload("c1", (16.0, 16.0, 64.0), 64, 128, grid_cell=0.25,
                        map_shape=(16, 16, 64), tol=1e-5, max_iter=500, seed=1)
    if name == "c2":   # 76,800 px survey chunk
        return Workload("c2", (30.0, 24.0, 440.0), 240, 320, noise="lognormal",
                        map_shape=(60, 48, 880), tol=1e-6, seed=2)
    if name == "c3":   # 100,000 px, 32 noise realisations
        return Workload("c3", (40.0, 40.0, 500.0), 400, 250,
                        map_shape=(40, 40, 500), tol=1e-6, seed=3)
    if name == "c5":   # 300,000 px, learned warm start; seeds 1000-1199
        # train the network, 2000-2019 are held out (2000 is the benchmark)
        return Workload("c5", (64.0, 64.0, 600.0), 1000, 300,
                        map_shape=(64, 64, 600), tol=1e-6, seed=2000)
    if name == "tiny":
        return Workload("tiny", (8.0, 8.0, 16.0), 12, 24, grid_cell=0.25,
                        map_shape=(4, 4, 8), tol=1e-8, seed=7)
    if name == "small":
        return Workload("small", (16.0, 16.0, 64.0), 40, 50, grid_cell=0.25,
                        map_shape=(8, 8, 32), tol=1e-6, seed=11)
    raise KeyError(name)


def power_spectrum(kx, ky, kz, var_s, l_perp, l_para):
    """Fourier transform of the anisotropic Gaussian covariance."""
    amp = var_s * (2.0 * math.pi) ** 1.5 * l_perp * l_perp * l_para
    return amp * np.exp(-0.5 * ((kx * kx + ky * ky) * l_perp ** 2
                                + kz * kz * l_para ** 2))


def gaussian_field(rng, box, shape, var_s, l_perp, l_para):
    """Periodic Gaussian random field on a grid of `shape` over `box`."""
    nx, ny, nz = shape
    lx, ly, lz = box
    dv = (lx / nx) * (ly / ny) * (lz / nz)
    white = rng.standard_normal(shape, dtype=np.float64)
    wk = _rfftn(white)
    kx = 2 * np.pi * np.fft.fftfreq(nx, d=lx / nx)
    ky = 2 * np.pi * np.fft.fftfreq(ny, d=ly / ny)
    kz = 2 * np.pi * np.fft.rfftfreq(nz, d=lz / nz)
    pk = power_spectrum(kx[:, None, None], ky[None, :, None],
                        kz[None, None, :], var_s, l_perp, l_para)
    wk *= np.sqrt(pk / dv)
    return _irfftn(wk, s=shape)


def trilinear_periodic(field, box, pts):
    """Interpolate a periodic grid (nodes at i*L/n) at points (m, 3)."""
    shape = np.array(field.shape)
    u = pts / np.asarray(box)[None, :] * shape[None, :]
    i0 = np.floor(u).astype(np.int64)
    t = u - i0
    out = np.zeros(len(pts))
    for cx in (0, 1):
        wx = t[:, 0] if cx else 1.0 - t[:, 0]
        ix = (i0[:, 0] + cx) % shape[0]
        for cy in (0, 1):
            wy = t[:, 1] if cy else 1.0 - t[:, 1]
            iy = (i0[:, 1] + cy) % shape[1]
            for cz in (0, 1):
                wz = t[:, 2] if cz else 1.0 - t[:, 2]
                iz = (i0[:, 2] + cz) % shape[2]
                out += wx * wy * wz * field[ix, iy, iz]
    return out


def _grid_shape(box, cell):
    """Grid with cells no larger than `cell`, sized for a fast FFT."""
    try:
        from scipy.fft import next_fast_len
    except ImportError:  # pragma: no cover
        def next_fast_len(m, real=True):
            return m
    return tuple(int(next_fast_len(max(8, int(math.ceil(b / cell - 1e-9))),
                                   real=True)) for b in box)


def skewer_geometry(rng, box, n_skewers, n_per_skewer):
    """Skewer (x, y) ~ U[0,L) and cell-centred z; pixel order is skewer-major."""
    lx, ly, lz = box
    xy = rng.uniform(0.0, 1.0, size=(n_skewers, 2)) * np.array([lx, ly])
    z = (np.arange(n_per_skewer) + 0.5) * (lz / n_per_skewer)
    pts = np.empty((n_skewers, n_per_skewer, 3))
    pts[:, :, 0] = xy[:, 0:1]
    pts[:, :, 1] = xy[:, 1:2]
    pts[:, :, 2] = z[None, :]
    return pts.reshape(-1, 3)


def noise_levels(rng, kind, n):
    if kind == "uniform":
        return rng.uniform(0.1, 0.3, size=n)
    if kind == "lognormal":
        return np.clip(rng.lognormal(math.log(0.2), 0.4, size=n), 0.05, 1.0)
    raise ValueError(kind)


def make_pixels(w: Workload, seed: int | None = None, n_real: int = 1):
    """Return Pixel records (n, 5) [x,y,z,n,d]; with n_real > 1 return
    (pixels of realisation 0, d of shape (n_real, n)) -- the noise amplitudes
    and the field are shared, the noise draws differ."""
    rng = np.random.default_rng(w.seed if seed is None else seed)
    shape = _grid_shape(w.box, w.grid_cell)
    field = gaussian_field(rng, w.box, shape, w.var_s, w.l_perp, w.l_para)
    pts = skewer_geometry(rng, w.box, w.n_skewers, w.n_per_skewer)
    sig = trilinear_periodic(field, w.box, pts)
    nl = noise_levels(rng, w.noise, len(pts))
    ds = np.stack([sig + nl * rng.standard_normal(len(pts))
                   for _ in range(n_real)])
    pix = np.empty((len(pts), 5))
    pix[:, 0:3] = pts
    pix[:, 3] = nl
    pix[:, 4] = ds[0]
    if n_real > 1:
        return pix, ds
    return pix


def batch_workloads(seed: int = 4, count: int = 512,
                    n_min: int = 2000, n_max: int = 40000):
    """Subvolumes of the batched case: pixel counts log-uniform in
    [n_min, n_max], skewer length and count drawn per subvolume, l_perp and
    l_para ~ U[1.5, 3.5], var_s 0.05, noise U[0.1, 0.3].  Skewers are spaced
    about 2 Mpc/h apart transversely and pixels 2 Mpc/h apart in z."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(count):
        n_target = math.exp(rng.uniform(math.log(n_min), math.log(n_max)))
        npz = int(rng.integers(100, 301))
        nsk = max(4, int(round(n_target / npz)))
        lt = 2.0 * math.sqrt(nsk)
        out.append(Workload(f"sub{k}", (lt, lt, 2.0 * npz), nsk, npz,
                            l_perp=float(rng.uniform(1.5, 3.5)),
                            l_para=float(rng.uniform(1.5, 3.5)),
                            grid_cell=0.5, map_shape=(4, 4, 16),
                            tol=1e-6, seed=int(rng.integers(1 << 30))))
    return out
