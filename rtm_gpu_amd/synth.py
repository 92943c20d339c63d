"""Seeded synthetic inputs for the benchmark workloads (no datasets are
available offline).  Models are returned column-major (x outer, z inner),
the on-disk layout the reference reads (GPU_velocity_real.cpp:12-18).

  layered_5     200 x 400 interior, 5 flat layers           (single-shot RTM, 10 m)
  faulted_rvsp  351 x 1701, 8 dipping layers, 3 normal faults, a 1200 m x 600 m
                high-velocity body                          (reverse VSP, 12.5 m)
  grf_4001      801 x 4001 Gaussian-random-field background + 6 reflectors (10 m)
  reflectors    256 x 1024 gradient + Gaussian-random-field background and
                10-40 dipping planar reflectors       (learned deblurring, 10 m)
"""
import numpy as np


def _sorted_vels(rng, n, lo, hi):
    return np.sort(rng.uniform(lo, hi, n)).astype(np.float32)


def layered_5(seed=1, mod_nz=200, mod_nx=400, vlo=1500.0, vhi=4500.0):
    rng = np.random.default_rng(seed)
    depths = np.sort(rng.choice(np.arange(10, mod_nz - 10), 4, replace=False))
    vels = _sorted_vels(rng, 5, vlo, vhi)
    z = np.arange(mod_nz)
    vz = vels[np.searchsorted(depths, z, side="right")]
    return np.ascontiguousarray(np.broadcast_to(vz[None, :], (mod_nx, mod_nz))).astype(np.float32)


def faulted_rvsp(seed=2, mod_nz=351, mod_nx=1701, h=12.5, vlo=1500.0, vhi=5500.0):
    """8 dipping layers, 3 normal faults (throw 20-80 m, 60 degree dip) and a
    96 x 48-cell (1200 m x 600 m) body at 4500 m/s."""
    rng = np.random.default_rng(seed)
    nlay = 8
    vels = _sorted_vels(rng, nlay, vlo, vhi)
    base = np.sort(rng.uniform(0.06, 0.95, nlay - 1)) * mod_nz  # interface depths at x = 0
    dips = np.deg2rad(rng.uniform(-4.0, 4.0, nlay - 1))
    x = np.arange(mod_nx)[:, None].astype(np.float64)
    z = np.arange(mod_nz)[None, :].astype(np.float64)
    shift = np.zeros((mod_nx, mod_nz))
    for _ in range(3):  # normal faults: hanging wall (right of the plane) moves down
        xf = rng.uniform(0.15, 0.85) * mod_nx
        throw = rng.uniform(20.0, 80.0) / h
        plane = xf + (z - mod_nz / 2) / np.tan(np.deg2rad(60.0))
        shift += np.where(x > plane, throw, 0.0)
    zz = z - shift
    lay = np.zeros((mod_nx, mod_nz), np.int64)
    for i in range(nlay - 1):
        iface = base[i] + np.tan(dips[i]) * x
        lay += (zz >= iface)
    v = vels[lay]
    bw, bh = int(round(1200 / h)), int(round(600 / h))
    bx = int(rng.uniform(0.2, 0.8) * (mod_nx - bw))
    bz = int(rng.uniform(0.3, 0.7) * (mod_nz - bh))
    v[bx:bx + bw, bz:bz + bh] = 4500.0
    return np.ascontiguousarray(v, np.float32)


def grf_4001(seed=3, mod_nz=801, mod_nx=4001, h=10.0, vlo=1500.0, vhi=4500.0, corr=500.0,
             nrefl=6):
    """Smooth Gaussian random field (correlation length corr metres) scaled
    to [vlo, vhi], plus nrefl sharp horizontal-ish reflectors (+-8%)."""
    rng = np.random.default_rng(seed)
    kx = np.fft.rfftfreq(mod_nx, d=h)
    kz = np.fft.fftfreq(mod_nz, d=h)
    k2 = kz[:, None] ** 2 + kx[None, :] ** 2
    spec = np.exp(-0.5 * k2 * (2 * np.pi * corr / 2.0) ** 2)
    noise = rng.normal(0, 1, (mod_nz, kx.size)) + 1j * rng.normal(0, 1, (mod_nz, kx.size))
    f = np.fft.irfft2(noise * spec, s=(mod_nz, mod_nx))
    f = (f - f.min()) / (f.max() - f.min())
    v = vlo + (vhi - vlo) * (0.15 + 0.7 * f)
    z = np.arange(mod_nz)[:, None]
    x = np.arange(mod_nx)[None, :]
    for d in np.sort(rng.uniform(0.15, 0.95, nrefl)) * mod_nz:
        amp = rng.uniform(60, 150) / h
        iface = d + amp * np.sin(2 * np.pi * x / rng.uniform(1500, 4000) + rng.uniform(0, 6.3))
        v = np.where(z >= iface, v * np.float32(1.08), v)
    v = np.clip(v, vlo, vhi)
    return np.ascontiguousarray(v.T, np.float32)


def smooth(raw_colmajor, sigma_cells):
    """Gaussian-smoothed migration velocity."""
    from scipy.ndimage import gaussian_filter
    return np.ascontiguousarray(gaussian_filter(raw_colmajor.astype(np.float64), sigma_cells,
                                                mode="nearest"), np.float32)


def _grf(rng, mod_nz, mod_nx, h, corr):
    """Zero-mean, unit-variance Gaussian random field [z][x] with a Gaussian
    spectrum of correlation length corr metres."""
    kx = np.fft.rfftfreq(mod_nx, d=h)
    kz = np.fft.fftfreq(mod_nz, d=h)
    k2 = kz[:, None] ** 2 + kx[None, :] ** 2
    spec = np.exp(-0.5 * k2 * (2 * np.pi * corr / 2.0) ** 2)
    noise = rng.normal(0, 1, (mod_nz, kx.size)) + 1j * rng.normal(0, 1, (mod_nz, kx.size))
    f = np.fft.irfft2(noise * spec, s=(mod_nz, mod_nx))
    return (f - f.mean()) / f.std()


def reflectors(seed, mod_nz=256, mod_nx=1024, h=10.0, vtop=1500.0, vbot=4000.0, corr=400.0,
               grf_std=200.0, nmin=10, nmax=40, dipmax=30.0, rmax=0.1):
    """A random reflectivity model for the deblurring network.

    Background: the linear gradient vtop (surface) .. vbot (bottom row) plus a
    Gaussian random field of standard deviation grf_std m/s and correlation
    length corr m, clipped to [vtop, vbot].  Reflectivity: nmin..nmax planar
    interfaces across the model, each through a seeded point (x0 anywhere, z0
    in the upper 95% of the depth range) with a dip uniform in +-dipmax
    degrees and a reflection strength uniform in +-rmax, drawn one cell thick
    (the strength split linearly between the two rows around the interface
    depth in each column); crossing interfaces add.

    Returns (background column-major [x][z], reflectivity [z][x]), float32."""
    rng = np.random.default_rng(seed)
    z = np.arange(mod_nz, dtype=np.float64)
    grad = vtop + (vbot - vtop) * z / (mod_nz - 1)
    v = np.clip(grad[:, None] + grf_std * _grf(rng, mod_nz, mod_nx, h, corr), vtop, vbot)
    m = np.zeros((mod_nz, mod_nx), np.float64)
    x = np.arange(mod_nx, dtype=np.float64)
    cols = np.arange(mod_nx)
    for _ in range(int(rng.integers(nmin, nmax + 1))):
        dip = np.deg2rad(rng.uniform(-dipmax, dipmax))
        x0 = rng.uniform(0.0, mod_nx - 1.0)
        z0 = rng.uniform(0.05, 0.95) * (mod_nz - 1)
        r = rng.uniform(-rmax, rmax)
        zi = z0 + np.tan(dip) * (x - x0)
        k = np.floor(zi).astype(np.int64)
        f = zi - k
        ok = (k >= 0) & (k < mod_nz - 1)
        np.add.at(m, (k[ok], cols[ok]), r * (1.0 - f[ok]))
        np.add.at(m, (k[ok] + 1, cols[ok]), r * f[ok])
    return np.ascontiguousarray(v.T, np.float32), m.astype(np.float32)
