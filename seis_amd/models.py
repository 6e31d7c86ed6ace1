"""Seeded synthetic models and acquisition geometry for the benchmark workloads.

The reference ships no models or data (its examples read absolute paths under
a private home directory, lsrtm.cpp:83-84), so the workloads are synthetic:
dipping layers cut by normal faults, Gaussian-smoothed, with a much smoother
migration velocity.  Arrays are [nx][nz] float32, z fastest, like the
reference's model files (inversion.cpp:370).
"""
import numpy as np
from scipy.ndimage import gaussian_filter


def layered_faulted(nx, nz, dx, dz, vmin=1500.0, vmax=5500.0, nlayers=40, max_dip_deg=15.0, nfaults=3,
                    throw=(20.0, 150.0), smooth_cells=2.0, seed=0):
    """True velocity of a layered-and-faulted section.

    `nlayers` interfaces with dips uniform in +-max_dip_deg; layer velocities rise
    from vmin to vmax with depth (+-3 % jitter); `nfaults` normal faults (60 deg
    planes, hanging wall dropped by a throw drawn in `throw` metres); then a
    Gaussian smoothing of `smooth_cells` cells."""
    rng = np.random.default_rng(seed)
    x = np.arange(nx, dtype=np.float64)[:, None] * dx
    z = np.arange(nz, dtype=np.float64)[None, :] * dz
    width, depth = nx * dx, nz * dz
    z0 = np.sort(rng.uniform(0.03 * depth, depth, nlayers))
    dips = np.tan(np.deg2rad(rng.uniform(-max_dip_deg, max_dip_deg, nlayers)))
    speeds = np.linspace(vmin, vmax, nlayers + 1)
    speeds[1:] *= 1.0 + rng.uniform(-0.03, 0.03, nlayers)
    speeds = np.clip(speeds, vmin, vmax)

    def section(xx, zz):
        k = np.zeros(np.broadcast(xx, zz).shape, np.int64)
        for i in range(nlayers):
            k += zz >= z0[i] + dips[i] * (xx - 0.5 * width)
        return speeds[k]

    # normal faults: points on the hanging-wall side take the section from
    # `throw` metres higher up
    shift = np.zeros(np.broadcast(x, z).shape)
    for _ in range(nfaults):
        xf = rng.uniform(0.15 * width, 0.85 * width)
        t = rng.uniform(*throw)
        side = rng.choice([-1.0, 1.0])
        plane = xf + side * z / np.tan(np.deg2rad(60.0))
        shift += np.where(side * (x - plane) > 0, t, 0.0)
    v = section(x + 0 * z, z - shift)
    if smooth_cells > 0:
        v = gaussian_filter(v, smooth_cells, mode="nearest")
    return np.clip(v, vmin, vmax).astype(np.float32)


def smooth_model(v, cells):
    """Migration velocity: Gaussian smoothing over `cells` cells."""
    return gaussian_filter(np.asarray(v, np.float64), cells, mode="nearest").astype(np.float32)


def sparse_spikes(nx, nz, density=0.02, seed=0):
    """Sparse spike reflectivity: N(0,1) amplitudes at a `density` fraction of cells."""
    rng = np.random.default_rng(seed)
    mask = rng.random((nx, nz)) < density
    return (rng.standard_normal((nx, nz)) * mask).astype(np.float32)


def three_layer(nx, nz, dz, interfaces=(600.0, 1400.0), speeds=(1500.0, 2200.0, 3000.0)):
    """Horizontally layered model (the single-shot forward modeling case)."""
    zz = np.arange(nz) * dz
    col = np.full(nz, speeds[0], np.float32)
    for zi, v in zip(interfaces, speeds[1:]):
        col[zz >= zi] = v
    return np.repeat(col[None, :], nx, axis=0)


def salt_body(v, dx, dz, speed=4500.0, seed=0):
    """Adds a salt-like body: the inside of a random smooth closed contour
    (a radius 1 + sum of 4 low harmonics around a centre in the middle third
    of the section) set to `speed`."""
    rng = np.random.default_rng(seed)
    nx, nz = v.shape
    width, depth = nx * dx, nz * dz
    cx, cz = rng.uniform(0.35, 0.65) * width, rng.uniform(0.35, 0.6) * depth
    r0 = rng.uniform(0.12, 0.18) * min(width, depth)
    amp = rng.uniform(0.05, 0.2, 4) / np.arange(1, 5)
    ph = rng.uniform(0, 2 * np.pi, 4)
    x = np.arange(nx)[:, None] * dx - cx
    z = np.arange(nz)[None, :] * dz - cz
    th = np.arctan2(z, x)
    r = r0 * (1.0 + sum(a * np.cos((k + 1) * th + p) for k, (a, p) in enumerate(zip(amp, ph))))
    out = np.array(v, np.float32, copy=True)
    out[np.hypot(x, z) < r] = speed
    return out


def spread_window(nx, sx, first, last, pad):
    """Model window of one shot: columns [x0, x0 + n) covering the source at
    column sx and receivers at columns first..last, widened by `pad` columns
    on each side and clipped to the model's nx columns (the aperture of
    getVel, lsrtm/inversion.cpp:552-598, without its edge replication)."""
    lo = max(0, min(sx, first) - pad)
    hi = min(nx - 1, max(sx, last) + pad)
    return lo, hi - lo + 1
