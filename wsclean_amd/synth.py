"""Seeded synthetic observations for the benchmark and parity workloads.

uvw follow the standard rotation of equatorial baseline vectors by hour angle
H and declination dec (Thompson, Moran & Swenson, eq. 4.1):
  u =  sin H X + cos H Y
  v = -sin dec cos H X + sin dec sin H Y + cos dec Z
  w =  cos dec cos H X - cos dec sin H Y + sin dec Z
with (X, Y, Z) from local east/north/up offsets at latitude phi:
  X = -sin phi N + cos phi U,  Y = E,  Z = cos phi N + sin phi U.
Rows are time-major, baseline-minor (the order a MeasurementSet holds).
Wavelengths are c / frequency in FP64.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

SPEED_OF_LIGHT = 299792458.0
EARTH_ROTATION = 7.2921150e-5  # rad / s
ARCSEC = np.pi / (180.0 * 3600.0)


@dataclass
class Observation:
    uvw_m: np.ndarray          # float64 [nrow, 3]
    frequencies: np.ndarray    # float64 [nchan]
    n_antennas: int
    n_times: int
    meta: dict = field(default_factory=dict)

    @property
    def wavelengths(self):
        return SPEED_OF_LIGHT / self.frequencies

    @property
    def nrow(self):
        return self.uvw_m.shape[0]

    @property
    def nchan(self):
        return self.frequencies.shape[0]


def disc_array(n_ant, radius_m, rng):
    r = radius_m * np.sqrt(rng.uniform(0, 1, n_ant))
    t = rng.uniform(0, 2 * np.pi, n_ant)
    return np.column_stack([r * np.cos(t), r * np.sin(t), np.zeros(n_ant)])


def log_spiral_array(n_ant, radius_m, rng, arms=3, r0=20.0):
    """n_ant stations on `arms` logarithmic spiral arms from r0 to radius_m."""
    idx = np.arange(n_ant)
    arm = idx % arms
    k = idx // arms
    per_arm = int(np.ceil(n_ant / arms))
    frac = k / max(per_arm - 1, 1)
    r = r0 * (radius_m / r0) ** frac
    theta = 2 * np.pi * arm / arms + 1.5 * np.log(r / r0)
    jitter = rng.normal(0, 0.02, n_ant) * r
    return np.column_stack([(r + jitter) * np.cos(theta), (r + jitter) * np.sin(theta),
                            np.zeros(n_ant)])


def observe(enu, latitude_deg, dec_deg, n_times, dt_s, frequencies, ha0=None):
    lat = np.deg2rad(latitude_deg)
    dec = np.deg2rad(dec_deg)
    n_ant = enu.shape[0]
    a1, a2 = np.triu_indices(n_ant, 1)
    d = enu[a2] - enu[a1]
    E, N, U = d[:, 0], d[:, 1], d[:, 2]
    X = -np.sin(lat) * N + np.cos(lat) * U
    Y = E
    Z = np.cos(lat) * N + np.sin(lat) * U
    t = (np.arange(n_times) - (n_times - 1) / 2.0) * dt_s
    ha = (ha0 if ha0 is not None else 0.0) + EARTH_ROTATION * t
    sh, ch = np.sin(ha)[:, None], np.cos(ha)[:, None]
    u = sh * X + ch * Y
    v = -np.sin(dec) * ch * X + np.sin(dec) * sh * Y + np.cos(dec) * Z
    w = np.cos(dec) * ch * X - np.cos(dec) * sh * Y + np.sin(dec) * Z
    uvw = np.stack([u, v, w], axis=-1).reshape(-1, 3)
    return Observation(np.ascontiguousarray(uvw), np.asarray(frequencies, np.float64),
                       n_ant, n_times, meta=dict(n_baselines=len(a1)))


def config1(scale=1.0, seed=101):
    """64 antennas in a 3 km disc (2016 baselines), dec -30 deg, 100 x 10 s,
    16 channels 120-168 MHz; latitude -26.70 deg. `scale` < 1 keeps a
    fraction of the timesteps (at least 2)."""
    rng = np.random.default_rng(seed)
    enu = disc_array(64, 3000.0, rng)
    nt = max(2, int(round(100 * scale)))
    freqs = np.linspace(120e6, 168e6, 16)
    return observe(enu, -26.70, -30.0, nt, 10.0, freqs)


def config2(scale=1.0, nchan=64, seed=202):
    """62 stations on a log spiral to 40 km (1891 baselines), dec +52 deg,
    1800 x 2 s, 64 channels 115-165 MHz; latitude +52.91 deg."""
    rng = np.random.default_rng(seed)
    enu = log_spiral_array(62, 40000.0, rng)
    nt = max(2, int(round(1800 * scale)))
    freqs = np.linspace(115e6, 165e6, nchan)
    return observe(enu, 52.91, 52.0, nt, 2.0, freqs)


def point_sources(n, field_rad, rng, flux_range=(0.01, 10.0)):
    """l, m uniform within +-field/2 (square), log-uniform flux."""
    lm = rng.uniform(-0.5, 0.5, (n, 2)) * field_rad
    lo, hi = np.log(flux_range[0]), np.log(flux_range[1])
    s = np.exp(rng.uniform(lo, hi, n))
    return lm[:, 0], lm[:, 1], s


def source_visibilities(uvw_lambda, l, m, flux):
    """V = sum_s S_s exp(+2 pi i (u l + v m + w (n - 1))) in FP64 -- the
    convention whose adjoint (the reference's invert) images the source at
    (l, m)."""
    n1 = np.sqrt(1.0 - l * l - m * m) - 1.0
    ph = 2 * np.pi * (np.outer(uvw_lambda[:, 0], l) + np.outer(uvw_lambda[:, 1], m) +
                      np.outer(uvw_lambda[:, 2], n1))
    return (np.exp(1j * ph) * flux[None, :]).sum(axis=1)


def pixel_of(l, m, width, height, dl, dm):
    """Image pixel (x, y) of (l, m): l = (W/2 - x) dl, m = (y - H/2) dm."""
    return width // 2 - l / dl, height // 2 + m / dm


def config4_facets(seed=404, n=16, lo=512, hi=2048, radius_deg=5.0):
    """The batching workload's facet set: n facets whose boxes are uniform in
    [lo, hi] pixels, each gridded at the next multiple of 64 with no prime
    factor above 7 (the low-prime rule of math/calculatefftsize.h:4-19), with
    centres uniform over the disc of radius_deg about the phase centre; the
    facet's l/m shift is the direction cosine of its centre."""
    from .facets import Facet, gridded_size
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        box = int(rng.integers(lo, hi + 1))
        size = gridded_size(box, 1.0, 64)
        r = np.deg2rad(radius_deg) * np.sqrt(rng.uniform())
        th = rng.uniform(0.0, 2.0 * np.pi)
        out.append(Facet(size, size, float(np.sin(r) * np.cos(th)),
                         float(np.sin(r) * np.sin(th)), (0, 0, box, box), f"f{i}"))
    return out
