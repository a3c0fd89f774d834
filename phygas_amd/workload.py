"""Synthetic work units of the Phygas trace-back path (seeded, deterministic, numpy only).

Records are numpy structured arrays with the exact byte layout of ``PhCluster`` / ``PhTarget``
(reference: src/phygas.h:19-41), so the same buffer can be handed to the HIP library and to the
CPU checker, and written to the reference's text formats (19-field cluster lines and 14-field target
lines, reference: src/phygas.c:462-506) with round-trip-exact numbers.

Distributions (all distances in pc, velocities in km/s):

* stars: Galactic longitude uniform; heliocentric distance log-normal (median 500, sigma_ln 0.7,
  clipped to 20..5000); height above the plane from an exponential disk of scale height 300,
  truncated to |z| < d; velocity = flat rotation of 220 plus an isotropic Gaussian of 25 per axis;
  errors log-normal inside their ranges (parallax 0.02..0.3 mas, proper motion 0.02..0.2 mas/yr,
  RV 0.5..10): ln(e) ~ N(mid, span/4) clipped to the range.  Observed values are the true ones
  (a parallax that is always > 0; draws around it may still fail, exercising the retry rule).
* clusters: uniform in a 3-kpc disk around the Sun, Galactocentric height exponential with scale
  80; rotation 220 plus 10 per axis; distance error 5 %, proper-motion errors 0.02..0.05, RV errors
  0.5..3; r50 log-normal (median 3, sigma_ln 0.5), rtot = 3 r50; age log-uniform 1..300 Myr,
  age16 = age/1.5, age84 = 1.5 age, tmax = min(age84, 100); 878 of 4087 (the same fraction at
  other sizes) have no RV ("nan nan").
* pulsars: distance from DM, log-normal (median 2000, sigma_ln 0.3) with e_lndist 0.3 (30 %),
  no parallax, no RV; velocity = rotation plus a Maxwellian of sigma 265 per axis; proper-motion
  errors 0.1 mas/yr.
"""
from __future__ import annotations

import math
import os

import numpy as np

KMS_PER_KPCMYR = 977.7922216807892
AUYR_KMS = 1.495978707e8 / 3.15576e7
SUN_R0, SUN_Z0 = 8.122, 0.0208
SUN_V = np.array([11.1, 229.0 + 12.24, 7.25])
ROT_G = np.array([
    [-0.0548755604162154, -0.8734370902348850, -0.4838350155487132],
    [0.4941094278755837, -0.4448296299600112, 0.7469822444972189],
    [-0.8676661490190047, -0.1980763734312015, 0.4559837761750669],
])

CLUSTER_DTYPE = np.dtype({
    "names": ["name", "kind", "ra", "dec", "dist_pc", "e_dist_pc", "pmra", "e_pmra", "pmdec", "e_pmdec",
              "rv", "e_rv", "has_rv", "r50_pc", "rtot_pc", "age16", "age50", "age84", "tmax_myr"],
    "formats": ["S40", "S1"] + ["<f8"] * 10 + ["<i4"] + ["<f8"] * 6,
    "offsets": [0, 40, 48, 56, 64, 72, 80, 88, 96, 104, 112, 120, 128, 136, 144, 152, 160, 168, 176],
    "itemsize": 184,
})

TARGET_DTYPE = np.dtype({
    "names": ["id", "kind", "ra", "dec", "plx", "e_plx", "dist_pc", "e_lndist", "pmra", "e_pmra", "pmdec",
              "e_pmdec", "rv", "e_rv", "has_rv"],
    "formats": ["S40", "S1"] + ["<f8"] * 12 + ["<i4"],
    "offsets": [0, 40, 48, 56, 64, 72, 80, 88, 96, 104, 112, 120, 128, 136, 144],
    "itemsize": 152,
})


def _lognormal_in(rng, n, lo, hi):
    a, b = math.log(lo), math.log(hi)
    return np.clip(np.exp(rng.normal(0.5 * (a + b), 0.25 * (b - a), n)), lo, hi)


def _helio_gal_to_observed(hx, hy, hz, vgal):
    """Heliocentric Galactic position (pc) and Galactocentric velocity (km/s) -> ra, dec (deg),
    distance (pc), pmra, pmdec (mas/yr), rv (km/s) in the reference's conventions
    (inverse of reference src/coords.c:77-124)."""
    d = np.sqrt(hx * hx + hy * hy + hz * hz)
    rg = np.stack([hx, hy, hz], axis=1) / d[:, None]
    ricrs = rg @ ROT_G            # rows: A_G^T r
    ra = np.mod(np.arctan2(ricrs[:, 1], ricrs[:, 0]), 2 * np.pi)
    dec = np.arcsin(np.clip(ricrs[:, 2], -1, 1))
    vh = vgal - SUN_V[None, :]
    vicrs = vh @ ROT_G
    ra_hat = np.stack([-np.sin(ra), np.cos(ra), np.zeros_like(ra)], axis=1)
    dec_hat = np.stack([-np.sin(dec) * np.cos(ra), -np.sin(dec) * np.sin(ra), np.cos(dec)], axis=1)
    r_hat = np.stack([np.cos(dec) * np.cos(ra), np.cos(dec) * np.sin(ra), np.sin(dec)], axis=1)
    k = AUYR_KMS * 1e-3 * d
    pmra = np.sum(vicrs * ra_hat, axis=1) / k
    pmdec = np.sum(vicrs * dec_hat, axis=1) / k
    rv = np.sum(vicrs * r_hat, axis=1)
    return np.degrees(ra), np.degrees(dec), d, pmra, pmdec, rv


def _rotation(gx, gy, vc):
    R = np.hypot(gx, gy)
    return np.stack([vc * gy / R, -vc * gx / R, np.zeros_like(gx)], axis=1)


def make_stars(n: int, seed: int = 0, prefix: str = "S") -> np.ndarray:
    rng = np.random.default_rng(seed)
    lon = rng.uniform(0, 2 * np.pi, n)
    d = np.clip(500.0 * np.exp(0.7 * rng.normal(size=n)), 20.0, 5000.0)
    h = 300.0
    u = rng.uniform(size=n)
    az = -h * np.log1p(-u * (-np.expm1(-d / h)))      # |z| from Exp(h) truncated to [0, d)
    z = np.where(rng.uniform(size=n) < 0.5, -az, az)
    sinb = np.clip(z / d, -0.999999, 0.999999)
    cosb = np.sqrt(1 - sinb * sinb)
    hx, hy, hz = d * cosb * np.cos(lon), d * cosb * np.sin(lon), d * sinb
    gx, gy = hx * 1e-3 - SUN_R0, hy * 1e-3
    v = _rotation(gx, gy, 220.0) + rng.normal(0.0, 25.0, (n, 3))
    ra, dec, dist, pmra, pmdec, rv = _helio_gal_to_observed(hx, hy, hz, v)
    out = np.zeros(n, TARGET_DTYPE)
    out["id"] = [f"{prefix}{i:07d}".encode() for i in range(n)]
    out["kind"] = b"S"
    out["ra"], out["dec"] = ra, dec
    out["plx"] = 1000.0 / dist
    out["e_plx"] = _lognormal_in(rng, n, 0.02, 0.3)
    out["dist_pc"], out["e_lndist"] = 0.0, 0.0
    out["pmra"], out["pmdec"] = pmra, pmdec
    out["e_pmra"] = _lognormal_in(rng, n, 0.02, 0.2)
    out["e_pmdec"] = _lognormal_in(rng, n, 0.02, 0.2)
    out["rv"] = rv
    out["e_rv"] = _lognormal_in(rng, n, 0.5, 10.0)
    out["has_rv"] = 1
    return out


def make_clusters(n: int = 4087, seed: int = 0, n_no_rv: int | None = None) -> np.ndarray:
    rng = np.random.default_rng(seed)
    rho = 3000.0 * np.sqrt(rng.uniform(size=n))
    phi = rng.uniform(0, 2 * np.pi, n)
    hx, hy = rho * np.cos(phi), rho * np.sin(phi)
    gz = rng.laplace(0.0, 80.0, n)                    # Galactocentric height, pc
    hz = gz - SUN_Z0 * 1e3
    dd = np.sqrt(hx * hx + hy * hy + hz * hz)
    s = np.where(dd < 20.0, 20.0 / np.maximum(dd, 1e-9), 1.0)   # keep every cluster >= 20 pc away
    hx, hy, hz = hx * s, hy * s, hz * s
    gx, gy = hx * 1e-3 - SUN_R0, hy * 1e-3
    v = _rotation(gx, gy, 220.0) + rng.normal(0.0, 10.0, (n, 3))
    ra, dec, dist, pmra, pmdec, rv = _helio_gal_to_observed(hx, hy, hz, v)
    out = np.zeros(n, CLUSTER_DTYPE)
    out["name"] = [f"C{i:05d}".encode() for i in range(n)]
    out["kind"] = b"o"
    out["ra"], out["dec"], out["dist_pc"] = ra, dec, dist
    out["e_dist_pc"] = 0.05 * dist
    out["pmra"], out["pmdec"] = pmra, pmdec
    out["e_pmra"] = _lognormal_in(rng, n, 0.02, 0.05)
    out["e_pmdec"] = _lognormal_in(rng, n, 0.02, 0.05)
    e_rv = _lognormal_in(rng, n, 0.5, 3.0)
    if n_no_rv is None:
        n_no_rv = int(round(n * 878 / 4087))
    no_rv = rng.permutation(n)[:n_no_rv]
    rvv = rv.copy()
    rvv[no_rv] = np.nan
    e_rv[no_rv] = np.nan
    out["rv"], out["e_rv"] = rvv, e_rv
    out["has_rv"] = (~np.isnan(rvv)).astype(np.int32)
    r50 = 3.0 * np.exp(0.5 * rng.normal(size=n))
    out["r50_pc"], out["rtot_pc"] = r50, 3.0 * r50
    age = np.exp(rng.uniform(math.log(1.0), math.log(300.0), n))
    out["age16"], out["age50"], out["age84"] = age / 1.5, age, 1.5 * age
    out["tmax_myr"] = np.minimum(1.5 * age, 100.0)
    return out


def make_pulsars(n: int = 1, seed: int = 1) -> np.ndarray:
    rng = np.random.default_rng(seed)
    lon = rng.uniform(0, 2 * np.pi, n)
    d = 2000.0 * np.exp(0.3 * rng.normal(size=n))
    h = 300.0
    u = rng.uniform(size=n)
    az = -h * np.log1p(-u * (-np.expm1(-d / h)))
    z = np.where(rng.uniform(size=n) < 0.5, -az, az)
    sinb = np.clip(z / d, -0.999999, 0.999999)
    cosb = np.sqrt(1 - sinb * sinb)
    hx, hy, hz = d * cosb * np.cos(lon), d * cosb * np.sin(lon), d * sinb
    gx, gy = hx * 1e-3 - SUN_R0, hy * 1e-3
    v = _rotation(gx, gy, 220.0) + rng.normal(0.0, 265.0, (n, 3))
    ra, dec, dist, pmra, pmdec, _ = _helio_gal_to_observed(hx, hy, hz, v)
    out = np.zeros(n, TARGET_DTYPE)
    out["id"] = [f"J{i:04d}+{(i * 7919) % 10000:04d}".encode() for i in range(n)]
    out["kind"] = b"P"
    out["ra"], out["dec"] = ra, dec
    out["plx"], out["e_plx"] = 0.0, 0.0
    out["dist_pc"], out["e_lndist"] = dist, 0.3
    out["pmra"], out["pmdec"] = pmra, pmdec
    out["e_pmra"], out["e_pmdec"] = 0.1, 0.1
    out["rv"], out["e_rv"] = np.nan, np.nan
    out["has_rv"] = 0
    return out


def shuffled_control(stars: np.ndarray, seed: int = 3) -> np.ndarray:
    """Chance-alignment control: each star keeps its position and parallax and takes the velocity
    measurement (pmra, pmdec, RV with their errors) of the star a seeded permutation assigns."""
    perm = np.random.default_rng(seed).permutation(len(stars))
    out = stars.copy()
    for f in ("pmra", "e_pmra", "pmdec", "e_pmdec", "rv", "e_rv", "has_rv"):
        out[f] = stars[f][perm]
    return out


def half_file_control(stars: np.ndarray) -> np.ndarray:
    """The control rule of phygas_screen (reference: src/phygas_screen.c:35-47): velocities of the
    star nt/2 lines away, proper motions and their errors scaled by the parallax ratio."""
    nt = len(stars)
    out = stars.copy()
    for j in range(nt):
        o = stars[(j + nt // 2) % nt]
        f = stars[j]["plx"] / o["plx"] if (stars[j]["plx"] > 0 and o["plx"] > 0) else 1.0
        out[j]["pmra"], out[j]["e_pmra"] = o["pmra"] * f, o["e_pmra"] * f
        out[j]["pmdec"], out[j]["e_pmdec"] = o["pmdec"] * f, o["e_pmdec"] * f
        out[j]["rv"], out[j]["e_rv"], out[j]["has_rv"] = o["rv"], o["e_rv"], o["has_rv"]
    return out


def _num(x: float) -> str:
    return "nan" if math.isnan(x) else repr(float(x))


def write_clusters(path: str | os.PathLike, cl: np.ndarray) -> None:
    """19-field cluster lines (reference: src/phygas.c:472-475)."""
    with open(path, "w") as f:
        for i, c in enumerate(cl):
            vals = [c["ra"], c["dec"], c["dist_pc"], c["e_dist_pc"], c["pmra"], c["e_pmra"], c["pmdec"],
                    c["e_pmdec"], c["rv"], c["e_rv"], c["r50_pc"], c["rtot_pc"], c["age16"], c["age50"],
                    c["age84"], c["tmax_myr"]]
            f.write(f"{i} {c['name'].decode()} {c['kind'].decode()} " + " ".join(_num(v) for v in vals) + "\n")


def write_targets(path: str | os.PathLike, tg: np.ndarray) -> None:
    """14-field target lines (reference: src/phygas.c:495-497)."""
    with open(path, "w") as f:
        for t in tg:
            vals = [t["ra"], t["dec"], t["plx"], t["e_plx"], t["dist_pc"], t["e_lndist"], t["pmra"], t["e_pmra"],
                    t["pmdec"], t["e_pmdec"], t["rv"], t["e_rv"]]
            f.write(f"{t['id'].decode()} {t['kind'].decode()} " + " ".join(_num(v) for v in vals) + "\n")


def work_unit(name: str) -> dict:
    """The named work units: clusters, targets and run parameters."""
    if name == "screen_fp64_1k":        # plain FP64 linear screening, 1,000 x 4,087 (seed 0)
        return dict(clusters=make_clusters(4087, 0), stars=make_stars(1000, 0), chi2_max=25.0, n_samp=64,
                    r_kernel=2.0, p_min=0.0)
    if name == "pair_1x300x200":        # Monte Carlo pairs, 1 pulsar x 300 clusters x 200 clones (seed 1)
        return dict(clusters=make_clusters(300, 1), targets=make_pulsars(1, 1), n_clones=200, dt_myr=0.05)
    if name == "screen_20k_control":    # batched work unit, 20,000 x 4,087 + shuffled control (seeds 2, 3)
        stars = make_stars(20000, 2)
        return dict(clusters=make_clusters(4087, 0), stars=stars, control=shuffled_control(stars, 3),
                    chi2_max=25.0, n_samp=64, r_kernel=2.0, p_min=0.001)
    raise KeyError(name)
