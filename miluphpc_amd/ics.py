"""Initial conditions of the five workloads (numpy, seeded), and their parameters.

  plummer           -- Plummer sphere (G = M = a = 1), truncated at 10 a, isotropic
                       velocities from the distribution function by Aarseth-Henon-Wielen
                       (1974) rejection, shifted to the centre-of-mass frame.
  sedov_lattice     -- Sedov-Taylor blast on a cell-centred cubic lattice in [-0.5, 0.5]^3,
                       rho = 1, total mass 1, e = 1e-6 plus E = 1 deposited with the
                       cubic-spline weight of support 2 h0 (testcases/sedov/initial_sedov.py
                       deposits within 2 sml).
  boss_bodenheimer  -- uniform sphere R = M = 1 with m_i ~ 1 + 0.1 cos(2 phi), solid-body
                       rotation omega = 0.72 about z, isothermal c_s = 0.167.

Particle masses are m = 1/N (reference input format /m, /x, /v; README "Input HDF5 file").
"""
from __future__ import annotations

import math

import numpy as np

from . import lib

SEED = 20261202


def _isotropic(rng: np.random.Generator, n: int) -> np.ndarray:
    cost = rng.uniform(-1.0, 1.0, n)
    phi = rng.uniform(0.0, 2.0 * math.pi, n)
    sint = np.sqrt(1.0 - cost * cost)
    return np.stack([sint * np.cos(phi), sint * np.sin(phi), cost], axis=1)


def plummer(n: int, seed: int = SEED, rmax: float = 10.0) -> dict:
    rng = np.random.default_rng(seed)
    r = np.empty(n)
    fill = 0
    while fill < n:
        u = rng.uniform(0.0, 1.0, 2 * (n - fill) + 16)
        u = u[u > 0.0]
        rr = 1.0 / np.sqrt(u ** (-2.0 / 3.0) - 1.0)
        rr = rr[rr <= rmax]
        take = min(len(rr), n - fill)
        r[fill:fill + take] = rr[:take]
        fill += take
    pos = _isotropic(rng, n) * r[:, None]
    # speed: q from g(q) = q^2 (1 - q^2)^(7/2) by von Neumann rejection (max g < 0.1)
    q = np.empty(n)
    fill = 0
    while fill < n:
        m = 2 * (n - fill) + 16
        x1 = rng.uniform(0.0, 1.0, m)
        x2 = rng.uniform(0.0, 0.1, m)
        ok = x2 < x1 * x1 * (1.0 - x1 * x1) ** 3.5
        qq = x1[ok]
        take = min(len(qq), n - fill)
        q[fill:fill + take] = qq[:take]
        fill += take
    v = q * math.sqrt(2.0) * (1.0 + r * r) ** (-0.25)
    vel = _isotropic(rng, n) * v[:, None]
    mass = np.full(n, 1.0 / n)
    pos -= (mass[:, None] * pos).sum(0) / mass.sum()
    vel -= (mass[:, None] * vel).sum(0) / mass.sum()
    return {"x": pos[:, 0].copy(), "y": pos[:, 1].copy(), "z": pos[:, 2].copy(),
            "vx": vel[:, 0].copy(), "vy": vel[:, 1].copy(), "vz": vel[:, 2].copy(), "mass": mass}


def cubic_spline_w(r: np.ndarray, h: float) -> np.ndarray:
    """kernel.cu:50-83 in 3D, support radius h."""
    q = r / h
    f = 8.0 / math.pi / h ** 3
    w = np.where(q <= 0.5, f * (6 * q ** 3 - 6 * q ** 2 + 1), 2 * f * (1 - q) ** 3)
    return np.where(q > 1.0, 0.0, w)


def h_for_neighbours(n_per_volume: float, n_nb: int) -> float:
    """Support radius holding n_nb neighbours at number density n_per_volume."""
    return (3.0 * n_nb / (4.0 * math.pi * n_per_volume)) ** (1.0 / 3.0)


def sedov_lattice(nside: int, n_nb: int = 64, e_bg: float = 1e-6, energy: float = 1.0) -> dict:
    n = nside ** 3
    c = -0.5 + (np.arange(nside) + 0.5) / nside
    z, y, x = np.meshgrid(c, c, c, indexing="ij")
    x, y, z = x.ravel().copy(), y.ravel().copy(), z.ravel().copy()
    mass = np.full(n, 1.0 / n)
    h0 = h_for_neighbours(float(n), n_nb)
    r = np.sqrt(x * x + y * y + z * z)
    w = cubic_spline_w(r, 2.0 * h0)
    e = np.full(n, e_bg)
    e += energy * w / float((mass * w).sum())
    zeros = np.zeros(n)
    return {"x": x, "y": y, "z": z, "vx": zeros.copy(), "vy": zeros.copy(), "vz": zeros.copy(),
            "mass": mass, "sml": np.full(n, h0), "e": e, "rho": np.ones(n), "p": zeros.copy(),
            "cs": zeros.copy()}


def boss_bodenheimer(n: int, seed: int = SEED, omega: float = 0.72, cs: float = 0.167,
                     n_nb: int = 64) -> dict:
    rng = np.random.default_rng(seed)
    pos = np.empty((n, 3))
    fill = 0
    while fill < n:
        p = rng.uniform(-1.0, 1.0, (2 * (n - fill) + 16, 3))
        p = p[(p * p).sum(1) <= 1.0]
        take = min(len(p), n - fill)
        pos[fill:fill + take] = p[:take]
        fill += take
    phi = np.arctan2(pos[:, 1], pos[:, 0])
    w = 1.0 + 0.1 * np.cos(2.0 * phi)
    mass = w / w.sum()
    vel = np.stack([-omega * pos[:, 1], omega * pos[:, 0], np.zeros(n)], axis=1)
    h0 = h_for_neighbours(n / (4.0 / 3.0 * math.pi), n_nb)
    zeros = np.zeros(n)
    return {"x": pos[:, 0].copy(), "y": pos[:, 1].copy(), "z": pos[:, 2].copy(),
            "vx": vel[:, 0].copy(), "vy": vel[:, 1].copy(), "vz": vel[:, 2].copy(),
            "mass": mass, "sml": np.full(n, h0), "e": zeros.copy(), "rho": np.ones(n),
            "p": zeros.copy(), "cs": np.full(n, cs)}


# ---- the five workloads' parameters ---------------------------------------------------
def params_plummer_small() -> lib.Params:
    """N = 65,536, theta 0.5, eps 1e-3, Hilbert keys, FP64, monopole (plain method)."""
    return lib.default_params(theta=0.5, smoothing=1e-6, curve=lib.CURVE_HILBERT,
                              mac=lib.MAC_CLASSIC, quadrupole=0, precision=lib.PREC_FP64)


def params_sedov(**overrides) -> lib.Params:
    """Ideal gas gamma 5/3, cubic spline, 64 neighbours (band 50-80), alpha 1, beta 2, FP64."""
    kw = dict(eos=lib.EOS_IDEAL_GAS, gamma=5.0 / 3.0, alpha=1.0, beta=2.0, n_target=64, band_lo=50,
              band_hi=80, max_counts=7, max_nb=180, sml_avg=1, precision=lib.PREC_FP64,
              curve=lib.CURVE_LEBESGUE)
    kw.update(overrides)
    return lib.default_params(**kw)


def params_boss_bodenheimer() -> lib.Params:
    """Self-gravity + SPH, theta 0.5, eps 5e-3, isothermal c_s 0.167, FP32 + Kahan pair sums."""
    return lib.default_params(theta=0.5, smoothing=25e-6, eos=lib.EOS_ISOTHERMAL_CONST_CS,
                              cs_const=0.167, alpha=1.0, beta=2.0, n_target=64, band_lo=50,
                              band_hi=80, max_nb=180, sml_avg=1,
                              precision=lib.PREC_FP32_KAHAN, curve=lib.CURVE_HILBERT)


def params_plummer_large() -> lib.Params:
    """N = 16,777,216, theta 0.6, eps 2e-4, monopole + quadrupole, FP32 + Kahan far field.
    Group acceptance by the group's bounding box: on this sphere it opens more cells than the
    per-member vote (2,648 against 2,085 interactions per particle) yet walks faster (34.6
    against 35.9 ms on one MI355X) with a slightly lower force error against direct
    summation (99th percentile 1.38e-3 against 1.44e-3)."""
    return lib.default_params(theta=0.6, smoothing=4e-8, curve=lib.CURVE_HILBERT,
                              mac=lib.MAC_CLASSIC, quadrupole=1, precision=lib.PREC_FP32_KAHAN,
                              bh_accept=lib.BH_BOX)
