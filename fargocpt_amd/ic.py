"""Seeded initial conditions and polar-grid geometry for the hydro step.

The radial mesh follows the reference's logarithmic spacing
(src/init.cpp:93-101): ``Radii[i] = Rmin * g**(i-1)`` with
``g = (Rmax/Rmin)**(1/(Nr-2))`` so that ``Radii[1] = Rmin`` and
``Radii[Nr-1] = Rmax``; ring 0 and ring Nr-1 are ghost rings.  ``Rmed`` is the
area-weighted ring centre of src/init.cpp:188-196.

The initial gas state is this package's own (the reference generates its
initial state with open-simplex noise, src/init.cpp:1016-1060, which is not part
of the hydro step):

* ``Sigma = sigma0 * Rmed**(-p) * (1 + amp * u)`` with ``u`` uniform in
  [-1, 1) drawn from SplitMix64(seed XOR cell index);
* locally isothermal sound speed ``c_s = h * Rmed**F * v_K``;
* ``v_phi`` in radial force balance with the unperturbed power-law pressure,
  minus the frame rotation ``Omega_F * Rmed``;
* ``v_r = 0``;
* adiabatic runs: ``E = Sigma * c_s**2 / (gamma - 1)`` (temperature from the
  aspect ratio).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

_MASK64 = (1 << 64) - 1


def splitmix64(x: np.ndarray) -> np.ndarray:
    """SplitMix64 finaliser applied element-wise to uint64 counters."""
    z = (x + np.uint64(0x9E3779B97F4A7C15)) & np.uint64(_MASK64)
    z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
    z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return z ^ (z >> np.uint64(31))


def uniform_pm1(seed: int, n: int) -> np.ndarray:
    """n values uniform in [-1, 1): top 53 bits of SplitMix64(seed ^ index)."""
    with np.errstate(over="ignore"):
        idx = np.arange(n, dtype=np.uint64) ^ np.uint64(seed & _MASK64)
        bits = splitmix64(idx) >> np.uint64(11)
    return bits.astype(np.float64) * (2.0 / 9007199254740992.0) - 1.0


def log_radii(nr: int, rmin: float, rmax: float) -> np.ndarray:
    """Radii[0..nr+1] of the logarithmic mesh (src/init.cpp:93-101).

    Radii[nr+1] lies one ring beyond the grid; the reference reads it only
    through Rmed[nr] (src/init.cpp:188-196 fills the radial arrays past the end).
    """
    g = math.pow(rmax / rmin, 1.0 / (float(nr) - 2.0))
    return np.array([rmin * math.pow(g, float(i) - 1.0) for i in range(nr + 2)])


def rmed_from_radii(radii: np.ndarray) -> np.ndarray:
    """Area-weighted ring centres (src/init.cpp:188-196), length len(radii)-1."""
    rinf, rsup = radii[:-1], radii[1:]
    out = np.empty(len(rinf))
    for i in range(len(rinf)):
        a = 2.0 / 3.0 * (math.pow(rsup[i], 3) - math.pow(rinf[i], 3))
        out[i] = a / (math.pow(rsup[i], 2) - math.pow(rinf[i], 2))
    return out


@dataclass
class DiskSetup:
    """Physical and numerical parameters of one disk (code units)."""

    nr: int = 128
    nphi: int = 384
    rmin: float = 0.4
    rmax: float = 2.5
    sigma0: float = 2.25e-5
    sigma_slope: float = 0.5
    aspect_ratio: float = 0.05
    flaring: float = 0.0
    G: float = 1.0
    mstar: float = 1.0
    omega_frame: float = 1.0
    adiabatic: bool = False
    gamma: float = 1.4
    noise_amp: float = 1e-4
    seed: int = 1337
    extra: dict = field(default_factory=dict)


def initial_state(s: DiskSetup) -> dict:
    """Return dict of float64 arrays: radii (nr+2), sigma, vrad, vazi, energy."""
    radii = log_radii(s.nr, s.rmin, s.rmax)
    rmed = rmed_from_radii(radii[: s.nr + 1])
    nr, nphi = s.nr, s.nphi
    u = uniform_pm1(s.seed, nr * nphi).reshape(nr, nphi)
    prof = np.array([s.sigma0 * math.pow(r, -s.sigma_slope) for r in rmed])
    sigma = prof[:, None] * (1.0 + s.noise_amp * u)
    vk = np.sqrt(s.G * s.mstar / rmed)
    h = s.aspect_ratio * rmed ** s.flaring
    cs = h * vk
    # d ln P / d ln r for P = Sigma c_s^2 with Sigma ~ r^-p, c_s^2 ~ r^(2F-1)
    dlnp = -s.sigma_slope + 2.0 * s.flaring - 1.0
    vphi_ring = vk * np.sqrt(1.0 + dlnp * h * h) - s.omega_frame * rmed
    vazi = np.repeat(vphi_ring[:, None], nphi, axis=1)
    vrad = np.zeros((nr + 1, nphi))
    if s.adiabatic:
        energy = sigma * (cs * cs)[:, None] / (s.gamma - 1.0)
    else:
        energy = np.zeros((nr, nphi))
    return {
        "radii": radii,
        "sigma": np.ascontiguousarray(sigma),
        "vrad": vrad,
        "vazi": np.ascontiguousarray(vazi),
        "energy": np.ascontiguousarray(energy),
    }
