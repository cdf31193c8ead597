"""The torus input bundle: seven .npy fields of a constant-angular-momentum
torus around a Kerr hole, the inputs every reference configuration names
(the reference generates them with create_example_data.py).

Physics (units G = c = M = 1 inside; lengths scale with the mass):
  * The torus gas rotates with constant specific angular momentum
    l = -u_phi / u_t and u^r = u^theta = 0 (Fishbone & Moncrief 1976;
    Abramowicz, Jaroszynski & Sikora 1978).  l is the value of the
    prograde circular equatorial geodesic at the pressure maximum r_pmax,
    whose angular velocity is Omega_K = 1 / (r^(3/2) + a) (Bardeen, Press &
    Teukolsky 1972, eq. 2.16); for that orbit l = -(g_tphi + Omega g_phiphi)
    / (g_tt + Omega g_tphi).
  * With u_mu = u_t (1, 0, 0, -l) the normalisation g^{mu nu} u_mu u_nu = -1
    gives u_t = -1 / sqrt(-(g^tt - 2 l g^tphi + l^2 g^phiphi)) from the
    inverse Boyer-Lindquist metric; the contravariant components follow by
    raising the index.  The gas exists where this is a future-directed
    timelike vector.
  * For constant l the Euler equation integrates to h u_t = const: the
    specific enthalpy is h = u_t(r_in, pi/2) / u_t(r, theta), and the torus
    is where r >= r_in and h >= 1.
  * A polytrope P = K rho^gamma (K = 1) has h = 1 + gamma/(gamma-1) rho^(gamma-1),
    so rho = ((gamma-1)/gamma (h-1))^(1/(gamma-1)), scaled to a peak of 1;
    the temperature is rho^(gamma-1) of the scaled density.
  * Outside the torus: density and temperature rho_floor, and the
    zero-angular-momentum observer u_mu = u_t (1, 0, 0, 0), i.e.
    u^t = sqrt(-g^tt), u^phi = u^t g^tphi / g^tt.
  * The stored velocity is (u^t, 0, 0, u^phi / M).  With u^r = 0 the
    Boyer-Lindquist and Kerr-Schild components agree, so the same array
    serves the Kerr-Schild path.

Grid: r geometric on [r_min, r_max] (r_min defaults to 1.2 horizon radii),
theta uniform on [1e-3, pi - 1e-3], phi = 2 pi k / nphi.  Arrays are C-order
float64, shapes (nr, nth, nph) and (nr, nth, nph, 4), the layout the
reference's .npy loader accepts.
"""
import os

import numpy as np

FIELD_FILES = ("r", "theta", "phi", "rho", "Tgas", "vel", "mag")
THETA_MARGIN = 1e-3


class KerrBL:
    """Boyer-Lindquist Kerr metric (M = 1) at points (r, theta): the t-phi
    block of the metric and of its inverse."""

    def __init__(self, r, theta, a):
        self.a = a
        self.s2 = np.sin(theta) ** 2
        self.sigma = r * r + a * a * np.cos(theta) ** 2
        self.delta = r * r - 2.0 * r + a * a
        self.big_a = (r * r + a * a) ** 2 - a * a * self.delta * self.s2
        self.r = r

    def lower(self):
        """(g_tt, g_tphi, g_phiphi)."""
        r, a, sg = self.r, self.a, self.sigma
        return (-(1.0 - 2.0 * r / sg), -2.0 * a * r * self.s2 / sg, self.big_a * self.s2 / sg)

    def upper(self):
        """(g^tt, g^tphi, g^phiphi)."""
        sd = self.sigma * self.delta
        return (-self.big_a / sd, -2.0 * self.a * self.r / sd,
                (self.delta - self.a * self.a * self.s2) / (sd * self.s2))


def keplerian_angular_momentum(r, a):
    """-u_phi/u_t of the prograde circular equatorial geodesic at radius r."""
    omega = 1.0 / (r ** 1.5 + a)
    g_tt, g_tp, g_pp = KerrBL(np.asarray(r, dtype=np.float64), np.pi / 2.0, a).lower()
    return -(g_tp + omega * g_pp) / (g_tt + omega * g_tp)


def constant_l_velocity(r, theta, a, ell):
    """(u_t, u^t, u^phi) of gas with -u_phi/u_t = ell at (r, theta); NaN
    where no future-directed timelike solution exists."""
    gtt, gtp, gpp = KerrBL(r, theta, a).upper()
    q = -(gtt - 2.0 * ell * gtp + ell * ell * gpp)
    with np.errstate(invalid="ignore", divide="ignore"):
        u_t = -1.0 / np.sqrt(q)
    ut = u_t * (gtt - ell * gtp)
    uph = u_t * (gtp - ell * gpp)
    bad = ~(ut > 0.0)
    return (np.where(bad, np.nan, u_t), np.where(bad, np.nan, ut), np.where(bad, np.nan, uph))


def zamo_velocity(r, theta, a):
    """(u^t, u^phi) of the zero-angular-momentum observer."""
    gtt, gtp, _ = KerrBL(r, theta, a).upper()
    ut = np.sqrt(-gtt)
    return ut, ut * gtp / gtt


def _check(nr, ntheta, nphi, mass, spin, gamma, r_in, r_pressure_max):
    if min(nr, ntheta, nphi) < 2:
        raise ValueError(f"every grid axis needs two or more points (nr={nr}, "
                         f"ntheta={ntheta}, nphi={nphi})")
    if not mass > 0.0:
        raise ValueError(f"the black-hole mass must be > 0 (mass={mass})")
    if abs(spin) > mass:
        raise ValueError(f"|spin| may not exceed the mass (spin={spin}, mass={mass})")
    if not gamma > 1.0:
        raise ValueError(f"the adiabatic index must be > 1 (gamma={gamma})")
    if not 0.0 < r_in < r_pressure_max:
        raise ValueError(f"the torus needs 0 < r_in < r_pressure_max (r_in={r_in}, "
                         f"r_pressure_max={r_pressure_max})")


def fishbone_moncrief(nr=64, ntheta=64, nphi=32, r_min=None, r_max=50.0, mass=1.0,
                      spin=0.9375, r_in=6.0, r_pressure_max=10.0, gamma=4.0 / 3.0,
                      rho_floor=1.0e-6):
    """The dict {name: array} of the seven input fields (arguments as the
    reference generator's command line: create_example_data.py:283-305)."""
    _check(nr, ntheta, nphi, mass, spin, gamma, r_in, r_pressure_max)
    r_horizon = mass + np.sqrt(mass * mass - spin * spin)
    if r_min is None:
        r_min = 1.2 * r_horizon
    if not r_horizon < r_min < r_max:
        raise ValueError(f"the radial grid must lie outside the horizon: need "
                         f"{r_horizon:.6g} < r_min ({r_min}) < r_max ({r_max})")

    r_axis = np.geomspace(r_min, r_max, nr)
    th_axis = np.linspace(THETA_MARGIN, np.pi - THETA_MARGIN, ntheta)
    ph_axis = np.arange(nphi) * (2.0 * np.pi / nphi)
    R, TH, PH = np.meshgrid(r_axis, th_axis, ph_axis, indexing="ij")

    # dimensionless (M = 1) geometry
    a = spin / mass
    x = R / mass
    ell = keplerian_angular_momentum(r_pressure_max / mass, a)
    u_t, ut, uph = constant_l_velocity(x, TH, a, ell)
    u_t_edge = constant_l_velocity(np.array(r_in / mass), np.array(np.pi / 2.0), a, ell)[0]
    with np.errstate(invalid="ignore", divide="ignore"):
        h = u_t_edge / u_t                    # h u_t = const, h = 1 at the inner edge
    torus = (x >= r_in / mass) & np.isfinite(h) & (h >= 1.0)

    excess = np.where(torus, (gamma - 1.0) / gamma * (h - 1.0), 0.0)
    rho = np.power(np.maximum(excess, 0.0), 1.0 / (gamma - 1.0))
    if rho.max() > 0.0:
        rho = rho / rho.max()
    temperature = np.power(np.maximum(rho, 1e-300), gamma - 1.0)

    zt, zph = zamo_velocity(x, TH, a)
    vel = np.zeros(R.shape + (4,))
    vel[..., 0] = np.where(torus, ut, zt)
    vel[..., 3] = np.where(torus, uph, zph) / mass
    out = dict(r=R, theta=TH, phi=PH,
               rho=np.where(torus, rho, rho_floor),
               Tgas=np.where(torus, temperature, rho_floor),
               vel=vel, mag=np.zeros_like(vel))
    return {k: np.ascontiguousarray(v, dtype=np.float64) for k, v in out.items()}


def save_bundle(fields, directory):
    """Write r.npy ... mag.npy into `directory` (the reference's file names)."""
    os.makedirs(directory, exist_ok=True)
    for name in FIELD_FILES:
        np.save(os.path.join(directory, name + ".npy"), fields[name])


def load_bundle(directory):
    """Read the seven float64 C-order fields written by save_bundle()."""
    out = {}
    for name in FIELD_FILES:
        a = np.load(os.path.join(directory, name + ".npy"), allow_pickle=False)
        if a.dtype != np.float64:
            raise ValueError(f"{name}.npy must be float64")
        out[name] = np.ascontiguousarray(a)
    return out
