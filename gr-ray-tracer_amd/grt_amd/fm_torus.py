"""Fishbone-Moncrief torus input bundle (the seven .npy fields).

Restates the closed forms of the reference generator
(create_example_data.py:54-248): a constant-angular-momentum torus around a
Kerr hole, with the specific angular momentum set to the prograde Keplerian
value at the pressure maximum (Bardeen, Press & Teukolsky 1972, eqs. 2.12-2.13),
the enthalpy from h * u_t = const, an ideal-gas polytrope normalised to peak
density 1, and a zero-angular-momentum-observer atmosphere at `rho_floor`.

Grid: r log-spaced on [r_min, r_max], theta uniform on [1e-3, pi - 1e-3],
phi uniform on [0, 2 pi) without the endpoint (create_example_data.py:166-169).
Arrays are C-order float64 with shapes (nr, nth, nph) and (nr, nth, nph, 4),
the layout the reference's .npy loader accepts.
"""
import os

import numpy as np

FIELD_FILES = ("r", "theta", "phi", "rho", "Tgas", "vel", "mag")


def _kepler_ell(r, a):
    # create_example_data.py:54-64
    sr = np.sqrt(r)
    return (r * r - 2.0 * a * sr + a * a) / (r ** 1.5 - 2.0 * sr + a)


def _bl_metric(r, th, a):
    # create_example_data.py:67-77 (mass = 1)
    s2 = np.sin(th) ** 2
    sigma = r * r + a * a * np.cos(th) ** 2
    delta = r * r - 2.0 * r + a * a
    big_a = (r * r + a * a) ** 2 - delta * a * a * s2
    return -(1.0 - 2.0 * r / sigma), -2.0 * r * a * s2 / sigma, big_a * s2 / sigma


def _rigid_rotation(r, th, a, ell):
    # create_example_data.py:80-99: u^t, u^phi, u_t for u_phi = -ell u_t
    gtt, gtp, gpp = _bl_metric(r, th, a)
    omega = -(gtp + ell * gtt) / (gpp + ell * gtp)
    norm = -(gtt + 2.0 * omega * gtp + omega * omega * gpp)
    with np.errstate(invalid="ignore", divide="ignore"):
        ut = 1.0 / np.sqrt(norm)
    return ut, omega * ut, ut * (gtt + omega * gtp)


def fishbone_moncrief(nr=64, ntheta=64, nphi=32, r_min=None, r_max=50.0, mass=1.0,
                      spin=0.9375, r_in=6.0, r_pressure_max=10.0, gamma=4.0 / 3.0,
                      rho_floor=1.0e-6):
    """Return the dict {name: array} of the seven input fields."""
    if min(nr, ntheta, nphi) < 2:
        raise ValueError("nr, ntheta, and nphi must each be at least 2")
    if mass <= 0.0:
        raise ValueError("mass must be positive")
    if abs(spin) > mass:
        raise ValueError("require |spin| <= mass")
    if gamma <= 1.0:
        raise ValueError("gamma must be greater than 1")
    if not (r_pressure_max > r_in > 0.0):
        raise ValueError("require 0 < r_in < r_pressure_max")
    horizon = mass + np.sqrt(mass * mass - spin * spin)
    if r_min is None:
        r_min = 1.2 * horizon
    if r_min <= horizon or r_max <= r_min:
        raise ValueError(f"require horizon ({horizon:.6g}) < r_min < r_max")

    r1 = np.logspace(np.log10(r_min), np.log10(r_max), nr)
    th1 = np.linspace(1e-3, np.pi - 1e-3, ntheta)
    ph1 = np.linspace(0.0, 2.0 * np.pi, nphi, endpoint=False)
    R, TH, PH = np.meshgrid(r1, th1, ph1, indexing="ij")

    a = spin / mass
    rh = R / mass
    ell = _kepler_ell(r_pressure_max / mass, a)
    ut, uph, u_t = _rigid_rotation(rh, TH, a, ell)
    _, _, u_t_in = _rigid_rotation(np.array(r_in / mass), np.array(np.pi / 2.0), a, ell)
    with np.errstate(invalid="ignore", divide="ignore"):
        h = u_t_in / u_t
    inside = (rh >= r_in / mass) & np.isfinite(h) & (h >= 1.0)

    rho = np.where(inside,
                   np.power(np.clip((gamma - 1.0) / gamma * (h - 1.0), 0.0, None),
                            1.0 / (gamma - 1.0)),
                   0.0)
    peak = rho.max()
    if peak > 0.0:
        rho /= peak
    tgas = np.where(inside, np.power(np.maximum(rho, 1e-300), gamma - 1.0), 0.0)
    rho = np.where(inside, rho, rho_floor)
    tgas = np.where(inside, tgas, rho_floor)

    # zero-angular-momentum observer outside the torus (create_example_data.py:217-227)
    gtt, gtp, gpp = _bl_metric(rh, TH, a)
    omega_z = -gtp / gpp
    lapse2 = gtp * gtp / gpp - gtt
    zamo_ut = 1.0 / np.sqrt(np.clip(lapse2, 1e-300, None))
    ut_f = np.where(inside, ut, zamo_ut)
    uph_f = np.where(inside, uph, omega_z * zamo_ut)

    vel = np.zeros((nr, ntheta, nphi, 4), dtype=np.float64)
    vel[..., 0] = ut_f
    vel[..., 3] = uph_f / mass
    mag = np.zeros_like(vel)
    out = dict(r=R, theta=TH, phi=PH, rho=rho, Tgas=tgas, vel=vel, mag=mag)
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
