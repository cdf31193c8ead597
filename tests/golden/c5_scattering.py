"""The batched scattering Monte Carlo workload (disk mode) and the histograms
its packets are compared by: a = 0.5 torus (r_in = 6, r_pmax = 12) on a
256 x 128 x 256 grid (r geometric on [2, 100]), packets emitted from the
torus cells (normalisation NORM, about 1e6 packets), isotropic elastic
grey scattering in the fluid tetrad with albedo 1, absorption and emission
on (coefficients 1, density scale 0.125), RK45 rtol = atol = 1e-7, packets
escaping at r = 1000.

Shared by make_c5_oracle_histograms.py (which writes the oracle's
histograms, with an independent seed) and the GPU test that compares the
device's packets with them.
"""
import numpy as np

GRID = dict(nr=256, ntheta=128, nphi=256, r_min=2.0, r_max=100.0, mass=1.0, spin=0.5,
            r_in=6.0, r_pressure_max=12.0)
NORM = 3.13          # superphoton normalisation: 1,003,008 packets on this grid
DENSITY_SCALE = 0.125
R_ESCAPE = 1000.0
MAX_STEPS = 20000


def params(seed):
    return dict(a_bh=0.5, m_bh=1.0, r_term_min=1.0 + np.sqrt(0.75), r_term_max=R_ESCAPE,
                atol=1e-7, rtol=1e-7, integrator=1, mode=1, absorption_enabled=1,
                emission_enabled=1, absorption_coefficient=1.0, emission_coefficient=1.0,
                scattering_enabled=1, albedo=1.0, seed=seed,
                superphoton_count_normalization=NORM)


COS_EDGES = np.linspace(-1.0, 1.0, 21)
PHI_EDGES = np.linspace(0.0, 2.0 * np.pi, 17)
# escaped packets hold log10 I in about [-0.38, 0.01] and log10 nu in about
# [9.3, 11.5]; the outer bins catch the rest
LOGI_EDGES = np.concatenate([[-np.inf], np.linspace(-0.45, 0.05, 26), [np.inf]])
LOGNU_EDGES = np.concatenate([[-np.inf], np.linspace(9.0, 11.8, 29), [np.inf]])


def outcome_histograms(ph, r_escape=R_ESCAPE):
    """Category counts (escaped, captured or zeroed, still live) and, for the
    escaped packets, histograms of the escape direction (cos theta and phi
    of the final position) and of log10 I and log10 frequency."""
    x, y, z = ph["x1"], ph["x2"], ph["x3"]
    r = np.sqrt(x * x + y * y + z * z)
    term = ph["terminate"] != 0
    esc = term & (r > r_escape) & (ph["I"] > 0)
    live = ~term
    other = term & ~esc
    out = dict(categories=np.array([esc.sum(), other.sum(), live.sum()]))
    cos_t = z[esc] / r[esc]
    phi = np.mod(np.arctan2(y[esc], x[esc]), 2 * np.pi)
    out["cos_theta"] = np.histogram(cos_t, COS_EDGES)[0]
    out["phi"] = np.histogram(phi, PHI_EDGES)[0]
    out["log_I"] = np.histogram(np.log10(ph["I"][esc]), LOGI_EDGES)[0]
    nu = ph["frequency"][esc]
    out["log_nu"] = np.histogram(np.log10(nu[nu > 0]), LOGNU_EDGES)[0]
    return out


def chi2_two_sample(a, b, min_count=5):
    """Two-sample chi-squared of histograms a, b (bins merged until each
    holds min_count of a + b): (statistic, degrees of freedom, p)."""
    from scipy import stats
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    ma, mb, ca, cb = [], [], 0.0, 0.0
    for x, y in zip(a, b):
        ca += x
        cb += y
        if ca + cb >= min_count:
            ma.append(ca)
            mb.append(cb)
            ca = cb = 0.0
    if ca + cb > 0 and ma:
        ma[-1] += ca
        mb[-1] += cb
    a, b = np.array(ma), np.array(mb)
    na, nb = a.sum(), b.sum()
    k1, k2 = np.sqrt(nb / na), np.sqrt(na / nb)
    chi2 = float(np.sum((k1 * a - k2 * b) ** 2 / (a + b)))
    dof = len(a) - 1
    return chi2, dof, float(stats.chi2.sf(chi2, dof)) if dof > 0 else 1.0
