"""The global loop's stop rule (integrate_cart_ks.hpp:468-527) pinned by
rays whose termination steps are known in closed form.

Flat space (M = 0, a = 0): the right-hand side is constant, every RK45
attempt is accepted with an error at the rounding level, so the step grows
by the maximum factor 10 after every step.  A ray leaving x = 1 along +x
with p = (-1, 1, 0, 0) and first step d0 is at x = 1 + d0 (10^j - 1) / 9
after j steps.  With d0 = 5 * 10^(7 - T) and r_max = 1e6 it is inside
(about 5.6e5) after T - 1 steps and outside (about 5.6e6) after T, so the
functor call of global step T terminates it.

For each (termination fraction, output interval, max_steps) the rule --
recount the terminated fraction only on steps with step % interval == 0 or
step == max_steps - 1, stop once it reaches the threshold -- gives the last
executed step by hand from the termination steps T below; the tables
are those hand results.  Also checked: which rays are still live, their
positions and steps, and the call / step counters.
"""
import numpy as np
import pytest

T_STEPS = np.array([1, 1, 2, 3, 3, 3, 5, 8, 8, 13])
R_MAX = 1.0e6

# terminated after step s: s=0:0, 1:2, 2:3, 3:6, 4:6, 5:7, 6-7:7, 8-12:9, 13:10
# (threshold, output interval, max_steps) -> last executed global step
STOP_TABLE = [
    ((1.0, 1, 20), 13),    # every step checked; all ten gone after step 13
    ((0.5, 1, 20), 3),     # 6/10 after step 3
    ((0.5, 4, 20), 4),     # checks at 0, 4: 6/10 at step 4
    ((0.7, 5, 20), 5),     # checks at 0, 5: 7/10 at step 5 (== threshold)
    ((0.95, 3, 11), 10),   # checks at 0, 3, 6, 9, 10: 9/10 at most -> all 11 steps
    ((0.2, 100, 12), 11),  # checks at 0 and max_steps - 1 only
    ((0.25, 2, 20), 2),    # 3/10 at step 2
    ((0.05, 1, 20), 1),    # 2/10 at step 1
    ((1.0, 7, 20), 14),    # checks at 0, 7, 14: the last ray left at 13, seen at 14
    ((1.0, 1, 10), 9),     # max_steps 10 < 13: runs out, one ray left
]


def flat_rays():
    """The ten rays as a numpy photon dict (disk phase, no fluid)."""
    import grt_amd
    n = len(T_STEPS)
    ph = grt_amd.empty_photons(n)
    ph["id"][:] = np.arange(n, dtype=np.uint64)
    ph["x1"][:] = 1.0
    ph["k0"][:] = -1.0
    ph["k1"][:] = 1.0
    ph["dlambda"][:] = 5.0 * 10.0 ** (7 - T_STEPS)
    ph["frequency"][:] = 1.0
    return ph


FLAT = dict(a_bh=0.0, m_bh=0.0, r_term_min=0.5, r_term_max=R_MAX, mode=1, integrator=1)


def expected(stop):
    """Per-ray live flag, accepted steps and calls when the loop's last
    executed step is `stop`."""
    live = T_STEPS > stop
    accepted = np.minimum(T_STEPS, stop + 1)
    calls = np.minimum(T_STEPS, stop) + 1
    return live, accepted, calls


def check_result(ph, st, stop, d0):
    live, accepted, calls = expected(stop)
    assert st["global_steps"] == stop + 1
    assert np.array_equal(ph["terminate"] == 0, live)
    assert st["terminated"] == int((~live).sum())
    assert st["accepted_steps"] == accepted.sum()
    assert st["functor_calls"] == calls.sum()
    assert st["rejected_attempts"] == 0 and st["failed_steps"] == 0
    # a live ray has taken stop + 1 steps: x = 1 + d0 (10^(stop+1) - 1) / 9
    j = stop + 1
    want_x = 1.0 + d0[live] * (10.0 ** j - 1.0) / 9.0
    assert np.allclose(ph["x1"][live], want_x, rtol=1e-12)
    assert np.allclose(ph["dlambda"][live], d0[live] * 10.0 ** j, rtol=1e-12)
    assert np.allclose(ph["x0"][live], ph["x1"][live] - 1.0, rtol=1e-12)   # t = x - 1
    # terminated rays stopped just outside r_max after T steps
    gone = ~live
    assert np.all(ph["x1"][gone] > R_MAX)
    assert np.allclose(ph["x1"][gone], 1.0 + d0[gone] * (10.0 ** T_STEPS[gone] - 1) / 9,
                       rtol=1e-12)


def test_flat_rays_close_form(oracle):
    """Without any stop, ray i terminates in the call of step T_i."""
    ph = flat_rays()
    d0 = ph["dlambda"].copy()
    st = oracle.trace(oracle.make_params(**FLAT), None, ph, 40, 1.0, 1, 1)
    assert np.all(ph["terminate"] == 1) and np.all(ph["phase"] == 4)
    check_result(ph, st, 13, d0)


@pytest.mark.parametrize("case,stop", STOP_TABLE)
def test_stop_rule(oracle, case, stop):
    thr, interval, max_steps = case
    ph = flat_rays()
    d0 = ph["dlambda"].copy()
    st = oracle.trace(oracle.make_params(**FLAT), None, ph, max_steps, thr, interval, 1)
    check_result(ph, st, stop, d0)
