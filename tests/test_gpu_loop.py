"""The device driver's global loop against the closed-form flat-space
rays of test_oracle_loop: the last executed step, the live set, positions,
step sizes and counters for every (threshold, interval, max_steps) of the
hand-derived table, under the default launch shape and under a step budget
of 3 calls per launch (so every epoch is split into several compacting
launches)."""
import numpy as np
import pytest

from test_oracle_loop import FLAT, STOP_TABLE, check_result, flat_rays

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("budget", [None, 3])
def test_stop_rule_on_the_device(gpu, budget):
    import torch
    import grt_amd
    ctx = grt_amd.Context(grt_amd.default_params(**FLAT))
    if budget:
        ctx.tuning = dict(sub_budget=budget)
    for (thr, interval, max_steps), stop in [((1.0, 1, 40), 13)] + STOP_TABLE:
        ph = flat_rays()
        d0 = ph["dlambda"].copy()
        g = grt_amd.to_device(ph)
        st = ctx.trace(g, max_steps, thr, interval)
        torch.cuda.synchronize()
        h = grt_amd.to_host(g)
        print((thr, interval, max_steps), st)
        check_result(h, st, stop, d0)
