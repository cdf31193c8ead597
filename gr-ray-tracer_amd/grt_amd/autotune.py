"""Launch-shape autotuner of the persistent tracing kernel.

Three knobs decide how the kernel spends the chip (grt_tuning):
  block          threads per workgroup (64 ... 256); every workgroup stages the
                 separable search tables in LDS once
  blocks_per_cu  resident workgroups per CU in the persistent grid; with the
                 kernel's register budget a CU holds 8 waves, so
                 block * blocks_per_cu = 512 fills it
  sub_budget     functor calls per photon per launch (the refill interval):
                 photons still live after it are compacted into the next
                 launch's work list, so long paths stop holding otherwise idle
                 wavefronts; each launch costs a host round trip

`autotune` traces copies of a representative batch with every candidate,
keeps the fastest and installs it on the context.  The photon results do not
depend on the launch shape (each photon's calls are the same whichever lane
and launch runs them).
"""
import itertools
import time

DEFAULT_SHAPES = ((256, 2), (128, 4), (64, 8), (256, 1), (128, 2))
DEFAULT_BUDGETS = (32, 64, 128, 256, 1 << 20)


def autotune(ctx, photons, max_steps, shapes=DEFAULT_SHAPES, budgets=DEFAULT_BUDGETS,
             repeats=2, termination_percent=1.0, output_interval=0):
    """Time ctx.trace on fresh copies of `photons` (dict of GPU tensors, not
    modified) for every (block, blocks_per_cu) x sub_budget candidate; install
    the fastest on ctx and return (best, table) where table lists
    (tuning, best milliseconds of `repeats` runs) from fastest to slowest."""
    import torch
    table = []
    for (block, per_cu), budget in itertools.product(shapes, budgets):
        tune = dict(block=block, blocks_per_cu=per_cu, sub_budget=budget)
        ctx.tuning = tune
        best = float("inf")
        for _ in range(repeats + 1):          # the first run warms up
            ph = {k: v.clone() for k, v in photons.items()}
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            ctx.trace(ph, max_steps, termination_percent, output_interval)
            torch.cuda.synchronize()
            best = min(best, time.perf_counter() - t0)
            del ph
        table.append((tune, best * 1e3))
    table.sort(key=lambda t: t[1])
    ctx.tuning = table[0][0]
    return table[0][0], table
