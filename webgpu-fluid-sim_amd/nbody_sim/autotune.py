"""Launch autotuner for the tree walk.

Times each candidate (walk variant: 0 LDS stack, 1 register stack;
workgroup size; stack entries) on the
handle's own bodies with the engine's HIP-event phase timers, restores the
bodies after every trial, and applies the fastest.  Every candidate computes
the same result (the tunables only change how wavefronts are laid out and
how deep the LDS stack is), so tuning never changes the physics.
"""
from __future__ import annotations

import itertools

DEFAULT_VARIANTS = (1, 0)
DEFAULT_XCD = (1, 0)
DEFAULT_BLOCKS = (64, 128, 256, 512)
DEFAULT_STACKS = (64, 128)


def autotune(sim, steps: int = 3, variants=DEFAULT_VARIANTS, blocks=DEFAULT_BLOCKS,
             stacks=DEFAULT_STACKS, xcd=DEFAULT_XCD, verbose: bool = False):
    """Pick the fastest walk configuration for `sim` (Simulation or Ensemble).

    Returns (best, table) where best = (variant, block, stack, xcd_order) and table maps
    every tried configuration to its mean walk time in ms."""
    mass, pos, vel = sim.getBodies()
    table = {}
    sim.setProfiling(True)
    try:
        for v, b, st, x in itertools.product(variants, blocks, stacks, xcd):
            sim.setTuning(walk_block=b, walk_stack=st, walk_variant=v, xcd_order=x)
            sim.setBodies(mass, pos, vel)
            sim.step(1)  # warm-up
            sim.synchronize()
            sim.phaseTimes(reset=True)
            sim.step(steps)
            sim.synchronize()
            ms, n = sim.phaseTimes(reset=True)
            table[(v, b, st, x)] = ms["walk"] / max(n, 1)
            if verbose:
                print(f"variant {v} block {b} stack {st} xcd {x}: walk "
                      f"{table[(v, b, st, x)]:.4f} ms", flush=True)
    finally:
        sim.setProfiling(False)
        sim.setBodies(mass, pos, vel)
    best = min(table, key=table.get)
    sim.setTuning(walk_block=best[1], walk_stack=best[2], walk_variant=best[0],
                  xcd_order=best[3])
    return best, table
