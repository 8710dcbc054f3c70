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

DEFAULT_VARIANTS = (0, 1)
DEFAULT_BLOCKS = (64, 128, 256, 512)
DEFAULT_STACKS = (64, 128)


def autotune(sim, steps: int = 3, variants=DEFAULT_VARIANTS, blocks=DEFAULT_BLOCKS,
             stacks=DEFAULT_STACKS, verbose: bool = False):
    """Pick the fastest walk configuration for `sim` (Simulation or Ensemble).

    Returns (best, table) where best = (variant, block, stack) and table maps
    every tried configuration to its mean walk time in ms."""
    mass, pos, vel = sim.getBodies()
    table = {}
    sim.setProfiling(True)
    try:
        for v, b, st in itertools.product(variants, blocks, stacks):
            sim.setTuning(walk_block=b, walk_stack=st, walk_variant=v)
            sim.setBodies(mass, pos, vel)
            sim.step(1)  # warm-up
            sim.synchronize()
            sim.phaseTimes(reset=True)
            sim.step(steps)
            sim.synchronize()
            ms, n = sim.phaseTimes(reset=True)
            table[(v, b, st)] = ms["walk"] / max(n, 1)
            if verbose:
                print(f"variant {v} block {b} stack {st}: walk {table[(v, b, st)]:.4f} ms")
    finally:
        sim.setProfiling(False)
        sim.setBodies(mass, pos, vel)
    best = min(table, key=table.get)
    sim.setTuning(walk_block=best[1], walk_stack=best[2], walk_variant=best[0])
    return best, table
