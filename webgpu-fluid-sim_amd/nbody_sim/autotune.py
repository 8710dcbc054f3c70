"""Launch autotuner for the tree walk.

Times each candidate (walk variant: 0 LDS stack, 1 register stack, 2 pair
walk; bodies per lane of the register-stack walk; workgroup size; stack
entries; XCD order) on the handle's own bodies with the engine's HIP-event
phase timers, restores the bodies after every trial, and applies the
fastest.  Every candidate takes the same opening decisions; variants 0 and 1
add each body's terms in the reference's order and agree bit for bit, the
pair walk (fast mode) adds the same terms in its own order, so tuning never
changes the physics beyond rounding.
"""
from __future__ import annotations

import itertools

DEFAULT_VARIANTS = (2, 1, 0)
DEFAULT_LANE_BODIES = (1, 2, 4)
DEFAULT_XCD = (1, 0)
DEFAULT_BLOCKS = (64, 128, 256, 512)
DEFAULT_STACKS = (64, 128)


def _candidates(variants, blocks, stacks, xcd, lane_bodies):
    for v, b, st, x in itertools.product(variants, blocks, stacks, xcd):
        # the LDS-stack and pair walks have one body per lane, and the pair
        # walk's stack is fixed
        if v == 2 and st != stacks[0]:
            continue
        for lb in (lane_bodies if v == 1 else (1,)):
            yield (v, b, st, x, lb)


def autotune(sim, steps: int = 3, variants=DEFAULT_VARIANTS, blocks=DEFAULT_BLOCKS,
             stacks=DEFAULT_STACKS, xcd=DEFAULT_XCD, lane_bodies=DEFAULT_LANE_BODIES,
             verbose: bool = False):
    """Pick the fastest walk configuration for `sim` (Simulation or Ensemble).

    Returns (best, table) where best = (variant, block, stack, xcd_order,
    lane_bodies) and table maps every tried configuration to its mean walk
    time in ms."""
    mass, pos, vel = sim.getBodies()
    table = {}
    sim.setProfiling(True)
    try:
        for cand in _candidates(variants, blocks, stacks, xcd, lane_bodies):
            v, b, st, x, lb = cand
            sim.setTuning(walk_block=b, walk_stack=st, walk_variant=v, xcd_order=x,
                          lane_bodies=lb)
            sim.setBodies(mass, pos, vel)
            sim.step(1)  # warm-up
            sim.synchronize()
            sim.phaseTimes(reset=True)
            sim.step(steps)
            sim.synchronize()
            ms, n = sim.phaseTimes(reset=True)
            table[cand] = ms["walk"] / max(n, 1)
            if verbose:
                print(f"variant {v} block {b} stack {st} xcd {x} lane_bodies {lb}: walk "
                      f"{table[cand]:.4f} ms", flush=True)
    finally:
        sim.setProfiling(False)
        sim.setBodies(mass, pos, vel)
    best = min(table, key=table.get)
    sim.setTuning(walk_block=best[1], walk_stack=best[2], walk_variant=best[0],
                  xcd_order=best[3], lane_bodies=best[4])
    return best, table
