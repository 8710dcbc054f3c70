"""Launch autotuner for the tree walk and the radix sort.

Times each candidate (walk variant: 0 LDS stack, 1 register stack, 2 pair
walk, 4 pair walk run from a persistent work queue; bodies per lane of the
register-stack walk; bodies per wavefront of the pair walks; workgroup size;
stack entries; XCD order; then the radix sort's digit width) on the handle's own bodies with the engine's
HIP-event phase timers, restores the bodies after every trial, and applies
the fastest.  Every candidate takes the same
opening decisions and the sort is stable for every width; walk variants 0
and 1 add each body's terms in the reference's order and agree bit for bit,
the pair walk (fast mode) adds the same terms in its own order, so tuning
never changes the physics beyond rounding.
"""
from __future__ import annotations

import itertools

DEFAULT_VARIANTS = (2, 4, 1, 0)
DEFAULT_WAVE_BODIES = (64, 32, 16, 8)
DEFAULT_LANE_BODIES = (1, 2, 4)
DEFAULT_XCD = (32, 1, 0, 16)
DEFAULT_BLOCKS = (64, 128, 256, 512)
DEFAULT_STACKS = (64, 128)


def _candidates(variants, blocks, stacks, xcd, lane_bodies, wave_bodies):
    for v, b, st, x in itertools.product(variants, blocks, stacks, xcd):
        # the LDS-stack and pair walks have one body per lane, and the pair
        # walk's stack is fixed; only the pair walks share a wavefront among
        # fewer than 64 bodies
        if v in (2, 4) and st != stacks[0]:
            continue
        for lb in (lane_bodies if v == 1 else (1,)):
            for wb in (wave_bodies if v in (2, 4) else (64,)):
                yield (v, b, st, x, lb, wb)


DEFAULT_RADIX = (10, 8, 11)


def _time(sim, mass, pos, vel, steps, phases):
    """Mean device time (ms) per step of the given phases, from the bodies
    (mass, pos, vel), after one warm-up step."""
    sim.setBodies(mass, pos, vel)
    sim.step(1)
    sim.synchronize()
    sim.phaseTimes(reset=True)
    sim.step(steps)
    sim.synchronize()
    ms, n = sim.phaseTimes(reset=True)
    return sum(ms[p] for p in phases) / max(n, 1)


def autotune(sim, steps: int = 3, variants=DEFAULT_VARIANTS, blocks=DEFAULT_BLOCKS,
             stacks=DEFAULT_STACKS, xcd=DEFAULT_XCD, lane_bodies=DEFAULT_LANE_BODIES,
             radix=DEFAULT_RADIX, wave_bodies=DEFAULT_WAVE_BODIES, verbose: bool = False):
    """Pick the fastest walk configuration and radix width for `sim`
    (Simulation or Ensemble).

    The walk tunables are timed on the walk phase, the radix width on the
    encode and sort phases (whose digit counts and passes it sets).  Returns
    (best, table) where best = (variant, block, stack, xcd_order,
    lane_bodies, wave_bodies, radix_bits) and table maps every tried walk
    configuration (variant, block, stack, xcd_order, lane_bodies,
    wave_bodies) to its mean walk time and every ("radix_bits", r) to its mean
    encode + sort time, in ms.  Ensembles keep 64 bodies per wavefront."""
    if getattr(sim, "numSims", None) is not None:
        wave_bodies = (64,)
    mass, pos, vel = sim.getBodies()
    table = {}
    sim.setProfiling(True)
    try:
        for cand in _candidates(variants, blocks, stacks, xcd, lane_bodies, wave_bodies):
            v, b, st, x, lb, wb = cand
            sim.setTuning(walk_block=b, walk_stack=st, walk_variant=v, xcd_order=x,
                          lane_bodies=lb, wave_bodies=wb)
            table[cand] = _time(sim, mass, pos, vel, steps, ("walk",))
            if verbose:
                print(f"variant {v} block {b} stack {st} xcd {x} lane_bodies {lb} "
                      f"wave_bodies {wb}: walk {table[cand]:.4f} ms", flush=True)
        walk = min(table, key=table.get)
        sim.setTuning(walk_block=walk[1], walk_stack=walk[2], walk_variant=walk[0],
                      xcd_order=walk[3], lane_bodies=walk[4], wave_bodies=walk[5])
        for r in radix:
            sim.setTuning(radix_bits=r)
            table[("radix_bits", r)] = _time(sim, mass, pos, vel, steps, ("encode", "sort"))
            if verbose:
                print(f"radix_bits {r}: encode + sort {table[('radix_bits', r)]:.4f} ms",
                      flush=True)
        rb = min(radix, key=lambda r: table[("radix_bits", r)])
        sim.setTuning(radix_bits=rb)
    finally:
        sim.setProfiling(False)
        sim.setBodies(mass, pos, vel)
    return walk + (rb,), table
