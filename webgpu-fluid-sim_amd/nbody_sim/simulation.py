"""Host mirror of the reference ``Simulation`` (src/simulation.ts:39-433).

Same members, same meaning, over the C ABI of libnbody_sim.so:

    constructor(config, device)   -> Simulation(config, device=0)        simulation.ts:55-66
    setNumBodies(n)               -> setNumBodies(n)                     simulation.ts:227-231
    setScenario(s)                -> setScenario(s, seed=None)           simulation.ts:233-374
    updateMetadataBuffer()        -> updateMetadataBuffer()              simulation.ts:212-225
    getCommands()                 -> getCommands()  (enqueues substeps)  simulation.ts:386-425
    getBuffers()                  -> getBuffers()   (host read-back)     simulation.ts:427-429
    getNumBodies()                -> getNumBodies()                      simulation.ts:430-432

An unknown scenario raises like simulation.ts:364.  The random draws of the
scenarios come from a seeded Philox stream instead of Math.random.
"""
from __future__ import annotations

import ctypes
import random

import numpy as np

from . import _lib
from .config import Config, getDefaultConfig

SCENARIOS = {
    "default": _lib.SCENARIO_DEFAULT,
    "two-clusters": _lib.SCENARIO_TWO_CLUSTERS,
    "planetary-disk": _lib.SCENARIO_PLANETARY_DISK,
    "uniform-disk": _lib.SCENARIO_UNIFORM_DISK,
    "exp-disk": _lib.SCENARIO_EXP_DISK,
    "collision": _lib.SCENARIO_COLLISION,
    "plummer": _lib.SCENARIO_PLUMMER,
}


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.c_void_p)


def generate_radius(scenario: str, n: int) -> np.ndarray:
    """setScenario's radiusMultiplierData (simulation.ts:240-241, 321, 330, 339)."""
    if scenario not in SCENARIOS:
        raise ValueError(f"Unknown scenario: {scenario}")
    r = np.zeros(n, np.float32)
    _lib.check(_lib.lib().nbody_generate_radius(SCENARIOS[scenario], n, _ptr(r)))
    return r


def generate(scenario: str, n: int, seed: int = 0, grav_constant: float = 1.0,
             scale: float = 0.0, dispersion: float = 0.0, mass_sigma: float = 0.0):
    """Seeded initial conditions: (mass (n,), pos (n,2), vel (n,2)) float32."""
    if scenario not in SCENARIOS:
        raise ValueError(f"Unknown scenario: {scenario}")
    mass = np.zeros(n, np.float32)
    pos = np.zeros((n, 2), np.float32)
    vel = np.zeros((n, 2), np.float32)
    p = _lib.NbodyIcParams(grav_constant, scale, dispersion, mass_sigma)
    _lib.check(_lib.lib().nbody_generate(SCENARIOS[scenario], n, seed & 0xFFFFFFFFFFFFFFFF,
                                         ctypes.byref(p), _ptr(mass), _ptr(pos), _ptr(vel)))
    return mass, pos, vel


class _Handle:
    """Owns one nbody_sim handle; shared by Simulation and Ensemble."""

    def __init__(self, handle):
        self._h = handle

    @property
    def handle(self):
        if not self._h:
            raise _lib.NbodyError(_lib.ERR_STATE, "simulation is closed")
        return self._h

    def _check(self, rc):
        _lib.check(rc, self._h)

    def close(self):
        if self._h:
            _lib.lib().nbody_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- engine controls -------------------------------------------------
    def setBodies(self, mass, pos, vel):
        mass = np.ascontiguousarray(mass, np.float32)
        pos = np.ascontiguousarray(pos, np.float32).reshape(-1, 2)
        vel = np.ascontiguousarray(vel, np.float32).reshape(-1, 2)
        n = mass.shape[0]
        if pos.shape[0] != n or vel.shape[0] != n:
            raise ValueError("mass, pos and vel must describe the same number of bodies")
        self._check(_lib.lib().nbody_set_bodies(self.handle, n, _ptr(mass), _ptr(pos), _ptr(vel)))

    def getBodies(self):
        n = self.getNumBodies()
        mass = np.zeros(n, np.float32)
        pos = np.zeros((n, 2), np.float32)
        vel = np.zeros((n, 2), np.float32)
        self._check(_lib.lib().nbody_get_bodies(self.handle, _ptr(mass), _ptr(pos), _ptr(vel)))
        return mass, pos, vel

    def getNumBodies(self) -> int:
        return int(_lib.lib().nbody_get_num_bodies(self.handle))

    def step(self, nsteps: int = 1):
        self._check(_lib.lib().nbody_step(self.handle, int(nsteps)))

    def synchronize(self):
        self._check(_lib.lib().nbody_synchronize(self.handle))

    def useGraph(self, enable: bool = True, steps: int = 1):
        """Replay `steps` steps per captured HIP graph (enable False: no graph)."""
        self._check(_lib.lib().nbody_use_graph(self.handle, int(steps) if enable else 0))

    def prepareGraph(self):
        """Capture the step graph now, so the next step() replays it at once."""
        self._check(_lib.lib().nbody_prepare_graph(self.handle))

    def setRadiusMultiplier(self, r):
        """updateRadiusMultiplierBuffer (simulation.ts:382-384)."""
        r = np.ascontiguousarray(r, np.float32)
        if r.shape != (self.getNumBodies(),):
            raise ValueError("one radius multiplier per body")
        self._check(_lib.lib().nbody_set_radius_multiplier(self.handle, _ptr(r)))

    def getRadiusMultiplier(self) -> np.ndarray:
        r = np.zeros(self.getNumBodies(), np.float32)
        self._check(_lib.lib().nbody_get_radius_multiplier(self.handle, _ptr(r)))
        return r

    def getStepCount(self) -> int:
        k = ctypes.c_int64(0)
        self._check(_lib.lib().nbody_get_step_count(self.handle, ctypes.byref(k)))
        return int(k.value)

    def saveSnapshot(self, path):
        """Write mass/pos/vel/radius_multiplier/perm .npy files and meta.json."""
        self._check(_lib.lib().nbody_save_snapshot(self.handle, str(path).encode()))

    def loadSnapshot(self, path):
        """Restore config, bodies, radius multipliers and step count."""
        self._check(_lib.lib().nbody_load_snapshot(self.handle, str(path).encode()))
        self._after_load()

    def _after_load(self):
        pass

    def setDebug(self, accel: bool = False, stats: bool = False, fill_climb: bool = False,
                 sort_radix: bool = False):
        """Record accelerations / walk statistics; fill_climb forces the tree
        fill's global-counter climb for its upper levels, sort_radix the
        radix passes of the sort for every step (test hooks)."""
        flags = ((_lib.DEBUG_ACCEL if accel else 0) | (_lib.DEBUG_STATS if stats else 0)
                 | (_lib.DEBUG_FILL_CLIMB if fill_climb else 0)
                 | (_lib.DEBUG_SORT_RADIX if sort_radix else 0))
        self._check(_lib.lib().nbody_set_debug(self.handle, flags))

    def getTuning(self) -> dict:
        t = _lib.NbodyTuning()
        self._check(_lib.lib().nbody_get_tuning(self.handle, ctypes.byref(t)))
        return {f: getattr(t, f) for f, _ in _lib.NbodyTuning._fields_}

    def setTuning(self, walk_block=None, walk_stack=None, walk_variant=None, xcd_order=None,
                  lane_bodies=None, radix_bits=None, wave_bodies=None):
        """Launch tunables (nbody_tuning); None keeps a value.  Every setting takes
        the same decisions; only walk_variant 2 (fast mode) sums in another order."""
        cur = self.getTuning()
        for k, v in (("walk_block", walk_block), ("walk_stack", walk_stack),
                     ("walk_variant", walk_variant), ("xcd_order", xcd_order),
                     ("lane_bodies", lane_bodies), ("radix_bits", radix_bits),
                     ("wave_bodies", wave_bodies)):
            if v is not None:
                cur[k] = int(v)
        t = _lib.NbodyTuning(**cur)
        self._check(_lib.lib().nbody_set_tuning(self.handle, ctypes.byref(t)))

    def setProfiling(self, enable: bool = True):
        self._check(_lib.lib().nbody_set_profiling(self.handle, int(bool(enable))))

    def phaseTimes(self, reset: bool = False):
        ms = (ctypes.c_float * 6)()
        steps = ctypes.c_int64(0)
        self._check(_lib.lib().nbody_phase_times(self.handle, ms, ctypes.byref(steps), int(reset)))
        return dict(zip(_lib.PHASES, list(ms))), int(steps.value)

    def walkStats(self, reset: bool = False):
        out = (ctypes.c_uint64 * 3)()
        self._check(_lib.lib().nbody_walk_stats(self.handle, out, int(reset)))
        return {"accepted": int(out[0]), "opened": int(out[1]), "overflowed": int(out[2])}

    # -- adjoint ------------------------------------------------------------
    def adjointForward(self, nsteps: int, every: int = 20):
        """Run nsteps steps keeping a checkpoint every `every` steps."""
        self._check(_lib.lib().nbody_adjoint_forward(self.handle, int(nsteps), int(every)))

    def adjointBackward(self, gpos_T=None, gvel_T=None):
        """dL/d(pos0, vel0, mass) from dL/d(pos_T, vel_T), decisions frozen."""
        n = self.getNumBodies()
        gp = np.zeros((n, 2), np.float32)
        gv = np.zeros((n, 2), np.float32)
        gm = np.zeros(n, np.float32)
        a = None if gpos_T is None else np.ascontiguousarray(gpos_T, np.float32).reshape(n, 2)
        b = None if gvel_T is None else np.ascontiguousarray(gvel_T, np.float32).reshape(n, 2)
        self._check(_lib.lib().nbody_adjoint_backward(
            self.handle, None if a is None else _ptr(a), None if b is None else _ptr(b),
            _ptr(gp), _ptr(gv), _ptr(gm)))
        return gp, gv, gm

    def directAccelF64(self, targets=None) -> np.ndarray:
        """Binary64 softened direct-sum accelerations (ntargets, 2) of the
        current bodies at `targets` (None: all): the walk's accuracy reference."""
        n = self.getNumBodies()
        if targets is None:
            out = np.zeros((n, 2), np.float64)
            self._check(_lib.lib().nbody_direct_accel_f64(self.handle, None, n, _ptr(out)))
            return out
        tg = np.ascontiguousarray(targets, np.int64)
        out = np.zeros((tg.shape[0], 2), np.float64)
        self._check(_lib.lib().nbody_direct_accel_f64(self.handle, _ptr(tg), tg.shape[0],
                                                      _ptr(out)))
        return out

    def debugBuffer(self, which: str) -> np.ndarray:
        """Intermediate buffers of the last step, as numpy arrays."""
        n = self.getNumBodies()
        nn = max(2 * n - 1, 0)
        ni = max(n - 1, 0)
        ens = getattr(self, "numSims", None)
        s = ens if ens is not None else 1
        kdt = np.uint64 if ens is not None else np.uint32
        spec = {
            "keys": (_lib.BUF_KEYS, np.zeros(n, kdt)),
            "body_indices": (_lib.BUF_BODY_INDICES, np.zeros(n, np.uint32)),
            "children": (_lib.BUF_CHILDREN, np.zeros((ni, 2), np.uint32)),
            "parent": (_lib.BUF_PARENT, np.zeros(nn, np.uint32)),
            "node_cm": (_lib.BUF_NODE_CM, np.zeros((nn, 4), np.float32)),
            "node_aabb": (_lib.BUF_NODE_AABB, np.zeros((nn, 4), np.float32)),
            "accel": (_lib.BUF_ACCEL, np.zeros((n, 2), np.float32)),
            "bbox": (_lib.BUF_BBOX, np.zeros((s, 4), np.float32)),
            "sim_roots": (_lib.BUF_SIM_ROOTS, np.zeros(s, np.uint32)),
            "walk_counts": (_lib.BUF_WALK_COUNTS, np.zeros((n, 2), np.uint32)),
            "tone_table": (_lib.BUF_TONE_TABLE, np.zeros(65536, np.uint32)),
            "node_quad": (_lib.BUF_NODE_QUAD, np.zeros((nn, 3), np.float32)),
        }
        if which not in spec:
            raise ValueError(f"unknown buffer {which}")
        code, arr = spec[which]
        self._check(_lib.lib().nbody_debug_get(self.handle, code, _ptr(arr), arr.nbytes))
        return arr


class Simulation(_Handle):
    """simulation.ts:39 -- one simulation on one GPU."""

    def __init__(self, config: Config | None = None, device: int = 0, capacity: int = 0,
                 seed: int = 0):
        self.config = config or getDefaultConfig()
        self.currentScenario = "default"   # simulation.ts:59
        self.numBodies = 50000             # simulation.ts:60
        self._seed = seed
        self._rng = random.Random(seed)
        err = ctypes.c_int(0)
        h = _lib.lib().nbody_create(ctypes.byref(self.config.to_c()), int(capacity),
                                    int(device), ctypes.byref(err))
        if not h:
            raise _lib.NbodyError(err.value, _lib.last_error(None))
        super().__init__(h)

    # -- reference surface -------------------------------------------------
    def updateMetadataBuffer(self):
        """simulation.ts:212-225; offset = floor(random * N) (simulation.ts:222)."""
        off = int(self._rng.random() * max(self.numBodies, 1))
        self.config = self.config.with_(offset=off)
        self._check(_lib.lib().nbody_set_config(self.handle, ctypes.byref(self.config.to_c())))

    def setConfig(self, config: Config):
        self.config = config
        self._check(_lib.lib().nbody_set_config(self.handle, ctypes.byref(config.to_c())))

    def setNumBodies(self, numBodies: int):
        """simulation.ts:227-231."""
        self.numBodies = int(numBodies)
        self.updateMetadataBuffer()
        self.setScenario(self.currentScenario)

    def setScenario(self, scenario: str, seed: int | None = None, **ic):
        """simulation.ts:233-374: regenerate bodies, upload, refresh metadata."""
        if scenario not in SCENARIOS:
            raise ValueError(f"Unknown scenario: {scenario}")
        self.currentScenario = scenario
        s = self._seed if seed is None else seed
        mass, pos, vel = generate(scenario, self.numBodies, s,
                                  grav_constant=self.config.gravConstant, **ic)
        self.updateMetadataBuffer()
        self.setBodies(mass, pos, vel)
        self.setRadiusMultiplier(generate_radius(scenario, self.numBodies))  # simulation.ts:373

    def getCommands(self):
        """simulation.ts:386-425 + queue.submit (main.ts:38-41): substeps steps."""
        self.step(self.config.substeps)

    def getBuffers(self):
        """simulation.ts:427-429, read back to host."""
        mass, pos, vel = self.getBodies()
        return {"mass": mass, "pos": pos, "vel": vel,
                "radiusMultiplier": self.getRadiusMultiplier()}

    def getNumBodies(self) -> int:
        return super().getNumBodies()

    def setBodies(self, mass, pos, vel):
        super().setBodies(mass, pos, vel)
        self.numBodies = int(np.asarray(mass).shape[0])

    def _after_load(self):
        self.numBodies = self.getNumBodies()
        c = _lib.NbodyConfig()
        self._check(_lib.lib().nbody_get_config(self.handle, ctypes.byref(c)))
        self.config = Config.from_c(c)


class Ensemble(_Handle):
    """Many independent simulations stepped by one sort, one build, one walk."""

    def __init__(self, config: Config, n_per_sim, device: int = 0):
        self.config = config
        n_per_sim = np.ascontiguousarray(n_per_sim, np.int64)
        self.numSims = int(n_per_sim.shape[0])
        self.nPerSim = n_per_sim.copy()
        self.simOffset = np.concatenate([[0], np.cumsum(n_per_sim)]).astype(np.int64)
        err = ctypes.c_int(0)
        h = _lib.lib().nbody_ensemble_create(
            ctypes.byref(config.to_c()), self.numSims,
            n_per_sim.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), int(device),
            ctypes.byref(err))
        if not h:
            raise _lib.NbodyError(err.value, _lib.last_error(None))
        super().__init__(h)

    def setConfig(self, config: Config):
        self.config = config
        self._check(_lib.lib().nbody_set_config(self.handle, ctypes.byref(config.to_c())))
