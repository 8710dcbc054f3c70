"""Config surface of the reference (src/config.ts:1-22).

``Config`` keeps the reference's field names (gravConstant, bhTheta,
substeps, deltaTime, epsilonMultiplier) and adds the optional fields of the
MI355X engine: a constant softening length, the Morton key mode and the walk
arithmetic mode.  Defaults reproduce getDefaultConfig() (config.ts:9-22).
"""
from __future__ import annotations

from dataclasses import dataclass, replace

from . import _lib


@dataclass(frozen=True)
class Config:
    gravConstant: float = 1.0
    bhTheta: float = 0.6
    substeps: int = 1
    deltaTime: float = 0.1 * 1.0 / (60.0 * 1)
    epsilonMultiplier: float = 1.0
    # engine extensions (not in config.ts)
    epsilon: float = 1e-3
    softeningMode: str = "reference"  # "reference" (barnes_hut.wgsl:90-91) | "constant"
    mortonMode: str = "reference"     # "reference" (lbvh.wgsl:48-65) | "bbox30"
    mathMode: str = "fast"            # "fast" | "exact"
    offset: int = 0                   # Metadata.offset (simulation.ts:222)
    expansion: str = "monopole"       # "monopole" (barnes_hut.wgsl:92-97) | "quadrupole"

    def to_c(self) -> _lib.NbodyConfig:
        soft = {"reference": _lib.SOFTEN_REFERENCE, "constant": _lib.SOFTEN_CONSTANT}
        mort = {"reference": _lib.MORTON_REFERENCE, "bbox30": _lib.MORTON_BBOX30}
        math = {"fast": _lib.MATH_FAST, "exact": _lib.MATH_EXACT}
        expn = {"monopole": _lib.EXPANSION_MONOPOLE, "quadrupole": _lib.EXPANSION_QUADRUPOLE}
        try:
            return _lib.NbodyConfig(self.gravConstant, self.bhTheta, int(self.substeps),
                                    self.deltaTime, self.epsilonMultiplier, self.epsilon,
                                    soft[self.softeningMode], mort[self.mortonMode],
                                    math[self.mathMode], int(self.offset) & 0xFFFFFFFF,
                                    expn[self.expansion])
        except KeyError as e:
            raise ValueError(f"unknown config mode {e}") from None

    @staticmethod
    def from_c(c: _lib.NbodyConfig) -> "Config":
        soft = {_lib.SOFTEN_REFERENCE: "reference", _lib.SOFTEN_CONSTANT: "constant"}
        mort = {_lib.MORTON_REFERENCE: "reference", _lib.MORTON_BBOX30: "bbox30"}
        math = {_lib.MATH_FAST: "fast", _lib.MATH_EXACT: "exact"}
        expn = {_lib.EXPANSION_MONOPOLE: "monopole", _lib.EXPANSION_QUADRUPOLE: "quadrupole"}
        return Config(gravConstant=c.grav_constant, bhTheta=c.bh_theta, substeps=c.substeps,
                      deltaTime=c.delta_time, epsilonMultiplier=c.epsilon_multiplier,
                      epsilon=c.epsilon, softeningMode=soft[c.softening_mode],
                      mortonMode=mort[c.morton_mode], mathMode=math[c.math_mode],
                      offset=c.offset, expansion=expn[c.expansion])

    def with_(self, **kw) -> "Config":
        return replace(self, **kw)


def getDefaultConfig() -> Config:
    """config.ts:9-22."""
    substeps = 1
    return Config(gravConstant=1.0, bhTheta=0.6, substeps=substeps,
                  deltaTime=0.1 * 1.0 / (60.0 * substeps), epsilonMultiplier=1.0)
