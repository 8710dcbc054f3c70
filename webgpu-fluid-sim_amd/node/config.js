// config.js -- the reference's Config surface (src/config.ts:1-22).
"use strict";

/** getDefaultConfig(): config.ts:9-22. */
function getDefaultConfig() {
    const gravConstant = 1.0;
    const bhTheta = 0.6;
    const substeps = 1;
    const deltaTime = 0.1 * 1.0 / (60.0 * substeps);
    const epsilonMultiplier = 1.0;
    return Object.freeze({ gravConstant, bhTheta, substeps, deltaTime, epsilonMultiplier });
}

// Engine extensions (not in config.ts): numeric codes of include/nbody_sim.h.
const SofteningMode = Object.freeze({ reference: 0, constant: 1 });
const MortonMode = Object.freeze({ reference: 0, bbox30: 1 });
const MathMode = Object.freeze({ fast: 0, exact: 1 });

module.exports = { getDefaultConfig, SofteningMode, MortonMode, MathMode };
