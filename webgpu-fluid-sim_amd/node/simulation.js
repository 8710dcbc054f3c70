// simulation.js -- the reference's Simulation class (src/simulation.ts:39-433)
// over the C ABI, through the N-API addon nbody_addon.node.
//
//   constructor(config, device)  simulation.ts:55-66
//   setNumBodies(n)              simulation.ts:227-231
//   setScenario(s)               simulation.ts:233-374 (seeded Philox draws)
//   updateMetadataBuffer()       simulation.ts:212-225
//   getCommands()                simulation.ts:386-425 + queue.submit (main.ts:38-41):
//                                enqueues config.substeps steps on the GPU
//   getBuffers()                 simulation.ts:427-429, read back as Float32Arrays
//                                (mass, pos, vel, radiusMultiplier)
//   getNumBodies()               simulation.ts:430-432
"use strict";

const path = require("path");
const { getDefaultConfig } = require("./config");

const addon = require(path.join(__dirname, "nbody_addon.node"));

const SCENARIOS = Object.freeze({
    "default": 0,
    "two-clusters": 1,
    "planetary-disk": 2,
    "uniform-disk": 3,
    "exp-disk": 4,
    "collision": 5,
    "plummer": 6,
});

/** Seeded initial conditions: {mass: Float32Array(n), pos/vel: Float32Array(2n)}. */
function generate(scenario, n, seed = 0, params = {}) {
    if (!(scenario in SCENARIOS)) throw new Error(`Unknown scenario: ${scenario}`);
    return addon.generate(SCENARIOS[scenario], n, seed, params);
}

class Simulation {
    constructor(config, device = 0, seed = 0) {
        this.config = config || getDefaultConfig();
        this.device = device;
        this.currentScenario = "default";  // simulation.ts:59
        this.numBodies = 50000;            // simulation.ts:60
        this.seed = seed;
        this.rngState = (seed >>> 0) || 1;
        this.handle = addon.create(this.config, 0, device);
    }

    // xorshift32 for the metadata offset (simulation.ts:222 uses Math.random)
    random() {
        let x = this.rngState;
        x ^= x << 13; x >>>= 0;
        x ^= x >>> 17;
        x ^= x << 5; x >>>= 0;
        this.rngState = x;
        return x / 4294967296;
    }

    updateMetadataBuffer() {
        const offset = Math.floor(this.random() * this.numBodies);
        addon.setConfig(this.handle, Object.assign({}, this.config, { offset }));
    }

    setNumBodies(numBodies) {
        this.numBodies = numBodies;
        this.updateMetadataBuffer();
        this.setScenario(this.currentScenario);
    }

    setScenario(scenario, seed) {
        if (!(scenario in SCENARIOS)) throw new Error(`Unknown scenario: ${scenario}`);
        this.currentScenario = scenario;
        const b = generate(scenario, this.numBodies, seed === undefined ? this.seed : seed,
                           { gravConstant: this.config.gravConstant });
        this.updateMetadataBuffer();
        addon.setBodies(this.handle, b.mass, b.pos, b.vel);
        // updateRadiusMultiplierBuffer (simulation.ts:373, 382-384)
        addon.setRadius(this.handle, addon.generateRadius(SCENARIOS[scenario], this.numBodies));
    }

    setRadiusMultiplier(r) {
        addon.setRadius(this.handle, r);
    }

    saveSnapshot(dir) {
        addon.saveSnapshot(this.handle, dir);
    }

    loadSnapshot(dir) {
        addon.loadSnapshot(this.handle, dir);
        this.numBodies = addon.getNumBodies(this.handle);
    }

    setBodies(mass, pos, vel) {
        addon.setBodies(this.handle, mass, pos, vel);
        this.numBodies = mass.length;
    }

    getCommands() {
        addon.step(this.handle, this.config.substeps);
    }

    step(nsteps = 1) {
        addon.step(this.handle, nsteps);
    }

    synchronize() {
        addon.synchronize(this.handle);
    }

    // Replay `stepsPerGraph` steps per captured HIP graph (false/0: no graph).
    useGraph(enable = true, stepsPerGraph = 1) {
        addon.useGraph(this.handle, enable ? stepsPerGraph : 0);
    }

    // Capture the step graph now, so the next step() replays it at once.
    prepareGraph() {
        addon.prepareGraph(this.handle);
    }

    getBuffers() {
        const b = addon.getBodies(this.handle);
        b.radiusMultiplier = addon.getRadius(this.handle);
        return b;
    }

    getNumBodies() {
        return addon.getNumBodies(this.handle);
    }
}

module.exports = { Simulation, generate, SCENARIOS, getDefaultConfig, version: addon.version };
