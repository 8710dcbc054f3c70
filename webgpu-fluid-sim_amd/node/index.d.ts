// Type declarations of the Node host (simulation.js, config.js), mirroring
// the reference's src/config.ts and src/simulation.ts.

export type Config = Readonly<{
    gravConstant: number;
    bhTheta: number;
    substeps: number;
    deltaTime: number;
    epsilonMultiplier: number;
    // engine extensions (optional)
    epsilon?: number;
    softeningMode?: 0 | 1;   // 0 reference (barnes_hut.wgsl:90-91), 1 constant
    mortonMode?: 0 | 1;      // 0 reference box (lbvh.wgsl:48-65), 1 bbox30
    mathMode?: 0 | 1;        // 0 fast, 1 exact
    expansion?: 0 | 1;       // 0 monopole (barnes_hut.wgsl:92-97), 1 quadrupole
}>;

export declare function getDefaultConfig(): Config;

export type SimScenario = "default" | "two-clusters" | "planetary-disk"
    | "uniform-disk" | "exp-disk" | "collision" | "plummer";

export type SimBuffers = {
    mass: Float32Array;   // [n]
    pos: Float32Array;    // [2n] x0,y0,x1,y1,...
    vel: Float32Array;    // [2n]
    radiusMultiplier?: Float32Array;  // [n] (getBuffers; simulation.ts:17)
};

export declare function generate(scenario: SimScenario, n: number, seed?: number,
    params?: { gravConstant?: number; scale?: number; dispersion?: number; massSigma?: number }): SimBuffers;

export declare class Simulation {
    constructor(config?: Config, device?: number, seed?: number);
    updateMetadataBuffer(): void;
    setNumBodies(numBodies: number): void;
    setScenario(scenario: SimScenario, seed?: number): void;
    setBodies(mass: Float32Array, pos: Float32Array, vel: Float32Array): void;
    getCommands(): void;
    step(nsteps?: number): void;
    synchronize(): void;
    useGraph(enable?: boolean, stepsPerGraph?: number): void;
    prepareGraph(): void;
    setRadiusMultiplier(r: Float32Array): void;
    saveSnapshot(dir: string): void;
    loadSnapshot(dir: string): void;
    getBuffers(): SimBuffers;
    getNumBodies(): number;
}

export declare function version(): string;

// renderer.js: the reference's Renderer (src/renderer.ts) over nbody_render.
export declare class Renderer {
    constructor(sim: Simulation, width?: number, height?: number);
    camCenter: [number, number];
    camHalfSize: [number, number];
    viewPort: [number, number];
    rebindSimBuffers(sim: Simulation): void;
    setNumBodies(numBodies: number): void;
    resize(width: number, height: number): void;
    deltaPxToDeltaWorld(deltaPx: number, deltaPy: number): [number, number];
    canvasPxToWorld(px: number, py: number): [number, number];
    panCamera(deltaPx: number, deltaPy: number): void;
    zoomCamera(zoomFactor: number, zoomCenterX: number, zoomCenterY: number): void;
    getCommands(): { rgba: Uint8Array; density: Uint16Array };
    writePng(file: string): void;
}
