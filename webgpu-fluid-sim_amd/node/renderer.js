// renderer.js -- the reference's Renderer (src/renderer.ts:22-329) over the
// C ABI: the same camera state and members, drawing with the library's HIP
// kernels (nbody_render: density splat + tone map) instead of WebGPU passes.
//
//   constructor(..., posBuffer, radiusBuffer)  -> new Renderer(sim, width, height)
//   rebindSimBuffers(pos, radius)              -> rebindSimBuffers(sim)
//   resizeCanvasToDisplaySize()                -> resize(width, height)
//   deltaPxToDeltaWorld / canvasPxToWorld / panCamera / zoomCamera (renderer.ts:250-289)
//   getCommands(numBodies)                     -> getCommands(): {rgba, density}
"use strict";

const path = require("path");

const addon = require(path.join(__dirname, "nbody_addon.node"));

class Renderer {
    constructor(sim, width = 1280, height = 720) {
        this.sim = sim;
        this.camCenter = [0.0, 0.0];     // renderer.ts:49
        this.camHalfSize = [10.0, 10.0]; // renderer.ts:50
        this.numBodies = 0;
        this.resize(width, height);
    }

    rebindSimBuffers(sim) { this.sim = sim; }

    setNumBodies(numBodies) { this.numBodies = numBodies; }

    resize(width, height) {
        this.viewPort = [Math.max(1, width | 0), Math.max(1, height | 0)];
        this.camHalfSize[0] = this.camHalfSize[1] * (this.viewPort[0] / this.viewPort[1]);
    }

    deltaPxToDeltaWorld(deltaPx, deltaPy) {
        const dnx = (2.0 * deltaPx) / this.viewPort[0];
        const dny = (2.0 * deltaPy) / this.viewPort[1];
        return [dnx * this.camHalfSize[0], dny * this.camHalfSize[1]];
    }

    canvasPxToWorld(px, py) {
        const nx = (2.0 * px) / this.viewPort[0] - 1.0;
        const ny = (2.0 * py) / this.viewPort[1] - 1.0;
        return [this.camCenter[0] + nx * this.camHalfSize[0], this.camCenter[1] + ny * this.camHalfSize[1]];
    }

    panCamera(deltaPx, deltaPy) {
        const [wx, wy] = this.deltaPxToDeltaWorld(deltaPx, deltaPy);
        this.camCenter[0] -= wx;
        this.camCenter[1] -= wy;
    }

    zoomCamera(zoomFactor, zoomCenterX, zoomCenterY) {
        const [zx, zy] = this.canvasPxToWorld(zoomCenterX, zoomCenterY);
        this.camHalfSize[0] *= zoomFactor;
        this.camHalfSize[1] *= zoomFactor;
        this.camCenter[0] = this.camCenter[0] + (zx - this.camCenter[0]) * (1.0 - zoomFactor);
        this.camCenter[1] = this.camCenter[1] + (zy - this.camCenter[1]) * (1.0 - zoomFactor);
    }

    camera() {
        return { centerX: this.camCenter[0], centerY: this.camCenter[1], halfX: this.camHalfSize[0],
                 halfY: this.camHalfSize[1], width: this.viewPort[0], height: this.viewPort[1] };
    }

    // One frame: {rgba: Uint8Array(4wh) canvas rows, density: Uint16Array(wh) fp16 bits}.
    getCommands() {
        this.last = addon.render(this.sim.handle, this.camera());
        return this.last;
    }

    writePng(file) {
        if (!this.last) this.getCommands();
        addon.writePng(file, this.viewPort[0], this.viewPort[1], this.last.rgba);
    }
}

module.exports = { Renderer };
