"""Host mirror of the reference ``Renderer`` (src/renderer.ts:22-329).

Same camera state and members, drawing with the library's HIP kernels
(nbody_render) instead of WebGPU render passes:

    constructor(device, canvas, ..., posBuffer, radiusBuffer)
                              -> Renderer(sim, width, height)         renderer.ts:43-61
    rebindSimBuffers(...)     -> rebindSimBuffers(sim)                renderer.ts:207-217
    setNumBodies(n)           -> setNumBodies(n)                      renderer.ts:219-222
    resizeCanvasToDisplaySize -> resize(width, height)                renderer.ts:224-248
    deltaPxToDeltaWorld       -> deltaPxToDeltaWorld(dx, dy)          renderer.ts:250-258
    canvasPxToWorld           -> canvasPxToWorld(px, py)              renderer.ts:260-268
    panCamera / zoomCamera    -> panCamera(dx, dy) / zoomCamera(f, px, py)  renderer.ts:270-289
    getCommands(numBodies)    -> getCommands() -> RGBA image (h, w, 4) uint8  renderer.ts:293-329

The frame is the density splat of the simulation's positions and radius
multipliers followed by the tone map; ``density()`` returns the fp16 density
texture and ``writePng(path)`` saves the last frame.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib


class Renderer:
    def __init__(self, sim, width: int = 1280, height: int = 720):
        self.sim = sim
        self.camCenter = [0.0, 0.0]   # renderer.ts:49
        self.camHalfSize = [10.0, 10.0]  # renderer.ts:50
        self.numBodies = 0
        self.resize(width, height)
        self._last = None

    # -- reference surface -------------------------------------------------
    def rebindSimBuffers(self, sim):
        """renderer.ts:207-217: draw another simulation's buffers."""
        self.sim = sim

    def setNumBodies(self, numBodies: int):
        """renderer.ts:219-222 (the kernels draw every body of the simulation)."""
        self.numBodies = int(numBodies)

    def resize(self, width: int, height: int):
        """resizeCanvasToDisplaySize (renderer.ts:224-248): new viewport, half
        width = half height * aspect."""
        self.viewPort = [max(1, int(width)), max(1, int(height))]
        self.camHalfSize[0] = self.camHalfSize[1] * (self.viewPort[0] / self.viewPort[1])

    def deltaPxToDeltaWorld(self, deltaPx: float, deltaPy: float):
        """renderer.ts:250-258."""
        dnx = 2.0 * deltaPx / self.viewPort[0]
        dny = 2.0 * deltaPy / self.viewPort[1]
        return dnx * self.camHalfSize[0], dny * self.camHalfSize[1]

    def canvasPxToWorld(self, px: float, py: float):
        """renderer.ts:260-268."""
        nx = 2.0 * px / self.viewPort[0] - 1.0
        ny = 2.0 * py / self.viewPort[1] - 1.0
        return self.camCenter[0] + nx * self.camHalfSize[0], self.camCenter[1] + ny * self.camHalfSize[1]

    def panCamera(self, deltaPx: float, deltaPy: float):
        """renderer.ts:270-275."""
        wx, wy = self.deltaPxToDeltaWorld(deltaPx, deltaPy)
        self.camCenter[0] -= wx
        self.camCenter[1] -= wy

    def zoomCamera(self, zoomFactor: float, zoomCenterX: float, zoomCenterY: float):
        """renderer.ts:277-289: zoom about a canvas point."""
        zx, zy = self.canvasPxToWorld(zoomCenterX, zoomCenterY)
        self.camHalfSize[0] *= zoomFactor
        self.camHalfSize[1] *= zoomFactor
        self.camCenter[0] = self.camCenter[0] + (zx - self.camCenter[0]) * (1.0 - zoomFactor)
        self.camCenter[1] = self.camCenter[1] + (zy - self.camCenter[1]) * (1.0 - zoomFactor)

    def camera(self) -> _lib.NbodyCamera:
        return _lib.NbodyCamera(self.camCenter[0], self.camCenter[1], self.camHalfSize[0],
                                self.camHalfSize[1], self.viewPort[0], self.viewPort[1])

    def getCommands(self):
        """renderer.ts:293-329: density splat + tone map; returns the RGBA
        frame (height, width, 4) uint8, canvas row 0 at the top."""
        w, h = self.viewPort
        rgba = np.zeros((h, w, 4), np.uint8)
        cam = self.camera()
        _lib.check(_lib.lib().nbody_render(self.sim.handle, ctypes.byref(cam), None,
                                           rgba.ctypes.data_as(ctypes.c_void_p)), self.sim.handle)
        self._last = rgba
        return rgba

    def density(self):
        """The fp16 density texture of a frame of the current state, as
        (height, width) float16, texture row 0 at NDC y = +1."""
        w, h = self.viewPort
        d = np.zeros((h, w), np.uint16)
        cam = self.camera()
        _lib.check(_lib.lib().nbody_render(self.sim.handle, ctypes.byref(cam),
                                           d.ctypes.data_as(ctypes.c_void_p), None),
                   self.sim.handle)
        return d.view(np.float16)

    def writePng(self, path, rgba=None):
        img = self._last if rgba is None else np.ascontiguousarray(rgba, np.uint8)
        if img is None:
            img = self.getCommands()
        h, w = img.shape[:2]
        _lib.check(_lib.lib().nbody_write_png(str(path).encode(), w, h,
                                              img.ctypes.data_as(ctypes.c_void_p)))


def write_png(path, rgba):
    """Save an (h, w, 4) uint8 RGBA array as PNG (nbody_write_png)."""
    img = np.ascontiguousarray(rgba, np.uint8)
    h, w = img.shape[:2]
    _lib.check(_lib.lib().nbody_write_png(str(path).encode(), w, h,
                                          img.ctypes.data_as(ctypes.c_void_p)))
