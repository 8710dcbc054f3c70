"""The renderer kernels (density splat + tone map, renderer.ts:293-329 with
density.wgsl and tone_map.wgsl) against the oracle's restatement: fp16
density texture and RGBA frame bit for bit, and the tone table at every fp16
value."""
import numpy as np
import pytest

import nbody_sim
from nbody_sim import render
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def gpu_frame(sim, r):
    rgba = r.getCommands()
    d16 = r.density().view(np.uint16)
    return d16, rgba


def test_tone_table_exhaustive():
    sim = nbody_sim.Simulation(nbody_sim.getDefaultConfig())
    sim.setNumBodies(10)
    sim.setScenario("default")
    render.Renderer(sim, 8, 8).getCommands()
    got = sim.debugBuffer("tone_table")
    sim.close()
    assert np.array_equal(got, O.tone_table())


@pytest.mark.parametrize("scenario,n,w,h,zoom", [
    ("planetary-disk", 50000, 640, 360, 1.0),    # radius multipliers 12, 5, 5, 1
    ("default", 50000, 320, 320, 0.5),           # the app's default scene (main.ts:20-21)
    ("two-clusters", 20000, 200, 120, 3.0),      # bodies off-screen, zoomed out
])
def test_frame_matches_oracle(scenario, n, w, h, zoom):
    sim = nbody_sim.Simulation(nbody_sim.getDefaultConfig(), seed=4)
    sim.setNumBodies(n)
    sim.setScenario(scenario)
    sim.step(2)
    r = render.Renderer(sim, w, h)
    r.zoomCamera(zoom, w * 0.4, h * 0.6)
    r.panCamera(7.0, -3.0)
    d16, rgba = gpu_frame(sim, r)
    b = sim.getBuffers()
    cam = np.array([r.camCenter[0], r.camCenter[1], r.camHalfSize[0], r.camHalfSize[1]],
                   np.float32)
    sim.close()
    od16, orgba = O.render(b["pos"], b["radiusMultiplier"], cam, w, h)
    assert np.count_nonzero(od16) > 100
    assert np.array_equal(d16, od16)
    assert np.array_equal(rgba, orgba)


def test_frame_800k_and_png(tmp_path):
    """The bench workload's frame (800k bodies) at 1920x1080, and its PNG."""
    from nbody_sim import workloads as W

    w = W.WORKLOADS["plain_800k"]
    sim = nbody_sim.Simulation(w.config("fast"))
    sim.setBodies(*w.bodies())
    sim.step(3)
    r = render.Renderer(sim, 1920, 1080)
    r.zoomCamera(0.08, 960, 540)  # the disk has radius ~1
    d16, rgba = gpu_frame(sim, r)
    b = sim.getBuffers()
    cam = np.array([r.camCenter[0], r.camCenter[1], r.camHalfSize[0], r.camHalfSize[1]],
                   np.float32)
    r.writePng(tmp_path / "frame.png", rgba)
    sim.close()
    od16, orgba = O.render(b["pos"], b["radiusMultiplier"], cam, 1920, 1080)
    assert np.array_equal(d16, od16) and np.array_equal(rgba, orgba)
    assert (tmp_path / "frame.png").stat().st_size > 10000
