"""The Node host (N-API addon + simulation.js) mirrors the reference's
TypeScript Simulation; its steps go through the same C ABI."""
import json
import os
import shutil
import subprocess

import numpy as np
import pytest

import nbody_sim
from conftest import PKG_DIR

NODE_DIR = os.path.join(PKG_DIR, "node")
NODE = shutil.which("node")

pytestmark = pytest.mark.skipif(NODE is None, reason="node is not installed")


def run_node(script, timeout=120):
    r = subprocess.run([NODE, "-e", script], cwd=NODE_DIR, capture_output=True, text=True,
                       timeout=timeout)
    assert r.returncode == 0, r.stderr
    return r.stdout


def test_node_config_and_generate_match_python():
    out = run_node("""
const s = require('./simulation.js');
const c = s.getDefaultConfig();
const b = s.generate('exp-disk', 1000, 5, {dispersion: 0.05, massSigma: 0.5});
process.stdout.write(JSON.stringify({c, m: Array.from(b.mass.slice(0, 5)),
  p: Array.from(b.pos.slice(0, 6)), v: s.version()}));
""")
    d = json.loads(out)
    ref = nbody_sim.getDefaultConfig()
    assert d["c"]["gravConstant"] == ref.gravConstant and d["c"]["bhTheta"] == ref.bhTheta
    assert d["c"]["deltaTime"] == ref.deltaTime and d["c"]["substeps"] == 1
    m, p, _ = nbody_sim.generate("exp-disk", 1000, seed=5, dispersion=0.05, mass_sigma=0.5)
    assert np.array_equal(np.float32(d["m"]), m[:5])
    assert np.array_equal(np.float32(d["p"]), p.reshape(-1)[:6])
    assert "gfx950" in d["v"]


def test_node_unknown_scenario_throws():
    out = run_node("""
const s = require('./simulation.js');
try { s.generate('nope', 10); console.log('no-throw'); } catch (e) { console.log(e.message); }
""")
    assert "Unknown scenario" in out


@pytest.mark.skipif(os.path.exists("/dev/kfd"), reason="a GPU is present")
def test_node_create_without_gpu_throws():
    out = run_node("""
const s = require('./simulation.js');
try { new s.Simulation(); console.log('created'); } catch (e) { console.log(e.message); }
""")
    assert "no HIP device" in out


@pytest.mark.gpu
def test_node_steps_match_python_path(tmp_path):
    """Same workload stepped from JS and from Python: bit-identical state."""
    n, steps = 4096, 3
    f = str(tmp_path / "state.bin")
    run_node(f"""
const fs = require('fs');
const s = require('./simulation.js');
const cfg = Object.assign({{}}, s.getDefaultConfig(), {{bhTheta: 0.5, deltaTime: 1e-4,
  epsilon: 1e-3, softeningMode: 1, mortonMode: 1, mathMode: 1}});
const sim = new s.Simulation(cfg, 0);
const b = s.generate('uniform-disk', {n}, 1);
sim.setBodies(b.mass, b.pos, b.vel);
sim.step({steps});
sim.synchronize();
const out = sim.getBuffers();
fs.writeFileSync('{f}', Buffer.concat([Buffer.from(out.pos.buffer), Buffer.from(out.vel.buffer)]));
""")
    raw = np.fromfile(f, np.float32)
    jpos, jvel = raw[: 2 * n].reshape(n, 2), raw[2 * n:].reshape(n, 2)
    cfg = nbody_sim.Config(bhTheta=0.5, deltaTime=1e-4, epsilon=1e-3, softeningMode="constant",
                           mortonMode="bbox30", mathMode="exact")
    sim = nbody_sim.Simulation(cfg)
    sim.setBodies(*nbody_sim.generate("uniform-disk", n, seed=1))
    sim.step(steps)
    _, pos, vel = sim.getBodies()
    sim.close()
    assert np.array_equal(jpos, pos) and np.array_equal(jvel, vel)


@pytest.mark.gpu
def test_node_radius_and_snapshot_resume_in_python(tmp_path):
    """The Node host's setScenario fills radiusMultiplier like simulation.ts:
    321-339; a snapshot saved from Node resumes in Python bit for bit."""
    snap = str(tmp_path / "snap")
    f = str(tmp_path / "after.bin")
    out = run_node(f"""
const fs = require('fs');
const s = require('./simulation.js');
const cfg = Object.assign({{}}, s.getDefaultConfig(), {{epsilon: 1e-2, softeningMode: 1}});
const sim = new s.Simulation(cfg, 0, 3);
sim.setNumBodies(2000);
sim.setScenario('planetary-disk');
const r = sim.getBuffers().radiusMultiplier;
sim.step(4);
sim.saveSnapshot('{snap}');
sim.step(3);
sim.synchronize();
const b = sim.getBuffers();
fs.writeFileSync('{f}', Buffer.concat([Buffer.from(b.pos.buffer), Buffer.from(b.vel.buffer)]));
process.stdout.write(JSON.stringify(Array.from(r.slice(0, 4))));
""")
    assert json.loads(out) == [12.0, 5.0, 5.0, 1.0]
    raw = np.fromfile(f, np.float32)
    sim = nbody_sim.Simulation()
    sim.loadSnapshot(snap)
    assert sim.getNumBodies() == 2000 and sim.getStepCount() == 4
    assert sim.getRadiusMultiplier()[:3].tolist() == [12.0, 5.0, 5.0]
    sim.step(3)
    _, pos, vel = sim.getBodies()
    sim.close()
    assert np.array_equal(raw[:4000].reshape(-1, 2), pos)
    assert np.array_equal(raw[4000:].reshape(-1, 2), vel)


@pytest.mark.gpu
def test_node_renderer_matches_python(tmp_path):
    """The Node Renderer (renderer.js) draws the same frame as the Python one."""
    f = str(tmp_path / "frame.bin")
    png = str(tmp_path / "frame.png")
    run_node(f"""
const fs = require('fs');
const s = require('./simulation.js');
const {{ Renderer }} = require('./renderer.js');
const sim = new s.Simulation(s.getDefaultConfig(), 0, 2);
sim.setNumBodies(5000);
sim.setScenario('planetary-disk');
const r = new Renderer(sim, 160, 90);
r.zoomCamera(2.0, 80, 45);
const fr = r.getCommands();
r.writePng('{png}');
fs.writeFileSync('{f}', Buffer.from(fr.rgba.buffer));
""")
    from nbody_sim import render

    sim = nbody_sim.Simulation(nbody_sim.getDefaultConfig(), seed=2)
    sim.setNumBodies(5000)
    sim.setScenario("planetary-disk")
    r = render.Renderer(sim, 160, 90)
    r.zoomCamera(2.0, 80, 45)
    rgba = r.getCommands()
    sim.close()
    assert np.array_equal(np.fromfile(f, np.uint8).reshape(90, 160, 4), rgba)
    assert os.path.getsize(png) > 100
