"""CPU checks of the C ABI library: it loads, exports every declared symbol,
its host-only entry points work, and device entry points fail cleanly
without a GPU."""
import ctypes
import os
import re

import numpy as np
import pytest

from conftest import ROOT

import nbody_sim
from nbody_sim import _lib

HEADER = os.path.join(ROOT, "include", "nbody_sim.h")


def declared_symbols():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"\b(nbody_[a-z0-9_]+)\s*\(", text)))


def test_header_matches_binding_list():
    assert declared_symbols() == sorted(_lib.EXPORTS)


def test_library_exports_every_declared_symbol():
    L = _lib.lib()
    for name in declared_symbols():
        assert hasattr(L, name), name


def test_version_and_default_config():
    L = _lib.lib()
    assert b"gfx950" in L.nbody_version()
    c = _lib.NbodyConfig()
    assert L.nbody_default_config(ctypes.byref(c)) == 0
    d = nbody_sim.getDefaultConfig()  # config.ts:9-22
    assert c.grav_constant == np.float32(d.gravConstant) == 1.0
    assert c.bh_theta == np.float32(0.6)
    assert c.substeps == 1
    assert c.delta_time == np.float32(0.1 / 60.0)
    assert c.epsilon_multiplier == 1.0
    assert c.softening_mode == 0 and c.morton_mode == 0


def test_philox_known_answers():
    """Random123 known-answer vectors for philox4x32-10."""
    L = _lib.lib()
    cases = [
        ((0, 0, 0, 0), (0, 0), (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)),
        ((0xFFFFFFFF,) * 4, (0xFFFFFFFF, 0xFFFFFFFF),
         (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)),
        ((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0),
         (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1)),
    ]
    for ctr, key, want in cases:
        c = (ctypes.c_uint32 * 4)(*ctr)
        k = (ctypes.c_uint32 * 2)(*key)
        o = (ctypes.c_uint32 * 4)()
        L.nbody_philox4x32(c, k, o)
        assert tuple(o) == want


def test_generate_deterministic_and_shaped():
    m1, p1, v1 = nbody_sim.generate("exp-disk", 5000, seed=3, dispersion=0.05, mass_sigma=0.5)
    m2, p2, v2 = nbody_sim.generate("exp-disk", 5000, seed=3, dispersion=0.05, mass_sigma=0.5)
    assert np.array_equal(p1, p2) and np.array_equal(v1, v2) and np.array_equal(m1, m2)
    assert abs(m1.astype(np.float64).sum() - 1.0) < 1e-4
    _, p3, _ = nbody_sim.generate("exp-disk", 5000, seed=4)
    assert not np.array_equal(p1, p3)
    m, p, v = nbody_sim.generate("uniform-disk", 4000, seed=1)
    assert np.all(np.hypot(p[:, 0], p[:, 1]) <= 1.0 + 1e-6)
    assert np.all(v == 0) and np.allclose(m, 1 / 4000)
    m, p, v = nbody_sim.generate("planetary-disk", 100, seed=1)  # simulation.ts:315-339
    assert m[0] == 1000.0 and m[1] == 10.0 and m[2] == 10.0
    assert tuple(p[1]) == (2.5, 0.0) and tuple(p[2]) == (12.0, 0.0)
    assert np.isclose(v[1, 1], 0.8 * np.sqrt(1000 / 2.5))
    assert np.allclose(m[3:], 1000.0 / (100.0 * 100))
    m, p, v = nbody_sim.generate("two-clusters", 1000, seed=1)
    assert np.all(p[:500, 0] < -11) and np.all(p[500:, 0] > 11)
    m, p, v = nbody_sim.generate("collision", 2000, seed=1)
    assert np.all(p[:1000, 0].mean() < 0) and p[1000:, 0].mean() > 0
    assert np.allclose(m, 1 / 2000)


def test_unknown_scenario_raises():
    with pytest.raises(ValueError):
        nbody_sim.generate("nope", 10)


def test_config_modes():
    c = nbody_sim.Config(softeningMode="constant", mortonMode="bbox30", mathMode="exact").to_c()
    assert (c.softening_mode, c.morton_mode, c.math_mode) == (1, 1, 1)
    with pytest.raises(ValueError):
        nbody_sim.Config(mortonMode="bogus").to_c()


@pytest.mark.skipif(os.path.exists("/dev/kfd"), reason="a GPU is present")
def test_create_without_gpu_fails_cleanly():
    L = _lib.lib()
    cfg = _lib.NbodyConfig()
    L.nbody_default_config(ctypes.byref(cfg))
    err = ctypes.c_int(0)
    h = L.nbody_create(ctypes.byref(cfg), 100, 0, ctypes.byref(err))
    assert not h
    assert err.value == _lib.ERR_NO_DEVICE
    assert "device" in _lib.last_error(None)
    with pytest.raises(nbody_sim.NbodyError):
        nbody_sim.Simulation(nbody_sim.getDefaultConfig())


def test_null_handle_errors():
    L = _lib.lib()
    assert L.nbody_step(None, 1) == _lib.ERR_INVALID_ARG
    assert L.nbody_get_num_bodies(None) == -1
    assert L.nbody_destroy(None) == 0


def test_static_step_rejects_sizes_beyond_its_scratch():
    """nbody_step_static refuses n above its fixed scratch before any device call."""
    L = _lib.lib()
    c = nbody_sim.Config().to_c()
    rc = L.nbody_step_static(ctypes.byref(c), _lib.STATIC_MAX_BODIES + 1, None, None, None, 1, None)
    assert rc == _lib.ERR_CAPACITY
    assert str(_lib.STATIC_MAX_BODIES) in _lib.last_error(None)
    assert L.nbody_step_static(ctypes.byref(c), -1, None, None, None, 1, None) == _lib.ERR_INVALID_ARG
    assert L.nbody_step_static(ctypes.byref(c), 0, None, None, None, 1, None) == _lib.OK


def test_generate_radius_known_answers():
    """setScenario's radiusMultiplierData (simulation.ts:240-241, 321, 330, 339)."""
    r = nbody_sim.generate_radius("planetary-disk", 6)
    assert r.dtype == np.float32 and r.tolist() == [12.0, 5.0, 5.0, 1.0, 1.0, 1.0]
    assert nbody_sim.generate_radius("planetary-disk", 2).tolist() == [12.0, 5.0]
    for scen in ("default", "two-clusters", "exp-disk"):
        assert np.all(nbody_sim.generate_radius(scen, 100) == 1.0)
    assert nbody_sim.generate_radius("default", 0).shape == (0,)
    with pytest.raises(ValueError):
        nbody_sim.generate_radius("nope", 3)
    assert _lib.lib().nbody_generate_radius(99, 3, None) == _lib.ERR_INVALID_ARG


def test_write_png_round_trip(tmp_path):
    """nbody_write_png: PNG signature, IHDR (8-bit RGBA), chunk CRCs, and the
    decompressed scanlines (filter byte 0) equal the image."""
    import struct
    import zlib

    from nbody_sim.render import write_png

    rgba = np.random.default_rng(0).integers(0, 256, (37, 53, 4), dtype=np.uint8)
    path = tmp_path / "x.png"
    write_png(path, rgba)
    data = path.read_bytes()
    assert data[:8] == b"\x89PNG\r\n\x1a\n"
    pos, chunks = 8, {}
    while pos < len(data):
        (length,) = struct.unpack(">I", data[pos:pos + 4])
        ctype = data[pos + 4:pos + 8]
        body = data[pos + 8:pos + 8 + length]
        (crc,) = struct.unpack(">I", data[pos + 8 + length:pos + 12 + length])
        assert crc == zlib.crc32(ctype + body) & 0xFFFFFFFF
        chunks.setdefault(ctype, b"")
        chunks[ctype] += body
        pos += 12 + length
    w, h, depth, ctype, comp, filt, inter = struct.unpack(">IIBBBBB", chunks[b"IHDR"])
    assert (w, h, depth, ctype, comp, filt, inter) == (53, 37, 8, 6, 0, 0, 0)
    raw = np.frombuffer(zlib.decompress(chunks[b"IDAT"]), np.uint8).reshape(37, 1 + 53 * 4)
    assert np.all(raw[:, 0] == 0)
    assert np.array_equal(raw[:, 1:].reshape(37, 53, 4), rgba)
    assert b"IEND" in chunks
    assert _lib.lib().nbody_write_png(str(tmp_path / "y.png").encode(), 0, 5, None) == _lib.ERR_INVALID_ARG


def test_default_camera():
    """renderer.ts:49-50, 240-243: centre 0, half height 10, half width 10 * aspect."""
    c = _lib.NbodyCamera()
    assert _lib.lib().nbody_default_camera(1920, 1080, ctypes.byref(c)) == 0
    assert (c.center_x, c.center_y, c.half_y, c.width, c.height) == (0, 0, 10, 1920, 1080)
    assert c.half_x == np.float32(10 * np.float32(1920 / 1080)) or abs(c.half_x - 10 * 1920 / 1080) < 1e-5
