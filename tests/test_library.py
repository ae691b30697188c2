"""CPU-side checks of the GPU library: it loads, exports every entry point include/keepsake_amd.h
declares, generates the benchmark scenes with the configured sizes, and fails loudly without a GPU."""

import re
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]


def _declared():
    text = (ROOT / "include" / "keepsake_amd.h").read_text()
    return sorted(set(re.findall(r"^(?:int|const char \*)\s*(ks_\w+)\(", text, re.M)))


def test_exports_every_declared_symbol(ks):
    lib = ks.load_library()
    names = _declared()
    assert len(names) >= 20
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(ks.EXPORTED)


def test_cornell_scene_layout(cornell):
    assert cornell.n_tris == 36
    assert cornell.materials.shape[0] == 4
    emissive = cornell.materials["has_emission"][cornell.geom_material[cornell.tri_geom]] != 0
    assert emissive.sum() == 2
    assert np.allclose(cornell.materials["emission"][3], [17, 12, 4])
    v = cornell.vertices
    assert v.min() >= 0.0 and v.max() <= 1.0
    assert cornell.camera["vfov"] == 40.0


def test_spheres_scene_sizes(ks):
    hs = ks.procedural_scene("spheres", 1)
    assert hs.n_tris == 64 * 20480 + 8 == 1310728
    assert hs.lights.shape[0] == 1 and hs.lights["type"][0] == ks.KS_LIGHT_SKY_AMBIENT
    assert np.allclose(hs.lights["color"][0], 0.5)
    assert (np.diff(hs.tri_geom.astype(np.int64)) >= 0).all()
    mats = hs.materials[:64]
    assert ((mats["basecolor"] >= 0.1) & (mats["basecolor"] <= 0.9)).all()
    assert ((mats["roughness"] >= 0.05) & (mats["roughness"] <= 1.0)).all()
    assert set(np.unique(mats["metallic"]).tolist()) <= {0.0, 1.0}
    # deterministic for a seed
    hs2 = ks.procedural_scene("spheres", 1)
    assert np.array_equal(hs.vertices, hs2.vertices)


def test_room_scene(ks):
    hs = ks.procedural_scene("room", 3)
    assert 190000 < hs.n_tris < 215000
    em = hs.materials["has_emission"][hs.geom_material[hs.tri_geom]] != 0
    assert em.sum() == 2
    assert np.allclose(hs.materials["emission"][hs.materials["has_emission"] != 0], 50.0)


def test_unknown_scene_raises(ks):
    with pytest.raises(ks.KsError):
        ks.procedural_scene("nope")


def test_device_open_fails_loudly_without_gpu(ks):
    if ks.device_count() > 0:
        pytest.skip("a GPU is present")
    with pytest.raises(ks.KsError) as e:
        ks.Device(0)
    assert e.value.status == ks.KS_ERR_NO_DEVICE


def test_entry_points_report_bad_arguments(ks):
    """Errors come back as status codes with a message, never as an exception or abort across the ABI."""
    import ctypes as C

    lib = ks.load_library()
    out = C.c_void_p()
    assert lib.ks_film_create(None, 4, 4, None, C.byref(out)) == ks.KS_ERR_INVALID
    assert b"null" in lib.ks_last_error()
    assert lib.ks_scene_create(None, None, None, C.byref(out)) == ks.KS_ERR_INVALID
    assert lib.ks_guide_set_mode(None, 0.5, 1) == ks.KS_ERR_INVALID
    assert lib.ks_host_scene_create(b"no-such-scene", 0, C.byref(out)) == ks.KS_ERR_INVALID
