"""CPU-side checks of liborbfe: it loads without a GPU and exports exactly the
entry points include/orbfe.h declares; host-only entry points work."""
import os
import re

import numpy as np

from tests.conftest import ROOT
from uni_slam_amd import orbfe


def _declared():
    text = open(os.path.join(ROOT, "include", "orbfe.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(orbfe_[a-z_0-9]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = orbfe.lib()
    declared = _declared()
    assert len(declared) >= 20
    for name in declared:
        assert hasattr(lib, name), name


def test_bindings_cover_the_header():
    from uni_slam_amd import serve
    assert sorted(orbfe.ENTRY_POINTS + serve.ENTRY_POINTS) == _declared()
    serve._lib()  # every server entry point binds


def test_config_default_matches_reference_constants():
    cfg = orbfe.default_config(752, 480)
    assert (cfg.levels, cfg.scale_factor, cfg.fast_min_run, cfg.orientation_radius) == (6, 1.2, 9, 3)
    assert np.float32(cfg.fast_threshold) == np.float32(0.15)
    assert np.float32(cfg.lowe_ratio) == np.float32(0.85)
    assert cfg.features_per_level == 0  # W*H/250 chosen at create time
    assert (cfg.detect_tile_rows, cfg.describe_lanes, cfg.match_lanes) == (0, 0, 0)  # defaults


def test_create_rejects_unsupported_launch_shapes():
    """Launch shapes are checked before any device call, so this runs
    without a GPU."""
    import ctypes as C
    for bad in ({"detect_tile_rows": 12}, {"detect_tile_rows": 64}, {"describe_lanes": 4},
                {"describe_lanes": 64}, {"match_lanes": 8}, {"match_lanes": 128}):
        cfg = orbfe.default_config(320, 240, **bad)
        h = C.c_void_p()
        assert orbfe.lib().orbfe_create(C.byref(cfg), 0, C.byref(h)) == orbfe.ERANGE, bad
        assert not h.value


def test_create_rejects_ratio_outside_unit_interval():
    """A Lowe ratio above 1 would let tied best distances pass, where the
    reference's choice depends on its unstable sort; 0, negatives and NaN
    are refused as well (checked before any device call)."""
    import ctypes as C
    for bad in (1.0001, 1.5, 0.0, -0.8, float("nan")):
        cfg = orbfe.default_config(320, 240, lowe_ratio=bad)
        h = C.c_void_p()
        assert orbfe.lib().orbfe_create(C.byref(cfg), 0, C.byref(h)) == orbfe.ERANGE, bad
        assert not h.value


def test_create_rejects_frames_past_32bit_offsets():
    """The kernels address a frame's pyramid, smoothed images and survivor
    words with 32-bit byte offsets; sizes whose buffers would pass 4 GiB
    are refused before any device call (so this runs without a GPU).
    30000 x 30000: about 1.1e10 bytes of smoothed images over 6 levels."""
    import ctypes as C
    for w, h in ((30000, 30000), (65535, 65535), (65535, 20000), (0, 480), (640, 0)):
        cfg = orbfe.default_config(w, h)
        hnd = C.c_void_p()
        assert orbfe.lib().orbfe_create(C.byref(cfg), 0, C.byref(hnd)) == orbfe.ERANGE, (w, h)
        assert not hnd.value


def test_load_features_host_entry_point():
    cap = 5
    feats = np.zeros(6 * (8 + 48 * cap), np.uint8)
    for lod, n in enumerate([7, 2, 0, 1, 0, 0]):
        blk = feats[lod * (8 + 48 * cap):(lod + 1) * (8 + 48 * cap)]
        blk[:4].view(np.uint32)[0] = n
        recs = blk[8:].view(orbfe.GPU_FEATURE_DTYPE)
        for i in range(min(n, cap)):
            recs[i]["x"] = 10 * lod + i
            recs[i]["y"] = i
            recs[i]["strength"] = 10 + i
            recs[i]["orientation"] = 0.5
            recs[i]["descriptor"] = np.arange(8) + i
    out, counts = orbfe.load_features(feats, 6, cap)
    assert counts == [5, 2, 0, 1, 0, 0]
    assert list(out["lod"]) == [0] * 5 + [1] * 2 + [3]
    assert list(out["x"]) == [0, 1, 2, 3, 4, 10, 11, 30]
    assert np.all(out["orientation"] == np.float32(0.5))


def test_level_dims_helper():
    assert orbfe.level_dims(752, 480)[5] == (302, 192)
