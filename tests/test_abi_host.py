"""C ABI surface and host-side parameter construction (CPU only: no kernel runs)."""
import ctypes as C
import math
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "lightgraph.h")


def _declared():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?[a-z_0-9]+\s*\*?\s*(lg_[a-z0-9_]+)\s*\(", text, re.M)))


def test_library_exports_every_declared_symbol(lga):
    names = _declared()
    assert len(names) > 40
    lib = C.CDLL(os.path.join(ROOT, "lightgraph_amd", "liblightgraph.so"))
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    from lightgraph_amd import _abi
    assert sorted(_abi.SIGNATURES) == names
    assert lga.lib().lg_abi_version() == 1


def test_struct_layouts_match_header(tmp_path):
    """Every ctypes mirror of a header struct has the C size and field offsets
    (gcc compiles a probe against include/lightgraph.h)."""
    import shutil
    import subprocess
    from lightgraph_amd import _abi
    if shutil.which("gcc") is None:
        pytest.skip("gcc not found")
    structs = {n: c for n, c in vars(_abi).items()
               if n.startswith("lg_") and isinstance(c, type) and issubclass(c, C.Structure)}
    assert len(structs) >= 15
    lines = ['#include <stddef.h>', '#include <stdio.h>', '#include "lightgraph.h"', "int main(void) {"]
    for n, c in sorted(structs.items()):
        lines.append(f'    printf("{n} size %zu\\n", sizeof({n}));')
        for f in c._fields_:
            lines.append(f'    printf("{n} {f[0]} %zu\\n", offsetof({n}, {f[0]}));')
    lines += ["    return 0;", "}"]
    src = tmp_path / "probe.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "probe"
    subprocess.run(["gcc", "-std=c11", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    got = {}
    for line in subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.splitlines():
        n, field, v = line.split()
        got[(n, field)] = int(v)
    for n, c in structs.items():
        assert got[(n, "size")] == C.sizeof(c), f"{n}: C {got[(n, 'size')]} B, ctypes {C.sizeof(c)} B"
        for f in c._fields_:
            assert got[(n, f[0])] == getattr(c, f[0]).offset, f"{n}.{f[0]}"


def test_shader_defaults_match_reference(lga):
    s = lga.default_shader()
    # Shader.h:123-138
    assert (s.density_scale, s.density_max, s.shadow_density_scale) == (1.0, 100.0, 1.0)
    assert (s.scatter_density_scale, s.scatter_density_min) == (0.35, 0.025)
    assert list(s.albedo) == [1.5 / 1.6] * 3 and list(s.extinction) == [-1.6] * 3
    assert (s.emission_scale, s.scatter_scale, s.scatter_density_mask_power) == (0.0, 4.0, 0.25)
    s2 = lga.default_shader(absorption=0.1, scattering=(1.0, 0.9, 0.8), volume_color=(1, 1, 0.5))
    np.testing.assert_allclose(list(s2.extinction), [-1.1, -1.0, -0.9], rtol=1e-15)
    np.testing.assert_allclose(list(s2.albedo), [1.0 / 1.1, 0.9 / 1.0, 0.5 * 0.8 / 0.9], rtol=1e-15)


def test_ramp_parse(lga):
    s = lga.default_shader(ramp="0.0,0.33,0.66,1.0/0,0,0:1.0,0.076,0.0:1.0,0.322,0.0:1.0,0.45,0.15/0.35/1.5")
    assert s.ramp_count == 4 and (s.ramp_in, s.ramp_out) == (0.35, 1.5)
    assert list(s.ramp_pos[:4]) == [0.0, 0.33, 0.66, 1.0]
    assert list(s.ramp_color[1]) == [1.0, 0.076, 0.0]
    with pytest.raises(Exception):
        lga.default_shader(ramp="0,1/1,1,1/0.5")


def test_march_defaults(lga):
    p = lga.march_params()
    assert p.cutoff == 0.005 and list(p.step) == [1.0, 3.0]


def test_light_parse_directions(lga):
    l = lga.parse_light("dir:0:1,0.89,0.79:0.7:0,0,0")
    assert l.type == 0 and l.samples == 1 and abs(l.intensity - 0.7) < 1e-7
    np.testing.assert_allclose(list(l.dir), [0, 0, -1], atol=1e-15)
    # a -90 degree X rotation turns (0,0,-1) into (0,-1,0) (row vectors, Rx: y' = cy - sz)
    np.testing.assert_allclose(list(lga.parse_light("dir:0:1,1,1:1:-90,0,0").dir), [0, -1, 0], atol=1e-15)
    # the world Y rotation: Ry(90) maps (0,0,-1) to (-1,0,0)
    np.testing.assert_allclose(list(lga.parse_light("dir:90:1,1,1:1:0,0,0").dir), [-1, 0, 0], atol=1e-15)
    p = lga.parse_light("point:90:1,1,1:5:1,2,3")
    np.testing.assert_allclose(list(p.pos), [3, 2, -1], atol=1e-14)
    with pytest.raises(Exception, match="not a valid light type"):
        lga.parse_light("spot:0:1,1,1:1:0,0,0")


def test_camera_setup_closed_form(lga):
    cam = lga.camera_setup(cam_translate=(1, 2, 3), width=960, height=540)
    M = np.array(cam.screen_to_world).reshape(4, 4)
    np.testing.assert_allclose(M[:3, :3], np.eye(3), atol=0)
    np.testing.assert_allclose(M[3, :3], [1, 2, 3])
    assert cam.scale_width == 0.5 * 36 / 50 and cam.scale_height == cam.scale_width * 540 / 960
    # camYRot 90 rotates both the orientation and the position about world Y (main.cc:98-102)
    cam = lga.camera_setup(cam_translate=(0, 0, 10), cam_y_rotate=90.0)
    M = np.array(cam.screen_to_world).reshape(4, 4)
    np.testing.assert_allclose(M[3, :3], [10, 0, 0], atol=1e-12)
    fwd = np.array([0, 0, -1.0]) @ M[:3, :3]
    np.testing.assert_allclose(fwd, [-1, 0, 0], atol=1e-12)  # looks back at the origin


def test_host_generator_properties(lga):
    spec = lga.scene_spec(lga.LG_SCENE_SPHERE_FBM, 64, seed=1, octaves=4, radius=24.0)
    g = lga.generate_host(spec, threads=8)
    assert g.dtype == np.float32 and g.min() >= 0.0 and g.max() <= 1.0
    z, y, x = np.meshgrid(*[np.arange(64)] * 3, indexing="ij")
    r = np.sqrt((x - 31.5) ** 2 + (y - 31.5) ** 2 + (z - 31.5) ** 2)
    assert np.all(g[r >= 25.0] == 0.0)
    assert np.mean(g[r < 22.0] > 0) > 0.99
    spec = lga.scene_spec(lga.LG_SCENE_BLOBS_FBM, 128, seed=2, octaves=6, blobs=20, axis_min=8, axis_max=30,
                          active_fraction=0.3)
    b = lga.generate_host(spec, threads=8)
    assert abs(np.mean(b > 0) - 0.30) < 0.02
    np.testing.assert_array_equal(b, lga.generate_host(spec, threads=3))
