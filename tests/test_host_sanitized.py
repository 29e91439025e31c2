"""Host code on the CPU, over a fake device (tests/host/fake_abi.cpp), and under
the address, undefined-behaviour and leak sanitizers:

* the serving loop (lightgraph_amd/csrc/lg_serve.cpp): HBM-budget admission
  against the fake allocator, an injected allocation failure that fails its
  job with LG_ERR_NOMEM while the others complete, refusal of an oversized job,
  continuous batching and overlap (tests/host/serve_driver.cpp), plain and
  sanitized;
* the command-line front end (host/main.cc, host/lightgraph.hpp, host/exr.cpp,
  csrc/lg_host.cpp) end to end on the fake device, sanitized: env light read
  from an EXR, point and directional lights, ramp, -v, -geoDump, the 13-channel
  EXR written and read back, the dart sampler, help and option errors;
* -vdbIn: density and temperature read from a .vdb file (host/vdb.cpp) through
  the same front end, sanitized;
* the oracle (oracle/lgoracle.cpp), sanitized, running the oracle's CPU tests.

Builds go to a temporary directory; g++ compiles the host sources directly."""
import os
import shutil
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HOST = os.path.join(ROOT, "tests", "host")
SAN = ["-fsanitize=address,undefined", "-fno-sanitize-recover=undefined", "-fno-omit-frame-pointer"]

pytestmark = pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not found")


def _build(out, sources, extra=(), libs=()):
    cmd = ["g++", "-std=c++17", "-O1", "-g", "-Wall", "-I", os.path.join(ROOT, "include"), "-I",
           os.path.join(ROOT, "lightgraph_amd", "host"), *extra, *sources, "-o", str(out), "-lpthread", *libs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    return str(out)


def _run(cmd, cwd=None, env=None, timeout=120):
    e = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1")
    if env:
        e.update(env)
    return subprocess.run(cmd, capture_output=True, text=True, cwd=cwd, env=e, timeout=timeout)


SERVE_SOURCES = [os.path.join(HOST, "fake_abi.cpp"), os.path.join(ROOT, "lightgraph_amd", "csrc", "lg_serve.cpp"),
                 os.path.join(ROOT, "lightgraph_amd", "csrc", "lg_host.cpp"), os.path.join(HOST, "serve_driver.cpp")]


@pytest.mark.parametrize("sanitized", [False, True])
def test_server_admission_and_failures(tmp_path, sanitized):
    exe = _build(tmp_path / "serve_driver", SERVE_SOURCES, SAN if sanitized else ())
    r = _run([exe])
    out = r.stdout + r.stderr
    assert r.returncode == 0, out[-4000:]
    for scenario in ("budget_admission", "injected_failure", "continuous_batching"):
        assert f"PASS {scenario}" in r.stdout, out[-4000:]
    assert "ERROR: AddressSanitizer" not in out and "runtime error" not in out


@pytest.fixture(scope="module")
def cli_san(tmp_path_factory):
    d = tmp_path_factory.mktemp("cli")
    src = [os.path.join(HOST, "fake_abi.cpp"), os.path.join(ROOT, "lightgraph_amd", "csrc", "lg_host.cpp"),
           os.path.join(ROOT, "lightgraph_amd", "host", "main.cc"), os.path.join(ROOT, "lightgraph_amd", "host", "exr.cpp"),
           os.path.join(ROOT, "lightgraph_amd", "host", "vdb.cpp")]
    return _build(d / "lightgraph_san", src, SAN, libs=["-lz", "-ldl"])


def _clean(r):
    out = r.stdout + r.stderr
    assert "ERROR: AddressSanitizer" not in out and "runtime error" not in out, out[-4000:]
    assert "LeakSanitizer" not in out, out[-4000:]
    return out


def test_cli_end_to_end_under_sanitizers(cli_san, tmp_path):
    sys.path.insert(0, os.path.dirname(__file__))
    import exrio
    sky = np.full((16, 32, 3), 0.5, np.float32)
    sky[3, 20] = 400.0
    env = str(tmp_path / "sky.exr")
    exrio.write_rgb(env, sky, compression=3)
    out_exr = str(tmp_path / "out.exr")
    ramp = "0.0,0.33,0.66,1.0/0,0,0:1.0,0.076,0.0:1.0,0.322,0.0:1.0,0.45,0.15/0.35/1.5"
    r = _run([cli_san, "-scene", "sphere", "-sceneSize", "48", "-points", "300", "-iter", "2", "-radius", "5",
              "-resWidth", "40", "-resHeight", "30", "-camPos", "23.5,23.5,120", "-focal", "60",
              "-light", f"env:0:1,1,1:0.5:1,{env}", "-light", "dir:0:1,1,1:1:-45,30,0",
              "-light", "point:0:1,0.8,0.6:2000:20,70,30", "-tempColorRamp", ramp, "-v", "-geoDump", "g",
              "-exrOut", out_exr], cwd=str(tmp_path))
    out = _clean(r)
    assert r.returncode == 0, out[-4000:]
    for line in ("Point Cloud Creation:", "Graph Creation:", "Raymarch Graph Edges:", "Vertex Lighting Calculation:",
                 "Shortest Path Finding:", "Radiance Diffusion:", "LightGraphs Calculation:",
                 "Rasterize LightGraphs to Grid:", "Render:"):
        assert line in out, line
    chans, hdr = exrio.read(out_exr)
    assert len(hdr["channels"]) == 13 and chans["R"].shape == (30, 40)
    assert len(os.listdir(tmp_path / "geoDump")) == 10
    r = _run([cli_san, "-scene", "blobs", "-sceneSize", "40", "-sampler", "dart", "-iter", "2", "-resWidth", "16",
              "-resHeight", "12", "-compression", "none", "-exrOut", str(tmp_path / "dart.exr")], cwd=str(tmp_path))
    assert r.returncode == 0, _clean(r)[-4000:]
    _clean(r)


@pytest.mark.parametrize("args,code", [(["-h"], 0), (["-notAnOption"], 1), (["-resWidth", "16"], 1),
                                       (["-scene", "cube"], 1), (["-light", "dir:bad"], 1)])
def test_cli_options_under_sanitizers(cli_san, tmp_path, args, code):
    r = _run([cli_san, *args], cwd=str(tmp_path))
    _clean(r)
    assert (r.returncode == 0) == (code == 0), r.stdout[-2000:]


def test_oracle_under_sanitizers(tmp_path):
    asan = subprocess.run(["gcc", "-print-file-name=libasan.so"], capture_output=True, text=True).stdout.strip()
    ubsan = subprocess.run(["gcc", "-print-file-name=libubsan.so"], capture_output=True, text=True).stdout.strip()
    if not (os.path.isabs(asan) and os.path.isabs(ubsan)):
        pytest.skip("sanitizer runtimes not found")
    lib = tmp_path / "liblgoracle_san.so"
    r = subprocess.run(["g++", "-std=c++17", "-O1", "-g", "-fPIC", "-shared", "-ffp-contract=off", "-fno-fast-math",
                        *SAN, "-I", os.path.join(ROOT, "include"), "-o", str(lib),
                        os.path.join(ROOT, "oracle", "lgoracle.cpp")], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    tests = [os.path.join(ROOT, "tests", t) for t in ("test_oracle_core.py", "test_oracle_render.py",
                                                      "test_oracle_pins.py")]
    r = _run([sys.executable, "-m", "pytest", "-x", "-q", "-s", "-p", "no:cacheprovider", *tests], cwd=ROOT,
             env={"LD_PRELOAD": f"{asan} {ubsan}", "ASAN_OPTIONS": "detect_leaks=0", "LG_ORACLE_LIB": str(lib)},
             timeout=600)
    out = r.stdout + r.stderr
    assert r.returncode == 0, out[-4000:]
    assert "runtime error" not in out and "ERROR: AddressSanitizer" not in out


def test_cli_vdb_input_under_sanitizers(cli_san, tmp_path):
    """-vdbIn: density + temperature FloatGrids read from a .vdb file
    (tests/vdbio.py writes it), the stages run, the image and the geometry dumps
    are named after the file's stem; a missing density grid and a missing file
    are errors."""
    sys.path.insert(0, os.path.dirname(__file__))
    import exrio
    import vdbio
    rng = np.random.default_rng(2)
    z, y, x = np.mgrid[0:24, 0:24, 0:24]
    r = np.sqrt((x - 11.5) ** 2 + (y - 11.5) ** 2 + (z - 11.5) ** 2)
    dens = np.where(r < 10, 0.5 + 0.5 * rng.random(r.shape), 0.0).astype(np.float32)
    temp = np.where(r < 10, 800.0 * (1 - r / 10), 0.0).astype(np.float32)
    g = vdbio.Grid.from_dense("density", dens, dens > 0, (-12, -12, -12), voxel_size=0.5, translation=(1, 2, 3))
    t = vdbio.Grid.from_dense("temperature", temp, temp > 0, (-12, -12, -12), voxel_size=0.5, translation=(1, 2, 3))
    path = vdbio.write(str(tmp_path / "ball.vdb"), [g, t])
    ramp = "0.0,0.33,0.66,1.0/0,0,0:1.0,0.076,0.0:1.0,0.322,0.0:1.0,0.45,0.15/0.35/1.5"
    r = _run([cli_san, "-vdbIn", path, "-points", "200", "-iter", "2", "-resWidth", "24", "-resHeight", "16",
              "-camPos", "1,2,30", "-focal", "50", "-light", "dir:0:1,1,1:1:-45,30,0", "-tempColorRamp", ramp,
              "-v", "-geoDump"], cwd=str(tmp_path))
    out = _clean(r)
    assert r.returncode == 0, out[-4000:]
    assert f"Reading in: {path}" in out and "Input Density Grid Voxel size: 0.5" in out
    assert "Input Temperature Grid: same transform" in out
    chans, _ = exrio.read(str(tmp_path / "ball.exr"))
    assert chans["R"].shape == (16, 24)
    assert all(f.startswith("ball_") for f in os.listdir(tmp_path / "geoDump"))
    # no density grid; no file; both inputs
    only_t = vdbio.write(str(tmp_path / "t.vdb"), [t])
    for args in (["-vdbIn", only_t], ["-vdbIn", str(tmp_path / "none.vdb")],
                 ["-vdbIn", path, "-scene", "sphere"]):
        r = _run([cli_san, *args, "-resWidth", "8", "-resHeight", "8", "-exrOut", str(tmp_path / "x.exr")],
                 cwd=str(tmp_path))
        _clean(r)
        assert r.returncode != 0, args
    assert not os.path.exists(tmp_path / "x.exr")
