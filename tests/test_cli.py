"""The command-line front end (lightgraph_amd/host/main.cc): the reference's
options and help aliases (main.cc:138-209, 359-470) and its error behaviour.
Runs on the CPU: no option here reaches the device."""
import os
import re
import subprocess

import pytest

CLI = os.path.join(os.path.dirname(__file__), "..", "lightgraph_amd", "lightgraph")

REFERENCE_OPTIONS = ["-vdbIn", "-exrOut", "-camPos", "-camRot", "-camYRot", "-resWidth", "-resHeight", "-focal",
                     "-samples", "-primaryStep", "-shadowStep", "-cutoff", "-light", "-iter", "-scatterGridResFactor",
                     "-ptsPerKernel", "-densityScale", "-densityMax", "-shadowDensityScale", "-scatterDensityScale",
                     "-scatterDensityMin", "-absorption", "-scattering", "-volumeColor", "-emissionScale",
                     "-scatterScale", "-scatterDensityMaskPower", "-tempColorRamp", "-v", "-geoDump"]
HELP_ALIASES = ["-h", "-help", "--help"]


def _run(*args):
    if not os.path.exists(CLI):
        pytest.fail("lightgraph CLI not built (run __graft_entry__.build())")
    return subprocess.run([CLI, *args], capture_output=True, text=True, timeout=60)


@pytest.mark.parametrize("alias", HELP_ALIASES)
def test_help_lists_every_reference_option(alias):
    r = _run(alias)
    assert r.returncode == 0
    listed = set(re.findall(r"^(-{1,2}[A-Za-z]+)\s", r.stdout, re.M))
    for opt in REFERENCE_OPTIONS + HELP_ALIASES + ["-scene", "-points", "-radius", "-maxConnect", "-precision"]:
        assert opt in listed, opt
    assert "point:yRot" in r.stdout


def test_unknown_option_fails():
    r = _run("-notAnOption")
    assert r.returncode == 1
    assert "-notAnOption is not a valid option" in r.stdout


def test_no_input_fails():
    r = _run("-resWidth", "16")
    assert r.returncode == 1
    r = _run("-vdbIn", "/nonexistent.vdb")
    assert r.returncode == 1 and "vdb: cannot open /nonexistent.vdb" in r.stderr
    r = _run("-vdbIn", "/nonexistent.vdb", "-scene", "sphere")
    assert r.returncode == 1 and "either -vdbIn or -scene" in r.stdout


def test_bad_values_fail_cleanly():
    r = _run("-scene", "cube")
    assert r.returncode != 0
    r = _run("-scene", "sphere", "-precision", "fp8")
    assert r.returncode != 0
