"""The C++ drop-in path: fs2d_run drives fluid_solver_hip through the
reference's fluid_solver& interface (include/fs2d/fluid_solver_hip.hpp) and
must reproduce configuration 1 of the reference bit for bit, both through
solve() with host grids every step and through the resident API."""
from __future__ import annotations

import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT, bits_equal, describe, golden

BIN = os.path.join(ROOT, "fluid-2d_amd", "bin", "fs2d_run")


def run(tmp_path, *extra):
    out = tmp_path / ("res" if "--resident" in extra else "solve")
    out.mkdir()
    cmd = [BIN, "--n", "128", "--steps", "200", "--dt", "0.1", "--diff", "0", "--visc", "0", "--iters", "20",
           "--sources", "disc", "--out", str(out), *extra]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    print(r.stdout.strip())
    return [np.fromfile(out / f"{n}.f32", dtype=np.float32).reshape(130, 130) for n in ("density", "u", "v")]


@pytest.mark.gpu
@pytest.mark.parametrize("mode", [(), ("--resident",)])
def test_cpp_adapter_config1(tmp_path, mode):
    g = golden("config1_R130.npz")
    got = run(tmp_path, *mode)
    for f, name in enumerate(("rho", "u", "v")):
        want = g[f"{name}_s200"]
        assert bits_equal(got[f], want), f"{name}: {describe(got[f], want)}"


def test_cpp_harness_builds_and_links():
    r = subprocess.run([BIN, "--bogus"], capture_output=True, text=True, timeout=30)
    assert r.returncode == 2 and "usage" in r.stdout
