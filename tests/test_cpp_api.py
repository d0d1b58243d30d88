"""The C++ host API (include/xitils/xitils.hpp) compiles against the C ABI
with a plain g++ and, on the GPU, renders the same image as the Python
binding of the same scene."""
import os
import subprocess
import tempfile

import numpy as np
import pytest

from helpers import rel_mse
from xitils_amd import api, scenes

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "cpp", "_build", "cornell_api")


def build_example():
    os.makedirs(os.path.dirname(EXE), exist_ok=True)
    libdir = os.path.dirname(api.LIB_PATH)
    subprocess.run(["g++", "-O2", "-std=c++17", "-I", os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "cpp", "cornell_api.cpp"), "-o", EXE,
                    "-L", libdir, "-lxitils_amd", f"-Wl,-rpath,{libdir}"], check=True)
    return EXE


def test_cpp_api_compiles_and_links():
    exe = build_example()
    assert os.path.exists(exe)


@pytest.mark.gpu
def test_cpp_api_render_matches_python_binding():
    exe = EXE if os.path.exists(EXE) else build_example()
    with tempfile.TemporaryDirectory() as d:
        out = os.path.join(d, "f.f32")
        print("[gpu] C++ API cornell render", flush=True)
        r = subprocess.run([exe, out, "64", "64", "4"], capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stdout + r.stderr
        fc = np.fromfile(out, np.float32).reshape(64, 64, 3)
    dev = api.Device(0)
    g = api.Scene(dev, scenes.cornell(64, 64))
    fp = g.render(64, 64, 4, seed=1, max_depth=8)
    # the two callers derive the camera's tan(fov/2) through different float
    # paths (std::tan on 40 * ToRad vs numpy on radians(40)), so rays can
    # differ in the last ulp: same image up to sampling-noise-free rounding
    assert rel_mse(fc / 4, fp / 4) < 1e-6
