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
NMAP_EXE = os.path.join(ROOT, "tests", "cpp", "_build", "nmap_api")


def build_example(exe=EXE):
    os.makedirs(os.path.dirname(exe), exist_ok=True)
    libdir = os.path.dirname(api.LIB_PATH)
    subprocess.run(["g++", "-O2", "-std=c++17", "-I", os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "cpp", os.path.basename(exe) + ".cpp"), "-o", exe,
                    "-L", libdir, "-lxitils_amd", f"-Wl,-rpath,{libdir}"], check=True)
    return exe


def test_cpp_api_compiles_and_links():
    for exe in (EXE, NMAP_EXE):
        assert os.path.exists(build_example(exe))


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


def _write_plane(d, desc):
    o = desc.objects[0]
    with open(os.path.join(d, "mesh.bin"), "wb") as f:
        np.array([o.positions.shape[0], o.indices.size], np.int32).tofile(f)
        for a in (o.positions, o.normals, o.texcoords, o.tangents, o.bitangents):
            np.ascontiguousarray(a, np.float32).tofile(f)
        np.ascontiguousarray(o.indices, np.int32).tofile(f)
    for name, img in (("tex.bin", desc.textures[0]), ("sky.bin", desc.env.map)):
        with open(os.path.join(d, name), "wb") as f:
            np.array([img.shape[1], img.shape[0]], np.int32).tofile(f)
            np.ascontiguousarray(img, np.float32).tofile(f)
    cam = desc.camera
    np.concatenate([cam.cam_to_world.astype(np.float32).reshape(16),
                    np.array([cam.fov, cam.aspect], np.float32)]).tofile(os.path.join(d, "cam.bin"))


@pytest.mark.gpu
def test_cpp_api_microfacet_normal_mapping_matches_python_binding():
    """MicrofacetNormalMapping + Material::normalmap + setGeometry with
    texCoords and tangents, through the C++ API, give the Python binding's
    image of the same arrays."""
    exe = NMAP_EXE if os.path.exists(NMAP_EXE) else build_example(NMAP_EXE)
    desc = scenes.displaced_plane(grid=48, width=48, height=32, normal_map_size=128, sky_size=(128, 64), variant=2)
    with tempfile.TemporaryDirectory() as d:
        _write_plane(d, desc)
        out = os.path.join(d, "f.f32")
        print("[gpu] C++ API normal-mapped plane render", flush=True)
        r = subprocess.run([exe, out, "48", "32", "4", d], capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stdout + r.stderr
        fc = np.fromfile(out, np.float32).reshape(32, 48, 3)
    dev = api.Device(0)
    fp = api.Scene(dev, desc).render(48, 32, 4, seed=1, max_depth=12, flags=scenes.FLAG_ENV_NEE)
    assert np.isfinite(fc).all()
    assert rel_mse(fc / 4, fp / 4) < 1e-6
