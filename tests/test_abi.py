"""C-ABI checks that need no GPU: the library loads, exports every entry point
include/xitils_amd.h declares, and its struct layouts match the header."""
import ctypes as C
import os
import re
import subprocess
import tempfile

import pytest

from xitils_amd import api

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "xitils_amd.h")


def declared_functions():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^(?:const char\*|xt_status|void)\s+(xt_\w+)\s*\(", text, re.M)))


def test_library_exports_every_declared_symbol():
    L = api.lib()
    names = declared_functions()
    assert len(names) >= 10
    for n in names:
        assert hasattr(L, n), n
    assert set(names) == set(api.EXPORTED)
    out = subprocess.run(["nm", "-D", "--defined-only", api.LIB_PATH], capture_output=True, text=True).stdout
    exported = set(re.findall(r" T (xt_\w+)", out))
    assert set(names) <= exported


def test_version_string():
    assert b"gfx950" in api.lib().xt_version()


STRUCTS = {
    "xt_material": api.xt_material,
    "xt_object_desc": api.xt_object_desc,
    "xt_env": api.xt_env,
    "xt_scene_desc": api.xt_scene_desc,
    "xt_texture": api.xt_texture,
    "xt_camera": api.xt_camera,
    "xt_render_params": api.xt_render_params,
    "xt_render_stats": api.xt_render_stats,
    "xt_guide_config": api.xt_guide_config,
    "xt_batch_item": api.xt_batch_item,
}


def test_struct_layouts_match_header():
    prog = ['#include <stdio.h>', '#include <stddef.h>', f'#include "{HEADER}"', "int main(void) {"]
    for name, cls in STRUCTS.items():
        prog.append(f'printf("{name} %zu\\n", sizeof({name}));')
        for f, _ in cls._fields_:
            prog.append(f'printf("{name}.{f} %zu\\n", offsetof({name}, {f}));')
    prog += ['printf("xt_ray %zu\\n", sizeof(xt_ray));', 'printf("xt_hit %zu\\n", sizeof(xt_hit));', "return 0; }"]
    with tempfile.TemporaryDirectory() as d:
        src = os.path.join(d, "l.c")
        open(src, "w").write("\n".join(prog))
        exe = os.path.join(d, "l")
        subprocess.run(["gcc", "-std=c99", "-o", exe, src], check=True)
        lines = subprocess.run([exe], capture_output=True, text=True, check=True).stdout.split("\n")
    got = dict(l.split() for l in lines if l)
    for name, cls in STRUCTS.items():
        assert int(got[name]) == C.sizeof(cls), name
        for f, _ in cls._fields_:
            assert int(got[f"{name}.{f}"]) == getattr(cls, f).offset, f"{name}.{f}"
    assert int(got["xt_ray"]) == 32 and int(got["xt_hit"]) == 16
    assert int(got["xt_material"]) == 32


def test_invalid_arguments_are_reported_without_a_device():
    L = api.lib()
    h = C.c_void_p()
    assert L.xt_scene_create(None, None, C.byref(h)) == 1
    assert b"NULL" in L.xt_last_error()
    assert L.xt_render(None, None, None, None, 0, 0) == 1
    assert L.xt_trace_closest(None, None, 0, None, 0) == 1
    assert L.xt_scene_destroy(None) == 0
    assert L.xt_device_close(None) == 0


def test_device_open_fails_cleanly_without_gpu():
    L = api.lib()
    h = C.c_void_p()
    st = L.xt_device_open(0, C.byref(h))
    if st == 0:
        L.xt_device_close(h)
        pytest.skip("a gfx950 device is present")
    assert st == 5   # XT_ERR_NO_DEVICE
    assert L.xt_last_error()
    with pytest.raises(api.XtError):
        api.Device(0)
