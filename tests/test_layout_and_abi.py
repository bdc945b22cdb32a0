"""Layouts, exported C symbols and CPU-side error paths of the C ABIs (no GPU needed)."""

import ctypes
import os
import re

import numpy as np
import pytest

from conftest import ROOT


def _declared_functions(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?[A-Za-z_][\w\s\*]*?\b(hvvr_\w+)\s*\(", text, flags=re.M)
    return sorted(set(names))


@pytest.mark.parametrize("header,libname", [("hvvr_gpu.h", "libhvvr_gpu.so"), ("hvvr_host.h", "libhvvr.so")])
def test_every_declared_symbol_is_exported(H, header, libname):
    lib = ctypes.CDLL(os.path.join(H.LIB_DIR, libname))
    names = _declared_functions(header)
    assert len(names) > 10
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, f"{libname} does not export {missing}"


def test_layouts_match_oracle(H, O):
    sizes = O.layout_sizes()
    # vertex, tri, trishade, node, material, sfrustum, beam, frustum, stack entry, light env
    assert list(sizes) == [32, 36, 16, 128, 64, 96, 36, 400, 8, 372]
    assert H.DTYPES["vertex"].itemsize == sizes[0]
    assert H.DTYPES["tri_intersect"].itemsize == sizes[1]
    assert H.DTYPES["tri_shade"].itemsize == sizes[2]
    assert H.DTYPES["node"].itemsize == sizes[3]
    assert H.DTYPES["material"].itemsize == sizes[4]
    assert H.DTYPES["sfrustum"].itemsize == sizes[5]
    assert H.DTYPES["beam"].itemsize == sizes[6]
    # BVHNode: bounds at 0..95, boxData at 96 (bvh_node.h:17-37, gpu_scene_state.cu:145-146)
    assert H.DTYPES["node"].fields["leafMask"][1] == 96


def test_gpu_version_string(H):
    assert H.load_gpu().hvvr_gpu_version().decode().startswith("hvvr_gpu")


def test_gpu_null_arguments_rejected(H):
    lib = H.load_gpu()
    assert lib.hvvr_gpu_context_synchronize(None) == -1
    assert lib.hvvr_gpu_scene_update(None) == -1
    assert lib.hvvr_gpu_view_render(None, None, None, None) == -1
    assert b"null" in lib.hvvr_gpu_last_error()


def test_host_errors_reported(H):
    lib = H.load_host()
    h = ctypes.c_void_p()
    assert lib.hvvr_host_scene_create(99, ctypes.c_uint64(0), 0, ctypes.byref(h)) < 0
    assert b"unknown scene kind" in lib.hvvr_host_last_error()
    with pytest.raises(H.HvvrError):
        H.subsample_offsets(64)


def test_context_without_device_reports_no_device(H):
    import subprocess
    import sys
    code = ("import sys; sys.path.insert(0, %r); import hvvr_amd as H, ctypes;"
            "h=ctypes.c_void_p(); rc=H.load_gpu().hvvr_gpu_context_create(0, ctypes.byref(h)); print(rc)") % ROOT
    env = dict(os.environ, HIP_VISIBLE_DEVICES="-1")
    out = subprocess.run([sys.executable, "-c", code], env=env, capture_output=True, text=True, timeout=120)
    assert out.stdout.strip() in ("-3",), out.stdout + out.stderr
