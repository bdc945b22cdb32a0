"""The hvvr::Raycaster C++ API (include/hvvr/raycaster.h) driven by the example program
hvvr_amd/bin/render_scene, as an application of the reference would drive it
(projects/hvvr_samples/modelviewer/modelviewer.cpp).

GPU: the images it writes match the oracle's frame for the same meshes, transforms,
lights and camera (colour within 1/255), before and after a transform-only update.
CPU: the scene file round trip and the error path without a GPU."""

import os
import struct
import subprocess

import numpy as np
import pytest

from conftest import ROOT, rgba_delta

BIN = os.path.join(ROOT, "hvvr_amd", "bin", "render_scene")
W, H_, MSAA = 160, 120, 4
VIEWPORT = (-0.6, -0.45, 0.6, 0.45)
CAMERA = ((0.0, 3.0, 9.0), (-0.1305262, 0.0, 0.0, 0.9914449), 1.0)
MOVED = ((0.5, -0.3, -1.0), (0.0, 0.2588190, 0.0, 0.9659258), 1.1)


def _grid_mesh(H, n=20, size=8.0):
    xs = np.linspace(-size / 2, size / 2, n + 1, dtype=np.float32)
    gx, gz = np.meshgrid(xs, xs)
    gy = (0.4 * np.sin(gx * 1.3) * np.cos(gz * 0.9)).astype(np.float32)
    v = np.zeros((n + 1) * (n + 1), H.DTYPES["vertex"])
    v["pos"] = np.stack([gx.ravel(), gy.ravel(), gz.ravel()], 1)
    v["normal"] = np.array([0, 1, 0, 0], np.float16).view(np.uint16)
    tris, mats = [], []
    for j in range(n):
        for i in range(n):
            a, b = j * (n + 1) + i, j * (n + 1) + i + 1
            c, d = a + n + 1, b + n + 1
            tris += [[a, c, b], [b, c, d]]
            mats += [(i + j) % 2] * 2
    return v, np.array(tris, np.uint32), np.array(mats, np.uint32)


def _box_mesh(H, lo, hi):
    lo, hi = np.asarray(lo, np.float32), np.asarray(hi, np.float32)
    corners = np.array([[x, y, z] for z in (lo[2], hi[2]) for y in (lo[1], hi[1]) for x in (lo[0], hi[0])],
                       np.float32)
    faces = [(0, 2, 3, 1, (0, 0, -1)), (4, 5, 7, 6, (0, 0, 1)), (0, 1, 5, 4, (0, -1, 0)),
             (2, 6, 7, 3, (0, 1, 0)), (0, 4, 6, 2, (-1, 0, 0)), (1, 3, 7, 5, (1, 0, 0))]
    v = np.zeros(24, H.DTYPES["vertex"])
    tris = []
    for f, (a, b, c, d, nrm) in enumerate(faces):
        v["pos"][4 * f:4 * f + 4] = corners[[a, b, c, d]]
        v["normal"][4 * f:4 * f + 4] = np.array([*nrm, 0], np.float16).view(np.uint16)
        tris += [[4 * f, 4 * f + 1, 4 * f + 2], [4 * f, 4 * f + 2, 4 * f + 3]]
    return v, np.array(tris, np.uint32), np.zeros(12, np.uint32)


def _materials(H, colors):
    m = np.zeros(len(colors), H.DTYPES["material"])
    for i, c in enumerate(colors):
        m["diffuse"][i] = [*c, 1]
        m["specular"][i] = [0.04, 0.04, 0.04, 1]
        m["glossiness"][i] = 0.4 + 0.1 * i
        m["opacity"][i] = 1
        m["textureIDsAndShadingModel"][i] = 1 | (0xFFFF << 16) | (0xFFFF << 32) | (0xFFFF << 48)
    return m


def _scene(H):
    models = [
        (_grid_mesh(H), _materials(H, [(0.7, 0.7, 0.6), (0.3, 0.4, 0.7)]), ((0.0, -0.5, 0.0), (0, 0, 0, 1), 1.0)),
        (_box_mesh(H, (-1, 0, -1), (1, 2, 1)), _materials(H, [(0.8, 0.2, 0.2)]),
         ((-1.5, -0.2, 0.5), (0.0, 0.3826834, 0.0, 0.9238795), 1.0)),
        (_box_mesh(H, (-0.5, 0, -0.5), (0.5, 3, 0.5)), _materials(H, [(0.2, 0.8, 0.3)]),
         ((2.0, -0.4, -1.5), (0, 0, 0, 1), 0.8)),
    ]
    lights = np.zeros(2, H.DTYPES["light"])
    lights["type"] = [1, 2]
    lights["body"][0, :6] = [0.3, 0.8, 0.5, 1.2, 1.1, 1.0]
    lights["body"][1, :8] = [0.0, 4.0, 3.0, 12.0, 3.0, 3.0, 2.5, 1 / 10.0]
    return models, lights


def _write_scene(path, models, lights, moved=None):
    with open(path, "wb") as f:
        f.write(b"HVSC" + struct.pack("<II", 1, len(models)))
        for (v, tris, mats), materials, (t, r, s) in models:
            f.write(struct.pack("<III", len(v), len(tris), len(materials)))
            f.write(struct.pack("<8f", *t, *r, s))
            f.write(v.tobytes() + tris.astype("<u4").tobytes() + mats.astype("<u4").tobytes() + materials.tobytes())
        f.write(struct.pack("<I", len(lights)) + lights.tobytes())
        f.write(struct.pack("<III", W, H_, MSAA) + struct.pack("<4f", *VIEWPORT))
        t, r, s = CAMERA
        f.write(struct.pack("<8f", *t, *r, s))
        if moved is None:
            f.write(struct.pack("<I", 0))
        else:
            t, r, s = moved
            f.write(struct.pack("<I8f", 1, *t, *r, s))


def _oracle_image(H, O, models, lights, transforms, spherical=None):
    meshes, mat_offset, all_mats = [], 0, []
    for ((v, tris, mats), materials, _), (t, r, s) in zip(models, transforms):
        meshes.append((v, tris, mats + mat_offset, H.transform_matrix(t, r, s)))
        all_mats.append(materials)
        mat_offset += len(materials)
    scene = H.HostScene.from_meshes(meshes, np.concatenate(all_mats), lights)
    view = H.HostView.from_samples(H.grid_samples(W, H_), W, H_, VIEWPORT, spherical=spherical)
    m = H.transform_matrix(*CAMERA)
    return O.render_view(scene, view, MSAA, camera_to_world=m)["image"]


def test_example_is_built():
    assert os.access(BIN, os.X_OK), f"{BIN} missing: run __graft_entry__.build()"


def test_scene_file_is_rejected_when_malformed(tmp_path):
    bad = tmp_path / "bad.bin"
    bad.write_bytes(b"NOPE")
    r = subprocess.run([BIN, str(bad), str(tmp_path / "o.rgba")], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "truncated header" in r.stderr
    bad.write_bytes(b"HVSC" + struct.pack("<I", 2))
    r = subprocess.run([BIN, str(bad), str(tmp_path / "o.rgba")], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "HVSC" in r.stderr


@pytest.mark.skipif(os.path.exists("/dev/kfd"), reason="a GPU driver is present")
def test_no_device_raises_raycaster_error(H, tmp_path):
    models, lights = _scene(H)
    path = tmp_path / "scene.bin"
    _write_scene(path, models, lights)
    r = subprocess.run([BIN, str(path), str(tmp_path / "o.rgba")], capture_output=True, text=True, timeout=60)
    assert r.returncode == 3 and "RaycasterError" in r.stderr


@pytest.mark.gpu
def test_cpp_api_matches_oracle(H, O, tmp_path):
    models, lights = _scene(H)
    path, out = tmp_path / "scene.bin", tmp_path / "out.rgba"
    _write_scene(path, models, lights, moved=MOVED)
    r = subprocess.run([BIN, str(path), str(out), "3"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    imgs = np.fromfile(out, np.uint32).reshape(2, H_, W)
    base = [m[2] for m in models]
    want0 = _oracle_image(H, O, models, lights, base).reshape(H_, W)
    want1 = _oracle_image(H, O, models, lights, [MOVED] + base[1:]).reshape(H_, W)
    assert not np.array_equal(want0, want1), "the move must change the image"
    assert (want0 != 0xFF000000).mean() > 0.3, "scene should cover a good part of the frame"
    assert rgba_delta(imgs[0], want0).max() <= 1
    assert rgba_delta(imgs[1], want1).max() <= 1


@pytest.mark.gpu
def test_cpp_api_foveated_matches_c_abi(H, ctx, tmp_path):
    # the Raycaster's foveated mode (eye transforms, screen cull planes, history) against
    # the same frames driven through the C ABI: bit-identical images
    models, lights = _scene(H)
    path, out = tmp_path / "scene.bin", tmp_path / "out.rgba"
    _write_scene(path, models, lights)
    r = subprocess.run([BIN, str(path), str(out), "2", "foveated"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    img = np.fromfile(out, np.uint32).reshape(H_, W)
    meshes, mat_offset, all_mats = [], 0, []
    for (v, tris, mats), materials, (t, rot, s) in models:
        meshes.append((v, tris, mats + mat_offset, H.transform_matrix(t, rot, s)))
        all_mats.append(materials)
        mat_offset += len(materials)
    scene = H.HostScene.from_meshes(meshes, np.concatenate(all_mats), lights)
    screen = H.HostView.from_samples(H.grid_samples(W, H_), W, H_, VIEWPORT)
    fov = H.HostFoveated(H.HostFoveated.FEB2017)
    gv = H.GpuFoveatedView(ctx, fov, W, H_, MSAA)
    gs = H.GpuScene(ctx, scene)
    hist = H.foveated_history()
    m = H.transform_matrix(*CAMERA)
    for _ in range(2):
        fr = H.foveated_frame(m, (0.0, 0.0, -1.0), screen.sample_to_camera, screen.sample_bounds, hist)
        gv.render_foveated(gs, fr)
    want = gv.read_image()
    assert (want != 0).mean() > 0.9
    assert np.array_equal(img, want)


@pytest.mark.gpu
def test_cpp_api_spherical_warp_and_shared_sample_data(H, O, tmp_path):
    # Camera::setSphericalWarpSettings(100, 75) before setSamples: the image is the
    # oracle's frame of the same spherical-section view; and a second camera given the
    # first one's sample hierarchy (getSampleData / setSampleData) renders the same image
    models, lights = _scene(H)
    path, out = tmp_path / "scene.bin", tmp_path / "out.rgba"
    _write_scene(path, models, lights)
    r = subprocess.run([BIN, str(path), str(out), "1", "spherical=100,75"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    img = np.fromfile(out, np.uint32).reshape(H_, W)
    base = [m[2] for m in models]
    want = _oracle_image(H, O, models, lights, base, spherical=(100.0, 75.0)).reshape(H_, W)
    assert not np.array_equal(want, _oracle_image(H, O, models, lights, base).reshape(H_, W))
    assert rgba_delta(img, want).max() <= 1
    r = subprocess.run([BIN, str(path), str(out), "1", "share-samples"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    imgs = np.fromfile(out, np.uint32).reshape(2, H_, W)
    assert (imgs[0] != 0xFF000000).mean() > 0.3
    assert np.array_equal(imgs[0], imgs[1])
