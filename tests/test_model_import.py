"""Scene files in the reference's binary format (samples_shared/model_import_bin.cpp:14-168;
include/hvvr/model_import.h, hvvr_amd/csrc/host/model_import.cpp).

CPU pins: a file the library writes for a seeded scene is read back here by a Python
parser written from the format's layout (16-byte header, 56-byte lights, 48-byte mesh
headers holding the transform as quaternion, translation and scale, then the mesh arrays):
every byte is accounted for, the lights are the scene's, each mesh's transform gives, by
the quaternion-to-matrix formula computed here, the scene's model-to-world matrix, and the
meshes' arrays concatenate to the flattened scene's. The library's reader returns the
same scene bit for bit, reads a file this test writes itself (meshes chosen and moved,
a texture before the lights), and rejects truncated, mis-versioned, inconsistent and
over-long files. GPU: the example program renders a .bin file through model_import::load
and createObjects, and its image is the oracle's frame of the same scene."""

import os
import struct
import subprocess

import numpy as np
import pytest

from conftest import ROOT, rgba_delta

BIN = os.path.join(ROOT, "hvvr_amd", "bin", "render_scene")


def _parse(H, data):
    off = 0

    def take(fmt):
        nonlocal off
        v = struct.unpack_from(fmt, data, off)
        off += struct.calcsize(fmt)
        return v

    def arr(dtype, n):
        nonlocal off
        a = np.frombuffer(data, dtype, n, off).copy()
        off += a.nbytes
        return a

    version, ntex, nlights, nmeshes = take("<4I")
    assert version == 1
    textures = []
    for _ in range(ntex):
        fmt, w, h, stride = take("<4I")
        bpp = {1: 4, 2: 4, 3: 8, 4: 16, 5: 8, 6: 4, 7: 4}[fmt]
        textures.append((fmt, w, h, stride, bytes(arr(np.uint8, stride * h * bpp))))
    lights = arr(H.DTYPES["light"], nlights)
    meshes = []
    for _ in range(nmeshes):
        q = take("<4f")
        t = take("<3f")
        (s,) = take("<f")
        nv, nt, nn, nm = take("<4I")
        meshes.append(dict(rotation=q, translation=t, scale=s, verts=arr(H.DTYPES["vertex"], nv),
                           tris=arr(np.dtype([("idx", "<u4", 3), ("material", "<u4")]), nt),
                           nodes=arr(np.dtype([("leafMask", "<u4"), ("data", "<u4", 7)]), nn),
                           materials=arr(H.DTYPES["material"], nm)))
    assert off == len(data), "bytes left after the last mesh"
    return textures, lights, meshes


def _write(H, path, textures, lights, meshes):
    out = [struct.pack("<4I", 1, len(textures), len(lights), len(meshes))]
    for fmt, w, h, stride, texels in textures:
        out += [struct.pack("<4I", fmt, w, h, stride), texels]
    out.append(np.ascontiguousarray(lights, H.DTYPES["light"]).tobytes())
    for m in meshes:
        out.append(struct.pack("<4f3ff4I", *m["rotation"], *m["translation"], m["scale"], len(m["verts"]),
                               len(m["tris"]), len(m["nodes"]), len(m["materials"])))
        out += [m[k].tobytes() for k in ("verts", "tris", "nodes", "materials")]
    with open(path, "wb") as f:
        f.write(b"".join(out))


def _matrix(q, t, s):
    # model-to-world of transform(t, q, s): translation * rotation(q) * uniform scale,
    # column-major (vector_math.h:1008-1025 for the rotation)
    x, y, z, w = (float(c) for c in q)
    r = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                  [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                  [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])
    m = np.eye(4)
    m[:3, :3] = r * s
    m[:3, 3] = t
    return m.T.reshape(16)  # column-major


@pytest.fixture(scope="module")
def room16(H, tmp_path_factory):
    path = tmp_path_factory.mktemp("bin") / "room16.bin"
    H.HostScene.save_bin(path, H.SCENE_ROOM, 0, 16)
    return path, H.HostScene(H.SCENE_ROOM, 0, 16)


def test_written_file_follows_the_reference_layout(H, room16):
    path, scene = room16
    textures, lights, meshes = _parse(H, open(path, "rb").read())
    assert textures == [] and len(meshes) == 16
    assert lights.tobytes() == scene.lights.tobytes()
    for i, m in enumerate(meshes):
        assert len(m["tris"]) == 256
        assert np.allclose(_matrix(m["rotation"], m["translation"], m["scale"]), scene.transforms[i], atol=1e-5)
    # the flattened scene concatenates the meshes: vertices as stored, triangle indices
    # offset by the vertices and materials before each mesh
    assert np.concatenate([m["verts"] for m in meshes]).tobytes() == scene.verts.tobytes()
    vo = np.cumsum([0] + [len(m["verts"]) for m in meshes])
    mo = np.cumsum([0] + [len(m["materials"]) for m in meshes])
    idx = np.concatenate([m["tris"]["idx"] + vo[i] for i, m in enumerate(meshes)])
    mat = np.concatenate([m["tris"]["material"] + mo[i] for i, m in enumerate(meshes)])
    assert np.array_equal(idx, scene.triShade["indices"]) and np.array_equal(mat, scene.triShade["material"])
    assert np.concatenate([m["materials"] for m in meshes]).tobytes() == scene.materials.tobytes()


def test_loaded_file_equals_generated_scene(H, room16):
    path, scene = room16
    loaded = H.HostScene.from_bin(path)
    for name in ("verts", "vertexToTransform", "triShade", "nodes", "materials", "transforms", "lights"):
        assert getattr(loaded, name).tobytes() == getattr(scene, name).tobytes(), name


def test_reader_takes_a_file_written_here(H, room16, tmp_path):
    path, scene = room16
    _, lights, meshes = _parse(H, open(path, "rb").read())
    picked = [dict(meshes[3]), dict(meshes[1])]
    picked[0]["translation"] = (1.0, 2.0, -3.0)
    tex = (1, 3, 2, 4, bytes(range(32)))  # r8g8b8a8_unorm_srgb, 3x2, stride 4
    out = tmp_path / "mine.bin"
    _write(H, out, [tex], lights[:2], picked)
    loaded = H.HostScene.from_bin(out)
    assert loaded.model_count == 2 and len(loaded.lights) == 2
    assert loaded.verts.tobytes() == np.concatenate([m["verts"] for m in picked]).tobytes()
    assert np.allclose(loaded.transforms[0], _matrix(picked[0]["rotation"], (1.0, 2.0, -3.0), picked[0]["scale"]),
                       atol=1e-5)
    assert loaded.transforms[1].tobytes() == scene.transforms[1].tobytes()


@pytest.mark.parametrize("damage", ["version", "truncated", "vertex_index", "child_offset", "trailing"])
def test_malformed_files_are_rejected(H, room16, tmp_path, damage):
    path, _ = room16
    _, lights, meshes = _parse(H, open(path, "rb").read())
    meshes = [dict(m) for m in meshes[:2]]
    data = None
    if damage == "vertex_index":
        meshes[0]["tris"] = meshes[0]["tris"].copy()
        meshes[0]["tris"]["idx"][5, 1] = len(meshes[0]["verts"])
    elif damage == "child_offset":
        nodes = meshes[1]["nodes"].copy()
        k = next(i for i in range(4) if not (nodes["leafMask"][0] >> i) & 1)
        nodes["data"][0, 1 + k] = len(nodes)
        meshes[1]["nodes"] = nodes
    out = tmp_path / "bad.bin"
    _write(H, out, [], lights, meshes)
    data = open(out, "rb").read()
    if damage == "version":
        data = struct.pack("<I", 2) + data[4:]
    elif damage == "truncated":
        data = data[:-7]
    elif damage == "trailing":
        data = data + b"\0" * 3
    open(out, "wb").write(data)
    want = {"version": "version 2", "truncated": "truncated mesh 1", "vertex_index": "vertex index",
            "child_offset": "child node offset", "trailing": "bytes after"}[damage]
    with pytest.raises(H.HvvrError, match=want):
        H.HostScene.from_bin(out)


@pytest.mark.gpu
def test_example_renders_a_bin_file(H, O, tmp_path):
    # the heightfield scene written as .bin, rendered by the example through
    # model_import::load + createObjects at 320x240, MSAA 4, from the heightfield camera
    path, out = tmp_path / "hf.bin", tmp_path / "hf.rgba"
    H.HostScene.save_bin(path, H.SCENE_HEIGHTFIELD_BOXES, 0)
    w, h = 320, 240
    t, q = (8.0, 8.0, 19.0), (-0.3420201, 0.0, 0.0, 0.9396926)  # 40 degrees down, looking at -z
    r = subprocess.run([BIN, str(path), str(out), "size=%dx%d" % (w, h), "msaa=4",
                        "camera=" + ",".join(str(x) for x in (*t, *q))], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    img = np.fromfile(out, np.uint32).reshape(h, w)
    scene = H.HostScene.from_bin(path)
    view = H.HostView.from_samples(H.grid_samples(w, h), w, h, (-w / h, -1.0, w / h, 1.0))
    want = O.render_view(scene, view, 4, camera_to_world=H.transform_matrix(t, q, 1.0))["image"].reshape(h, w)
    assert (want != 0xFF000000).mean() > 0.2
    assert rgba_delta(img, want).max() <= 1


@pytest.mark.gpu
def test_example_renders_a_textured_bin_file(H, O, tmp_path):
    # the heightfield scene with one sRGB texture (rows padded: stride 40 for width 32)
    # named by every other material: the example uploads it through createTexture and the
    # frame matches the oracle reading the same texture
    src, path, out = tmp_path / "hf.bin", tmp_path / "hft.bin", tmp_path / "hft.rgba"
    H.HostScene.save_bin(src, H.SCENE_HEIGHTFIELD_BOXES, 0)
    _, lights, meshes = _parse(H, open(src, "rb").read())
    rng = np.random.default_rng(8)
    padded = rng.integers(0, 256, (32, 40, 4)).astype(np.uint8)
    for m in meshes:
        codes = m["materials"]["textureIDsAndShadingModel"]
        codes[::2] = 1 | (0 << 48) | (0xFFFF << 32) | (0xFFFF << 16)
        m["verts"]["uv"] *= 5.0
    _write(H, path, [(1, 32, 32, 40, padded.tobytes())], lights, meshes)
    w, h = 320, 240
    t, q = (8.0, 8.0, 19.0), (-0.3420201, 0.0, 0.0, 0.9396926)
    r = subprocess.run([BIN, str(path), str(out), "size=%dx%d" % (w, h), "msaa=4",
                        "camera=" + ",".join(str(x) for x in (*t, *q))], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    img = np.fromfile(out, np.uint32).reshape(h, w)
    scene = H.HostScene.from_bin(path)
    view = H.HostView.from_samples(H.grid_samples(w, h), w, h, (-w / h, -1.0, w / h, 1.0))
    m = H.transform_matrix(t, q, 1.0)
    want = O.render_view(scene, view, 4, camera_to_world=m, textures=[("srgb8", padded[:, :32])])["image"]
    plain = O.render_view(scene, view, 4, camera_to_world=m)["image"]
    assert (want.reshape(h, w) != plain.reshape(h, w)).mean() > 0.1  # the texture shows
    assert rgba_delta(img, want.reshape(h, w)).max() <= 1
