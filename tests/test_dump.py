"""Scene and ray dumps (Raycaster::dumpSceneAndRays; raycaster/render.cpp:28-201,
DumpRaysKernel intersect.cu:407-441).

GPU, through the example program after one frame of the C++ API: the binary OFF file holds
the scene's world-space vertices (each model's vertices through its model-to-world
transform, computed here in double) and its triangles with vertex indices offset per
model; the BFF file holds each model's BVH topology, whose leaf masks and leaf triangle
ranges match the flattened scene's nodes; the RFF file holds one ray per subsample in
image order, origin at the camera and direction normalize(R (c + du a_x + dv a_y)) of the
sample's beam and the subsample offsets a, computed here in double. The text forms carry
the same numbers."""

import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT
from test_raycaster_api import CAMERA, H_, MSAA, VIEWPORT, W, _scene, _write_scene

BIN = os.path.join(ROOT, "hvvr_amd", "bin", "render_scene")
pytestmark = pytest.mark.gpu


def _ints(data, off, n):
    return np.frombuffer(data, np.int32, n, off)


def _flat_scene(H, models, lights):
    meshes, mat_offset, all_mats = [], 0, []
    for (v, tris, mats), materials, (t, r, s) in models:
        meshes.append((v, tris, mats + mat_offset, H.transform_matrix(t, r, s)))
        all_mats.append(materials)
        mat_offset += len(materials)
    return H.HostScene.from_meshes(meshes, np.concatenate(all_mats), lights)


def test_dumps_match_scene_and_rays(H, tmp_path):
    models, lights = _scene(H)
    path = tmp_path / "scene.bin"
    _write_scene(path, models, lights)
    for form in ("", ",text"):
        r = subprocess.run([BIN, str(path), str(tmp_path / "o.rgba"), "dump=%s/%s" % (tmp_path, form)],
                           capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stderr
    scene = _flat_scene(H, models, lights)

    # OFF: world-space vertices and triangles
    off = open(tmp_path / "scene_dump.offb", "rb").read()
    assert off[:4] == b"OFFb"
    nv, nt, ne = _ints(off, 4, 3)
    assert (nv, nt, ne) == (len(scene.verts), len(scene.triShade), 0)
    verts = np.frombuffer(off, np.float32, 3 * nv, 16).reshape(nv, 3)
    tris = np.frombuffer(off, np.uint32, 3 * nt, 16 + 12 * nv).reshape(nt, 3)
    assert 16 + 12 * nv + 12 * nt == len(off)
    xf = scene.transforms.reshape(-1, 4, 4).astype(np.float64)  # column-major: row k = column k
    p = np.concatenate([scene.verts["pos"].astype(np.float64), np.ones((nv, 1))], 1)
    world = np.einsum("vk,vkj->vj", p, xf[scene.vertexToTransform])[:, :3]
    assert np.abs(verts - world).max() < 1e-4
    assert np.array_equal(tris, scene.triShade["indices"])
    text = open(tmp_path / "scene_dump.off").read().split()
    assert text[0] == "OFF" and [int(x) for x in text[1:4]] == [nv, nt, 0]
    tv = np.array(text[4:4 + 3 * nv], float).reshape(nv, 3)
    assert np.abs(tv - verts).max() < 1e-5
    tt = np.array(text[4 + 3 * nv:], int).reshape(nt, 4)
    assert (tt[:, 0] == 3).all() and np.array_equal(tt[:, 1:], tris)

    # BFF: the models' topology nodes, leaf masks and triangle ranges as flattened
    bff = open(tmp_path / "bvh_dump.bffb", "rb").read()
    assert bff[:4] == b"BFFb"
    (n,) = _ints(bff, 4, 1)
    nodes = np.frombuffer(bff, np.uint32, 6 * n, 8).reshape(n, 6)
    assert 8 + 24 * n == len(bff)
    flat = scene.nodes[len(scene.nodes) - n:]
    assert np.array_equal(nodes[:, 0], flat["leafMask"])
    tri_base = np.repeat(np.cumsum([0] + [len(m[0][1]) for m in models])[:-1],
                         np.diff(np.concatenate([[0], np.cumsum(_node_counts(nodes, len(models)))])))
    for k in range(4):
        leaf = (nodes[:, 0] >> k) & 1 == 1
        assert np.array_equal(nodes[leaf, 2 + k] + tri_base[leaf], flat["data"][leaf, 1 + k])
    btext = open(tmp_path / "bvh_dump.bff").read().split()
    assert btext[0] == "BFF" and int(btext[1]) == n
    assert np.array_equal(np.array(btext[2:], np.uint64).reshape(n, 6), nodes)

    # RFF: one ray per subsample in image order
    rff = open(tmp_path / "ray_dump.rffb", "rb").read()
    assert rff[:4] == b"RFFb"
    (nr,) = _ints(rff, 4, 1)
    rays = np.frombuffer(rff, np.float32, 6 * nr, 8).reshape(nr, 6)
    assert nr == W * H_ * MSAA and 8 + 24 * nr == len(rff)
    view = H.HostView.from_samples(H.grid_samples(W, H_), W, H_, VIEWPORT)
    m = H.transform_matrix(*CAMERA).reshape(4, 4).astype(np.float64)
    rot = m[:3, :3].T  # column-major storage
    b = view.beams[view.remap]
    c, du, dv = (b[k].astype(np.float64) @ rot.T for k in ("centerRay", "du", "dv"))
    off = H.subsample_offsets(MSAA).astype(np.float64)
    d = c[:, None, :] + du[:, None, :] * off[None, :, :1] + dv[:, None, :] * off[None, :, 1:]
    d /= np.linalg.norm(d, axis=-1, keepdims=True)
    assert np.abs(rays[:, 3:] - d.reshape(-1, 3)).max() < 2e-6
    assert np.abs(rays[:, :3] - m[3, :3]).max() == 0
    rtext = open(tmp_path / "ray_dump.rff").read().split()
    assert rtext[0] == "RFF" and int(rtext[1]) == nr
    assert np.array_equal(np.array(rtext[2:], np.float32).reshape(nr, 6), rays)


def _node_counts(nodes, models):
    # nodes per model: a model's topology is contiguous, depth first from its root, and an
    # internal child's data[1 + k] (column 2 + k) is its offset from the parent
    counts, start = [], 0
    for _ in range(models):
        end, i = start + 1, start
        while i < end:
            for k in range(4):
                if not (nodes[i, 0] >> k) & 1:
                    end = max(end, i + int(nodes[i, 2 + k]) + 1)
            i += 1
        counts.append(end - start)
        start = end
    assert start == len(nodes)
    return counts
