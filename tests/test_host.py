"""Host-side building blocks: seeded generators, the BVH builder, scene flattening and
the camera sample hierarchy (CPU only)."""

import numpy as np
import pytest


def test_heightfield_scene_counts(scene1):
    assert len(scene1.triShade) == 8192 + 256 * 12 == 11264
    assert scene1.model_count == 257
    assert len(scene1.materials) == 11264  # per-triangle GGX materials
    assert len(scene1.lights) == 1 and scene1.lights["type"][0] == 1


def test_room_scene_counts(scene2):
    assert len(scene2.triShade) == 1024 * 256 == 262144
    assert scene2.model_count == 1024
    assert scene2.lights["type"].tolist().count(2) == 4 and scene2.lights["type"].tolist().count(1) == 1


def test_generators_are_seeded(H):
    a = H.HostScene(H.SCENE_ROOM, 5, 16)
    b = H.HostScene(H.SCENE_ROOM, 5, 16)
    c = H.HostScene(H.SCENE_ROOM, 6, 16)
    assert np.array_equal(a.verts.view(np.uint8), b.verts.view(np.uint8))
    assert np.array_equal(a.transforms, b.transforms)
    assert not np.array_equal(a.transforms, c.transforms)


def _world_positions(scene):
    M = scene.transforms.reshape(-1, 4, 4).transpose(0, 2, 1).astype(np.float64)
    p = np.concatenate([scene.verts["pos"], np.ones((len(scene.verts), 1))], 1)
    w = np.einsum("vij,vj->vi", M[scene.vertexToTransform], p)
    return w[:, :3] / w[:, 3:]


@pytest.mark.parametrize("which", ["scene1", "scene2"])
def test_closed_meshes_wound_outward(request, which):
    scene = request.getfixturevalue(which)
    P = _world_positions(scene)
    idx = scene.triShade["indices"].astype(np.int64)
    model = scene.vertexToTransform[idx[:, 0]]
    a, b, c = P[idx[:, 0]], P[idx[:, 1]], P[idx[:, 2]]
    vol = np.einsum("ij,ij->i", a, np.cross(b, c)) / 6  # signed volume per triangle
    per_model = np.bincount(model, weights=vol)
    closed = per_model[1:] if which == "scene1" else per_model  # the heightfield is open
    assert np.all(closed > 0), "closed meshes must be wound counter-clockwise seen from outside"


@pytest.mark.parametrize("which", ["scene1", "scene2"])
def test_bvh_topology_invariants(request, which):
    scene = request.getfixturevalue(which)
    nodes = scene.nodes
    T = len(scene.triShade)
    seen = np.zeros(T, np.int32)
    for n in range(len(nodes)):
        lm = nodes["leafMask"][n]
        d = nodes["data"][n]
        kinds = [(lm >> k) & 1 for k in range(4)]
        # leaves first within every node (bvh.cpp:435-439), so leaf k spans data[k]..data[k+1]
        assert kinds == sorted(kinds, reverse=True)
        for k in range(4):
            if kinds[k]:
                assert d[k] <= d[k + 1] <= T
                seen[d[k]:d[k + 1]] += 1
            else:
                child = n + d[1 + k]
                assert n < child < len(nodes), "children follow their parent (depth-first order)"
    assert np.all(seen == 1), "every triangle is in exactly one leaf"


def test_bvh_bounds_contain_triangles(O, scene2):
    wv, _, nodes = O.refit(scene2)
    v = wv["pos"][scene2.triShade["indices"]]  # [t, 3 vertices, xyz]
    lo, hi = v.min(1), v.max(1)
    # walk the tree and check containment child-in-parent and triangle-in-leaf
    stack = [(0, None, None)]
    while stack:
        n, plo, phi = stack.pop()
        b = nodes["bounds"][n]
        for k in range(4):
            clo = -b[[1, 3, 5], k]
            chi = b[[0, 2, 4], k]
            if plo is not None and np.all(np.isfinite(clo)):
                assert np.all(clo >= plo) and np.all(chi <= phi)
            if scene2.nodes["leafMask"][n] >> k & 1:
                s, e = scene2.nodes["data"][n][k], scene2.nodes["data"][n][k + 1]
                if e > s:
                    assert np.all(lo[s:e] >= clo) and np.all(hi[s:e] <= chi)
            else:
                stack.append((n + scene2.nodes["data"][n][1 + k], clo, chi))


def test_sah_builder_on_tiny_meshes(H, O):
    # a single triangle, two triangles and a strip: builds, covers every triangle once
    for ntri in (1, 2, 7, 33):
        verts = np.zeros(ntri + 2, H.DTYPES["vertex"])
        verts["pos"][:, 0] = np.arange(ntri + 2)
        verts["pos"][:, 1] = np.arange(ntri + 2) % 2
        tris = [[i, i + 1, i + 2] if i % 2 == 0 else [i, i + 2, i + 1] for i in range(ntri)]
        mat = np.zeros(1, H.DTYPES["material"])
        s = H.HostScene.from_meshes([(verts, tris, [0] * ntri, None)], mat)
        assert len(s.triShade) == ntri
        seen = np.zeros(ntri, int)
        for n in range(len(s.nodes)):
            for k in range(4):
                if s.nodes["leafMask"][n] >> k & 1:
                    seen[s.nodes["data"][n][k]:s.nodes["data"][n][k + 1]] += 1
        assert np.all(seen == 1)


def test_unreferenced_vertex_rejected(H):
    verts = np.zeros(4, H.DTYPES["vertex"])
    verts["pos"] = [[0, 0, 0], [1, 0, 0], [0, 1, 0], [5, 5, 5]]
    with pytest.raises(H.HvvrError, match="vertices no triangle references"):
        H.HostScene.from_meshes([(verts, [[0, 1, 2]], [0], None)], np.zeros(1, H.DTYPES["material"]))


def test_scene_merge_root_fixups(H, scene1):
    # 257 models -> (257 + 1) / 3 = 86 fixup nodes linking the model roots (scene_update.cpp:39)
    nodes = scene1.nodes
    reached = set()
    stack = [0]
    while stack:
        n = stack.pop()
        reached.add(n)
        for k in range(4):
            if not nodes["leafMask"][n] >> k & 1:
                stack.append(n + nodes["data"][n][1 + k])
    assert len(reached) == len(nodes)
    assert all(nodes["leafMask"][i] != 0xF or i == 0 for i in range(86))


def test_grid_samples(H):
    s = H.grid_samples(4, 2)
    assert s["pixel"].tolist() == [[x, y] for y in range(2) for x in range(4)]
    assert np.allclose(s["position"][:, 0], (s["pixel"][:, 0] + 0.5) / 4)
    assert np.allclose(s["position"][:, 1], (s["pixel"][:, 1] + 0.5) / 2)


def test_barrel_distortion_is_radial(H):
    plain = H.grid_samples(64, 48)
    bent = H.grid_samples(64, 48, 0.22, 0.24)
    u = np.stack([2 * plain["position"][:, 0] - 1, (2 * plain["position"][:, 1] - 1) * 48 / 64], 1)
    r2 = (u ** 2).sum(1)
    f = 1 + 0.22 * r2 + 0.24 * r2 ** 2
    want = np.stack([0.5 + 0.5 * u[:, 0] * f, 0.5 + 0.5 * u[:, 1] * f * 64 / 48], 1)
    assert np.allclose(bent["position"], want, atol=1e-6)


def test_sample_hierarchy(H, view1):
    v = view1
    assert v.valid_sample_count == 640 * 480 and v.sample_count % 8192 == 0
    assert v.tile_count == v.sample_count // 128 and v.block_count == v.sample_count // 8192
    # the remap is a permutation of the valid samples
    assert np.array_equal(np.sort(v.remap), np.arange(v.valid_sample_count))
    # every sample's centre ray lies inside its tile frustum (frustumTestPoint, traversal.h:49-55)
    for t in range(0, v.tile_count, 7):
        f = v.tile_frusta[t]
        d = f["directions"].astype(np.float64)
        planes = np.cross(d, np.roll(d, -1, axis=0))
        c = v.beams["centerRay"][t * 128:(t + 1) * 128].astype(np.float64)
        assert np.all(c @ planes.T <= 1e-7), f"tile {t}: sample outside its frustum"


def test_distorted_view_frusta_contain_samples(H):
    v = H.HostView(H.VIEW_PRESET_ROOM_EYE, 1, 216, 240)
    for t in range(0, v.tile_count, 5):
        d = v.tile_frusta[t]["directions"].astype(np.float64)
        planes = np.cross(d, np.roll(d, -1, axis=0))
        c = v.beams["centerRay"][t * 128:(t + 1) * 128].astype(np.float64)
        assert np.all(c @ planes.T <= 1e-7)
