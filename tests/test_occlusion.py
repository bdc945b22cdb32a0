"""The intersect kernel's occlusion skip (hvvr_amd/csrc/gpu/intersect.hip): a wave skips a
survivor only when no subsample ray of the wave could pass its depth test, so the results
must not change. Stress scene: a wall in front of half the view and, behind and around
it, triangles chosen to sit near the bounds' edge cases -- slivers, triangles almost
parallel to the rays, triangles crossing the camera's plane, exact duplicates of the wall
(depth ties, which the later list entry wins, prim_tests.h:209-225), triangles a hair
behind the wall, huge triangles far off -- rendered on the GPU against the oracle: triangle
ids and subsample masks bit-exact, per-subsample depth and barycentrics bit-exact."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

W, H_ = 160, 120


def _stress_scene(H, seed):
    rng = np.random.default_rng(seed)
    tris = []

    def add(a, b, c):
        tris.append((np.asarray(a, np.float64), np.asarray(b, np.float64), np.asarray(c, np.float64)))

    # the wall: two triangles at z = -5 over the left part of the view
    wall = [((-12, -9, -5), (1, -9, -5), (1, 9, -5)), ((-12, -9, -5), (1, 9, -5), (-12, 9, -5))]
    for t in wall:
        add(*t)
    for _ in range(600):  # random triangles behind and beside it
        c = rng.uniform([-8, -6, -30], [8, 6, -5.2])
        s = 10 ** rng.uniform(-1.5, 0.5)
        add(c + rng.normal(size=3) * s, c + rng.normal(size=3) * s, c + rng.normal(size=3) * s)
    for _ in range(150):  # slivers: a vertex on (or almost on) the opposite edge
        a, b = rng.uniform([-8, -6, -25], [8, 6, -4]), rng.uniform([-8, -6, -25], [8, 6, -4])
        c = a + (b - a) * rng.uniform(0.2, 0.8) + rng.normal(size=3) * 10 ** rng.uniform(-7, -3)
        add(a, b, c)
    for _ in range(150):  # nearly parallel to the rays: long in z, thin across
        p = rng.uniform([-6, -5, -8], [6, 5, -6])
        d = p / np.linalg.norm(p)
        side = np.cross(d, rng.normal(size=3))
        side *= 10 ** rng.uniform(-4, -1) / np.linalg.norm(side)
        add(p, p + d * rng.uniform(2, 20), p + side)
    for _ in range(60):  # crossing the camera's plane (some vertices behind the camera)
        add(rng.uniform([-3, -3, 1], [3, 3, 4]), rng.uniform([-6, -4, -12], [6, 4, -6]),
            rng.uniform([-6, -4, -12], [6, 4, -6]))
    for t in wall:  # exact duplicates: depth ties
        add(*t)
    for t in wall:  # a hair behind
        add(*[np.asarray(v) - (0, 0, 1e-4) for v in t])
    for _ in range(20):  # huge and far
        c = rng.uniform([-50, -50, -200], [50, 50, -100])
        add(c + rng.normal(size=3) * 80, c + rng.normal(size=3) * 80, c + rng.normal(size=3) * 80)
    meshes, mats = [], np.zeros(1, H.DTYPES["material"])
    mats["diffuse"] = (0.6, 0.5, 0.4, 1.0)
    mats["specular"] = (0.04, 0.04, 0.04, 1.0)
    mats["glossiness"] = 0.5
    mats["textureIDsAndShadingModel"] = 1 | (0xFFFF << 16) | (0xFFFF << 32) | (0xFFFF << 48)
    for a, b, c in tris:
        v = np.zeros(3, H.DTYPES["vertex"])
        v["pos"] = np.stack([a, b, c]).astype(np.float32)
        n = np.cross(b - a, c - a)
        n = n / (np.linalg.norm(n) + 1e-30)
        v["normal"] = np.array([*n, 0.0], np.float16).view(np.uint16)
        meshes.append((v, [[0, 1, 2]], [0], None))
    return H.HostScene.from_meshes(meshes, mats)


def _view(H):
    t = np.tan(np.radians(35.0))
    return H.HostView.from_samples(H.grid_samples(W, H_), W, H_, (-t * W / H_, -t, t * W / H_, t))


@pytest.mark.parametrize("seed", [0, 1])
@pytest.mark.parametrize("A", [1, 4, 8])
def test_occlusion_skip_changes_nothing(H, O, ctx, seed, A):
    scene = _stress_scene(H, seed)
    view = _view(H)
    m = np.eye(4, dtype=np.float32).reshape(16)
    ref = O.render_view(scene, view, A, camera_to_world=m, with_colors=False)
    gs = H.GpuScene(ctx, scene)
    gv = H.GpuView(ctx, view, A)
    gv.set_hit_output(True)
    gv.render(gs, m)
    counts, lists = gv.read_tile_lists()
    assert np.array_equal(counts, ref["counts"]) and np.array_equal(lists, ref["lists"])
    tri, mask = gv.read_gbuffer()
    rtri, rmask = O.gbuffer_slots(ref["tri_ids"])
    bad = np.nonzero((tri != rtri).any(1) | (mask != rmask).any(1))[0]
    assert len(bad) == 0, f"{len(bad)} samples differ, first {bad[:5]}"
    hit = ref["tri_ids"].reshape(-1, A) != 0xFFFFFFFF
    assert hit.mean() > 0.5
    got = gv.read_hits()
    want_b = O.hits_view(view.beams, m, ref["offsets"], A, ref["tri_ids"], ref["tris"], ref["counts"], None)
    assert np.array_equal(got[..., 0].view(np.uint32), ref["depth"].reshape(-1, A).view(np.uint32))
    assert np.array_equal(got[..., 1:].view(np.uint32), want_b.view(np.uint32))
    st = gv.stats()
    assert st.survivorPairs == ref["survivors"]
    # the skip did engage (most of the scene is behind the wall or the first hits)
    assert 0 < st.testedPairs < 0.8 * st.survivorPairs, (st.testedPairs, st.survivorPairs)
    gv.close()
    gs.close()
