"""Pins the CPU oracle against mathematics that fixes its results independently:
closed forms, brute force on small inputs and invariants. (The reference ships no
golden vectors or tests; each test's docstring says what it pins.)"""

import itertools

import numpy as np
import pytest

BAD = 0xFFFFFFFF


# ---------------------------------------------------------------------------------------
# subsample pattern (gpu_samples.h:35-61): alpha_i = sqrt((i+.5)/A) (cos t, sin t),
# t = (i+.5)/A * tau[A] * 6.28
# ---------------------------------------------------------------------------------------
TAU = {1: 1, 2: 1, 4: 3, 8: 3, 16: 7, 32: 7}  # Mara et al. 2016 table entries


@pytest.mark.parametrize("A", [1, 2, 4, 8, 16, 32])
def test_subsample_offsets_closed_form(H, A):
    got = H.subsample_offsets(A).astype(np.float64)
    i = np.arange(A)
    a = (i + 0.5) / A
    # the spiral-turn product is a float product in the reference (tau[A] * 6.28f)
    t = a * np.float64(np.float32(TAU[A]) * np.float32(6.28))
    want = np.stack([np.cos(t), np.sin(t)], 1) * np.sqrt(a)[:, None]
    assert np.allclose(got, want, rtol=0, atol=2e-7)
    assert np.all(np.hypot(got[:, 0], got[:, 1]) < 1.0)


def test_subsample_spin_rotates(H):
    base = H.subsample_offsets(8).astype(np.float64)
    spun = H.subsample_offsets(8, 0.5).astype(np.float64)
    c, s = np.cos(0.5), np.sin(0.5)
    rot = np.stack([base[:, 0] * c - base[:, 1] * s, base[:, 0] * s + base[:, 1] * c], 1)
    assert np.allclose(spun, rot, atol=3e-7)


# ---------------------------------------------------------------------------------------
# GBuffer compression (intersect.cu:28-89): group subsamples by triangle id, ascending,
# misses last, masks partition the full mask
# ---------------------------------------------------------------------------------------
def _group(ids):
    out = []
    for t in sorted(set(int(x) for x in ids)):
        m = 0
        for i, x in enumerate(ids):
            if int(x) == t:
                m |= 1 << i
        out.append((t, m))
    return out


@pytest.mark.parametrize("A", [1, 2, 4])
def test_gbuffer_compression_exhaustive(O, A):
    # 2^26 - 1 is the largest id the (id << 5 | subsample) key round-trips: the key is
    # decoded with a sign-filling shift (intersect.cu:67,71), so id bit 26 would read as a miss
    for ids in itertools.product([0, 5, 7, (1 << 26) - 1, BAD], repeat=A):
        t, m = O.emit_gbuffer(np.array(ids, np.uint32))
        assert list(zip(t.tolist(), m.tolist())) == _group(ids)


@pytest.mark.parametrize("A", [8, 16, 32])
def test_gbuffer_compression_random(O, A):
    rng = np.random.default_rng(A)
    for _ in range(300):
        ids = rng.choice(np.array([3, 9, 11, 200, BAD], np.uint32), size=A)
        t, m = O.emit_gbuffer(ids)
        assert list(zip(t.tolist(), m.tolist())) == _group(ids)
        full = (1 << A) - 1
        assert np.bitwise_or.reduce(m) == full and sum(bin(x).count("1") for x in m) == A


# ---------------------------------------------------------------------------------------
# frustum vs AABB (traverse_avx.h:104-217): conservative - never culls a box that a ray
# of the frustum enters
# ---------------------------------------------------------------------------------------
def _frustum(rng):
    fr = np.zeros(1, dtype=[("origins", "<f4", (4, 3)), ("directions", "<f4", (4, 3))])
    c = rng.normal(size=3) * 3
    fwd = rng.normal(size=3)
    fwd /= np.linalg.norm(fwd)
    up = np.cross(fwd, rng.normal(size=3))
    up /= np.linalg.norm(up)
    right = np.cross(fwd, up)
    w, h = rng.uniform(0.02, 0.4, 2)
    corners = [(-w, -h), (w, -h), (w, h), (-w, h)]
    d = np.array([fwd + x * right + y * up for x, y in corners])
    d /= np.linalg.norm(d, axis=1)[:, None]
    fr["origins"][0] = np.tile(c, (4, 1))
    fr["directions"][0] = d
    return fr, c, d


def _node_with_boxes(H, lo, hi):
    node = np.zeros(1, H.DTYPES["node"])
    b = node["bounds"][0]
    for a in range(3):
        b[2 * a] = hi[:, a]
        b[2 * a + 1] = -lo[:, a]
    return node


def test_frustum_aabb_conservative(H, O):
    rng = np.random.default_rng(1)
    eye = np.eye(4, dtype=np.float32).reshape(16)
    culled = 0
    for trial in range(200):
        fr, c, d = _frustum(rng)
        _, pre = O.frustum_setup(fr, eye)
        # 4 boxes: two around points inside the frustum, two random
        pts = []
        for _ in range(2):
            wts = rng.dirichlet(np.ones(4))
            ray = (wts[:, None] * d).sum(0)
            pts.append(c + ray * rng.uniform(0.5, 20))
        pts += [c + rng.normal(size=3) * 10 for _ in range(2)]
        pts = np.array(pts)
        ext = rng.uniform(0.01, 1.0, size=(4, 3))
        lo, hi = (pts - ext).astype(np.float32), (pts + ext).astype(np.float32)
        node = _node_with_boxes(H, lo, hi)
        mask = O.test_children_refine(pre, node, O.test_children(pre, node))
        # brute force: does any ray of the frustum (convex combos of corner dirs) enter box k?
        wts = rng.dirichlet(np.ones(4), size=400)
        rays = wts @ d
        ts = np.linspace(0, 40, 400)
        P = c[None, None, :] + rays[:, None, :] * ts[None, :, None]
        for k in range(4):
            inside = np.all((P >= lo[k] - 1e-5) & (P <= hi[k] + 1e-5), axis=2).any()
            if inside:
                assert mask & (1 << k), f"trial {trial}: box {k} hit by a frustum ray but culled"
            culled += 0 if mask & (1 << k) else 1
    assert culled > 100  # the test also culls


def test_degenerate_frustum_culls_everything(H, O, scene1):
    fr = np.zeros(1, H.DTYPES["sfrustum"])
    fr["origins"][0] = np.inf
    _, pre = O.frustum_setup(fr, np.eye(4, dtype=np.float32).reshape(16))
    assert pre["degenerate"][0] == 1
    _, _, nodes = O.refit(scene1)
    assert len(O.block_traverse(nodes, pre)) == 0


# ---------------------------------------------------------------------------------------
# traversal + intersection vs brute force in double precision
# ---------------------------------------------------------------------------------------
def _rays(H, view, m, A, tiles):
    M = np.asarray(m, np.float64).reshape(4, 4).T  # column-major -> row-major
    R = M[:3, :3]
    off = H.subsample_offsets(A).astype(np.float64)
    out = []
    for t in tiles:
        b = view.beams[t * 128:(t + 1) * 128]
        c = b["centerRay"].astype(np.float64) @ R.T
        du = b["du"].astype(np.float64) @ R.T
        dv = b["dv"].astype(np.float64) @ R.T
        dirs = c[:, None, :] + du[:, None, :] * off[None, :, 0, None] + dv[:, None, :] * off[None, :, 1, None]
        out.append(dirs.reshape(-1, 3))
    return M[:3, 3], out


def _brute(origin, dirs, tris):
    """Moller-Trumbore in double, front faces only: (hit[ray, tri], t[ray, tri], u, v)."""
    v0 = tris["v0"].astype(np.float64)
    e1 = tris["edge0"].astype(np.float64)
    e2 = tris["edge1"].astype(np.float64)
    n = np.cross(e1, e2)
    p = np.cross(dirs[:, None, :], e2[None])
    det = np.einsum("tk,rtk->rt", e1, p)
    s = origin[None, :] - v0
    with np.errstate(divide="ignore", invalid="ignore"):
        u = np.einsum("tk,rtk->rt", s, p) / det
        q = np.cross(s, e1)
        v = np.einsum("rk,tk->rt", dirs, q) / det
        t = np.einsum("tk,tk->t", e2, q)[None, :] / det
    front = (np.einsum("rk,tk->rt", dirs, n) < 0)
    hit = front & (u >= 0) & (v >= 0) & (u + v <= 1) & (t > 0)
    return hit, t, u, v


def test_lists_contain_every_hit_and_ids_match_double(H, O, scene1, view1_small):
    A = 4
    m = view1_small.camera_to_world
    ref = O.render_view(scene1, view1_small, A, with_colors=False)
    counts, lists = ref["counts"], ref["lists"]
    starts = np.concatenate([[0], np.cumsum(counts)]).astype(np.int64)
    rng = np.random.default_rng(3)
    occupied = np.nonzero(counts)[0]
    tiles = list(rng.choice(occupied, size=min(12, len(occupied)), replace=False))
    tiles += list(np.nonzero(counts == 0)[0][:4])
    origin, dirs = _rays(H, view1_small, m, A, tiles)
    tris = ref["tris"]
    world_tiles, _ = O.frustum_setup(view1_small.tile_frusta, m)
    normals = np.cross(tris["edge0"].astype(np.float64), tris["edge1"].astype(np.float64))
    checked = agree = skipped = 0
    for t, d in zip(tiles, dirs):
        hit, tt, u, v = _brute(origin, d, tris)
        # TestTriangleFrustaUVW (prim_tests.h:89-116) culls on the barycentrics of the four
        # tile-corner rays, which bound the tile's rays only while the triangle faces every
        # corner ray; at grazing angles the reference culls visible triangles, and so does
        # the oracle. Those triangles are excluded from the id comparison below.
        corner = world_tiles["directions"][t].astype(np.float64)
        faces_corners = (corner @ normals.T < 0).all(0)
        listed = set(lists[starts[t]:starts[t + 1]].tolist())
        # every triangle a subsample ray of the tile hits (with margin) is in the tile's list
        w = 1 - u - v
        firm = hit & (u > 1e-6) & (v > 1e-6) & (w > 1e-6)
        missing = set(np.nonzero(firm.any(0))[0].tolist()) - listed
        assert not missing, f"tile {t}: hit triangles {sorted(missing)[:5]} missing from its list"
        # nearest hit id agrees with double precision where the answer is unambiguous
        ids = ref["tri_ids"][t * 128:(t + 1) * 128].reshape(-1)
        depth = ref["depth"][t * 128:(t + 1) * 128].reshape(-1)
        tmask = np.where(hit, tt, np.inf)
        order = np.argsort(tmask, axis=1)
        for r in range(len(d)):
            best = order[r, 0]
            if not np.isfinite(tmask[r, best]):
                if not hit[r].any():
                    assert ids[r] == BAD
                continue
            second = tmask[r, order[r, 1]] if hit.shape[1] > 1 else np.inf
            margin = min(u[r, best], v[r, best], w[r, best])
            if not faces_corners[best]:
                skipped += 1
                continue
            if margin > 1e-4 and second > tmask[r, best] * (1 + 1e-4):
                checked += 1
                agree += ids[r] == best
                assert ids[r] == best, f"tile {t} ray {r}: id {ids[r]} vs double {best}"
                assert abs(depth[r] - tmask[r, best]) <= 1e-5 * tmask[r, best]
    assert checked > 1000 and agree == checked and skipped < checked // 10


def test_tile_lists_have_no_duplicates(O, scene1, view1_small):
    ref = O.render_view(scene1, view1_small, 1, with_colors=False)
    starts = np.concatenate([[0], np.cumsum(ref["counts"])]).astype(np.int64)
    for t in np.nonzero(ref["counts"])[0]:
        seg = ref["lists"][starts[t]:starts[t + 1]]
        assert len(np.unique(seg)) == len(seg)


def test_block_stack_sorted_closest_on_top(O, scene1, view1):
    _, _, nodes = O.refit(scene1)
    _, pre = O.frustum_setup(view1.block_frusta, view1.camera_to_world)
    for b in range(len(pre)):
        st = O.block_traverse(nodes, pre[b:b + 1])
        assert len(st) <= 64
        assert np.all(np.diff(st["tMin"]) <= 0), "stack must be sorted by distance, closest on top"
        assert np.all((st["nodeMask"] & 0x0FFFFFFF) < len(nodes))


# ---------------------------------------------------------------------------------------
# shading closed form: camera, light and normal aligned (N = V = L = H)
# ---------------------------------------------------------------------------------------
def _aces(x):
    return np.clip((x * (2.51 * x + 0.03)) / (x * (2.43 * x + 0.59) + 0.14), 0, 1)


def _srgb(x):
    return np.where(x <= 0.0031308, x * 12.92, np.power(x, 1 / 2.4) * 1.055 - 0.055)


def _facing_scene(H, roughness, albedo, light):
    v = np.zeros(3, H.DTYPES["vertex"])
    v["pos"] = [[-50, -50, -5], [50, -50, -5], [0, 50, -5]]
    v["normal"] = np.array([0, 0, 1, 0], np.float16).view(np.uint16)
    mat = np.zeros(1, H.DTYPES["material"])
    mat["diffuse"] = [*albedo, 1]
    mat["specular"] = [0.04, 0.04, 0.04, 1]
    mat["glossiness"] = 1 - roughness
    mat["textureIDsAndShadingModel"] = 1 | (0xFFFF << 16) | (0xFFFF << 32) | (0xFFFF << 48)
    return H.HostScene.from_meshes([(v, [[0, 1, 2]], [0], None)], mat, light)


@pytest.mark.parametrize("roughness", [0.1, 0.5, 0.9])
def test_ggx_closed_form_aligned(H, O, roughness):
    albedo = np.array([0.8, 0.3, 0.5])
    power = np.array([1.5, 1.0, 0.5])
    light = np.zeros(1, H.DTYPES["light"])
    light["type"] = 1
    light["body"][0, :6] = [0, 0, 1, *power]
    scene = _facing_scene(H, roughness, albedo, light)
    wv, tris, nodes = O.refit(scene)
    env, dropped = O.lighting_env(scene.lights)
    assert dropped == 0
    beams = np.zeros(1, H.DTYPES["beam"])
    beams["centerRay"] = [0, 0, -1]
    beams["du"] = [1e-3, 0, 0]
    beams["dv"] = [0, 1e-3, 0]
    A = 1
    off = np.zeros((1, 2), np.float32)  # centroid at the centre ray
    m = np.eye(4, dtype=np.float32).reshape(16)
    ids = np.zeros((1, A), np.uint32)
    # closed form with N = V = L = H: D = 1/(pi a^2), F = F0, Vis = 1, spec = F0 D / 4
    a = np.float64(np.float32(roughness)) ** 2
    D = 1 / (np.pi * a * a)
    spec = 0.04 * D / 4
    rad = albedo * np.array([0.15, 0.15, 0.165]) + albedo * power * (1 - spec) + power * spec
    want = (_srgb(_aces(rad)) * 255).astype(int)
    # float32 radiance near the GGX peak (D ~ 1/alpha^2) moves ~1e-3 per ulp of NdotH
    for precision, rtol in ((0, 2e-3), (1, 1e-6)):
        col, lin = O.resolve_view(beams, m, off, A, ids, tris, scene.triShade, wv, scene.materials, env, precision)
        assert np.allclose(lin[0], rad, rtol=rtol)
        got = col.view(np.uint8).reshape(-1, 4)[0, :3].astype(int)
        assert np.max(np.abs(got - want)) <= 1
        assert col.view(np.uint8).reshape(-1, 4)[0, 3] == 255


def test_point_light_falloff_closed_form(H, O):
    albedo = np.array([0.5, 0.5, 0.5])
    light = np.zeros(1, H.DTYPES["light"])
    light["type"] = 2
    # position (0,0,-1) is 4 m in front of the hit point (0,0,-5); falloff from 2 to 10 m
    light["body"][0, :8] = [0, 0, -1, 10.0, 2.0, 2.0, 2.0, 1 / 8.0]
    scene = _facing_scene(H, 0.5, albedo, light)
    wv, tris, _ = O.refit(scene)
    env, _ = O.lighting_env(scene.lights)
    beams = np.zeros(1, H.DTYPES["beam"])
    beams["centerRay"] = [0, 0, -1]
    col, lin = O.resolve_view(beams, np.eye(4, dtype=np.float32).reshape(16), np.zeros((1, 2), np.float32), 1,
                              np.zeros((1, 1), np.uint32), tris, scene.triShade, wv, scene.materials, env, 1)
    f = np.clip((10.0 - 4.0) / 8.0, 0, 1)
    a = 0.25
    spec = 0.04 * (1 / (np.pi * a * a)) / 4
    rad = albedo * np.array([0.15, 0.15, 0.165]) + albedo * 2 * f * (1 - spec) + 2 * f * spec
    assert np.allclose(lin[0], rad, rtol=1e-6)


def test_lights_past_maxima_are_dropped(H, O):
    lights = np.zeros(7, H.DTYPES["light"])
    lights["type"] = [1, 1, 2, 2, 2, 2, 2]
    env, dropped = O.lighting_env(lights)
    assert dropped == 2


# ---------------------------------------------------------------------------------------
# refit vs a numpy recomputation; remap vs fancy indexing
# ---------------------------------------------------------------------------------------
def test_refit_matches_numpy(O, scene1):
    wv, tris, nodes = O.refit(scene1)
    M = scene1.transforms.reshape(-1, 4, 4).transpose(0, 2, 1).astype(np.float64)
    p = np.concatenate([scene1.verts["pos"], np.ones((len(wv), 1))], 1)
    w = np.einsum("vij,vj->vi", M[scene1.vertexToTransform], p)
    assert np.allclose(wv["pos"], w[:, :3] / w[:, 3:], atol=1e-5)
    idx = scene1.triShade["indices"]
    P = wv["pos"][idx]  # [t, 3, 3]
    assert np.array_equal(tris["v0"], P[:, 0])
    lo_tri, hi_tri = P.min(1), P.max(1)
    for n in range(len(nodes)):
        for k in range(4):
            b = nodes["bounds"][n][:, k]
            if scene1.nodes["leafMask"][n] >> k & 1:
                s, e = scene1.nodes["data"][n][k], scene1.nodes["data"][n][k + 1]
                if e == s:
                    assert np.all(b == -np.inf)
                    continue
                lo, hi = lo_tri[s:e].min(0), hi_tri[s:e].max(0)
            else:
                c = n + scene1.nodes["data"][n][1 + k]
                cb = nodes["bounds"][c]
                hi = cb[[0, 2, 4]].max(1)
                lo = -cb[[1, 3, 5]].max(1)
            assert np.array_equal(b[[0, 2, 4]], hi) and np.array_equal(-b[[1, 3, 5]], lo)


def test_remap_is_fancy_indexing(O, view1_small):
    rng = np.random.default_rng(0)
    res = rng.integers(0, 2**32, size=view1_small.sample_count, dtype=np.uint32)
    img = O.remap(res, view1_small.remap, view1_small.width, view1_small.height)
    assert np.array_equal(img.reshape(-1), res[view1_small.remap])
