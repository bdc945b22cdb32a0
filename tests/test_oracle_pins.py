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


def test_hit_barycentrics_and_depth_match_double(H, O, scene1, view1_small):
    # per-subsample barycentrics (or_hits_view: calcUVW on the subsample's own ray) against
    # Moller-Trumbore in double: the reference's v and w weigh vertices 1 and 2 (edge0,
    # edge1), u = 1 - v - w vertex 0; and the barycentric point is the ray point at depth t
    A = 4
    m = view1_small.camera_to_world
    ref = O.render_view(scene1, view1_small, A, with_colors=False)
    uvw = O.hits_view(view1_small.beams, m, ref["offsets"], A, ref["tri_ids"], ref["tris"], ref["counts"])
    rng = np.random.default_rng(5)
    occupied = np.nonzero(ref["counts"])[0]
    tiles = list(rng.choice(occupied, size=min(10, len(occupied)), replace=False))
    origin, dirs = _rays(H, view1_small, m, A, tiles)
    tris = ref["tris"]
    v0 = tris["v0"].astype(np.float64)
    v1 = v0 + tris["edge0"].astype(np.float64)
    v2 = v0 + tris["edge1"].astype(np.float64)
    checked = 0
    errs = []
    for t, d in zip(tiles, dirs):
        ids = ref["tri_ids"][t * 128:(t + 1) * 128].reshape(-1)
        depth = ref["depth"][t * 128:(t + 1) * 128].reshape(-1).astype(np.float64)
        b = uvw[t * 128:(t + 1) * 128].reshape(-1, 3).astype(np.float64)
        hit, tt, u_mt, v_mt = _brute(origin, d, tris)
        for r in np.nonzero(ids != BAD)[0]:
            k = ids[r]
            if not hit[r, k]:
                continue  # a grazing hit the double test rounds the other way
            w_mt = 1 - u_mt[r, k] - v_mt[r, k]
            if min(u_mt[r, k], v_mt[r, k], w_mt) < 1e-4:
                continue
            want = np.array([w_mt, u_mt[r, k], v_mt[r, k]])
            errs.append(np.abs(b[r] - want).max())
            assert abs(depth[r] - tt[r, k]) <= 1e-5 * tt[r, k]
            p_bary = b[r, 0] * v0[k] + b[r, 1] * v1[k] + b[r, 2] * v2[k]
            p_ray = origin + depth[r] * d[r]
            assert np.linalg.norm(p_bary - p_ray) <= 1e-5 * max(1.0, np.linalg.norm(p_ray - origin))
            checked += 1
    assert checked > 2000
    errs = np.asarray(errs)
    print("bary err max %.3g p99 %.3g median %.3g" % (errs.max(), np.percentile(errs, 99), np.median(errs)))
    # float32 barycentrics from the differential form lose a few bits to cancellation; a
    # swapped or dropped term is off by far more
    assert errs.max() <= 1e-4 and np.percentile(errs, 99) <= 1e-5
    # misses carry zero barycentrics
    miss = ref["tri_ids"] == BAD
    assert np.all(uvw[miss] == 0)


def test_tile_lists_have_no_duplicates(O, scene1, view1_small):
    ref = O.render_view(scene1, view1_small, 1, with_colors=False)
    starts = np.concatenate([[0], np.cumsum(ref["counts"])]).astype(np.int64)
    for t in np.nonzero(ref["counts"])[0]:
        seg = ref["lists"][starts[t]:starts[t + 1]]
        assert len(np.unique(seg)) == len(seg)


# ---------------------------------------------------------------------------------------
# tile-list emission order on a hand-built BVH (traverse_avx.h:412-527)
# ---------------------------------------------------------------------------------------
def _hand_bvh(H):
    """Eight nodes in depth-first order, leaf children in the low slots (a leaf child k
    spans data[k]..data[k+1], an internal child k sits at offset data[1 + k]). Triangle
    ranges follow node order, so the traversal order shows up as a permutation.

      N0: leaf A [0,2) z[-3,-2]   leaf B [2,3) z[-9,-8]   N1 z[-6,-4]     N2 z[-3,-1]
      N1: leaf H [3,5)            leaf I [5,6)            N6 z[-6,-5]     N7 z[-5,-4.5]
      N6: leaf K [6,8)            N7: leaf J [8,9)
      N2: leaf C [9,11) z[-2,-1.5]  N3 z[-3,-2.5]   N4 z[-3,-2.5] (the same distance key)
      N3: leaf G [11,12)
      N4: leaf D [12,13)          N5 z[-2.9,-2.6] (its only internal child)
      N5: leaf E [13,15)          leaf F [15,16)
    Every box spans x, y in [-0.2, 0.2]; unused slots are empty boxes (-inf)."""
    nodes = np.zeros(8, H.DTYPES["node"])
    nodes["bounds"] = -np.inf

    def child(n, k, z0, z1):
        b = nodes["bounds"][n]
        b[0, k], b[1, k], b[2, k], b[3, k] = 0.2, 0.2, 0.2, 0.2  # max x, -min x, max y, -min y
        b[4, k], b[5, k] = z1, -z0

    def node(n, leaves, internals):
        """leaves: [(z0, z1, first, end)] in slots 0..; internals: [(z0, z1, child)] after them."""
        data = np.zeros(5, np.uint32)
        k = 0
        for z0, z1, first, end in leaves:
            child(n, k, z0, z1)
            data[k], data[k + 1] = first, end
            nodes["leafMask"][n] |= 1 << k
            k += 1
        for z0, z1, c in internals:
            child(n, k, z0, z1)
            data[1 + k] = c - n
            k += 1
        nodes["data"][n] = data

    node(0, [(-3, -2, 0, 2), (-9, -8, 2, 3)], [(-6, -4, 1), (-3, -1, 4)])
    node(1, [(-6, -5.5, 3, 5), (-6, -5.5, 5, 6)], [(-6, -5, 2), (-5, -4.5, 3)])
    node(2, [(-6, -5, 6, 8)], [])
    node(3, [(-5, -4.5, 8, 9)], [])
    node(4, [(-2, -1.5, 9, 11)], [(-3, -2.5, 5), (-3, -2.5, 6)])
    node(5, [(-3, -2.5, 11, 12)], [])
    node(6, [(-2.9, -2.6, 12, 13)], [(-2.9, -2.6, 7)])
    node(7, [(-2.9, -2.8, 13, 15), (-2.7, -2.6, 15, 16)], [])
    return nodes


def _wide_tile(H, O):
    fr = np.zeros(1, H.DTYPES["sfrustum"])
    fr["directions"] = [[-0.5, -0.5, -1], [0.5, -0.5, -1], [0.5, 0.5, -1], [-0.5, 0.5, -1]]
    _, pre = O.frustum_setup(fr, np.eye(4, dtype=np.float32).reshape(16))
    return pre


def _entry(node, key, mask=0):
    return (np.float32(key), np.uint32(node | (mask << 28)))


def test_tile_emission_order_hand_built_bvh(H, O):
    # Distance keys along -z are -zMax of a child box, and the stack pops the smallest key
    # first; children are inserted in slot order and a later insert goes above an equal
    # key. Expected, by hand: N0 emits A (0 1) and B (2) in slot order, pushes N1 (key 4)
    # and N2 (key 1); N2 emits C (9 10) and pushes N3, then N4 (both key 2.5: N4 above);
    # N4 emits D (12) and, N5 being its only internal child, tests N5 at once (E 13 14,
    # F 15) although N3's key (2.5) is below N5's (2.6); then N3 (G 11), then N1 (H 3 4,
    # I 5), which pushes N6 (key 5) then N7 (key 4.5): N7 (J 8) before N6 (K 6 7).
    nodes = _hand_bvh(H)
    pre = _wide_tile(H, O)
    assert O.test_children(pre[0:1], nodes[0:1]) == 0b1111
    got = O.tile_traverse(nodes, np.array([_entry(0, 0.0)], O.O_STACK), pre[0:1])
    assert got.tolist() == [0, 1, 2, 9, 10, 12, 13, 14, 15, 11, 3, 4, 5, 8, 6, 7]


def test_tile_emission_order_leaf_entry_suppression(H, O):
    # A block stack as block traversal leaves it for N0 (traverse_avx.h:358-391): one entry
    # for N0's leaf children (mask = leaf slots) and one per internal child, sorted with the
    # closest on top: N1 (4), N0's leaves (2), N2 (1). Popping N0's leaf entry emits A and B
    # but must not push N1 and N2 again (traverse_avx.h:477-483). N2's children (2.5) are
    # inserted below N0's leaf entry (2), so N0's leaves come right after C:
    # C (9 10), A B (0 1 2), D E F (12 13 14 15), G (11), H I (3 4 5), J (8), K (6 7).
    nodes = _hand_bvh(H)
    pre = _wide_tile(H, O)
    stack = np.array([_entry(1, 4.0), _entry(0, 2.0, 0b0011), _entry(4, 1.0)], O.O_STACK)
    got = O.tile_traverse(nodes, stack, pre[0:1])
    assert got.tolist() == [9, 10, 0, 1, 2, 12, 13, 14, 15, 11, 3, 4, 5, 8, 6, 7]
    assert len(set(got.tolist())) == 16


def test_exact_depth_tie_later_list_entry_wins(H, O):
    # Two coplanar copies of one triangle: the second test sees depth * denom - t == +0
    # when 1/denom is exact (denom a power of two), passes the sign test
    # (prim_tests.h:209-225) and takes the sample: the later list entry wins a tie.
    v0 = np.array([-8.0, -8.0, -4.0], np.float32)
    tris = np.zeros(3, H.DTYPES["tri_intersect"])
    for i in range(3):
        tris["v0"][i] = v0
        tris["edge0"][i] = [16, 0, 0]
        tris["edge1"][i] = [0, 16, 0]
    tris["v0"][2] = [-8.0, -8.0, -6.0]  # a farther parallel copy
    beams = np.zeros(128, H.DTYPES["beam"])
    beams["centerRay"] = [0, 0, -1]
    fr = np.zeros(1, H.DTYPES["sfrustum"])
    fr["directions"] = [[-0.01, -0.01, -1], [0.01, -0.01, -1], [0.01, 0.01, -1], [-0.01, 0.01, -1]]
    m = np.eye(4, dtype=np.float32).reshape(16)
    world, _ = O.frustum_setup(fr, m)
    off = np.zeros((1, 2), np.float32)
    # denom = -dir . (e0 x e1) = 256, t = 4 * 256: depth = 4 exactly
    for order, want in (([0, 1, 2], 1), ([1, 0, 2], 0), ([2, 0, 1], 1), ([0, 2, 1], 1)):
        ids, depth, _ = O.intersect_view(beams, world, m, off, 1, np.array([3], np.uint32),
                                         np.array(order, np.uint32), tris)
        assert np.all(ids == want), (order, ids[0])
        assert np.all(depth == 4.0)


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


# ---------------------------------------------------------------------------------------
# shading closed forms away from N = V = L = H, written from the published formulas:
#   D = a^2 / (pi ((n.h)^2 (a^2 - 1) + 1)^2)            GGX / Trowbridge-Reitz (Walter et al. 2007)
#   F = F0 + (1 - F0) (1 - l.h)^5                        Schlick
#   G1(x) = x / (x (1 - k) + k), k = a / 2               Schlick's Smith approximation for GGX
#   f = D F G1(n.l) G1(n.v) / (4 (n.l) (n.v))            Cook-Torrance
# combined as PBR_BRDF1 does (shading_helpers.h:115-122): light * (base (1 - f) n.l + f), the
# specular term without the n.l cosine, with a = roughness^2 and roughness = 1 - glossiness
# ---------------------------------------------------------------------------------------
def _cook_torrance(n, l, v, rough, f0=0.04):
    a = rough * rough
    k = a / 2
    h = (l + v) / np.linalg.norm(l + v)
    sat = lambda x: min(max(x, 0.0), 1.0)
    nh, lh, nl, nv = sat(n @ h), sat(l @ h), sat(n @ l), sat(n @ v)
    D = a * a / (np.pi * (nh * nh * (a * a - 1) + 1) ** 2)
    F = f0 + (1 - f0) * (1 - lh) ** 5
    # G1(nl) G1(nv) / (4 nl nv) without the division by zero at a grazing light
    vis = 1 / ((nl * (1 - k) + k) * (nv * (1 - k) + k))
    return D * F * vis / 4, nl


def _big_floor(H, roughness, albedo, lights):
    v = np.zeros(3, H.DTYPES["vertex"])
    v["pos"] = [[-500, -500, -5], [500, -500, -5], [0, 500, -5]]
    v["normal"] = np.array([0, 0, 1, 0], np.float16).view(np.uint16)
    mat = np.zeros(1, H.DTYPES["material"])
    mat["diffuse"] = [*albedo, 1]
    mat["specular"] = [0.04, 0.04, 0.04, 1]
    mat["glossiness"] = 1 - roughness
    mat["textureIDsAndShadingModel"] = 1 | (0xFFFF << 16) | (0xFFFF << 32) | (0xFFFF << 48)
    return H.HostScene.from_meshes([(v, [[0, 1, 2]], [0], None)], mat, lights)


def _shade_ray(H, O, scene, ray, precision):
    wv, tris, _ = O.refit(scene)
    env, dropped = O.lighting_env(scene.lights)
    assert dropped == 0
    beams = np.zeros(1, H.DTYPES["beam"])
    beams["centerRay"] = ray
    beams["du"] = [1e-3, 0, 0]
    beams["dv"] = [0, 1e-3, 0]
    col, lin = O.resolve_view(beams, np.eye(4, dtype=np.float32).reshape(16), np.zeros((1, 2), np.float32), 1,
                              np.zeros((1, 1), np.uint32), tris, scene.triShade, wv, scene.materials, env, precision)
    return col[0], lin[0].astype(np.float64)


def _unit(x):
    x = np.asarray(x, np.float64)
    return x / np.linalg.norm(x)


# (camera ray, directional light direction): V along N with L at 60 deg; V and L at +-45 deg
# (H = N); a grazing view; an out-of-plane pair; a light below the horizon (n.l = 0: the
# reference still adds the specular term, whose visibility factor is then 1/k)
GGX_CASES = [((0, 0, -1), _unit([np.sin(np.pi / 3), 0, 0.5])),
             ((1, 0, -1), _unit([1, 0, 1])),
             ((8, 0, -1), (0.0, 0.0, 1.0)),
             ((-0.4, 0.2, -1), _unit([0.3, 0.5, 0.8])),
             ((0.3, 0, -1), _unit([np.sin(np.radians(100)), 0, np.cos(np.radians(100))]))]


@pytest.mark.parametrize("case", range(len(GGX_CASES)))
@pytest.mark.parametrize("roughness", [0.2, 0.6])
def test_ggx_closed_form_general_geometry(H, O, case, roughness):
    ray, L = GGX_CASES[case]
    albedo = np.array([0.7, 0.4, 0.2])
    power = np.array([1.2, 1.0, 0.8])
    light = np.zeros(1, H.DTYPES["light"])
    light["type"] = 1
    light["body"][0, :6] = [*L, *power]
    scene = _big_floor(H, roughness, albedo, light)
    ray = np.asarray(ray, np.float64)
    P = ray * (5.0 / -ray[2])
    n = np.array([0.0, 0.0, 1.0])
    v = _unit(-P)
    Lf = np.asarray(light["body"][0, :3], np.float64)  # the float32 light direction
    rough = 1 - np.float64(np.float32(1 - roughness))
    f, nl = _cook_torrance(n, Lf, v, rough)
    rad = albedo * np.array([0.15, 0.15, 0.165]) + albedo * power * (1 - f) * nl + power * f
    col, lin = _shade_ray(H, O, scene, ray.astype(np.float32), precision=1)
    assert np.allclose(lin, rad, rtol=1e-5), (lin, rad)
    col32, lin32 = _shade_ray(H, O, scene, ray.astype(np.float32), precision=0)
    assert np.allclose(lin32, rad, rtol=2e-3)
    want = (_srgb(_aces(rad)) * 255).astype(int)
    got = np.frombuffer(np.uint32(col32).tobytes(), np.uint8)[:3].astype(int)
    assert np.max(np.abs(got - want)) <= 1


@pytest.mark.parametrize("angle,expect", [(0.0, 1.0), (25.0, None), (40.0, 0.0)])
def test_spot_light_cone_closed_form(H, O, angle, expect):
    # spot 4 m above the hit point (0, 0, -5), aimed `angle` degrees off the vertical;
    # cone: outer 30 deg, inner 20 deg, f_angle = sat((cos a - cos outer) / (cos inner - cos outer))
    # (SpotLightContribution, shading_helpers.h:139-152); distance falloff 2..10 m
    albedo = np.array([0.5, 0.6, 0.7])
    color = np.array([3.0, 2.0, 1.0])
    cos_out, cos_in = np.cos(np.radians(30)), np.cos(np.radians(20))
    d = [np.sin(np.radians(angle)), 0.0, -np.cos(np.radians(angle))]
    light = np.zeros(1, H.DTYPES["light"])
    light["type"] = 3
    light["body"][0, :13] = [0, 0, -1, 10.0, *d, 1 / 8.0, *color, cos_out, 1 / (cos_in - cos_out)]
    scene = _big_floor(H, 0.5, albedo, light)
    b = light["body"][0].astype(np.float64)
    cos_a = np.asarray(b[4:7]) @ np.array([0, 0, -1.0])
    f_angle = min(max((cos_a - b[11]) * b[12], 0.0), 1.0)
    if expect is not None:
        assert f_angle == expect
    else:
        assert 0.3 < f_angle < 0.7
    f_dist = (10.0 - 4.0) / 8.0
    rough = 1 - np.float64(np.float32(0.5))
    f, nl = _cook_torrance(np.array([0, 0, 1.0]), np.array([0, 0, 1.0]), np.array([0, 0, 1.0]), rough)
    lc = color * f_dist
    rad = albedo * np.array([0.15, 0.15, 0.165]) + (albedo * lc * (1 - f) * nl + lc * f) * f_angle
    _, lin = _shade_ray(H, O, scene, np.array([0, 0, -1], np.float32), precision=1)
    assert np.allclose(lin, rad, rtol=1e-6), (lin, rad)


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
