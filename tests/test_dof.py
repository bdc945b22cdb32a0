"""Depth of field (thin lens): the DoF variants of intersect and resolve
(raycaster/prim_tests.h:247-392, tile_data.h:37-60, gpu_samples.h:84-103,
resolve.cu:161-284).

CPU pins of the oracle: every subsample ray of a sample passes through the same focal
point whatever the lens radius, and lens origins stay within the lens; the oracle's
per-subsample triangle ids equal a double-precision brute-force nearest hit along its own
DoF rays; a vanishing lens gives the pinhole ids. GPU: ids and subsample masks bit-exact
against the oracle, colours within 1/255."""

import numpy as np
import pytest

from conftest import rgba_delta

W, H_ = 160, 120


def _view(H, radius, focal, w=W, h=H_):
    t = np.tan(np.radians(30.0))
    vp = (-t * w / h, -t, t * w / h, t)
    return H.HostView.from_samples(H.grid_samples(w, h), w, h, vp, lens_radius=radius, focal_distance=focal)


def _camera(H):
    return H.HostView(H.VIEW_PRESET_HEIGHTFIELD, 0, W, H_).camera_to_world


def _lists(O, scene, view, m):
    wv, tris, nodes = O.refit(scene)
    world, tpre = O.frustum_setup(view.tile_frusta, m)
    _, bpre = O.frustum_setup(view.block_frusta, m)
    counts, lists = O.build_tile_lists(nodes, bpre, tpre)
    return tris, world, counts, lists


def test_dof_rays_focus_independent_of_lens_radius(H, O, scene1):
    A = 4
    m = _camera(H)
    offsets = H.subsample_offsets(A)
    table = H.lens_table(A)
    pts, lives = [], []
    for radius in (0.02, 0.2):
        view = _view(H, radius, 12.0)
        _, _, counts, _ = _lists(O, scene1, view, m)
        o, d = O.dof_rays(view.beams, m, offsets, A, (radius, 12.0, table), counts)
        live = np.repeat(np.repeat(counts > 0, 128), A)[: len(o)]
        # lens origins lie on the lens (radius r around the camera position)
        r = np.linalg.norm(o[live] - m[12:15], axis=1)
        assert r.max() <= radius * 1.0001
        assert r.max() > radius * 0.5
        pts.append((o + d).astype(np.float64))
        lives.append(live)
    # the focal points (origin + direction) do not depend on the lens radius (nor on
    # the lens-table row, which moves with the set of occupied tiles)
    both = lives[0] & lives[1]
    assert both.sum() > 10000
    scale = np.abs(pts[0][both]).max()
    assert np.abs(pts[0][both] - pts[1][both]).max() <= 1e-5 * scale


def _brute(o, d, tris):
    v0 = tris["v0"].astype(np.float64)
    e0 = tris["edge0"].astype(np.float64)
    e1 = tris["edge1"].astype(np.float64)
    out_id = np.full(len(o), 0xFFFFFFFF, np.int64)
    out_t = np.full(len(o), np.inf)
    second = np.full(len(o), np.inf)
    edge = np.zeros(len(o), bool)
    for k in range(len(o)):
        p = np.cross(d[k], e1)
        det = (e0 * p).sum(1)
        ok = np.abs(det) > 1e-12
        inv = np.where(ok, 1.0 / np.where(ok, det, 1), 0)
        s = o[k] - v0
        u = (s * p).sum(1) * inv
        q = np.cross(s, e0)
        v = (q * d[k]).sum(1) * inv
        t = (e1 * q).sum(1) * inv
        hit = ok & (u >= 0) & (v >= 0) & (u + v <= 1) & (t > 0)
        # front faces only (dot(d, e0 x e1) < 0, i.e. det > 0): the reference culls back faces
        hit &= det > 0
        if not hit.any():
            continue
        idx = np.nonzero(hit)[0]
        order = np.argsort(t[idx])
        out_id[k] = idx[order[0]]
        out_t[k] = t[idx[order[0]]]
        if len(order) > 1:
            second[k] = t[idx[order[1]]]
        j = idx[order[0]]
        edge[k] = min(u[j], v[j], 1 - u[j] - v[j]) < 1e-4
    return out_id, out_t, second, edge


def test_dof_ids_match_brute_force(H, O, scene1):
    A = 4
    radius, focal = 0.1, 12.0
    m = _camera(H)
    view = _view(H, radius, focal)
    tris, world, counts, lists = _lists(O, scene1, view, m)
    offsets = H.subsample_offsets(A)
    dof = (radius, focal, H.lens_table(A))
    ids, depth, _ = O.intersect_view(view.beams, world, m, offsets, A, counts, lists, tris, dof)
    o, d = O.dof_rays(view.beams, m, offsets, A, dof, counts)
    rng = np.random.default_rng(3)
    live = np.nonzero(np.repeat(np.repeat(counts > 0, 128), A)[: len(o)])[0]
    pick = rng.choice(live, 2500, replace=False)
    bid, bt, second, edge = _brute(o[pick].astype(np.float64), d[pick].astype(np.float64), tris)
    got = ids.reshape(-1)[pick].astype(np.int64)
    # near-ties and edge grazes are decided by float rounding; everything else must agree
    gap = np.where(np.isfinite(bt), second - np.where(np.isfinite(bt), bt, 0), np.inf)
    clear = ~edge & (gap > 1e-4 * np.maximum(np.where(np.isfinite(bt), bt, 1), 1))
    assert clear.sum() > 2000
    bad = clear & (got != bid)
    assert bad.sum() <= 2, f"{bad.sum()} of {clear.sum()} clear subsamples disagree"
    hits = np.isfinite(bt) & clear
    assert hits.sum() > 1000
    # depth = distance along the unnormalised direction
    assert np.allclose(depth.reshape(-1)[pick][hits & (got == bid)], bt[hits & (got == bid)], rtol=1e-4)
    # barycentrics on the lens ray (or_hits_view): the barycentric point is the lens ray's
    # point at the hit depth
    uvw = O.hits_view(view.beams, m, offsets, A, ids, tris, counts, dof).reshape(-1, 3).astype(np.float64)
    sel = pick[hits & (got == bid)]
    k = ids.reshape(-1)[sel]
    v0 = tris["v0"][k].astype(np.float64)
    pb = v0 + uvw[sel, 1:2] * tris["edge0"][k] + uvw[sel, 2:3] * tris["edge1"][k]
    assert np.allclose(uvw[sel].sum(1), 1, atol=1e-5)
    pr = o[sel].astype(np.float64) + depth.reshape(-1)[sel, None].astype(np.float64) * d[sel].astype(np.float64)
    err = np.linalg.norm(pb - pr, axis=1) / np.maximum(1, np.linalg.norm(pr - o[sel], axis=1))
    assert err.max() <= 1e-4, err.max()


def test_vanishing_lens_gives_pinhole_ids(H, O, scene1):
    # TileData<true>::load rotates the (already world-space) tile beam by the camera
    # rotation once more (tile_data.h:43-46), which also turns the subsample directions;
    # with an unrotated camera a vanishing lens must reproduce the pinhole ids
    A = 4
    m = H.transform_matrix((8.0, 3.0, 26.0), (0.0, 0.0, 0.0, 1.0))
    pin = _view(H, 0.0, 12.0)
    tiny = _view(H, 1e-6, 12.0)
    tris, world, counts, lists = _lists(O, scene1, pin, m)
    offsets = H.subsample_offsets(A)
    ids_pin, _, _ = O.intersect_view(pin.beams, world, m, offsets, A, counts, lists, tris)
    tris2, world2, counts2, lists2 = _lists(O, scene1, tiny, m)
    ids_dof, _, _ = O.intersect_view(tiny.beams, world2, m, offsets, A, counts2, lists2, tris2,
                                     (1e-6, 12.0, H.lens_table(A)))
    assert (ids_pin != 0xFFFFFFFF).mean() > 0.1
    assert (ids_pin == ids_dof).mean() > 0.995


# ---------------------------------------------------------------------------------------
# GPU parity
# ---------------------------------------------------------------------------------------
@pytest.mark.gpu
@pytest.mark.parametrize("A,radius,focal", [(4, 0.1, 12.0), (8, 0.05, 6.0), (1, 0.2, 14.0)])
def test_dof_heightfield_gpu_matches_oracle(H, O, ctx, scene1, A, radius, focal):
    m = _camera(H)
    view = _view(H, radius, focal)
    gs, gv = H.GpuScene(ctx, scene1), H.GpuView(ctx, view, A)
    gv.set_hit_output(True)
    gv.render(gs, m)
    ref = O.render_view(scene1, view, A, camera_to_world=m)
    counts, lists = gv.read_tile_lists()
    assert np.array_equal(counts, ref["counts"]) and np.array_equal(lists, ref["lists"])
    tri, mask = gv.read_gbuffer()
    rtri, rmask = O.gbuffer_slots(ref["tri_ids"])
    bad = np.nonzero((tri != rtri).any(1) | (mask != rmask).any(1))[0]
    assert len(bad) == 0, f"{len(bad)} samples differ, first {bad[:5]}"
    assert gv.stats().survivorPairs == ref["survivors"]
    assert rgba_delta(gv.read_sample_colors(), ref["colors"]).max() <= 1
    assert rgba_delta(gv.read_image(), ref["image"]).max() <= 1
    # per-subsample depth and barycentrics along the lens rays
    from test_gpu_parity import check_hits
    same = check_hits(O, gv, ref, view, A, m, dof=(radius, focal, H.lens_table(A)))
    assert same == view.sample_count * A * 4


@pytest.mark.gpu
def test_dof_room_gpu_matches_oracle(H, O, ctx, scene2):
    # the foveated/DoF workload's lens: radius 1 cm, focal distance inside the room
    A = 4
    eye = H.HostView(H.VIEW_PRESET_ROOM_EYE, 0, 270, 300)
    view = H.HostView.from_samples(H.grid_samples(270, 300, 0.22, 0.24), 270, 300, _room_viewport(270, 300),
                                   lens_radius=0.01, focal_distance=4.0)
    m = eye.camera_to_world
    gs, gv = H.GpuScene(ctx, scene2), H.GpuView(ctx, view, A)
    gv.render(gs, m)
    ref = O.render_view(scene2, view, A, camera_to_world=m)
    tri, mask = gv.read_gbuffer()
    rtri, rmask = O.gbuffer_slots(ref["tri_ids"])
    assert np.array_equal(tri, rtri) and np.array_equal(mask, rmask)
    assert rgba_delta(gv.read_image(), ref["image"]).max() <= 1


def _room_viewport(w, h):
    sx = np.tan(np.radians(50.0)) / 1.46
    sy = sx * h / w
    return (-sx, -sy, sx, sy)
