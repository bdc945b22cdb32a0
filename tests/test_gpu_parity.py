"""GPU render path vs the CPU oracle on the same seeded scenes and cameras.

Bar: BVH refit, tile triangle lists, per-subsample triangle ids and subsample masks
bit-exact; per-subsample depth and barycentrics within 1e-5 relative (the same float
operations, so they come out equal); colours within 1/255 per channel (libm pow/sqrt may
differ by an ulp)."""

import numpy as np
import pytest

from conftest import rgba_delta

pytestmark = pytest.mark.gpu


def _render(H, ctx, scene, view, A, limit=None, m=None, hits=False):
    gs = H.GpuScene(ctx, scene)
    gv = H.GpuView(ctx, view, A)
    if limit:
        gv.set_traversal_stack_limit(limit)
    if hits:
        gv.set_hit_output(True)
    gv.render(gs, m)
    return gs, gv


def check_hits(O, gv, ref, view, A, m=None, dof=None):
    """GPU per-subsample (t, u, v, w) against the oracle's depth and barycentrics:
    within 1e-5 relative (1e-7 absolute for barycentrics near 0); returns the count of
    bit-identical values."""
    m = view.camera_to_world if m is None else m
    got = gv.read_hits()
    n = view.sample_count
    want_t = ref["depth"].reshape(n, A)
    want_b = O.hits_view(view.beams, m, ref["offsets"], A, ref["tri_ids"], ref["tris"], ref["counts"], dof)
    hit = ref["tri_ids"].reshape(n, A) != 0xFFFFFFFF
    assert np.array_equal(got[..., 0] != np.float32(np.finfo(np.float32).max), hit), "hit/miss pattern differs"
    assert np.allclose(got[..., 0][hit], want_t[hit], rtol=1e-5, atol=0), "depth differs"
    assert np.all(got[..., 0][~hit] == want_t[~hit])
    assert np.allclose(got[..., 1:], want_b, rtol=1e-5, atol=1e-7), "barycentrics differ"
    assert np.all(got[..., 1:][~hit] == 0)
    # barycentrics of a hit sum to 1 and lie in [0, 1] up to rounding
    b = got[..., 1:][hit].astype(np.float64)
    assert np.all(np.abs(b.sum(1) - 1) < 1e-4) and b.min() > -1e-4 and b.max() < 1 + 1e-4
    return int((got[..., 0].view(np.uint32) == want_t.view(np.uint32)).sum() +
               (got[..., 1:].view(np.uint32) == want_b.view(np.uint32)).sum())


def _check_frame(O, gs, gv, scene, view, A, m=None, colors=True, hits=False):
    ref = O.render_view(scene, view, A, camera_to_world=m, with_colors=colors)
    assert np.array_equal(gs.read_nodes().view(np.uint8), ref["nodes"].view(np.uint8)), "refit bounds differ"
    assert np.array_equal(gs.read_triangles().view(np.uint8), ref["tris"].view(np.uint8)), "triangles differ"
    counts, lists = gv.read_tile_lists()
    assert np.array_equal(counts, ref["counts"]), "tile list lengths differ"
    assert np.array_equal(lists, ref["lists"]), "tile list contents/order differ"
    tri, mask = gv.read_gbuffer()
    rtri, rmask = O.gbuffer_slots(ref["tri_ids"])
    bad = np.nonzero((tri != rtri).any(1) | (mask != rmask).any(1))[0]
    assert len(bad) == 0, f"{len(bad)} samples differ, first {bad[:5]}"
    st = gv.stats()
    assert st.survivorPairs == ref["survivors"] and st.overflow == 0
    if colors:
        col = gv.read_sample_colors()
        assert rgba_delta(col, ref["colors"]).max() <= 1
        img = gv.read_image()
        assert rgba_delta(img, ref["image"]).max() <= 1
    if hits:
        same = check_hits(O, gv, ref, view, A, m)
        assert same == view.sample_count * A * 4, "depth and barycentrics are expected bit for bit"
    return ref


@pytest.mark.parametrize("A", [1, 4])
def test_heightfield_640x480(H, O, ctx, scene1, view1, A):
    # the CPU reference configuration at full size (640x480, A = 1), with per-subsample
    # depth and barycentrics
    gs, gv = _render(H, ctx, scene1, view1, A, hits=True)
    _check_frame(O, gs, gv, scene1, view1, A, hits=True)


@pytest.mark.parametrize("A", [2, 8, 16, 32])
def test_heightfield_msaa_rates(H, O, ctx, scene1, view1_small, A):
    gs, gv = _render(H, ctx, scene1, view1_small, A)
    _check_frame(O, gs, gv, scene1, view1_small, A)


def test_room_262k_reduced_eye(H, O, ctx, scene2, view2_small):
    gs, gv = _render(H, ctx, scene2, view2_small, 4)
    _check_frame(O, gs, gv, scene2, view2_small, 4)


@pytest.mark.parametrize("limit", [1, 60, 64])
def test_deep_stack_fallback_same_lists(H, O, ctx, scene2, view2_small, limit):
    # a stack budget below a tile's need sends that tile past every LDS configuration to
    # the LDS-stack walk: 1 sends every tile, 60 about a third of them (block stacks hold
    # ~60 entries); at 64 the large configuration takes tiles with more than 768 visited
    # nodes
    gs, gv = _render(H, ctx, scene2, view2_small, 4, limit=limit)
    _check_frame(O, gs, gv, scene2, view2_small, 4, colors=False)


def test_room_full_eye_1080x1200(H, O, ctx, scene2):
    # the benchmark's view at full size: eye 0, 1080x1200, barrel distortion, A = 4
    view = H.HostView(H.VIEW_PRESET_ROOM_EYE, 0, 1080, 1200)
    gs, gv = _render(H, ctx, scene2, view, 4, hits=True)
    _check_frame(O, gs, gv, scene2, view, 4, hits=True)


def test_hit_output_off_by_default_and_readback_guarded(H, ctx, scene1, view1_small):
    gs, gv = _render(H, ctx, scene1, view1_small, 2)
    with pytest.raises(H.HvvrError, match="hit output"):
        gv.read_hits()
    gv.set_hit_output(True)
    gv.render(gs)
    a = gv.read_hits()
    gv.set_hit_output(False)
    gv.render(gs)
    with pytest.raises(H.HvvrError, match="hit output"):
        gv.read_hits()
    # the frame's ids do not depend on the hit output
    gv.set_hit_output(True)
    gv.render(gs)
    assert np.array_equal(a, gv.read_hits())


def test_render_is_deterministic(H, ctx, scene2, view2_small):
    gs, gv = _render(H, ctx, scene2, view2_small, 4)
    a = gv.read_gbuffer()
    c = gv.read_sample_colors()
    gv.render(gs)
    b = gv.read_gbuffer()
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    assert np.array_equal(c, gv.read_sample_colors())


def test_profiling_levels(H, ctx, scene2, view2_small):
    # level 1: per-stage events and the tile-list kernel's own dispatch timing; level 2: only
    # the dispatch timing (stage times are then refused); 0: neither; results unchanged
    gs, gv = _render(H, ctx, scene2, view2_small, 4)
    want = gv.read_gbuffer()
    try:
        ctx.set_profiling(1)
        gv.render(gs)
        stages = gv.stage_times()
        k1 = gv.tile_list_kernel_ms()
        assert 0.0 < k1 <= stages[2] * 1.05 + 0.01, (k1, stages)
        ctx.set_profiling(2)
        gv.render(gs)
        assert gv.tile_list_kernel_ms() > 0.0
        with pytest.raises(H.HvvrError):
            gv.stage_times()
        with pytest.raises(H.HvvrError):
            ctx.set_profiling(3)
        ctx.set_profiling(0)
        gv.render(gs)
        with pytest.raises(H.HvvrError):
            gv.tile_list_kernel_ms()
    finally:
        ctx.set_profiling(0)
    got = gv.read_gbuffer()
    assert np.array_equal(got[0], want[0]) and np.array_equal(got[1], want[1])


def test_animated_transforms_refit(H, O, ctx, scene1, view1_small):
    gs, gv = _render(H, ctx, scene1, view1_small, 4)
    rng = np.random.default_rng(7)
    xf = scene1.transforms.copy()
    moved = rng.choice(len(xf), size=len(xf) // 10, replace=False)
    xf[moved, 12:15] += rng.normal(0, 0.5, size=(len(moved), 3)).astype(np.float32)
    gs.set_transforms(xf)
    gv.render(gs)
    animated = H.HostScene.__new__(H.HostScene)
    animated.__dict__.update(scene1.__dict__)
    animated.transforms = xf
    _check_frame(O, gs, gv, animated, view1_small, 4)


def test_vertex_displacement_refit(H, O, ctx, scene2, view2_small):
    gs, gv = _render(H, ctx, scene2, view2_small, 2)
    rng = np.random.default_rng(11)
    verts = scene2.verts.copy()
    first, n = 256 * 190, 50 * 190
    verts["pos"][first:first + n] += rng.normal(0, 0.02, size=(n, 3)).astype(np.float32)
    gs.set_vertices(first, verts[first:first + n])
    gv.render(gs)
    displaced = H.HostScene.__new__(H.HostScene)
    displaced.__dict__.update(scene2.__dict__)
    displaced.verts = verts
    _check_frame(O, gs, gv, displaced, view2_small, 2, colors=False)


def test_camera_looking_away_renders_clear(H, O, ctx, scene1, view1_small):
    # camera below the ground looking down: every tile list is empty
    m = view1_small.camera_to_world.copy().reshape(4, 4)
    m[3, :3] = [8.0, -50.0, 8.0]
    m[:3, :3] = np.diag([1.0, -1.0, -1.0]).astype(np.float32)
    gs, gv = _render(H, ctx, scene1, view1_small, 4, m=m.reshape(16))
    counts, _ = gv.read_tile_lists()
    assert counts.sum() == 0
    assert np.all(gv.read_sample_colors() == 0xFF000000)
    _check_frame(O, gs, gv, scene1, view1_small, 4, m=m.reshape(16))


def test_invalid_view_rejected(H, ctx, view1_small):
    with pytest.raises(H.HvvrError, match="msaaRate"):
        H.GpuView(ctx, view1_small, 3)


def test_batched_views_match_solo_renders(H, ctx, scene1, scene2, view1_small, view2_small):
    # several views of one scene in one submission run on separate streams; each result
    # must equal the same view rendered alone (a view listed twice shares a stream)
    eye1 = H.HostView(H.VIEW_PRESET_ROOM_EYE, 1, 270, 300)
    views = [H.GpuView(ctx, view2_small, 4), H.GpuView(ctx, eye1, 4), H.GpuView(ctx, view2_small, 2)]
    gs = H.GpuScene(ctx, scene2)
    solo = []
    for gv in views:
        gv.render(gs)
        solo.append((gv.read_gbuffer(), gv.read_sample_colors(), gv.read_image()))
    m = view2_small.camera_to_world
    H.render_views(gs, views + [views[0]], cameras=[m, None, m, m])
    for gv, (g, c, im) in zip(views, solo):
        g2 = gv.read_gbuffer()
        assert np.array_equal(g[0], g2[0]) and np.array_equal(g[1], g2[1])
        assert np.array_equal(c, gv.read_sample_colors())
        assert np.array_equal(im, gv.read_image())


@pytest.mark.parametrize("split", [2, 3])
def test_split_colour_remap(H, O, ctx, scene2, split):
    # split colour samples (camera.cpp:21-64, remap.cu:35-77): `split` samples per pixel,
    # sample n feeding channel n % split, side by side across the pixel; split 2 merges
    # pentile-style by pixel parity, split 3 takes R, G and B from their own samples
    w, h = 270, 300
    base = H.grid_samples(w, h, 0.22, 0.24)
    samples = np.repeat(base, split)
    for c in range(split):
        samples["position"][c::split, 0] += np.float32((c - (split - 1) / 2) / (split * w))
    sx = np.tan(np.radians(50.0)) / 1.46
    view = H.HostView.from_samples(samples, w, h, (-sx, -sx * h / w, sx, sx * h / w), split=split)
    assert view.split == split and view.valid_sample_count == w * h * split
    m = H.HostView(H.VIEW_PRESET_ROOM_EYE, 0, w, h).camera_to_world
    gs, gv = _render(H, ctx, scene2, view, 4, m=m)
    ref = _check_frame(O, gs, gv, scene2, view, 4, m=m)
    # every pixel draws on all of its channel samples
    assert (view.remap >= 0).all() and len(np.unique(view.remap)) == w * h * split
    assert (ref["image"] != 0).mean() > 0.9
