"""Batched multi-view rendering (hvvr_gpu_batch_*): every kernel of the frame launched once
for many views, optionally captured as a HIP graph and replayed, with the BVH refit in the
chain and an autotuned launch shape.

GPU: a batch of views of different sizes, MSAA rates and shading precisions gives, for
every view, bit for bit the tile lists, GBuffer, sample colours and image of the same view
rendered alone; graph replays with new cameras match eager batch renders frame by frame;
a refit inside the chain matches a solo render after the scene update; the autotuned shape
leaves results unchanged."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _moved(base, k):
    m = np.array(base, np.float32).reshape(4, 4).copy()
    m[3, :3] += np.float32(0.07) * np.float32(k) * np.array([1.0, -0.2, 0.4], np.float32)
    return m.reshape(16)


def _views(H, ctx, half=True):
    room = [H.HostView(H.VIEW_PRESET_ROOM_EYE, e, w, h) for e, (w, h) in enumerate([(270, 300), (128, 128)])]
    room.append(H.HostView(H.VIEW_PRESET_ROOM_EYE, 0, 360, 200))
    specs = [(room[0], 1, False), (room[1], 2, False), (room[2], 4, half), (room[0], 8, False), (room[1], 4, half)]
    out = []
    for hv, a, h16 in specs:
        gv = H.GpuView(ctx, hv, a)
        if h16:
            gv.set_shading_precision(16)
        out.append(gv)
    return out


def _results(gv):
    counts, lists = gv.read_tile_lists()
    st = gv.stats()
    return dict(counts=counts, lists=lists, gbuffer=gv.read_gbuffer(), colors=gv.read_sample_colors(),
                image=gv.read_image(), work=(st.survivorPairs, st.listEntries, st.nodeVisits))


def _same(a, b, what):
    for k in a:
        if isinstance(a[k], tuple):
            ok = all(np.array_equal(x, y) for x, y in zip(a[k], b[k])) if isinstance(a[k][0], np.ndarray) \
                else a[k] == b[k]
        else:
            ok = np.array_equal(a[k], b[k])
        assert ok, f"{what}: {k} differs"


@pytest.mark.parametrize("streams", [1, 3])
def test_batch_matches_solo_renders(H, ctx, scene2, streams):
    gs = H.GpuScene(ctx, scene2)
    views = _views(H, ctx)
    cams = [_moved(v.view.camera_to_world, i) for i, v in enumerate(views)]
    solo = []
    for v, m in zip(views, cams):
        v.render(gs, m)
        solo.append(_results(v))
    batch = H.GpuBatch(ctx, views)
    shape = batch.launch_shape()
    shape.streams = streams
    batch.set_launch_shape(shape)
    batch.render(gs, cams)
    ms = batch.synchronize()
    assert ms > 0
    for i, v in enumerate(views):
        _same(solo[i], _results(v), f"view {i}")
    batch.close()


def test_batch_graph_replays_match_eager(H, ctx, scene2):
    gs = H.GpuScene(ctx, scene2)
    va, vb = _views(H, ctx), _views(H, ctx)
    graph, eager = H.GpuBatch(ctx, va), H.GpuBatch(ctx, vb)
    shape = graph.launch_shape()
    shape.streams = 2  # the captured graph forks onto a second stream and joins back
    graph.set_launch_shape(shape)
    graph.set_graph(True)
    for k in range(5):
        cams = [_moved(v.view.camera_to_world, k + i) for i, v in enumerate(va)]
        graph.render(gs, cams)
        graph.synchronize()
        eager.render(gs, cams)
        eager.synchronize()
        for i in range(len(va)):
            _same(_results(vb[i]), _results(va[i]), f"replay {k} view {i}")
    graph.close()
    eager.close()


def test_batch_refit_in_chain_matches_solo(H, ctx, scene2):
    # 10% of the meshes displaced every frame, refit inside the (graph-captured) chain
    scene_b = H.GpuScene(ctx, scene2)
    scene_s = H.GpuScene(ctx, scene2)
    views = _views(H, ctx, half=False)[:3]
    solo_views = _views(H, ctx, half=False)[:3]
    batch = H.GpuBatch(ctx, views)
    batch.set_graph(True)
    batch.set_refit(True)
    rng = np.random.default_rng(5)
    verts = scene2.verts.copy()
    per_mesh = 256 * 3 // 2  # vertices per mesh of the room generator (boxes and cylinders differ; any run works)
    for frame in range(3):
        first = int(rng.integers(0, len(verts) - 100 * per_mesh))
        n = 100 * per_mesh
        verts["pos"][first:first + n] += rng.normal(0, 0.02, size=(n, 3)).astype(np.float32)
        scene_b.set_vertices(first, verts[first:first + n])
        scene_s.set_vertices(first, verts[first:first + n])
        batch.render(scene_b)
        batch.synchronize()
        for v, s in zip(views, solo_views):
            s.render(scene_s)
            _same(_results(s), _results(v), f"frame {frame}")
        assert np.array_equal(scene_b.read_nodes().view(np.uint8), scene_s.read_nodes().view(np.uint8))
    batch.close()


def test_batch_autotune_keeps_results(H, ctx, scene2):
    gs = H.GpuScene(ctx, scene2)
    views = _views(H, ctx)
    batch = H.GpuBatch(ctx, views)
    batch.render(gs)
    batch.synchronize()
    before = [_results(v) for v in views]
    best, ms = batch.autotune(gs, frames_per_trial=2)
    assert ms > 0 and best.largeGroups >= 1
    assert batch.launch_shape().as_dict() == best.as_dict()
    # the tuner compares per-view survivors, deepest stack, list entries and node visits,
    # folded from the per-tile work of the frame; a batch frame leaves them nonzero
    for i, v in enumerate(views):
        st = v.stats()
        assert st.survivorPairs > 0 and st.nodeVisits > 0 and st.listEntries > 0, f"view {i}: no work counted"
    batch.render(gs)
    batch.synchronize()
    for i, v in enumerate(views):
        _same(before[i], _results(v), f"tuned view {i}")
    batch.close()


def test_batch_rejects_foveated_and_dof_views(H, ctx, scene2):
    gs = H.GpuScene(ctx, scene2)
    hv = H.HostView(H.VIEW_PRESET_ROOM_EYE, 0, 128, 128)
    gv = H.GpuView(ctx, hv, 1)
    with pytest.raises(H.HvvrError):
        H.GpuBatch(ctx, [gv, gv])
    hv_dof = H.HostView(H.VIEW_PRESET_ROOM_EYE, 0, 128, 128)
    hv_dof.lens = np.array([0.01, 3.0], np.float32)
    dof = H.GpuView(ctx, hv_dof, 1)
    batch = H.GpuBatch(ctx, [gv, dof])
    with pytest.raises(H.HvvrError):
        batch.render(gs)
    batch.close()


def test_batch_1m_scene_refit_graph_matches_oracle(H, O, ctx):
    # the batched multi-view configuration at reduced view sizes: 16 views of the
    # 1M-triangle room (4,096 meshes), areas log-uniform between 128x128 and 480x270, MSAA
    # 1/2/4/8, half of them FP16-shaded, every frame 10% of the vertices displaced by
    # N(0, 0.02 m) and refit inside the replayed graph. After the last frame every view
    # matches the oracle of the displaced scene: refit bounds, tile lists, triangle ids and
    # masks bit for bit, FP32 colours within 1/255 of the FP32 oracle and FP16 colours
    # within 2/255 of the FP64 one.
    from concurrent.futures import ThreadPoolExecutor

    from conftest import rgba_delta

    rng = np.random.default_rng(4)
    scene = H.HostScene(H.SCENE_ROOM, 0, 4096)
    lo, hi = np.log(128 * 128), np.log(480 * 270)
    specs = []
    for k in range(16):
        area = np.exp(rng.uniform(lo, hi))
        h = int(round(np.sqrt(area / (16 / 9))))
        specs.append((int(round(h * 16 / 9)), h, [1, 2, 4, 8][k % 4], k % 2 == 1))
    hviews = [H.HostView(H.VIEW_PRESET_ROOM_EYE, k & 1, w, h) for k, (w, h, _, _) in enumerate(specs)]
    gs = H.GpuScene(ctx, scene)
    views = []
    for hv, (_, _, a, half) in zip(hviews, specs):
        gv = H.GpuView(ctx, hv, a)
        if half:
            gv.set_shading_precision(16)
        views.append(gv)
    cams = [_moved(hv.camera_to_world, k % 5) for k, hv in enumerate(hviews)]
    batch = H.GpuBatch(ctx, views)
    batch.set_refit(True)
    # size every view's lists for this scene first (growth allocates, outside the graph)
    batch.render(gs, cams)
    batch.synchronize()
    for gv in views:
        need, cap, of = gv.list_demand()
        if of:
            gv.reserve_lists(need + need // 4)
    batch.set_graph(True)
    verts = scene.verts.copy()
    n_move = len(verts) // 10
    for frame in range(3):
        first = int(rng.integers(0, len(verts) - n_move))
        verts["pos"][first:first + n_move] += rng.normal(0, 0.02, size=(n_move, 3)).astype(np.float32)
        gs.set_vertices(first, verts[first:first + n_move])
        batch.render(gs, cams)
        batch.synchronize()
    displaced = H.HostScene.__new__(H.HostScene)
    displaced.__dict__.update(scene.__dict__)
    displaced.verts = verts

    def oracle(i):
        return O.render_view(displaced, hviews[i], specs[i][2], camera_to_world=cams[i],
                             precision=1 if specs[i][3] else 0)

    with ThreadPoolExecutor(max_workers=8) as pool:
        refs = list(pool.map(oracle, range(len(views))))
    assert np.array_equal(gs.read_nodes().view(np.uint8), refs[0]["nodes"].view(np.uint8)), "refit bounds differ"
    for i, (gv, ref) in enumerate(zip(views, refs)):
        w, h, a, half = specs[i]
        what = f"view {i} ({w}x{h}, A={a}, {'fp16' if half else 'fp32'})"
        assert not gv.list_demand()[2], f"{what}: lists overflowed"
        counts, lists = gv.read_tile_lists()
        assert np.array_equal(counts, ref["counts"]) and np.array_equal(lists, ref["lists"]), f"{what}: lists differ"
        tri, mask = gv.read_gbuffer()
        rtri, rmask = O.gbuffer_slots(ref["tri_ids"])
        assert np.array_equal(tri, rtri) and np.array_equal(mask, rmask), f"{what}: ids or masks differ"
        d = rgba_delta(gv.read_sample_colors(), ref["colors"])
        assert d.max() <= (2 if half else 1), f"{what}: colour differs by {d.max()}/255"
        assert rgba_delta(gv.read_image(), ref["image"]).max() <= (2 if half else 1), f"{what}: image differs"
    batch.close()
