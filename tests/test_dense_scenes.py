"""Dense scenes: tiles beyond the standard and large LDS configurations, and frames whose
triangle lists outgrow the view's list storage.

The room scene at 4,096 meshes (1,048,576 triangles) seen at 360x400 gives tiles of
~2,700 list entries on average and several thousand visited nodes, past the large
configuration's 2,048 records, and ~3.1M list entries in all, past the default storage
of 1M entries. The reference's tile walk (traverse_avx.h:412-527) fixes the lists these
tiles must get; the oracle restates it. GPU: the first frame reports the overflow and the
entries it needed, the view grows its storage to that, and the frame rendered again
matches the oracle bit for bit (lists, triangle ids, masks), with every tile taken by the
huge LDS configuration or the LDS-stack walk as the stack limit decides; on the second
frame the queued configurations send a tile straight to the one the first frame needed
(per-tile hints)."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dense(H):
    return H.HostScene(H.SCENE_ROOM, 0, 4096)


@pytest.fixture(scope="module")
def dense_view(H):
    return H.HostView(H.VIEW_PRESET_ROOM_EYE, 0, 360, 400)


@pytest.fixture(scope="module")
def dense_ref(O, dense, dense_view):
    return O.render_view(dense, dense_view, 1, with_colors=False)


@pytest.mark.parametrize("limit", [None, 1])
def test_dense_scene_lists_grow_and_match(H, O, ctx, dense, dense_view, dense_ref, limit):
    gs = H.GpuScene(ctx, dense)
    gv = H.GpuView(ctx, dense_view, 1)
    if limit:
        gv.set_traversal_stack_limit(limit)
    gv.render(gs)
    first = gv.stats()  # no configuration hints yet: every tile starts at the standard one
    need, cap, overflowed = gv.list_demand()
    total = int(dense_ref["counts"].sum())
    print(f"limit {limit}: demand {need}, capacity {cap}, oracle total {total}", flush=True)
    assert overflowed and need > cap, "the scene is meant to outgrow the default list storage"
    assert need == total, "the demand is the frame's exact list length"
    with pytest.raises(H.HvvrError):
        gv.read_tile_lists()  # reads of the overflowed frame fail
    gv.reserve_lists(need)
    gv.render(gs)
    need2, cap2, overflowed2 = gv.list_demand()
    assert not overflowed2 and need2 == need and cap2 == need
    counts, lists = gv.read_tile_lists()
    assert np.array_equal(counts, dense_ref["counts"]), "tile list lengths differ"
    assert np.array_equal(lists, dense_ref["lists"]), "tile list contents/order differ"
    tri, mask = gv.read_gbuffer()
    rtri, rmask = O.gbuffer_slots(dense_ref["tri_ids"])
    bad = np.nonzero((tri != rtri).any(1) | (mask != rmask).any(1))[0]
    assert len(bad) == 0, f"{len(bad)} samples differ, first {bad[:5]}"
    st = gv.stats()
    print(f"tiles {st.tileCount}: large {st.largeTiles}, huge {st.hugeTiles}, giant {st.giantTiles}, "
          f"colossal {st.colossalTiles}, walk {st.walkTiles}, "
          f"deepest stack {st.maxTileStack}", flush=True)
    assert st.overflow == 0 and st.survivorPairs == dense_ref["survivors"]
    if limit is None:
        # tiles past the large configuration exist and the huge one holds most of them
        assert first.hugeTiles > 0 and first.walkTiles < first.hugeTiles
        assert st.hugeTiles > 0
    else:
        # every tile passes every queue to the walk; on the second frame the large
        # configuration's kernel sends it there directly (the hint from the first)
        assert first.walkTiles == first.largeTiles == first.hugeTiles == first.giantTiles == first.colossalTiles > 0
        assert st.largeTiles == st.walkTiles == first.walkTiles
        assert st.hugeTiles == st.giantTiles == st.colossalTiles == 0
    gv.close()
    gs.close()


def test_dense_scene_server_grows_lists(H, ctx, dense, dense_view):
    # the serving loop renders an overflowed frame again after growing the view's lists,
    # and new views of the shape start at the grown size
    gs = H.GpuScene(ctx, dense)
    srv = H.Server(ctx, 8 << 30, 2)
    sid = srv.add_shape(dense_view, 1)
    before = srv.shape_bytes(sid)
    rid, img = srv.submit(gs, sid, dense_view.camera_to_world, None)
    srv.drain()
    done, lat = srv.done(rid)
    st = srv.stats()
    assert done and st.listGrowths == 1 and st.failed == 0
    assert srv.shape_bytes(sid) > before
    # two at once: the second view is new and starts grown, so nothing renders twice
    r2 = [srv.submit(gs, sid, dense_view.camera_to_world, None) for _ in range(2)]
    srv.drain()
    st = srv.stats()
    assert st.listGrowths == 1 and st.viewsCreated == 2
    gv = H.GpuView(ctx, dense_view, 1)
    gv.reserve_lists(1 << 23)
    gv.render(gs)
    ref = gv.read_image()
    for _, im in [(rid, img)] + r2:
        assert np.array_equal(im.reshape(ref.shape), ref)
    srv.close()
    gv.close()
    gs.close()


def test_giant_configuration_lists_match(H, O, ctx):
    # the 2M-triangle room (8,192 meshes) at 256x256: some tiles need more than the huge
    # configuration's 4,608 records or 512 stack entries and are built by the giant one
    # (10,240 records, a sixteen-register stack); lists, ids and masks match the oracle bit
    # for bit on the first frame after the list growth and again on a frame that follows
    # the per-tile hints
    scene = H.HostScene(H.SCENE_ROOM, 0, 8192)
    view = H.HostView(H.VIEW_PRESET_ROOM_EYE, 0, 256, 256)
    ref = O.render_view(scene, view, 1, with_colors=False)
    gs = H.GpuScene(ctx, scene)
    gv = H.GpuView(ctx, view, 1)
    gv.render(gs)
    need, cap, overflowed = gv.list_demand()
    assert need == int(ref["counts"].sum())
    if overflowed:
        gv.reserve_lists(need)
    rtri, rmask = O.gbuffer_slots(ref["tri_ids"])
    for frame in range(2):
        gv.render(gs)
        st = gv.stats()
        print(f"frame {frame}: large {st.largeTiles}, huge {st.hugeTiles}, giant {st.giantTiles}, "
              f"colossal {st.colossalTiles}, walk {st.walkTiles}, deepest stack {st.maxTileStack}", flush=True)
        assert st.overflow == 0 and st.giantTiles > 0, "no tile reached the giant configuration"
        counts, lists = gv.read_tile_lists()
        assert np.array_equal(counts, ref["counts"]), f"frame {frame}: tile list lengths differ"
        assert np.array_equal(lists, ref["lists"]), f"frame {frame}: tile list contents/order differ"
        tri, mask = gv.read_gbuffer()
        bad = np.nonzero((tri != rtri).any(1) | (mask != rmask).any(1))[0]
        assert len(bad) == 0, f"frame {frame}: {len(bad)} samples differ, first {bad[:5]}"
        assert st.survivorPairs == ref["survivors"]
    gv.close()
    gs.close()


def test_colossal_configuration_lists_match(H, O, ctx):
    # the 2M-triangle room at 96x96: each tile's frustum is wide, and some tiles need more
    # than the giant configuration's 10,240 records (up to ~19k here) and are built by the
    # colossal one (20,480 records, a sixteen-register stack, a CU's whole LDS) instead
    # of the one-wave walk; lists, ids and masks match the oracle bit for bit on the first
    # frame after the list growth and on a frame that follows the per-tile hints
    scene = H.HostScene(H.SCENE_ROOM, 0, 8192)
    view = H.HostView(H.VIEW_PRESET_ROOM_EYE, 0, 96, 96)
    ref = O.render_view(scene, view, 1, with_colors=False)
    gs = H.GpuScene(ctx, scene)
    gv = H.GpuView(ctx, view, 1)
    gv.render(gs)
    need, cap, overflowed = gv.list_demand()
    assert need == int(ref["counts"].sum())
    if overflowed:
        gv.reserve_lists(need)
    rtri, rmask = O.gbuffer_slots(ref["tri_ids"])
    for frame in range(2):
        gv.render(gs)
        st = gv.stats()
        print(f"frame {frame}: large {st.largeTiles}, huge {st.hugeTiles}, giant {st.giantTiles}, "
              f"colossal {st.colossalTiles}, walk {st.walkTiles}, deepest stack {st.maxTileStack}", flush=True)
        assert st.overflow == 0 and st.colossalTiles > 0, "no tile reached the colossal configuration"
        counts, lists = gv.read_tile_lists()
        assert np.array_equal(counts, ref["counts"]), f"frame {frame}: tile list lengths differ"
        assert np.array_equal(lists, ref["lists"]), f"frame {frame}: tile list contents/order differ"
        tri, mask = gv.read_gbuffer()
        bad = np.nonzero((tri != rtri).any(1) | (mask != rmask).any(1))[0]
        assert len(bad) == 0, f"frame {frame}: {len(bad)} samples differ, first {bad[:5]}"
        assert st.survivorPairs == ref["survivors"]
    gv.close()
    gs.close()


def test_colossal_configuration_in_a_batch(H, O, ctx):
    # the same 2M-triangle room at 96x96, both eyes in one batch (each kernel launched once
    # for both views, the queued configurations taking every view's queue in turn): tiles
    # past the giant configuration are built by the batch's colossal kernel, and each view's
    # lists, ids and masks match the oracle bit for bit, eagerly and replayed as a graph
    scene = H.HostScene(H.SCENE_ROOM, 0, 8192)
    hviews = [H.HostView(H.VIEW_PRESET_ROOM_EYE, e, 96, 96) for e in range(2)]
    refs = [O.render_view(scene, hv, 1, with_colors=False) for hv in hviews]
    gs = H.GpuScene(ctx, scene)
    views = [H.GpuView(ctx, hv, 1) for hv in hviews]
    for gv, ref in zip(views, refs):
        gv.reserve_lists(int(ref["counts"].sum()))
    batch = H.GpuBatch(ctx, views)
    for graph in (False, True):
        batch.set_graph(graph)
        for _ in range(2):  # the second frame follows the per-tile hints
            batch.render(gs)
            batch.synchronize()
        for e, (gv, ref) in enumerate(zip(views, refs)):
            st = gv.stats()
            print(f"graph {graph} eye {e}: large {st.largeTiles}, huge {st.hugeTiles}, giant {st.giantTiles}, "
                  f"colossal {st.colossalTiles}, walk {st.walkTiles}", flush=True)
            assert st.overflow == 0
            counts, lists = gv.read_tile_lists()
            assert np.array_equal(counts, ref["counts"]), f"eye {e}: tile list lengths differ"
            assert np.array_equal(lists, ref["lists"]), f"eye {e}: tile list contents/order differ"
            tri, mask = gv.read_gbuffer()
            rtri, rmask = O.gbuffer_slots(ref["tri_ids"])
            bad = np.nonzero((tri != rtri).any(1) | (mask != rmask).any(1))[0]
            assert len(bad) == 0, f"eye {e}: {len(bad)} samples differ, first {bad[:5]}"
    # the colossal configuration took part (on the hinted frames a tile goes to it
    # straight from the large one's queue)
    st0 = [gv.stats() for gv in views]
    batch.close()
    for gv in views:
        gv.close()
    gs.close()
    assert sum(s.colossalTiles for s in st0) > 0, "no tile reached the colossal configuration"

