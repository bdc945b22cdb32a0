"""The allocation-free render entry point (hvvr_gpu_view_render_shared_scratch).

It renders a view like hvvr_gpu_view_render, with the tile-list kernels' record scratch
and tile queues in the library's per-GPU shared array. GPU: on the 1M-triangle room, whose
tiles go to the queued tile-list configurations (the ones that use that scratch), two calls
in a row, the second on a smaller view, give the same tile lists, GBuffer, sample colours
and image as the existing entry point, bit for bit; a view beyond the array's 65,536 tiles
is rejected before any work is queued."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _results(gv):
    counts, lists = gv.read_tile_lists()
    tri, mask = gv.read_gbuffer()
    st = gv.stats()
    return dict(counts=counts, lists=lists, tri=tri, mask=mask, colors=gv.read_sample_colors(),
                image=gv.read_image()), st


def _render_grown(gv, gs, render):
    # the first frame may outgrow the default list storage; grow it (an allocation, made
    # by its own entry point) and render again
    render(gs)
    need, cap, overflowed = gv.list_demand()
    if overflowed:
        gv.reserve_lists(need)
        render(gs)
    assert not gv.list_demand()[2]


def test_shared_scratch_render_matches_render(H, ctx):
    scene = H.HostScene(H.SCENE_ROOM, 0, 4096)
    gs = H.GpuScene(ctx, scene)
    shapes = [(360, 400), (256, 256)]  # the second call's view is the smaller one
    for w, h in shapes:
        hv = H.HostView(H.VIEW_PRESET_ROOM_EYE, 0, w, h)
        ref_view = H.GpuView(ctx, hv, 1)
        _render_grown(ref_view, gs, ref_view.render)
        want, want_st = _results(ref_view)
        gv = H.GpuView(ctx, hv, 1)
        _render_grown(gv, gs, gv.render_shared_scratch)
        got, st = _results(gv)
        print(f"{w}x{h}: tiles {st.tileCount}, queued large {st.largeTiles}, huge {st.hugeTiles}, "
              f"giant {st.giantTiles}, colossal {st.colossalTiles}, walk {st.walkTiles}", flush=True)
        assert st.largeTiles > 0, "no tile used the queued configurations' scratch"
        for k in want:
            assert np.array_equal(got[k], want[k]), f"{w}x{h}: {k} differs from hvvr_gpu_view_render"
        assert (st.largeTiles, st.hugeTiles, st.giantTiles, st.colossalTiles, st.walkTiles, st.listEntries,
                st.survivorPairs) == \
            (want_st.largeTiles, want_st.hugeTiles, want_st.giantTiles, want_st.colossalTiles, want_st.walkTiles,
             want_st.listEntries, want_st.survivorPairs)
        gv.close()
        ref_view.close()
    gs.close()


def test_shared_scratch_rejects_views_beyond_its_size(H, ctx, scene1):
    gs = H.GpuScene(ctx, scene1)
    hv = H.HostView(H.VIEW_PRESET_HEIGHTFIELD, 0, 4096, 2056)  # 65,792 tiles
    assert hv.tile_count > 65536
    gv = H.GpuView(ctx, hv, 1)
    with pytest.raises(H.HvvrError, match="65536 tiles"):
        gv.render_shared_scratch(gs)
    gv.close()
    gs.close()
