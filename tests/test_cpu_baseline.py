"""The CPU baseline bench.py reports (oracle/cpu_baseline.cpp): the reference's AVX2
traversal (traverse_avx.h:104-527) restated on __m256 lanes with FTZ/DAZ, then the scalar
oracle's intersect and resolve, on a thread pool.

CPU: its tile lists equal the scalar oracle's list for list, in order, on the heightfield
scene at 640x480 and on a full 1080x1200 eye of the 262k-triangle room; a short timed run
renders blocks on the requested threads and counts their subsamples."""

import numpy as np
import pytest


@pytest.mark.parametrize("case", ["heightfield_640x480", "room_1080x1200"])
def test_avx2_lists_equal_scalar_oracle(H, O, scene1, scene2, case):
    if case == "heightfield_640x480":
        scene, view = scene1, H.HostView(H.VIEW_PRESET_HEIGHTFIELD, 0, 640, 480)
    else:
        scene, view = scene2, H.HostView(H.VIEW_PRESET_ROOM_EYE, 0, 1080, 1200)
    _, _, nodes = O.refit(scene)
    _, tile_pre = O.frustum_setup(view.tile_frusta, view.camera_to_world)
    _, block_pre = O.frustum_setup(view.block_frusta, view.camera_to_world)
    counts, lists = O.build_tile_lists(nodes, block_pre, tile_pre)
    acounts, alists = O.avx_build_tile_lists(nodes, block_pre, tile_pre)
    assert lists.size > 100000
    assert np.array_equal(acounts, counts), "tile list lengths differ"
    assert np.array_equal(alists, lists), "tile list contents or order differ"


def test_baseline_run_counts_work(H, O, scene1):
    view = H.HostView(H.VIEW_PRESET_HEIGHTFIELD, 0, 640, 480)
    st = O.cpu_baseline_run(scene1, [view], 1, threads=2, budget_s=1.0)
    assert st.status == 0 and st.threads == 2 and st.hardwareConcurrency >= 1
    assert st.blocks > 0 and st.subsamples > 0 and 0 < st.seconds <= 1.0
    assert st.traverseSeconds > 0 and st.intersectSeconds > 0 and st.resolveSeconds > 0
