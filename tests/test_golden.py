"""Committed golden fixtures (tests/golden/*.npz, written by tests/golden/make_golden.py).

CPU: the oracle still reproduces every fixture bit for bit (a change to the oracle's
results fails here). GPU: the HIP path matches the fixtures without an oracle render
(the oracle only regroups the fixture's subsample ids into GBuffer slots): tile list
lengths and list digest, triangle ids and subsample masks bit-exact, colours and image
within 1/255."""

import os
import sys

import numpy as np
import pytest

from conftest import ROOT, rgba_delta

GOLDEN = os.path.join(ROOT, "tests", "golden")
sys.path.insert(0, GOLDEN)

import make_golden  # noqa: E402

NAMES = sorted(make_golden.CASES)


def _load(name):
    with np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


@pytest.mark.parametrize("name", NAMES)
def test_oracle_reproduces_fixture(H, O, name):
    want = _load(name)
    scene, view, A = make_golden.build_case(H, name)
    got = make_golden.compute(O, scene, view, A)
    for key in want:
        assert np.array_equal(got[key], want[key]), f"{name}: {key} differs"


@pytest.mark.gpu
@pytest.mark.parametrize("name", NAMES)
def test_gpu_matches_fixture(H, O, ctx, name):
    want = _load(name)
    scene, view, A = make_golden.build_case(H, name)
    gs, gv = H.GpuScene(ctx, scene), H.GpuView(ctx, view, A)
    gv.render(gs)
    counts, lists = gv.read_tile_lists()
    assert np.array_equal(counts, want["counts"])
    assert np.array_equal(make_golden.lists_digest(lists), want["lists_sha256"])
    tri, mask = gv.read_gbuffer()
    rtri, rmask = O.gbuffer_slots(want["tri_ids"])
    assert np.array_equal(tri, rtri) and np.array_equal(mask, rmask)
    assert rgba_delta(gv.read_sample_colors(), want["colors"]).max() <= 1
    assert rgba_delta(gv.read_image(), want["image"]).max() <= 1
