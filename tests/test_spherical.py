"""Spherical-section ray generation (Camera::setSphericalWarpSettings, camera.cpp:206-224;
sphericalUVToDirection and its uses, sample_hierarchy.cpp:27-35, 68-80, 144-150).

CPU pins: every beam's centre ray is the unit direction of the closed form yaw =
(u - 1/2) fovX, pitch = -(v - 1/2) fovY at its sample position (u, v), computed here in
double from the grid's positions; du and dv lead to directions on the unit sphere; each
tile frustum's corner rays are unit vectors from the lens square whose yaw and pitch bound
the tile's samples (the frusta are the sample-space rectangle's corners mapped onto the
sphere, the reference's construction, which fits no far plane). GPU: a spherical-section
view of the room scene matches the oracle (ids and masks bit-exact, colours within 1/255);
through the C++ API, Camera::setSphericalWarpSettings gives the image of the same view
through the C ABI, and a camera given another's sample hierarchy (getSampleData /
setSampleData) renders the same image."""

import os
import subprocess

import numpy as np
import pytest

W, H_, FX, FY = 320, 160, 150.0, 75.0


def _view(H, lens=(0.0, 1.0)):
    return H.HostView.from_samples(H.grid_samples(W, H_), W, H_, (-1.0, -0.5, 1.0, 0.5), spherical=(FX, FY),
                                   lens_radius=lens[0], focal_distance=lens[1])


def _yaw_pitch(d):
    d = d.astype(np.float64)
    d = d / np.linalg.norm(d, axis=-1, keepdims=True)
    return np.arctan2(d[..., 0], -d[..., 2]), np.arcsin(np.clip(d[..., 1], -1, 1))


def test_spherical_beams_closed_form(H):
    view = _view(H)
    idx = view.remap.reshape(H_, W)
    assert (idx >= 0).all()
    ys, xs = np.mgrid[0:H_, 0:W]
    u = (xs.astype(np.float32) + np.float32(0.5)) * np.float32(1.0 / W)
    v = (ys.astype(np.float32) + np.float32(0.5)) * np.float32(1.0 / H_)
    yaw = (u.astype(np.float64) - 0.5) * np.radians(FX)
    pitch = -(v.astype(np.float64) - 0.5) * np.radians(FY)
    want = np.stack([np.sin(yaw) * np.cos(pitch), np.sin(pitch), -np.cos(yaw) * np.cos(pitch)], -1)
    beams = view.beams[idx]
    c = beams["centerRay"].astype(np.float64)
    assert np.abs(c - want).max() < 2e-6
    for k in ("du", "dv"):
        e = c + beams[k].astype(np.float64)
        assert np.abs(np.linalg.norm(e, axis=-1) - 1.0).max() < 2e-6  # a point of the sphere
        ang = np.arccos(np.clip((e * c).sum(-1), -1, 1))
        assert (ang > 0).all() and ang.max() < 2 * np.radians(FX) / W  # within a couple of pixels


def test_spherical_frusta_bound_their_samples(H):
    view = _view(H, lens=(0.01, 3.0))
    n = view.valid_sample_count
    yaw, pitch = _yaw_pitch(view.beams["centerRay"][:n])
    fr = view.tile_frusta
    o = fr["origins"]
    r = 0.01
    assert np.allclose(o[:, :, 2], 0) and np.allclose(np.abs(o[:, :, :2]), r)  # the lens square
    assert np.allclose(np.linalg.norm(fr["directions"], axis=-1), 1.0, atol=1e-6)
    cy, cp = _yaw_pitch(fr["directions"])
    for t in range(view.tile_count):
        lo, hi = t * 128, min((t + 1) * 128, n)
        if lo >= n:
            break
        assert yaw[lo:hi].min() >= cy[t].min() - 1e-6 and yaw[lo:hi].max() <= cy[t].max() + 1e-6
        assert pitch[lo:hi].min() >= cp[t].min() - 1e-6 and pitch[lo:hi].max() <= cp[t].max() + 1e-6


@pytest.mark.gpu
def test_spherical_view_matches_oracle(H, O, ctx, scene2):
    from conftest import rgba_delta

    view = _view(H)
    m = H.HostView(H.VIEW_PRESET_ROOM_EYE, 0, W, H_).camera_to_world
    gs, gv = H.GpuScene(ctx, scene2), H.GpuView(ctx, view, 4)
    gv.render(gs, m)
    ref = O.render_view(scene2, view, 4, camera_to_world=m)
    counts, lists = gv.read_tile_lists()
    assert np.array_equal(counts, ref["counts"]) and np.array_equal(lists, ref["lists"])
    tri, mask = gv.read_gbuffer()
    rtri, rmask = O.gbuffer_slots(ref["tri_ids"])
    assert np.array_equal(tri, rtri) and np.array_equal(mask, rmask)
    assert rgba_delta(gv.read_image(), ref["image"]).max() <= 1
    assert (ref["image"] != 0).mean() > 0.9
