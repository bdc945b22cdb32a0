"""Scene textures: storage, mip generation and tex2DGrad sampling in the resolve
(raycaster/texture.cu:78-272, shading.h:45-47, 93-127, resolve.cu:236-275).

CPU pins of the oracle (oracle/or_texture.c) from what the definitions fix: the tables
against their formulas, level counts and the chain layout, format conversions on known
values, the level of detail against log2 in double, constant textures, bilinear known
answers (texel centres, midpoints, the wrap seam, the 8-bit weight), periodicity, sRGB
decoded before filtering, anisotropic probes on a quadratic ramp, trilinear level
selection on a chain whose levels are constants, mip known answers (truncation, the clamp
reads of d_mipmap), and the textured resolve on a tilted plane, whose UV derivatives have a
closed form. GPU: mip chains built on the device byte-equal to the oracle's, the device
sampler and level of detail (texture.h through a probe) bit-equal to the oracle on random
queries, textured frames (pinhole, depth of field, batched with a graph) within 1/255 of
the oracle, untextured materials unchanged by a texture upload."""

import ctypes
import os

import numpy as np
import pytest

from conftest import ROOT, rgba_delta

PROBE = os.path.join(ROOT, "tests", "gpu_units", "_build", "libhvvr_gpu_units.so")


def _srgb_to_linear(x):
    x = np.asarray(x, np.float64)
    return np.where(x <= 0.04045, x / 12.92, ((x + 0.055) / 1.055) ** 2.4)


def _lod_double(x, levels):
    """(level, fraction, distance of 256 * fraction from a rounding boundary) in double"""
    if not x >= 1.0:
        return 0, 0.0, 1.0
    lam = np.log2(np.float64(x))
    if lam >= levels - 1:
        return levels - 1, 0.0, 1.0
    l0 = np.floor(lam)
    q = (lam - l0) * 256.0
    return int(l0), np.floor(q + 0.5) / 256.0, abs(q - np.floor(q) - 0.5)


def _constant_levels(O, size, levels_value=None):
    """float4 texture size x size whose level l holds the constant l"""
    info, chain = O.texture_chain("float4", np.zeros((size, size, 4), np.float32))
    for l in range(int(info[0]["levels"])):
        n = max(size >> l, 1) ** 2
        off = int(info[0]["offset"][l])
        chain[off: off + 16 * n] = np.full(4 * n, l if levels_value is None else levels_value, np.float32).view(np.uint8)
    return info, chain


def _level(info, chain, l):
    """texels of level l (bytes for the 8-bit kinds, floats for kind 2)"""
    t = info[0]
    n = max(int(t["width"]) >> l, 1) * max(int(t["height"]) >> l, 1) * (16 if t["kind"] == 2 else 4)
    lv = chain[int(t["offset"][l]):][:n]
    return lv.view(np.float32) if t["kind"] == 2 else lv


# ---- oracle pins (CPU) ----

def test_tables(O):
    t = O.texture_tables()
    x = (np.arange(256, dtype=np.float32) / np.float32(255)).astype(np.float64)
    assert np.allclose(t[:256], _srgb_to_linear(x), rtol=2e-6, atol=1e-9)
    j = np.arange(256)
    assert np.allclose(t[256:], 2.0 ** ((j + 0.5) / 256), rtol=1e-7, atol=0)
    assert np.all(np.diff(t[256:]) > 0) and t[256] > 1 and t[511] < 2


@pytest.mark.parametrize("w,h", [(1, 1), (2, 1), (1, 7), (37, 19), (640, 480), (5, 3)])
def test_levels_and_layout(O, w, h):
    info, chain = O.texture_chain("unorm8", np.zeros((h, w, 4), np.uint8))
    levels = int(info[0]["levels"])
    assert levels == int(np.floor(np.log2(max(w, h)))) + 1
    at = 0
    for l in range(levels):
        assert int(info[0]["offset"][l]) == at and at % 16 == 0
        at += -(-max(w >> l, 1) * max(h >> l, 1) * 4 // 16) * 16
    assert chain.size == at


def test_format_conversions(O):
    k, v = O.texture_level0("unorm16", np.array([[[0, 65535, 32768, 1]]], np.uint16))
    assert k == 2 and v[0, 0, 0] == 0 and v[0, 0, 1] == 1
    assert v[0, 0, 2] == np.float32(32768) / np.float32(65535) and v[0, 0, 3] == np.float32(1 / 65535)
    h = np.array([[[0.5, -2.0, 65504.0, 2.0 ** -24]]], np.float16)
    assert np.array_equal(O.texture_level0("half4", h)[1], h.astype(np.float32))
    # r11g11b10: 1.0 and 0.5 as 11-bit floats (exponent 15 / 14, bias 15), 2.0 as a 10-bit one
    c = np.array([[960 | (896 << 11) | (512 << 22), 1]], np.uint32)
    _, v = O.texture_level0("r11g11b10", c)
    assert v[0, 0].tolist() == [1.0, 0.5, 2.0, 1.0]
    assert v[0, 1].tolist() == [2.0 ** -20, 0.0, 0.0, 1.0]  # smallest denormal of the red channel
    _, v = O.texture_level0("r32", np.array([[3.25]], np.float32))
    assert v[0, 0].tolist() == [3.25, 0.0, 0.0, 1.0]
    k, v = O.texture_level0("srgb8", np.full((1, 1, 4), 7, np.uint8))
    assert k == 1 and v.dtype == np.uint8


def test_lod_matches_log2(O):
    rng = np.random.default_rng(5)
    xs = np.concatenate([2.0 ** rng.uniform(-3, 12, 6000), 2.0 ** np.arange(-2, 13),
                         [0.0, -1.0, 1.0, 1.9999999, 1e-45, 3e38]]).astype(np.float32)
    levels, checked = 10, 0
    for x in xs:
        lvl, f = O.texture_lod(x, levels)
        wl, wf, margin = _lod_double(x, levels)
        if margin < 0.02:
            continue
        assert (lvl, f) == (wl, wf), x
        checked += 1
    assert checked > 5800
    assert O.texture_lod(np.nan, levels) == (0, 0.0)
    assert O.texture_lod(np.inf, levels) == (levels - 1, 0.0)
    assert O.texture_lod(5.0, 1) == (0, 0.0)  # a one-level texture


def _queries(rng, n, scale=(-4, 1)):
    uv = rng.uniform(-3, 3, (n, 2))
    d = rng.normal(size=(n, 4)) * 10.0 ** rng.uniform(*scale, (n, 1))
    q = np.concatenate([uv, d], 1).astype(np.float32)
    q[: n // 8, 2:] = 0  # no footprint
    q[n // 8: n // 4, 4:] = 0  # a degenerate minor axis
    return q


def test_constant_texture(O):
    rng = np.random.default_rng(1)
    c = np.array([0.25, 0.5, 0.75, 1.0], np.float32)
    info, chain = O.texture_chain("float4", np.broadcast_to(c, (8, 8, 4)))
    for q in _queries(rng, 400):
        assert np.allclose(O.texture_sample(info, chain, q[:2], q[2:4], q[4:]), c, rtol=1e-6, atol=0)
    for byte in (0, 255):
        info, chain = O.texture_chain("unorm8", np.full((6, 10, 4), byte, np.uint8))
        for l in range(int(info[0]["levels"])):  # every mip level keeps the value exactly
            assert np.all(_level(info, chain, l) == byte)
        for q in _queries(rng, 100):
            assert np.all(O.texture_sample(info, chain, q[:2], q[2:4], q[4:]) == byte / 255.0)


def test_bilinear_known_answers(O):
    t = np.zeros((2, 2, 4), np.float32)
    t[..., 0] = [[1, 2], [4, 8]]  # row y, column x
    t[..., 1] = [[16, 0], [0, 0]]
    info, chain = O.texture_chain("float4", t)
    s = lambda u, v: O.texture_sample(info, chain, (u, v))  # noqa: E731
    assert np.array_equal(s(0.25, 0.25), t[0, 0]) and np.array_equal(s(0.75, 0.25), t[0, 1])
    assert np.array_equal(s(0.25, 0.75), t[1, 0]) and np.array_equal(s(0.75, 0.75), t[1, 1])
    assert np.array_equal(s(0.5, 0.25), (t[0, 0] + t[0, 1]) / 2)
    assert np.array_equal(s(0.0, 0.25), (t[0, 1] + t[0, 0]) / 2)  # wrap seam: last and first column
    assert np.array_equal(s(0.5, 0.5), t.reshape(4, 4).mean(0))
    assert np.array_equal(s(0.0, 0.0), t.reshape(4, 4).mean(0))  # both seams
    # x = 0.3 * 2 - 0.5 = 0.1: the weight of the second texel is 26/256 (8-bit fraction)
    a = 26 / 256
    assert np.allclose(s(0.3, 0.25), t[0, 0] * (1 - a) + t[0, 1] * a, rtol=1e-7, atol=0)


def test_wrap_periodic(O):
    rng = np.random.default_rng(2)
    info, chain = O.texture_chain("float4", rng.random((3, 5, 4)).astype(np.float32))
    for _ in range(200):
        uv = rng.integers(0, 64, 2) / 64.0
        d = (rng.normal(size=4) * 0.05).astype(np.float32)
        base = O.texture_sample(info, chain, uv, d[:2], d[2:])
        for sh in ((1, 0), (-2, 0), (0, 3), (4, -1)):
            assert np.array_equal(O.texture_sample(info, chain, uv + sh, d[:2], d[2:]), base)


def test_srgb_decoded_before_filtering(O):
    info, chain = O.texture_chain("srgb8", np.array([[[0, 10, 128, 128]]], np.uint8))
    got = O.texture_sample(info, chain, (0.3, 0.8))
    want = list(_srgb_to_linear(np.array([0, 10, 128]) / 255.0)) + [128 / 255.0]  # alpha stays linear
    assert np.allclose(got, want, rtol=2e-6, atol=0)
    # a 2x1 texture (0, 255) read between the texels: the mean of the decoded values, 0.5,
    # not the decode of the mean byte (0.2158)
    info, chain = O.texture_chain("srgb8", np.array([[[0, 0, 0, 0], [255, 255, 255, 255]]], np.uint8))
    assert np.allclose(O.texture_sample(info, chain, (0.5, 0.5))[:3], 0.5, rtol=1e-6)


def test_anisotropic_probes(O):
    ramp = np.zeros((4, 64, 4), np.float32)
    ramp[..., 0] = np.arange(64, dtype=np.float32) ** 2  # quadratic in the column
    info, chain = O.texture_chain("float4", ramp)
    uv = (0.5, 0.5)
    # 8 texels along u, 1 along v: 8 probes at the centres of columns 28..35
    assert O.texture_sample(info, chain, uv, (8 / 64, 0), (0, 1 / 4))[0] == 997.5
    # a degenerate minor axis takes the most probes
    assert O.texture_sample(info, chain, uv, (8 / 64, 0), (0, 0))[0] == 997.5
    # 3 texels along u: 3 probes between columns 30|31, 31|32 and 32|33
    assert np.isclose(O.texture_sample(info, chain, uv, (3 / 64, 0), (0, 1 / 4))[0], 5959 / 6, rtol=1e-7)
    # isotropic: one bilinear read between columns 31 and 32
    assert O.texture_sample(info, chain, uv, (1 / 64, 0), (0, 1 / 4))[0] == (31 ** 2 + 32 ** 2) / 2
    # the major axis is v when it is longer
    info, chain = O.texture_chain("float4", np.ascontiguousarray(ramp.transpose(1, 0, 2)))
    assert O.texture_sample(info, chain, uv, (1 / 4, 0), (0, 8 / 64))[0] == 997.5


def test_level_selection(O):
    info, chain = _constant_levels(O, 256)
    levels = int(info[0]["levels"])
    assert levels == 9
    checked = 0
    for s in np.concatenate([2.0 ** np.linspace(-1, 10, 300), [1.0, 2.0, 1e6]]):
        for ratio in (1.0, 2.5):
            d = np.float32(s / 256)
            got = O.texture_sample(info, chain, (0.3, 0.6), (d, 0), (0, np.float32(d / ratio)))[0]
            n = min(int(np.ceil(ratio)), 8)
            l0, f, margin = _lod_double(np.float32(s) / n, levels)
            if margin < 0.02:
                continue
            assert np.isclose(got, l0 + f, rtol=0, atol=1e-5), (s, ratio)
            checked += 1
    assert checked > 550


def test_mip_known_answers(O):
    t = np.zeros((2, 2, 4), np.uint8)
    t[..., 0] = [[0, 255], [255, 0]]
    t[..., 1] = [[255, 255], [255, 0]]
    t[..., 2] = 255
    info, chain = O.texture_chain("unorm8", t)
    top = chain[int(info[0]["offset"][1]):][:4]
    assert top.tolist() == [127, 191, 255, 0]  # 127.5 and 191.25 truncate
    # 4x4 -> 2x2 reads at (x, y) / 2 with clamp addressing: one lit column (x = 1) of four
    # gives 0.25 * 255 = 63.75 -> 63 in both columns of level 1
    t = np.zeros((4, 4, 4), np.uint8)
    t[:, 1, 0] = 255
    info, chain = O.texture_chain("unorm8", t)
    lv1 = chain[int(info[0]["offset"][1]):][:16].reshape(2, 2, 4)
    assert np.all(lv1[..., 0] == 63) and np.all(lv1[..., 1:] == 0)
    lv2 = chain[int(info[0]["offset"][2]):][:4]
    assert lv2[0] in (62, 63)
    # float kinds keep the float mean
    f = np.zeros((2, 2, 4), np.float32)
    f[..., 0] = [[1, 2], [4, 8]]
    info, chain = O.texture_chain("float4", f)
    assert chain[int(info[0]["offset"][1]):][:16].view(np.float32)[0] == 3.75
    info, chain = O.texture_chain("float4", np.full((5, 3, 4), 0.375, np.float32))
    for l in range(int(info[0]["levels"])):
        assert np.all(_level(info, chain, l) == 0.375)


def _tilted_plane_scene(H, d, slope, k, textured=True):
    """two triangles on the plane z = -d - slope * y, UV = k * (x, y) at every point;
    an emissive material whose colour is texture 0"""
    c = [(-40.0, -30.0), (40.0, -30.0), (40.0, 30.0), (-40.0, 30.0)]
    verts = np.zeros(4, H.DTYPES["vertex"])
    n = np.array([0.0, slope, 1.0]) / np.hypot(slope, 1.0)
    for i, (x, y) in enumerate(c):
        verts[i]["pos"] = (x, y, -d - slope * y)
        verts[i]["normal"] = np.array(list(n) + [0.0], np.float16).view(np.uint16)
        verts[i]["uv"] = (k * x, k * y)
    mat = np.zeros(1, H.DTYPES["material"])
    mat["emissive"] = (0.5, 0.5, 0.5, 1.0)
    tex = 0 if textured else 0xFFFF
    mat["textureIDsAndShadingModel"] = 2 | (tex << 48) | (0xFFFF << 32) | (0xFFFF << 16)
    return H.HostScene.from_meshes([(verts, [[0, 1, 2], [0, 2, 3]], [0, 0], None)], mat)


def _plane_view(H, w=64, h=48):
    t = np.tan(np.radians(30.0))
    return H.HostView.from_samples(H.grid_samples(w, h), w, h, (-t * w / h, -t, t * w / h, t))


def test_textured_resolve_footprint_closed_form(H, O):
    # On the plane, a subsample offset (a, b) hits at p = t D with D = c + a du + b dv and
    # t = -d / (n . D); UV = k (p.x, p.y). The derivatives the resolve takes are the UV
    # differences one unit along each offset axis, so with texture levels holding their
    # own index, the emissive colour is level + mip fraction of that footprint.
    d, slope, k = 6.0, 0.9, 0.37
    scene = _tilted_plane_scene(H, d, slope, k)
    view = _plane_view(H)
    info, chain = _constant_levels(O, 256)
    ref = O.render_view(scene, view, 1, camera_to_world=np.eye(4, dtype=np.float32).reshape(16),
                        textures=[(info, chain)])
    o = H.subsample_offsets(1)[0].astype(np.float64)
    b = view.beams
    cr, du, dv = (b[f].astype(np.float64) for f in ("centerRay", "du", "dv"))
    nrm = np.array([0.0, slope, 1.0])

    def uv(a, bb):
        D = cr + a * du + bb * dv
        return k * (-d / (D @ nrm))[:, None] * D[:, :2]

    base = uv(o[0], o[1])
    ddx, ddy = uv(o[0] + 1, o[1]) - base, uv(o[0], o[1] + 1) - base
    lx, ly = np.linalg.norm(ddx * 256, axis=1), np.linalg.norm(ddy * 256, axis=1)
    hit = ref["tri_ids"] != 0xFFFFFFFF
    got = ref["linear"][:, 0]
    checked, levels_seen = 0, set()
    for s in np.nonzero(hit)[0]:
        pmax, pmin = max(lx[s], ly[s]), min(lx[s], ly[s])
        r = pmax / pmin
        if abs(r - np.round(r)) < 1e-3:
            continue
        n = min(int(np.ceil(r)), 8)
        l0, f, margin = _lod_double(pmax / n, 9)
        if margin < 0.05:
            continue
        assert abs(got[s] - (l0 + f)) < 1e-5, (s, got[s], l0 + f)
        levels_seen.add(l0)
        checked += 1
    assert checked > 0.9 * hit.sum() and hit.sum() > 2500
    assert len(levels_seen) >= 2  # the tilt spreads the footprint over several levels


def test_textured_resolve_constant_texture_equals_constant_material(H, O, scene1, view1_small):
    # a diffuse texture holding the material's own colour shades exactly like the colour
    # (0.5 is exact through every probe average), so the textured path changes nothing else
    s = _with_materials(H, scene1, lambda m: _set_diffuse_tex(m, 0, every=1, colour=0.5))
    plain = _with_materials(H, scene1, lambda m: _set_colour(m, 0.5))
    tex = [("float4", np.full((16, 16, 4), 0.5, np.float32))]
    a = O.render_view(s, view1_small, 4, textures=tex)
    b = O.render_view(plain, view1_small, 4)
    assert np.array_equal(a["colors"], b["colors"]) and np.array_equal(a["linear"], b["linear"])
    c = O.render_view(s, view1_small, 4)  # no textures uploaded: the constants
    assert np.array_equal(c["colors"], b["colors"])


class _Scene:
    def __init__(self, base, materials, uv_scale=1.0):
        for f in ("vertexToTransform", "triShade", "nodes", "transforms", "lights"):
            setattr(self, f, getattr(base, f))
        self.verts = np.array(base.verts)
        self.verts["uv"] *= uv_scale
        self.materials = materials


def _with_materials(H, base, edit, uv_scale=1.0):
    m = np.array(base.materials)
    edit(m)
    return _Scene(base, m, uv_scale)


def _code(model, diffuse=0xFFFF, specular=0xFFFF, glossy=0xFFFF):
    return model | (diffuse << 48) | (specular << 32) | (glossy << 16)


def _set_colour(m, c):
    m["diffuse"][:, :3] = c


def _set_diffuse_tex(m, tex, every=1, colour=None):
    if colour is not None:
        m["diffuse"][:, :3] = colour
    m["textureIDsAndShadingModel"][::every] = _code(1, diffuse=tex)


def _mixed_materials(m):
    n = len(m)
    idx = np.arange(n)
    codes = np.full(n, _code(1), np.uint64)
    codes[idx % 6 == 0] = _code(1, diffuse=0)
    codes[idx % 6 == 1] = _code(1, diffuse=0, specular=1, glossy=2)
    codes[idx % 6 == 2] = _code(1, glossy=2)
    codes[idx % 6 == 3] = _code(2, diffuse=3)
    codes[idx % 6 == 4] = _code(1, diffuse=9)  # beyond the texture count: the constants
    m["textureIDsAndShadingModel"] = codes
    m["emissive"][idx % 6 == 3] = (0.3, 0.2, 0.1, 1.0)


def _textures(seed=0):
    rng = np.random.default_rng(seed)
    checker = ((np.indices((64, 64)).sum(0) // 8) % 2).astype(np.uint8)
    srgb = np.stack([checker * 200 + 30, 255 - checker * 180, rng.integers(0, 256, (64, 64)),
                     np.full((64, 64), 255)], -1).astype(np.uint8)
    spec = rng.integers(0, 64, (32, 48, 4)).astype(np.uint8)
    gloss = rng.random((37, 19, 4)).astype(np.float32)
    emis = (rng.random((16, 16, 4)) * 2).astype(np.float16)
    return [("srgb8", srgb), ("unorm8", spec), ("float4", gloss), ("half4", emis)]


def test_textured_resolve_reads_textures(H, O, scene1, view1_small):
    s = _with_materials(H, scene1, _mixed_materials, uv_scale=6.0)
    a = O.render_view(s, view1_small, 4, textures=_textures())
    b = O.render_view(s, view1_small, 4)
    changed = (a["colors"] != b["colors"]).mean()
    assert changed > 0.2  # most hit samples read a texture


# ---- GPU ----

ALL_FORMATS = [
    ("srgb8", lambda r: r.integers(0, 256, (64, 64, 4)).astype(np.uint8)),
    ("unorm8", lambda r: r.integers(0, 256, (19, 37, 4)).astype(np.uint8)),
    ("unorm16", lambda r: r.integers(0, 65536, (8, 16, 4)).astype(np.uint16)),
    ("float4", lambda r: r.normal(size=(33, 9, 4)).astype(np.float32)),
    ("half4", lambda r: r.normal(size=(20, 20, 4)).astype(np.float16)),
    ("r11g11b10", lambda r: r.integers(0, 2 ** 32, (7, 13), dtype=np.uint64).astype(np.uint32) & 0x7BDEF7BD),
    ("r32", lambda r: r.normal(size=(5, 5)).astype(np.float32)),
    ("unorm8", lambda r: r.integers(0, 256, (1, 1, 4)).astype(np.uint8)),
]


def _gpu_texture_set(rng):
    return [(f, g(rng)) for f, g in ALL_FORMATS]


@pytest.mark.gpu
def test_gpu_texture_chains_match_oracle(H, O, ctx, scene1):
    rng = np.random.default_rng(11)
    texs = _gpu_texture_set(rng)
    gs = H.GpuScene(ctx, scene1)
    gs.set_textures(texs)
    for i, (fmt, data) in enumerate(texs):
        info, chain = O.texture_chain(fmt, data)
        w, h, levels, kind, nbytes = gs.texture_info(i)
        assert (w, h, levels, kind) == tuple(int(info[0][f]) for f in ("width", "height", "levels", "kind"))
        assert nbytes == chain.size
        got = gs.read_texture(i)
        assert np.array_equal(got, chain), f"texture {i} ({fmt}): chain differs at {np.nonzero(got != chain)[0][:5]}"
    gs.set_textures([])  # back to none
    with pytest.raises(RuntimeError):
        gs.texture_info(0)


@pytest.mark.gpu
def test_gpu_texture_errors(H, ctx, scene1):
    gs = H.GpuScene(ctx, scene1)
    with pytest.raises(ValueError):
        gs.set_textures([("srgb8", np.zeros((4, 4), np.uint8))])  # missing channels
    d = H._TextureDesc(99, 4, 4, 4, None)
    assert ctx.lib.hvvr_gpu_scene_set_textures(gs.handle, ctypes.c_uint32(1), ctypes.byref(d)) != 0


def _probe():
    lib = ctypes.CDLL(PROBE)
    return lib


@pytest.mark.gpu
@pytest.mark.parametrize("which", range(len(ALL_FORMATS)))
def test_gpu_sampler_matches_oracle(O, which):
    rng = np.random.default_rng(100 + which)
    fmt, gen = ALL_FORMATS[which]
    info, chain = O.texture_chain(fmt, gen(rng))
    q = np.concatenate([_queries(rng, 3000, (-4, 1)), _queries(rng, 1000, (-1, 3))])
    q[::97, 0] = 1e30  # far outside: still a wrapped texel
    tables = O.texture_tables()
    out = np.zeros((len(q), 4), np.float32)
    rc = _probe().texture_probe_sample(O._p(chain), ctypes.c_uint64(chain.size), O._p(info), O._p(tables), O._p(q),
                                       ctypes.c_uint32(len(q)), O._p(out))
    assert rc == 0
    want = np.stack([O.texture_sample(info, chain, x[:2], x[2:4], x[4:]) for x in q])
    nan = np.isnan(want)
    assert np.array_equal(np.isnan(out), nan)
    bad = np.nonzero((out.view(np.uint32) != want.view(np.uint32)) & ~nan)[0]
    assert len(bad) == 0, f"{len(bad)} values differ, first query {q[bad[0] // 4]}: {out[bad[0] // 4]} vs {want[bad[0] // 4]}"


@pytest.mark.gpu
def test_gpu_lod_matches_oracle(O):
    rng = np.random.default_rng(3)
    x = np.concatenate([2.0 ** rng.uniform(-3, 16, 20000), 2.0 ** np.arange(-3, 17), [0, -1, np.inf, np.nan]])
    x = np.ascontiguousarray(x, np.float32)
    tables = O.texture_tables()
    for levels in (1, 9, 16):
        lvl = np.zeros(len(x), np.uint32)
        f = np.zeros(len(x), np.float32)
        assert _probe().texture_probe_lod(O._p(tables), O._p(x), ctypes.c_uint32(len(x)), ctypes.c_uint32(levels),
                                          O._p(lvl), O._p(f)) == 0
        want = [O.texture_lod(v, levels) for v in x]
        assert lvl.tolist() == [w[0] for w in want]
        assert f.tolist() == [w[1] for w in want]


def _gpu_colors(H, ctx, scene, view, A, textures=None, bits=32):
    gs = H.GpuScene(ctx, scene)
    if textures is not None:
        gs.set_textures(textures)
    gv = H.GpuView(ctx, view, A)
    gv.set_shading_precision(bits)
    gv.render(gs)
    return gs, gv


@pytest.mark.gpu
@pytest.mark.parametrize("A", [1, 4])
def test_gpu_textured_frame_matches_oracle(H, O, ctx, scene1, view1_small, A):
    s = _with_materials(H, scene1, _mixed_materials, uv_scale=6.0)
    texs = _textures()
    ref = O.render_view(s, view1_small, A, textures=texs)
    _, gv = _gpu_colors(H, ctx, s, view1_small, A, texs)
    d = rgba_delta(gv.read_sample_colors(), ref["colors"])
    assert d.max() <= 1, f"max {d.max()}"
    assert rgba_delta(gv.read_image(), ref["image"]).max() <= 1
    # reduced-precision shading of the same textured frame against the FP64 oracle
    ref64 = O.render_view(s, view1_small, A, precision=1, textures=texs)
    _, gv16 = _gpu_colors(H, ctx, s, view1_small, A, texs, bits=16)
    d16 = rgba_delta(gv16.read_sample_colors(), ref64["colors"]).reshape(-1, 4)[:, :3]
    assert d16.max() <= 2 and d16.mean() <= 0.05


@pytest.mark.gpu
def test_gpu_textured_frame_dof_matches_oracle(H, O, ctx, scene1):
    A, w, h = 4, 160, 120
    t = np.tan(np.radians(30.0))
    view = H.HostView.from_samples(H.grid_samples(w, h), w, h, (-t * w / h, -t, t * w / h, t), lens_radius=0.05,
                                   focal_distance=12.0)
    m = H.HostView(H.VIEW_PRESET_HEIGHTFIELD, 0, w, h).camera_to_world
    s = _with_materials(H, scene1, _mixed_materials, uv_scale=6.0)
    texs = _textures(1)
    ref = O.render_view(s, view, A, camera_to_world=m, textures=texs)
    gs = H.GpuScene(ctx, s)
    gs.set_textures(texs)
    gv = H.GpuView(ctx, view, A)
    gv.render(gs, m)
    assert rgba_delta(gv.read_sample_colors(), ref["colors"]).max() <= 1


@pytest.mark.gpu
def test_gpu_footprint_plane_matches_oracle(H, O, ctx):
    # the tilted plane with a random texture: footprints over several mip levels
    scene = _tilted_plane_scene(H, 6.0, 0.9, 0.37)
    view = _plane_view(H)
    rng = np.random.default_rng(4)
    data = rng.random((256, 256, 4)).astype(np.float32) * 0.8
    m = np.eye(4, dtype=np.float32).reshape(16)
    ref = O.render_view(scene, view, 1, camera_to_world=m, textures=[("float4", data)])
    gs = H.GpuScene(ctx, scene)
    gs.set_textures([("float4", data)])
    gv = H.GpuView(ctx, view, 1)
    gv.render(gs, m)
    assert rgba_delta(gv.read_sample_colors(), ref["colors"]).max() <= 1


@pytest.mark.gpu
def test_gpu_untextured_materials_unchanged_by_upload(H, O, ctx, scene1, view1_small):
    _, plain = _gpu_colors(H, ctx, scene1, view1_small, 4)
    _, up = _gpu_colors(H, ctx, scene1, view1_small, 4, _textures())
    assert np.array_equal(plain.read_sample_colors(), up.read_sample_colors())


@pytest.mark.gpu
def test_gpu_batch_graph_textured_matches_solo(H, O, ctx, scene1):
    s = _with_materials(H, scene1, _mixed_materials, uv_scale=6.0)
    gs = H.GpuScene(ctx, s)
    views = [H.GpuView(ctx, H.HostView(H.VIEW_PRESET_HEIGHTFIELD, 0, 160, 120), 4) for _ in range(2)]
    cams = [v.view.camera_to_world for v in views]
    cams[1] = cams[1].copy()
    cams[1][12] += 0.5
    batch = H.GpuBatch(ctx, views)
    batch.set_graph(True)
    solo = H.GpuView(ctx, views[0].view, 4)
    for texs in (None, _textures(), _textures(2)):
        if texs is not None:
            gs.set_textures(texs)
        for _ in range(2):  # capture, then replay
            batch.render(gs, cams)
            batch.synchronize()
            for v, m in zip(views, cams):
                solo.render(gs, m)
                assert np.array_equal(v.read_sample_colors(), solo.read_sample_colors())
    batch.close()
