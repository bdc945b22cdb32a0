"""ctypes binding of the CPU oracle (oracle/_build/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg, never by the product package.
"""

from __future__ import annotations

import ctypes
import os
import subprocess
from ctypes import c_int, c_uint32, c_uint64, c_void_p

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(_HERE, "_build", "liboracle.so")

O_FRUSTUM = np.dtype([("planeNeg", "<f4", (4, 4)), ("planePos", "<f4", (4, 4)), ("projLo", "<f4", 3),
                      ("projHi", "<f4", 3), ("edge", "<f4", (12, 3)), ("edgeMin", "<f4", 12),
                      ("edgeMax", "<f4", 12), ("axis", "<i4"), ("degenerate", "<i4")])
O_STACK = np.dtype([("tMin", "<f4"), ("nodeMask", "<u4")])
O_LIGHTENV = np.dtype([("raw", "u1", 372)])
assert O_FRUSTUM.itemsize == 400

_lib = None


def build():
    subprocess.check_call(["make", "-s", "-C", _HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = ctypes.CDLL(LIB)
        L.or_tile_traverse.restype = ctypes.c_int64
        L.or_build_tile_lists.restype = ctypes.c_int64
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(c_void_p) if a is not None else None


def layout_sizes():
    out = np.zeros(10, np.uint32)
    lib().or_layout_sizes(_p(out))
    return out


def frustum_setup(frusta, m, cull_planes=None):
    """World-space frusta + test data; cull_planes (16 floats, 4 x (normal, dist)) culls
    frusta wholly outside one plane (the foveated path's tiles)."""
    from hvvr_amd import DTYPES
    frusta = np.ascontiguousarray(frusta)
    m = np.ascontiguousarray(m, np.float32).reshape(16)
    world = np.zeros(len(frusta), DTYPES["sfrustum"])
    pre = np.zeros(len(frusta), O_FRUSTUM)
    planes = None if cull_planes is None else np.ascontiguousarray(cull_planes, np.float32).reshape(16)
    lib().or_frustum_setup_cull(_p(frusta), c_uint32(len(frusta)), _p(m), _p(planes), _p(world), _p(pre))
    return world, pre


def test_children(pre_one, node_one):
    return lib().or_test_children(_p(pre_one), _p(node_one))


def test_children_refine(pre_one, node_one, prev):
    return lib().or_test_children_refine(_p(pre_one), _p(node_one), c_uint32(prev))


def block_traverse(nodes, pre_one):
    out = np.zeros(64, O_STACK)
    n = lib().or_block_traverse(_p(nodes), _p(pre_one), _p(out))
    return out[:n]


def tile_traverse(nodes, stack, pre_one, cap=1 << 22):
    out = np.zeros(cap, np.uint32)
    stack = np.ascontiguousarray(stack, O_STACK)
    n = lib().or_tile_traverse(_p(nodes), _p(stack), c_uint32(len(stack)), _p(pre_one), _p(out), c_uint64(cap))
    if n < 0:
        raise RuntimeError("oracle tile traversal overflowed")
    return out[:n]


def build_tile_lists(nodes, block_pre, tile_pre):
    nodes = np.ascontiguousarray(nodes)
    counts = np.zeros(len(tile_pre), np.uint32)
    total = lib().or_build_tile_lists(_p(nodes), _p(block_pre), _p(tile_pre), c_uint32(len(block_pre)),
                                      _p(counts), None, c_uint64(0))
    if total < 0:
        raise RuntimeError("oracle traversal overflowed")
    lists = np.zeros(max(total, 1), np.uint32)
    total2 = lib().or_build_tile_lists(_p(nodes), _p(block_pre), _p(tile_pre), c_uint32(len(block_pre)),
                                       _p(counts), _p(lists), c_uint64(len(lists)))
    assert total2 == total
    return counts, lists[:total]


class _Dof(ctypes.Structure):
    _fields_ = [("radius", ctypes.c_float), ("focalDistance", ctypes.c_float), ("lensTable", c_void_p)]


def _dof(dof):
    """dof = (radius, focal_distance, lens_table) or None"""
    if dof is None:
        return None, None
    table = np.ascontiguousarray(dof[2], np.float32)
    return _Dof(float(dof[0]), float(dof[1]), table.ctypes.data_as(c_void_p)), table


def intersect_view(beams, world_tiles, m, offsets, A, counts, lists, tris, dof=None):
    beams = np.ascontiguousarray(beams)
    m = np.ascontiguousarray(m, np.float32).reshape(16)
    offsets = np.ascontiguousarray(offsets, np.float32)
    tri_out = np.zeros(len(beams) * A, np.uint32)
    depth = np.zeros(len(beams) * A, np.float32)
    surv = np.zeros(1, np.uint64)
    lists = np.ascontiguousarray(lists, np.uint32)
    d, keep = _dof(dof)
    rc = lib().or_intersect_view_dof(_p(beams), c_uint32(len(beams)), _p(world_tiles), c_uint32(len(world_tiles)),
                                     _p(m), _p(offsets), c_uint32(A), _p(np.ascontiguousarray(counts, np.uint32)),
                                     _p(lists if lists.size else np.zeros(1, np.uint32)), _p(tris),
                                     ctypes.byref(d) if d else None, _p(tri_out), _p(depth), _p(surv))
    if rc:
        raise RuntimeError("oracle intersect failed")
    return tri_out.reshape(-1, A), depth.reshape(-1, A), int(surv[0])


def dof_rays(beams, m, offsets, A, dof, counts):
    """World-space depth-of-field subsample rays (origins, directions), [samples*A, 3] each."""
    beams = np.ascontiguousarray(beams)
    m = np.ascontiguousarray(m, np.float32).reshape(16)
    offsets = np.ascontiguousarray(offsets, np.float32)
    o = np.zeros((len(beams) * A, 3), np.float32)
    d = np.zeros((len(beams) * A, 3), np.float32)
    dd, keep = _dof(dof)
    rc = lib().or_dof_rays(_p(beams), c_uint32(len(beams)), _p(m), _p(offsets), c_uint32(A), ctypes.byref(dd),
                           _p(np.ascontiguousarray(counts, np.uint32)), _p(o), _p(d))
    if rc:
        raise RuntimeError("oracle dof rays failed")
    return o, d


def emit_gbuffer(sample_tris):
    A = len(sample_tris)
    st = np.ascontiguousarray(sample_tris, np.uint32)
    t = np.zeros(A, np.uint32)
    m = np.zeros(A, np.uint32)
    n = lib().or_emit_gbuffer(_p(st), c_uint32(A), _p(t), _p(m))
    return t[:n], m[:n]


def gbuffer_slots(tri_ids):
    """Logical GBuffer of every sample: (tri[s, A], mask[s, A]) padded with (0xffffffff, 0)."""
    S, A = tri_ids.shape
    tri = np.full((S, A), 0xFFFFFFFF, np.uint32)
    mask = np.zeros((S, A), np.uint32)
    for s in range(S):
        t, m = emit_gbuffer(tri_ids[s])
        tri[s, :len(t)] = t
        mask[s, :len(m)] = m
    return tri, mask


def lighting_env(lights):
    env = np.zeros(1, O_LIGHTENV)
    lights = np.ascontiguousarray(lights)
    dropped = lib().or_lighting_env(_p(lights), c_uint32(len(lights)), _p(env))
    return env, dropped


def resolve_view(beams, m, offsets, A, tri_ids, tris, shade, verts, materials, env, precision=0, dof=None,
                 counts=None, textures=None):
    """Resolve of every sample; `textures` = [(fmt, data), ...] the materials' texture
    indices name (None: untextured)."""
    if textures:
        set_textures(textures)
        try:
            return resolve_view(beams, m, offsets, A, tri_ids, tris, shade, verts, materials, env, precision, dof,
                                counts)
        finally:
            set_textures(None)
    beams = np.ascontiguousarray(beams)
    m = np.ascontiguousarray(m, np.float32).reshape(16)
    offsets = np.ascontiguousarray(offsets, np.float32)
    colors = np.zeros(len(beams), np.uint32)
    linear = np.zeros((len(beams), 3), np.float32)
    tri_ids = np.ascontiguousarray(tri_ids, np.uint32)
    d, keep = _dof(dof)
    cnt = None if counts is None else np.ascontiguousarray(counts, np.uint32)
    rc = lib().or_resolve_view_dof(_p(beams), c_uint32(len(beams)), _p(m), _p(offsets), c_uint32(A), _p(tri_ids),
                                   _p(tris), _p(shade), _p(verts), _p(materials), _p(env), c_int(precision),
                                   ctypes.byref(d) if d else None, _p(cnt), _p(colors), _p(linear))
    if rc:
        raise RuntimeError("oracle resolve failed")
    return colors, linear


def remap(results, remap_table, width, height, split=1):
    img = np.zeros((height, width), np.uint32)
    lib().or_remap(_p(np.ascontiguousarray(results, np.uint32)), _p(np.ascontiguousarray(remap_table, np.int32)),
                   c_uint32(width), c_uint32(height), c_uint32(split), c_uint32(width), _p(img))
    return img


def refit(scene):
    """World vertices, triangles and node bounds of a HostScene, from scratch."""
    from hvvr_amd import DTYPES
    verts = np.ascontiguousarray(scene.verts)
    wv = np.zeros_like(verts)
    tris = np.zeros(len(scene.triShade), DTYPES["tri_intersect"])
    nodes = np.zeros_like(scene.nodes)
    rc = lib().or_refit(_p(verts), c_uint32(len(verts)), _p(np.ascontiguousarray(scene.vertexToTransform)),
                        _p(np.ascontiguousarray(scene.transforms, np.float32)), _p(scene.triShade),
                        c_uint32(len(scene.triShade)), _p(scene.nodes), c_uint32(len(scene.nodes)), _p(wv), _p(tris),
                        _p(nodes))
    if rc:
        raise RuntimeError("oracle refit failed")
    return wv, tris, nodes


def render_beams(scene, beams, tile_frusta, block_frusta, A, m, lens=(0.0, 1.0), cull_planes=None, precision=0,
                 offsets=None, textures=None):
    """Oracle frame of a beam set (no remap): refit, frustum setup (tile frusta outside a
    cull plane skipped), traversal, intersect, resolve with empty tiles cleared, and the
    per-sample t-max of the foveated path. `offsets`: the frame's subsample offsets (the
    unspun spiral when None)."""
    m = np.asarray(m, np.float32).reshape(16)
    wv, tris, nodes = refit(scene)
    world_tiles, tile_pre = frustum_setup(tile_frusta, m, cull_planes)
    _, block_pre = frustum_setup(block_frusta, m)
    counts, lists = build_tile_lists(nodes, block_pre, tile_pre)
    from hvvr_amd import lens_table, subsample_offsets
    offsets = subsample_offsets(A) if offsets is None else np.asarray(offsets, np.float32)
    radius, focal = float(lens[0]), float(lens[1])
    dof = (radius, focal, lens_table(A)) if radius > 0 else None
    tri_ids, depth, surv = intersect_view(beams, world_tiles, m, offsets, A, counts, lists, tris, dof)
    env, _ = lighting_env(scene.lights)
    colors, linear = resolve_view(beams, m, offsets, A, tri_ids, tris, scene.triShade, wv, scene.materials, env,
                                  precision, dof, counts, textures)
    colors[np.repeat(counts == 0, 128)[: len(colors)]] = 0xFF000000
    tmax = tmax_view(beams, m, offsets, A, tri_ids, tris, counts, dof)
    return dict(nodes=nodes, tris=tris, counts=counts, lists=lists, tri_ids=tri_ids, depth=depth, survivors=surv,
                colors=colors, linear=linear, tmax=tmax, offsets=offsets)


def render_view(scene, view, A, camera_to_world=None, precision=0, with_colors=True, textures=None):
    """Full oracle frame: refit, frustum setup, traversal, intersect, resolve, remap."""
    m = view.camera_to_world if camera_to_world is None else np.asarray(camera_to_world, np.float32)
    wv, tris, nodes = refit(scene)
    world_tiles, tile_pre = frustum_setup(view.tile_frusta, m)
    _, block_pre = frustum_setup(view.block_frusta, m)
    counts, lists = build_tile_lists(nodes, block_pre, tile_pre)
    from hvvr_amd import lens_table, subsample_offsets
    offsets = subsample_offsets(A)
    radius, focal = float(view.lens[0]), float(view.lens[1])
    dof = (radius, focal, lens_table(A)) if radius > 0 else None
    tri_ids, depth, surv = intersect_view(view.beams, world_tiles, m, offsets, A, counts, lists, tris, dof)
    out = dict(nodes=nodes, tris=tris, world_verts=wv, counts=counts, lists=lists, tri_ids=tri_ids, depth=depth,
               survivors=surv, offsets=offsets)
    if with_colors:
        env, _ = lighting_env(scene.lights)
        colors, linear = resolve_view(view.beams, m, offsets, A, tri_ids, tris, scene.triShade, wv, scene.materials,
                                      env, precision, dof, counts, textures)
        # empty tiles: clear colour (ClearEmptyKernel, resolve.cu:527-542)
        empty = np.repeat(counts == 0, 128)
        colors[empty[: len(colors)]] = 0xFF000000
        out["colors"] = colors
        out["linear"] = linear
        out["image"] = remap(colors, view.remap, view.width, view.height, view.split)
    return out


# ---------------------------------------------------------------------------------------
# foveated path
# ---------------------------------------------------------------------------------------
def tmax_view(beams, m, offsets, A, tri_ids, tris, counts, dof=None):
    beams = np.ascontiguousarray(beams)
    m = np.ascontiguousarray(m, np.float32).reshape(16)
    out = np.zeros(len(beams), np.float32)
    d, keep = _dof(dof)
    rc = lib().or_tmax_view(_p(beams), c_uint32(len(beams)), _p(m), _p(np.ascontiguousarray(offsets, np.float32)),
                            c_uint32(A), _p(np.ascontiguousarray(tri_ids, np.uint32)), _p(tris),
                            ctypes.byref(d) if d else None, _p(np.ascontiguousarray(counts, np.uint32)), _p(out))
    assert rc == 0
    return out


def hits_view(beams, m, offsets, A, tri_ids, tris, counts, dof=None):
    """Barycentrics (u, v, w) of every subsample's hit on its own ray: [samples, A, 3]."""
    beams = np.ascontiguousarray(beams)
    m = np.ascontiguousarray(m, np.float32).reshape(16)
    out = np.zeros((len(beams), A, 3), np.float32)
    d, keep = _dof(dof)
    rc = lib().or_hits_view(_p(beams), c_uint32(len(beams)), _p(m), _p(np.ascontiguousarray(offsets, np.float32)),
                            c_uint32(A), _p(np.ascontiguousarray(tri_ids, np.uint32)), _p(tris),
                            ctypes.byref(d) if d else None, _p(np.ascontiguousarray(counts, np.uint32)), _p(out))
    assert rc == 0
    return out


def srgb_tables():
    lin = np.zeros(256, np.float32)
    rt = np.zeros(256, np.uint8)
    lib().or_srgb_tables(_p(lin), _p(rt))
    return lin, rt


def polar_remap(results, tmax, polar_remap_xy, spr, rings):
    color = np.zeros(spr * rings, np.uint32)
    depth = np.zeros(spr * rings, np.float32)
    xy = np.ascontiguousarray(polar_remap_xy, np.uint32).reshape(-1)
    lib().or_polar_remap(_p(np.ascontiguousarray(results, np.uint32)), _p(np.ascontiguousarray(tmax, np.float32)),
                         _p(xy), c_uint32(len(xy) // 2), c_uint32(spr), _p(color), _p(depth))
    return color.reshape(rings, spr), depth.reshape(rings, spr)


def polar_moments(polar_color):
    pc = np.ascontiguousarray(polar_color, np.uint32)
    h, w = pc.shape
    m1 = np.zeros((h, w), np.uint64)
    m2 = np.zeros((h, w), np.uint64)
    lib().or_polar_moments(_p(pc), c_uint32(w), c_uint32(h), _p(m1), _p(m2))
    return m1, m2


class _Reproject(ctypes.Structure):
    _fields_ = [("polarColor", c_void_p), ("polarDepth", c_void_p), ("moment1", c_void_p), ("moment2", c_void_p),
                ("previous", c_void_p), ("polarW", c_uint32), ("polarH", c_uint32), ("width", c_uint32),
                ("height", c_uint32), ("sampleToEye", ctypes.c_float * 9), ("eyeToSamplePrevious", ctypes.c_float * 16),
                ("emap", ctypes.c_float * 6), ("maxEccentricity", ctypes.c_float), ("alpha", ctypes.c_float),
                ("stddevMultiplier", ctypes.c_float)]


def foveated_reproject(polar_color, polar_depth, m1, m2, previous, width, height, sample_to_eye,
                       eye_to_sample_previous, emap, max_ecc, alpha=0.05, stddev=1.0):
    arrays = [np.ascontiguousarray(polar_color, np.uint32), np.ascontiguousarray(polar_depth, np.float32),
              np.ascontiguousarray(m1, np.uint64), np.ascontiguousarray(m2, np.uint64),
              np.ascontiguousarray(previous, np.uint32)]
    h, w = arrays[0].shape
    a = _Reproject(_p(arrays[0]), _p(arrays[1]), _p(arrays[2]), _p(arrays[3]), _p(arrays[4]), w, h, width, height,
                   (ctypes.c_float * 9)(*np.asarray(sample_to_eye, np.float32).reshape(9)),
                   (ctypes.c_float * 16)(*np.asarray(eye_to_sample_previous, np.float32).reshape(16)),
                   (ctypes.c_float * 6)(*np.asarray(emap, np.float32).reshape(6)), max_ecc, alpha, stddev)
    out = np.zeros((height, width), np.uint32)
    lib().or_foveated_reproject(ctypes.byref(a), _p(out))
    return out


def foveated_blur(img, step, sample_to_eye):
    img = np.ascontiguousarray(img, np.uint32)
    h, w = img.shape
    out = np.zeros_like(img)
    s2e = np.ascontiguousarray(sample_to_eye, np.float32).reshape(9)
    lib().or_foveated_blur(_p(img), _p(out), c_uint32(w), c_uint32(h), c_int(step[0]), c_int(step[1]), _p(s2e))
    return out


def foveated_contrast(unfiltered, filtered, sample_to_eye, max_ecc, fe=1.0):
    u = np.ascontiguousarray(unfiltered, np.uint32)
    f = np.ascontiguousarray(filtered, np.uint32)
    h, w = u.shape
    out = np.zeros_like(u)
    s2e = np.ascontiguousarray(sample_to_eye, np.float32).reshape(9)
    lib().or_foveated_contrast(_p(u), _p(f), _p(out), c_uint32(w), c_uint32(h), _p(s2e), ctypes.c_float(max_ecc),
                               ctypes.c_float(fe))
    return out


def emap_apply(emap, i):
    L = lib()
    L.or_emap_apply.restype = ctypes.c_float
    e = (ctypes.c_float * 6)(*np.asarray(emap, np.float32))
    return L.or_emap_apply(e, ctypes.c_float(i))


def emap_apply_inverse(emap, e_rad):
    L = lib()
    L.or_emap_apply_inverse.restype = ctypes.c_float
    e = (ctypes.c_float * 6)(*np.asarray(emap, np.float32))
    return L.or_emap_apply_inverse(e, ctypes.c_float(e_rad))


# ---------------------------------------------------------------------------------------
# CPU baseline (oracle/_build/libcpu_baseline.so): the AVX2 traversal on a thread pool
# ---------------------------------------------------------------------------------------
BASELINE_LIB = os.path.join(_HERE, "_build", "libcpu_baseline.so")
_baseline = None


class CbView(ctypes.Structure):
    _fields_ = [("beams", c_void_p), ("tileFrusta", c_void_p), ("blockFrusta", c_void_p), ("blockCount", c_uint32),
                ("validSampleCount", c_uint32), ("cameraToWorld", ctypes.c_float * 16)]


class CbStats(ctypes.Structure):
    _fields_ = [("seconds", ctypes.c_double), ("subsamples", c_uint64), ("blocks", c_uint64),
                ("frustumSeconds", ctypes.c_double), ("traverseSeconds", ctypes.c_double),
                ("intersectSeconds", ctypes.c_double), ("resolveSeconds", ctypes.c_double), ("threads", c_uint32),
                ("hardwareConcurrency", c_uint32), ("status", ctypes.c_int32)]


def baseline_lib():
    global _baseline
    if _baseline is None:
        lib()  # liboracle.so first (libcpu_baseline.so links it)
        if not os.path.exists(BASELINE_LIB):
            build()
        L = ctypes.CDLL(BASELINE_LIB)
        L.cb_build_tile_lists.restype = ctypes.c_int64
        _baseline = L
    return _baseline


def avx_build_tile_lists(nodes, block_pre, tile_pre):
    """build_tile_lists through the AVX2 traversal of the CPU baseline."""
    nodes = np.ascontiguousarray(nodes)
    counts = np.zeros(len(tile_pre), np.uint32)
    L = baseline_lib()
    total = L.cb_build_tile_lists(_p(nodes), _p(block_pre), _p(tile_pre), c_uint32(len(block_pre)), _p(counts), None,
                                  c_uint64(0))
    if total < 0:
        raise RuntimeError("AVX2 traversal overflowed")
    lists = np.zeros(max(total, 1), np.uint32)
    total2 = L.cb_build_tile_lists(_p(nodes), _p(block_pre), _p(tile_pre), c_uint32(len(block_pre)), _p(counts),
                                   _p(lists), c_uint64(len(lists)))
    assert total2 == total
    return counts, lists[:total]


def cpu_baseline_run(scene, views, A, threads=0, budget_s=20.0, seed=0):
    """Times the CPU baseline over the views' blocks (tasks of 2 blocks, seeded order,
    repeated) for budget_s seconds on `threads` threads (0 = hardware_concurrency).
    Returns the CbStats."""
    from hvvr_amd import subsample_offsets
    wv, tris, nodes = refit(scene)
    env, _ = lighting_env(scene.lights)
    offsets = np.ascontiguousarray(subsample_offsets(A), np.float32)
    keep = [wv, tris, nodes, env, offsets]
    cviews = (CbView * len(views))()
    for cv, v in zip(cviews, views):
        arrs = [np.ascontiguousarray(a) for a in (v.beams, v.tile_frusta, v.block_frusta)]
        keep += arrs
        cv.beams, cv.tileFrusta, cv.blockFrusta = (a.ctypes.data for a in arrs)
        cv.blockCount, cv.validSampleCount = v.block_count, v.valid_sample_count
        cv.cameraToWorld[:] = [float(x) for x in np.asarray(v.camera_to_world, np.float32).reshape(16)]
    shade = np.ascontiguousarray(scene.triShade)
    mats = np.ascontiguousarray(scene.materials)
    keep += [shade, mats]
    st = CbStats()
    rc = baseline_lib().cb_run(_p(nodes), _p(tris), _p(shade), _p(wv), _p(mats), _p(env), cviews,
                               c_uint32(len(views)), _p(offsets), c_uint32(A), c_uint32(threads),
                               ctypes.c_double(budget_s), c_uint32(seed), ctypes.byref(st))
    if rc != 0:
        raise RuntimeError("CPU baseline: a tile traversal overflowed its stack")
    return st


# ---- textures (or_texture.c) ----
O_TEXINFO = np.dtype([("width", "<u4"), ("height", "<u4"), ("levels", "<u4"), ("kind", "<u4"),
                      ("offset", "<u8", 16)])
assert O_TEXINFO.itemsize == 144
# TextureFormat (include/hvvr/raycaster.h): the natural numpy layout of each format's texels
TEX_FORMATS = {"srgb8": 1, "unorm8": 2, "unorm16": 3, "float4": 4, "half4": 5, "r11g11b10": 6, "r32": 7}


def _tex_lib():
    L = lib()
    if not getattr(L, "_tex_ready", False):
        L.or_texture_sample.argtypes = [c_void_p, c_void_p] + [ctypes.c_float] * 6 + [c_void_p]
        L.or_texture_lod.argtypes = [ctypes.c_float, c_uint32, c_void_p, c_void_p]
        L.or_texture_layout.restype = c_uint64
        L.or_texture_levels.restype = c_uint32
        L._tex_ready = True
    return L


def _small_float(bits, mbits):
    """unsigned float of 5 exponent and `mbits` mantissa bits (r11g11b10), as float64"""
    e = (bits >> mbits).astype(np.int64)
    m = (bits & ((1 << mbits) - 1)).astype(np.float64)
    out = np.where(e == 0, m / (1 << mbits) * 2.0 ** -14, (1.0 + m / (1 << mbits)) * np.exp2(e - 15.0))
    return np.where(e == 31, np.where(m > 0, np.nan, np.inf), out)


def texture_level0(fmt, data):
    """(kind, level 0 in the stored kind): RGBA8 bytes for srgb8 / unorm8 (kind 1 / 0),
    RGBA32F (kind 2) for the rest; `data` is [h, w, 4] (r11g11b10 and r32: [h, w])."""
    data = np.asarray(data)
    if fmt in ("srgb8", "unorm8"):
        return (1 if fmt == "srgb8" else 0), np.ascontiguousarray(data, np.uint8)
    h, w = data.shape[:2]
    out = np.zeros((h, w, 4), np.float32)
    if fmt == "unorm16":
        out[...] = data.astype(np.float32) / np.float32(65535.0)
    elif fmt == "float4":
        out[...] = data
    elif fmt == "half4":
        out[...] = data.astype(np.float16).astype(np.float32)
    elif fmt == "r11g11b10":
        c = data.astype(np.uint32)
        out[..., 0] = _small_float(c & 0x7FF, 6)
        out[..., 1] = _small_float((c >> 11) & 0x7FF, 6)
        out[..., 2] = _small_float(c >> 22, 5)
        out[..., 3] = 1.0
    elif fmt == "r32":
        out[..., 0] = data
        out[..., 3] = 1.0
    else:
        raise ValueError(fmt)
    return 2, out


def texture_chain(fmt, data):
    """(info, chain bytes) of one texture: level 0 converted, mips built (d_mipmap)."""
    L = _tex_lib()
    kind, level0 = texture_level0(fmt, data)
    info = np.zeros(1, O_TEXINFO)
    info["width"], info["height"] = level0.shape[1], level0.shape[0]
    info["levels"] = L.or_texture_levels(c_uint32(level0.shape[1]), c_uint32(level0.shape[0]))
    info["kind"] = kind
    nbytes = L.or_texture_layout(_p(info))
    chain = np.zeros(nbytes, np.uint8)
    raw = level0.reshape(-1).view(np.uint8)
    chain[: raw.size] = raw
    L.or_texture_build_mips(_p(chain), _p(info))
    return info, chain


def texture_sample(info, chain, uv, ddx=(0.0, 0.0), ddy=(0.0, 0.0)):
    out = np.zeros(4, np.float32)
    _tex_lib().or_texture_sample(_p(chain), _p(info), float(uv[0]), float(uv[1]), float(ddx[0]), float(ddx[1]),
                                 float(ddy[0]), float(ddy[1]), _p(out))
    return out


def texture_lod(x, levels):
    lvl = np.zeros(1, np.uint32)
    f = np.zeros(1, np.float32)
    _tex_lib().or_texture_lod(float(x), c_uint32(levels), _p(lvl), _p(f))
    return int(lvl[0]), float(f[0])


def texture_tables():
    out = np.zeros(512, np.float32)
    lib().or_texture_tables(_p(out))
    return out


_tex_keep = None


def set_textures(textures):
    """Make [(fmt, data), ...] the textures later resolves read (None / [] = none); an
    entry may also be a built (info, chain) pair of texture_chain. Returns [(info, chain)]
    per texture."""
    global _tex_keep
    L = _tex_lib()
    if not textures:
        L.or_set_textures(None, None, c_uint32(0))
        _tex_keep = None
        return []
    chains = [(f, d) if isinstance(f, np.ndarray) else texture_chain(f, d) for f, d in textures]
    base, total = [], 0
    for _, c in chains:
        base.append(total)
        total += (c.size + 255) & ~255
    buf = np.zeros(max(total, 1), np.uint8)
    info = np.zeros(len(chains), O_TEXINFO)
    for i, (inf, c) in enumerate(chains):
        buf[base[i]: base[i] + c.size] = c
        info[i] = inf[0]
        info["offset"][i, : int(inf[0]["levels"])] += np.uint64(base[i])
    L.or_set_textures(_p(buf), _p(info), c_uint32(len(chains)))
    _tex_keep = (buf, info)
    return chains
