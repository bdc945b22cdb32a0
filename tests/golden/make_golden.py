"""Writes the golden fixtures of tests/golden/ from the CPU oracle.

The reference ships no golden vectors for its render path (no test data, no known-answer
tests), so these fixtures are the oracle's outputs on seeded scenes and cameras, kept
so that (a) any change to the oracle's results is caught on the CPU and (b) the GPU
path can be checked against files without running the oracle.

    python tests/golden/make_golden.py

Fixtures (numpy .npz, no pickles):
  heightfield_160x120_a4.npz  heightfield + boxes scene (seed 0), pinhole 160x120, MSAA 4
  room_270x300_a4.npz         room scene (seed 0, 1,024 meshes), eye 0 at 270x300, MSAA 4
Each holds the tile list lengths, the SHA-256 of the concatenated lists (uint32 little
endian, tile order), the per-subsample triangle ids, the per-sample RGBA8 colours and
the remapped image.
"""

import hashlib
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [ROOT, os.path.join(ROOT, "oracle")]

CASES = {
    "heightfield_160x120_a4": dict(scene=("heightfield", 0, 0), view=("heightfield", 0, 160, 120), A=4),
    "room_270x300_a4": dict(scene=("room", 0, 1024), view=("room", 0, 270, 300), A=4),
}


def build_case(H, name):
    c = CASES[name]
    kind, seed, meshes = c["scene"]
    scene = H.HostScene(H.SCENE_HEIGHTFIELD_BOXES if kind == "heightfield" else H.SCENE_ROOM, seed, meshes)
    vkind, eye, w, h = c["view"]
    view = H.HostView(H.VIEW_PRESET_HEIGHTFIELD if vkind == "heightfield" else H.VIEW_PRESET_ROOM_EYE, eye, w, h)
    return scene, view, c["A"]


def compute(O, scene, view, A):
    ref = O.render_view(scene, view, A)
    lists = np.ascontiguousarray(ref["lists"], dtype="<u4")
    digest = np.frombuffer(hashlib.sha256(lists.tobytes()).digest(), np.uint8)
    return dict(counts=ref["counts"].astype(np.uint32), lists_sha256=digest,
                tri_ids=ref["tri_ids"].astype(np.uint32), colors=ref["colors"].astype(np.uint32),
                image=ref["image"].astype(np.uint32))


def main():
    import hvvr_amd as H
    import pyoracle as O

    for name in CASES:
        scene, view, A = build_case(H, name)
        out = compute(O, scene, view, A)
        path = os.path.join(HERE, name + ".npz")
        np.savez_compressed(path, **out)
        print(f"{path}: {os.path.getsize(path)} bytes, {int(out['counts'].sum())} list entries")


if __name__ == "__main__":
    main()


def lists_digest(lists):
    return np.frombuffer(hashlib.sha256(np.ascontiguousarray(lists, dtype="<u4").tobytes()).digest(), np.uint8)
