"""GPU unit probes of wave-level building blocks, against Python models.

The traversal stacks (hvvr_amd/csrc/gpu/reg_stack.h), one register (64 entries), four
eight and sixteen registers (256, 512, 1024) and LDS (2048 here): random sequences of sorted inserts and pops that keep
them shallow, past each register boundary and at capacity; every popped record and the
final contents must match a list model of the reference's insertion rule
(traverse_avx.h:499-520: a new entry goes below the run of top entries whose key is
smaller).

The wave prefix sum and sum (hvvr_amd/csrc/gpu/wave_traverse.h, DPP row shifts and row
broadcasts): against numpy's cumulative sum per 64 values, modulo 2^32."""

import ctypes
import os

import numpy as np
import pytest

from conftest import ROOT

LIB = os.path.join(ROOT, "tests", "gpu_units", "_build", "libhvvr_gpu_units.so")


def _model(ops, keys, cap):
    # stored bottom-up (the end of the lists is the top); the insertion rule reads top-down:
    # the new entry goes below the run of top entries whose key is smaller
    tk, tr = [], []
    popped = np.full(len(ops), 0xFFFFFFFF, np.uint32)
    for i, (op, k) in enumerate(zip(ops, keys)):
        if op == 0:
            if tk:
                tk.pop()
                popped[i] = tr.pop()
        elif len(tk) < cap:
            q = len(tk)
            while q > 0 and tk[q - 1] < k:
                q -= 1
            tk.insert(q, float(k))
            tr.insert(q, i)
    return popped, list(zip(tk[::-1], tr[::-1]))  # index 0 = top


def _sequence(rng, n, push_bias, key_pool):
    ops = (rng.random(n) < push_bias).astype(np.uint32)
    keys = rng.choice(key_pool, n).astype(np.float32)
    return ops, keys


def test_probe_library_built():
    assert os.path.exists(LIB), f"{LIB} missing: run __graft_entry__.build()"


CASES = {"small": (300, 0.55), "deep": (400, 0.6), "full": (1500, 0.75), "ties": (1200, 0.7)}


@pytest.mark.gpu
@pytest.mark.parametrize("cap", [64, 256, 512, 1024, 2048])
@pytest.mark.parametrize("case", list(CASES))
def test_register_stack_matches_model(case, cap):
    lib = ctypes.CDLL(LIB)
    rng = np.random.default_rng(list(CASES).index(case) + cap)
    n, bias = CASES[case]
    if cap in (256, 512, 1024):
        n *= cap // 85  # enough pushes to cross every register boundary
    elif cap == 2048:
        n *= 24  # the LDS stack: deep enough for many 64-entry chunks
    pool = np.arange(8, dtype=np.float32) if case == "ties" else rng.normal(size=4096).astype(np.float32)
    ops, keys = _sequence(rng, n, bias, pool)
    popped = np.zeros(n, np.uint32)
    ft = np.zeros(cap, np.float32)
    fs = np.zeros(cap, np.uint32)
    size = np.zeros(1, np.uint32)
    p = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    rc = lib.reg_stack_probe_run(ctypes.c_uint32(cap), p(ops), p(keys), ctypes.c_uint32(n), p(popped), p(ft), p(fs),
                                 p(size))
    assert rc == 0
    want_popped, want_stack = _model(ops, keys, cap)
    if case != "small":
        assert max(_depths(ops, cap)) > cap * 3 // 4
    assert int(size[0]) == len(want_stack)
    pops = ops == 0
    assert np.array_equal(popped[pops], want_popped[pops]), "popped records differ"
    # the stacks are bottom-up: the model's top (index 0) is position size-1
    top = len(want_stack) - 1
    got = [(float(ft[top - d]), int(fs[top - d])) for d in range(len(want_stack))]
    assert got == [(float(k), r) for k, r in want_stack], "final stack differs"


def _depths(ops, cap):
    d, out = 0, []
    for o in ops:
        d = max(d - 1, 0) if o == 0 else min(d + 1, cap)
        out.append(d)
    return out


@pytest.mark.gpu
def test_wave_scan_matches_cumsum():
    lib = ctypes.CDLL(LIB)
    rng = np.random.default_rng(11)
    groups = [rng.integers(0, 1000, 64), np.zeros(64, np.int64), np.ones(64, np.int64),
              rng.integers(0, 2**32, 64),  # sums wrap modulo 2^32
              np.where(np.arange(64) % 17 == 0, 5, 0), np.eye(64, dtype=np.int64)[63] * 7,
              np.eye(64, dtype=np.int64)[0] * 9] + [rng.integers(0, 64, 64) for _ in range(57)]
    x = np.concatenate(groups).astype(np.uint32)
    scan = np.zeros_like(x)
    total = np.zeros_like(x)
    p = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    assert lib.wave_scan_probe_run(p(x), ctypes.c_uint32(len(x)), p(scan), p(total)) == 0
    g = x.reshape(-1, 64).astype(np.uint64)
    want = (np.cumsum(g, axis=1) % 2**32).astype(np.uint32)
    assert np.array_equal(scan.reshape(-1, 64), want)
    assert np.array_equal(total.reshape(-1, 64), np.repeat(want[:, 63:], 64, axis=1))
