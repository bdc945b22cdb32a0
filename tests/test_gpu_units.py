"""GPU unit probes of wave-level building blocks, against Python models.

The VGPR-resident traversal stack (hvvr_amd/csrc/gpu/reg_stack.h): random sequences of
sorted inserts and pops that grow it past 64, 128, 192 entries and to its 256-entry
capacity; every popped record and the final contents must match a list model of the
reference's insertion rule (traverse_avx.h:499-520: a new entry goes below the run of
top entries whose key is smaller)."""

import ctypes
import os

import numpy as np
import pytest

from conftest import ROOT

LIB = os.path.join(ROOT, "tests", "gpu_units", "_build", "libhvvr_gpu_units.so")


def _model(ops, keys):
    stack = []  # index 0 = top
    popped = np.full(len(ops), 0xFFFFFFFF, np.uint32)
    for i, (op, k) in enumerate(zip(ops, keys)):
        if op == 0:
            if stack:
                popped[i] = stack.pop(0)[1]
        elif len(stack) < 256:
            r = 0
            while r < len(stack) and stack[r][0] < k:
                r += 1
            stack.insert(r, (k, i))
    return popped, stack


def _sequence(rng, n, push_bias, key_pool):
    ops = (rng.random(n) < push_bias).astype(np.uint32)
    keys = rng.choice(key_pool, n).astype(np.float32)
    return ops, keys


def test_probe_library_built():
    assert os.path.exists(LIB), f"{LIB} missing: run __graft_entry__.build()"


@pytest.mark.gpu
@pytest.mark.parametrize("case", ["small", "past64", "past128", "full", "ties"])
def test_register_stack_matches_model(case):
    lib = ctypes.CDLL(LIB)
    rng = np.random.default_rng(["small", "past64", "past128", "full", "ties"].index(case))
    n, bias = {"small": (300, 0.55), "past64": (600, 0.62), "past128": (900, 0.66), "full": (1500, 0.75),
               "ties": (1200, 0.7)}[case]
    pool = np.arange(8, dtype=np.float32) if case == "ties" else rng.normal(size=4096).astype(np.float32)
    ops, keys = _sequence(rng, n, bias, pool)
    popped = np.zeros(n, np.uint32)
    ft = np.zeros(256, np.float32)
    fs = np.zeros(256, np.uint32)
    size = np.zeros(1, np.uint32)
    p = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    rc = lib.reg_stack_probe_run(p(ops), p(keys), ctypes.c_uint32(n), p(popped), p(ft), p(fs), p(size))
    assert rc == 0
    want_popped, want_stack = _model(ops, keys)
    assert int(size[0]) == len(want_stack)
    pops = ops == 0
    assert np.array_equal(popped[pops], want_popped[pops]), "popped records differ"
    got = [(float(ft[d]), int(fs[d])) for d in range(len(want_stack))]
    assert got == [(float(k), r) for k, r in want_stack], "final stack differs"
