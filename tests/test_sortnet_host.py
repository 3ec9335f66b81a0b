"""Host (g++) test of the kNN's in-register sorting networks: the odd-even
merge sort of 32 values against numpy's sort (random values, ties, +inf
padding, the 0-1 principle over every 0/1 input of a 16-wide slice), and the
merge of a sorted 32 into a sorted 33-slot list against the 33 smallest of
the union."""
import ctypes
import os
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def net(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("sortnet") / "libsortnet.so")
    subprocess.run(["g++", "-O1", "-std=c++17", "-shared", "-fPIC", "-o", out,
                    os.path.join(HERE, "host", "sortnet_harness.cpp")], check=True)
    L = ctypes.CDLL(out)
    f = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
    L.sort32.argtypes = [f]
    L.merge33.argtypes = [f, f]
    return L


def test_sort32_matches_numpy(net):
    rng = np.random.default_rng(0)
    assert net.comparators32() == 191
    for trial in range(3000):
        v = rng.random(32).astype(np.float32)
        if trial % 3 == 1:
            v = np.round(v * 4).astype(np.float32)  # many ties
        if trial % 5 == 2:
            v[rng.random(32) < 0.3] = np.inf  # padding slots
        w = v.copy()
        net.sort32(w)
        assert np.array_equal(w, np.sort(v)), trial


def test_sort32_zero_one_principle(net):
    # a comparator network sorts every input iff it sorts every 0/1 input;
    # all 2^16 patterns in the upper half with the lower half random 0/1
    rng = np.random.default_rng(1)
    for bits in range(1 << 16):
        v = np.zeros(32, np.float32)
        v[16:] = [(bits >> b) & 1 for b in range(16)]
        v[:16] = rng.integers(0, 2, 16)
        w = v.copy()
        net.sort32(w)
        assert (np.diff(w) >= 0).all()


def test_merge33_keeps_the_33_smallest(net):
    rng = np.random.default_rng(2)
    for trial in range(3000):
        A = np.sort(rng.random(33)).astype(np.float32)
        S = np.sort(rng.random(32)).astype(np.float32)
        if trial % 4 == 1:
            A[rng.integers(10, 33):] = np.inf
        if trial % 4 == 2:
            S = np.round(S * 3).astype(np.float32)
            A = np.round(A * 3).astype(np.float32)
            A.sort()
            S.sort()
        ref = np.sort(np.concatenate([A, S]))[:33]
        a = A.copy()
        net.merge33(a, S)
        assert np.array_equal(a, ref), trial
