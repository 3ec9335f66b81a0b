"""Host (g++) test of the device linear algebra header: the in-register LDL^T
solve of the least-squares moment system against numpy.linalg.solve, and the
polynomial basis ordering of polyfit_leastsq_matrices (derivatives.py:83-92)."""
import ctypes
import os
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def harness(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("lin") / "libharness.so")
    subprocess.run(["g++", "-O1", "-std=c++17", "-shared", "-fPIC", "-o", out,
                    os.path.join(HERE, "host", "linalg_harness.cpp")], check=True)
    L = ctypes.CDLL(out)
    d = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
    L.fit.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, d, d, d, d]
    return L


def basis_np(dx, order):
    dim = dx.shape[1]
    cols = [dx[:, i] for i in range(dim)]
    if order == 2:
        cols += [0.5 * dx[:, i] ** 2 for i in range(dim)]
        nc = {1: 0, 2: 1, 3: 3}[dim]
        cols += [dx[:, c % dim] * dx[:, (c + 1) % dim] for c in range(nc)]
    return np.stack(cols, 1)


@pytest.mark.parametrize("dim,order", [(1, 1), (2, 1), (3, 1), (1, 2), (2, 2), (3, 2)])
def test_ldlt_fit_matches_numpy(harness, dim, order):
    rng = np.random.default_rng(dim * 7 + order)
    for _ in range(20):
        k = 32
        dx = rng.standard_normal((k, dim)) * 1e-3
        w = rng.random(k)
        df = rng.standard_normal(k)
        out = np.zeros(9)
        rc = harness.fit(dim, order, k, np.ascontiguousarray(dx), w, df, out)
        A = basis_np(dx, order)
        E = (A * w[:, None]).T @ A
        b = (A * w[:, None]).T @ df
        ref = np.linalg.solve(E, b)
        nd = A.shape[1]
        assert rc == 0
        np.testing.assert_allclose(out[:nd], ref, rtol=1e-8, atol=1e-8 * np.abs(ref).max())


def test_quadratic_exactness_3d(harness):
    rng = np.random.default_rng(0)
    dx = rng.standard_normal((40, 3)) * 0.01
    g = np.array([0.5, -2.0, 1.0])
    Hm = np.array([[2.0, 1.0, 0.0], [1.0, -1.0, 0.5], [0.0, 0.5, 3.0]])
    df = dx @ g + 0.5 * np.einsum("ji,ik,jk->j", dx, Hm, dx)
    out = np.zeros(9)
    assert harness.fit(3, 2, 40, np.ascontiguousarray(dx), np.ones(40), df, out) == 0
    np.testing.assert_allclose(out[:3], g, atol=1e-9)
    # [xx, yy, zz, xy, yz, zx]
    np.testing.assert_allclose(out[3:], [2.0, -1.0, 3.0, 1.0, 0.5, 0.0], atol=1e-7)


def test_singular_detected(harness):
    dx = np.zeros((10, 3))
    dx[:, 0] = np.arange(10)
    out = np.zeros(9)
    assert harness.fit(3, 1, 10, dx, np.ones(10), np.arange(10.0), out) == 1
