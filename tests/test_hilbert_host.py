"""Pins the host restatement of the sort key as a Hilbert curve: every
aligned block of 2^b cells per axis is one contiguous run of keys (the
curve's recursive structure), and consecutive keys within it are
face-adjacent cells (the curve is continuous).  A Z-order (Morton) key, a
wrong interleave order or a missing Gray step fails both."""
import itertools

import numpy as np
import pytest

from hilbert_host import hilbert_keys


def _block(dim, b, origin):
    g = np.array(list(itertools.product(range(1 << b), repeat=dim)), dtype=np.uint64)
    return g + np.asarray(origin, dtype=np.uint64)


@pytest.mark.parametrize("dim", [2, 3])
def test_aligned_blocks_are_contiguous_adjacent_runs(dim):
    rng = np.random.default_rng(dim)
    for b in (1, 2, 3):
        origins = [np.zeros(dim, np.uint64)] + [
            (rng.integers(0, 1 << (21 - b), dim).astype(np.uint64) << np.uint64(b)) for _ in range(6)]
        for o in origins:
            cells = _block(dim, b, o)
            k = hilbert_keys(cells, dim)
            order = np.argsort(k)
            ks = k[order]
            # one contiguous run of 2^(dim*b) keys, aligned to its length
            assert np.array_equal(ks - ks[0], np.arange(len(ks), dtype=np.uint64))
            assert int(ks[0]) % (1 << (dim * b)) == 0
            # consecutive cells differ by one step along one axis
            c = cells[order].astype(np.int64)
            step = np.abs(np.diff(c, axis=0)).sum(axis=1)
            assert (step == 1).all()


def test_curve_starts_at_origin_and_is_a_bijection_on_small_grids():
    for dim in (2, 3):
        cells = _block(dim, 2, np.zeros(dim))
        k = hilbert_keys(cells, dim)
        assert len(np.unique(k)) == len(k)
        assert k[np.all(cells == 0, axis=1)][0] == 0


def test_morton_order_fails_the_adjacency_check():
    # the check above is not vacuous: a Z-order of the same block jumps
    cells = _block(3, 2, np.zeros(3))
    z = np.zeros(len(cells), np.uint64)
    for b in range(2):
        for i in range(3):
            z |= ((cells[:, i] >> np.uint64(b)) & np.uint64(1)) << np.uint64(3 * b + i)
    c = cells[np.argsort(z)].astype(np.int64)
    assert (np.abs(np.diff(c, axis=0)).sum(axis=1) > 1).any()


@pytest.fixture(scope="module")
def sfc(tmp_path_factory):
    import ctypes
    import os
    import subprocess

    here = os.path.dirname(os.path.abspath(__file__))
    out = str(tmp_path_factory.mktemp("sfc") / "libsfc.so")
    subprocess.run(["g++", "-O1", "-std=c++17", "-shared", "-fPIC", "-o", out,
                    os.path.join(here, "host", "sfc_harness.cpp")], check=True)
    return ctypes.CDLL(out)


@pytest.mark.parametrize("dim", [1, 2, 3])
def test_device_key_header_matches_host_restatement(sfc, dim):
    import ctypes

    rng = np.random.default_rng(10 + dim)
    cells = np.zeros((20000, 3), np.uint32)
    cells[:, :dim] = rng.integers(0, 1 << 21, (20000, dim), dtype=np.uint32)
    cells[:50, :dim] = (1 << 21) - 1 - rng.integers(0, 3, (50, dim))  # the far corner
    out = np.zeros(len(cells), np.uint64)
    sfc.keys(cells.ctypes.data_as(ctypes.c_void_p), ctypes.c_long(len(cells)), ctypes.c_int(dim),
             out.ctypes.data_as(ctypes.c_void_p))
    assert np.array_equal(out, hilbert_keys(cells[:, :dim].astype(np.uint64), dim))
    assert int(out.max()) < (1 << (21 * dim))


def test_device_cell_quantisation_matches_host(sfc):
    import ctypes

    from hilbert_host import frame_cells

    v = np.array([-1.0, 0.0, 1e-30, 0.5, 9.999, 10.0, 11.0, np.nan, -np.inf, np.inf], np.float32)
    out = np.zeros(len(v), np.uint32)
    s = np.float32(2097151.0) / np.float32(10.0)
    sfc.cells_of(v.ctypes.data_as(ctypes.c_void_p), ctypes.c_long(len(v)), ctypes.c_float(0.0),
                 ctypes.c_float(s), out.ctypes.data_as(ctypes.c_void_p))
    with np.errstate(invalid="ignore", over="ignore"):
        ref = frame_cells(np.nan_to_num(v, nan=-1.0)[:, None], [0, 0, 0, 10, 10, 10], 1)[:, 0]
    assert np.array_equal(out.astype(np.uint64), ref)


def test_sorted_key_bits_match_host_table(sfc):
    """The bits the index leaves unsorted (header) are the ones the GPU tests
    drop before checking the order."""
    from hilbert_host import SORT_LO_BIT

    for dim in (1, 2, 3):
        assert sfc.sort_lo_bit(dim) == SORT_LO_BIT[dim]
        assert SORT_LO_BIT[dim] % 8 == 0 and SORT_LO_BIT[dim] < 21 * dim
