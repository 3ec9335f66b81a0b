"""Host (numpy) restatement of the index's sort key: the Hilbert index of a
particle's cell on the 2^21-per-axis grid of the bounding box's cubic frame
(J. Skilling, "Programming the Hilbert curve", AIP Conf. Proc. 707, 381
(2004): axes -> transposed index, then bit interleave)."""
import numpy as np

BITS = 21
# key bits the index leaves unsorted, per dimension (3D: the lowest 8-bit digit)
SORT_LO_BIT = {1: 0, 2: 0, 3: 8}


def hilbert_keys(cells, dim, bits=BITS):
    """cells: (n, dim) integer cell coordinates < 2^bits -> uint64 keys of
    dim*bits bits (dim 1: the coordinate)."""
    X = [np.asarray(cells[:, a], dtype=np.uint64).copy() for a in range(dim)]
    if dim == 1:
        return X[0]
    one = np.uint64(1)
    q = one << np.uint64(bits - 1)
    while q > one:
        p = q - one
        for i in range(dim):
            setb = (X[i] & q) != 0
            sw = (X[0] ^ X[i]) & p
            X[0] = np.where(setb, X[0] ^ p, X[0] ^ sw)
            if i > 0:
                X[i] = np.where(setb, X[i], X[i] ^ sw)
        q >>= one
    for i in range(1, dim):
        X[i] = X[i] ^ X[i - 1]
    t = np.zeros_like(X[0])
    q = one << np.uint64(bits - 1)
    while q > one:
        t = np.where((X[dim - 1] & q) != 0, t ^ (q - one), t)
        q >>= one
    X = [x ^ t for x in X]
    key = np.zeros_like(X[0])
    for b in range(bits):
        for i in range(dim):
            key |= ((X[i] >> np.uint64(b)) & one) << np.uint64(dim * b + dim - 1 - i)
    return key


def frame_cells(pos, bbox, dim):
    """Cells of the index's quantisation frame: t = (x - lo) * (2097151 /
    longest side), float32, clamped, truncated (as the device forms them)."""
    lo = np.asarray(bbox[:3], np.float32)
    ext = np.float32(max(np.float32(bbox[3 + c]) - lo[c] for c in range(dim)))
    scale = np.float32(2097151.0) / ext
    out = np.zeros((len(pos), dim), np.uint64)
    for c in range(dim):
        t = (pos[:, c].astype(np.float32) - lo[c]) * scale
        t = np.minimum(np.maximum(t, np.float32(0)), np.float32(2097151.0))
        out[:, c] = t.astype(np.uint64)
    return out
