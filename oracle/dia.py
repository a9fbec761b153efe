"""Sparse-DIA storage and products in complex128 (oracle).

Restates ``dynamiqs/qarrays/sparsedia_primitives.py``:

* storage (``_sparsedia_slice`` :11-15, ``_construct_diags`` :161-169): a matrix
  ``A`` of size n x n is ``offsets`` (tuple of ints) and ``diags[..., D, n]`` with
  ``A[i, j] = diags[k, j]`` for ``j - i == offsets[k]``; entries whose row index
  falls outside ``[0, n)`` are stored as zeros. This is the
  ``scipy.sparse.dia_matrix`` convention (``data[k, j] = A[j - offsets[k], j]``).
* ``matmul_sparsedia_array`` :232-245 -> :func:`dia_matmul_dense`
* ``matmul_array_sparsedia`` :248-261 -> :func:`dense_matmul_dia`
* ``matmul_sparsedia_sparsedia`` :199-229 -> :func:`dia_matmul_dia`
* ``transpose_sparsedia`` :18-29 (+ conj) -> :func:`dia_dag`

Pinned in ``tests/test_oracle.py`` against ``scipy.sparse.dia_matrix`` and
against dense products on the fixture patterns of the reference's DIA tests
(offsets (-2, 0, 2) and (-2, -1, 1, 3) at N=10, ``tests/qarrays/
test_sparse_dia_qarray.py:32-46``).
"""

from __future__ import annotations

import numpy as np


def _valid_cols(offset: int, n: int) -> slice:
    # columns j with a row i = j - offset inside [0, n)
    return slice(offset, None) if offset >= 0 else slice(None, n + offset)


def dia_from_dense(x: np.ndarray, offsets=None, tol: float = 0.0):
    """Return ``(offsets, diags)`` of a dense square matrix (all nonzero diagonals
    when ``offsets`` is None)."""
    n = x.shape[-1]
    if offsets is None:
        offsets = tuple(
            k for k in range(-(n - 1), n) if np.max(np.abs(np.diagonal(x, offset=k)), initial=0.0) > tol
        )
    diags = np.zeros((*x.shape[:-2], len(offsets), n), dtype=np.result_type(x.dtype, np.complex128))
    for i, k in enumerate(offsets):
        diags[..., i, _valid_cols(k, n)] = np.diagonal(x, offset=k, axis1=-2, axis2=-1)
    return tuple(int(k) for k in offsets), diags


def dia_to_dense(offsets, diags: np.ndarray) -> np.ndarray:
    n = diags.shape[-1]
    out = np.zeros((*diags.shape[:-2], n, n), dtype=diags.dtype)
    rows = np.arange(n)
    for i, k in enumerate(offsets):
        cols = rows + k
        m = (cols >= 0) & (cols < n)
        out[..., rows[m], cols[m]] = diags[..., i, cols[m]]
    return out


def dia_matmul_dense(offsets, diags: np.ndarray, x: np.ndarray) -> np.ndarray:
    """``A @ x``: ``out[r, :] += diags[k, r + o] * x[r + o, :]``."""
    n = diags.shape[-1]
    out = np.zeros(np.broadcast_shapes(diags.shape[:-2] + (n, x.shape[-1]), x.shape), dtype=np.complex128)
    for i, o in enumerate(offsets):
        if o >= 0:
            out[..., : n - o, :] += diags[..., i, o:, None] * x[..., o:, :]
        else:
            out[..., -o:, :] += diags[..., i, : n + o, None] * x[..., : n + o, :]
    return out


def dense_matmul_dia(x: np.ndarray, offsets, diags: np.ndarray) -> np.ndarray:
    """``x @ A``: ``out[:, c] += x[:, c - o] * diags[k, c]``."""
    n = diags.shape[-1]
    out = np.zeros(np.broadcast_shapes(x.shape, diags.shape[:-2] + (x.shape[-2], n)), dtype=np.complex128)
    for i, o in enumerate(offsets):
        if o >= 0:
            out[..., :, o:] += x[..., :, : n - o] * diags[..., i, None, o:]
        else:
            out[..., :, : n + o] += x[..., :, -o:] * diags[..., i, None, : n + o]
    return out


def dia_dag(offsets, diags: np.ndarray):
    """Conjugate transpose: offset ``o`` becomes ``-o`` and values move to the
    other end of the stored row."""
    n = diags.shape[-1]
    out = np.zeros_like(diags)
    for i, o in enumerate(offsets):
        if o >= 0:
            out[..., i, : n - o] = diags[..., i, o:].conj()
        else:
            out[..., i, -o:] = diags[..., i, : n + o].conj()
    return tuple(-o for o in offsets), out


def dia_matmul_dia(lo, ld: np.ndarray, ro, rd: np.ndarray):
    """Product of two DIA matrices; output offsets are the sorted sums."""
    n = ld.shape[-1]
    acc: dict[int, np.ndarray] = {}
    for i, a in enumerate(lo):
        for j, b in enumerate(ro):
            o = a + b
            if abs(o) > n - 1:
                continue
            # out[r, r+a+b] = L[r, r+a] * R[r+a, r+a+b] = ld[i, r+a] * rd[j, r+a+b]
            d = np.zeros(np.broadcast_shapes(ld.shape[:-2], rd.shape[:-2]) + (n,), dtype=np.complex128)
            for c in range(n):
                mid = c - b
                row = c - o
                if 0 <= mid < n and 0 <= row < n:
                    d[..., c] = ld[..., i, mid] * rd[..., j, c]
            acc[o] = acc.get(o, 0) + d
    offs = tuple(sorted(acc))
    if not offs:
        return (), np.zeros((0, n), dtype=np.complex128)
    return offs, np.stack([acc[o] for o in offs], axis=-2)
