"""Operator builders mirroring ``dynamiqs/utils/operators.py`` (host side)."""

from __future__ import annotations

import numpy as np

from ..qarrays import QArray, _dia_kron, cdtype, dense, get_layout


def _dia_single(offsets, diags_list, n) -> QArray:
    d = np.zeros((len(offsets), n), dtype=cdtype())
    for i, vals in enumerate(diags_list):
        d[i] = vals
    return QArray(offsets=offsets, diags=d, dims=(n,))


def _embed(singles, dims, layout):
    if len(dims) == 1:
        return singles[0] if layout is not dense else singles[0].asdense()
    out = []
    for k in range(len(dims)):
        op = None
        for i, d in enumerate(dims):
            f = singles[k] if i == k else eye(d, layout=layout)
            f = f.asdia()
            op = f if op is None else _dia_kron(op, f)
        op.dims = tuple(dims)
        out.append(op if layout is not dense else op.asdense())
    return tuple(out)


def eye(*dims: int, layout=None) -> QArray:
    layout = get_layout(layout)
    n = int(np.prod(dims))
    q = _dia_single((0,), [np.ones(n)], n)
    q.dims = tuple(dims)
    return q if layout is not dense else q.asdense()


def destroy(*dims: int, layout=None):
    layout = get_layout(layout)
    singles = []
    for d in dims:
        v = np.zeros(d)
        v[1:] = np.sqrt(np.arange(1, d))
        singles.append(_dia_single((1,), [v], d))
    return _embed(singles, dims, layout)


def create(*dims: int, layout=None):
    a = destroy(*dims, layout=layout)
    return tuple(x.dag() for x in a) if isinstance(a, tuple) else a.dag()


def number(*dims: int, layout=None):
    layout = get_layout(layout)
    singles = [_dia_single((0,), [np.arange(d, dtype=float)], d) for d in dims]
    return _embed(singles, dims, layout)


def displace(dim: int, alpha) -> QArray:
    alpha = np.asarray(alpha, dtype=np.complex128)[..., None, None]
    a = destroy(dim, layout=dense).to_numpy().astype(np.complex128)
    gen = alpha * a.conj().T - alpha.conj() * a
    return QArray(gen, dims=(dim,)).expm().astype(cdtype())


def position(dim: int, *, layout=None, hbar: float = 0.5) -> QArray:
    a = destroy(dim, layout=layout)
    return np.sqrt(hbar / 2) * (a.dag() + a)


def momentum(dim: int, *, layout=None, hbar: float = 0.5) -> QArray:
    a = destroy(dim, layout=layout)
    return np.sqrt(hbar / 2) * (1j * (a.dag() - a))


def _pauli(m, layout) -> QArray:
    q = QArray(np.asarray(m, dtype=cdtype()), dims=(2,))
    return q.asdia() if get_layout(layout) is not dense else q


def sigmax(*, layout=None) -> QArray:
    return _pauli([[0, 1], [1, 0]], layout)


def sigmay(*, layout=None) -> QArray:
    return _pauli([[0, -1j], [1j, 0]], layout)


def sigmaz(*, layout=None) -> QArray:
    return _pauli([[1, 0], [0, -1]], layout)


def sigmap(*, layout=None) -> QArray:
    return _pauli([[0, 1], [0, 0]], layout)


def sigmam(*, layout=None) -> QArray:
    return _pauli([[0, 0], [1, 0]], layout)
