"""Quantum utilities mirroring ``dynamiqs/utils/general.py`` (host side)."""

from __future__ import annotations

import numpy as np

from ..qarrays import QArray, _dia_kron, asqarray, dense


def dag(x) -> QArray:
    return asqarray(x).dag()


def isket(x) -> bool:
    return asqarray(x).isket()


def todm(x) -> QArray:
    return asqarray(x).todm()


def tensor(*args) -> QArray:
    qs = [asqarray(a) for a in args]
    if all(q.layout is not dense for q in qs):
        out = qs[0]
        for q in qs[1:]:
            out = _dia_kron(out, q)
        return out
    out = qs[0].to_numpy()
    dims = list(qs[0].dims)
    for q in qs[1:]:
        x = q.to_numpy()
        out = np.einsum('...ij,...kl->...ikjl', out, x).reshape(
            *np.broadcast_shapes(out.shape[:-2], x.shape[:-2]), out.shape[-2] * x.shape[-2], out.shape[-1] * x.shape[-1]
        )
        dims += list(q.dims)
    return QArray(out, dims=tuple(dims))


def tracemm(x, y) -> np.ndarray:
    """Tr(x @ y) as sum(x * y^T) (``general.py:241-276``)."""
    return (asqarray(x).to_numpy() * np.swapaxes(asqarray(y).to_numpy(), -1, -2)).sum((-2, -1))


def expect(O, x) -> np.ndarray:
    """``general.py:420-474``."""
    if isinstance(O, (list, tuple)):
        return np.stack([expect(o, x) for o in O])
    O = asqarray(O)
    if O.ndim > 2:
        return np.stack([expect(O[i], x) for i in range(O.shape[0])])
    x = asqarray(x)
    Om, xm = O.to_numpy(), x.to_numpy()
    if x.isket():
        return (np.swapaxes(xm.conj(), -1, -2) @ Om @ xm)[..., 0, 0]
    if x.isbra():
        return (xm @ Om @ np.swapaxes(xm.conj(), -1, -2))[..., 0, 0]
    return tracemm(Om, xm)


def dissipator(L, rho) -> QArray:
    L = asqarray(L).to_numpy()
    rho = asqarray(rho).to_numpy()
    Ld = np.swapaxes(L.conj(), -1, -2)
    LdL = Ld @ L
    return QArray(L @ rho @ Ld - 0.5 * LdL @ rho - 0.5 * rho @ LdL)


def lindbladian(H, jump_ops, rho) -> QArray:
    H = asqarray(H).to_numpy()
    r = asqarray(rho).to_numpy()
    out = -1j * H @ r + 1j * r @ H
    for L in jump_ops:
        out = out + dissipator(L, r).to_numpy()
    return QArray(out)
