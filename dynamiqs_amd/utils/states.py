"""State builders mirroring ``dynamiqs/utils/states.py`` (host side)."""

from __future__ import annotations

import numpy as np

from ..qarrays import QArray, cdtype
from .general import tensor
from .operators import displace


def fock(dim, number) -> QArray:
    if isinstance(dim, (tuple, list)):
        return tensor(*[fock(d, k) for d, k in zip(dim, number)])
    number = np.asarray(number)
    out = np.zeros((*number.shape, dim, 1), dtype=cdtype())
    idx = np.indices(number.shape)
    out[(*idx, number, np.zeros_like(number))] = 1
    return QArray(out, dims=(dim,))


def fock_dm(dim, number) -> QArray:
    return fock(dim, number).todm()


def coherent(dim, alpha) -> QArray:
    if isinstance(dim, (tuple, list)):
        return tensor(*[coherent(d, a) for d, a in zip(dim, alpha)])
    D = displace(int(dim), alpha).to_numpy()
    return QArray(D[..., :, :1].astype(cdtype()), dims=(int(dim),))


def coherent_dm(dim, alpha) -> QArray:
    return coherent(dim, alpha).todm()


def ground() -> QArray:
    return QArray(np.array([[0], [1]], dtype=cdtype()), dims=(2,))


def excited() -> QArray:
    return QArray(np.array([[1], [0]], dtype=cdtype()), dims=(2,))


def ground_dm() -> QArray:
    return ground().todm()


def excited_dm() -> QArray:
    return excited().todm()
