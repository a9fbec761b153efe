"""Fock-space operators and states in complex128 (oracle).

Restates ``dynamiqs/utils/operators.py`` (``destroy`` :242-307, ``number``
:378-444, ``displace`` :476-505, ``position``/``momentum`` :576-622, Pauli
matrices :624-740) and ``dynamiqs/utils/states.py`` (``fock`` :33,
``coherent`` :196-264 = displace(alpha) @ fock(0)). Multi-mode operators use
the Kronecker ordering of ``dynamiqs/utils/general.py:tensor`` (first mode is
the most significant index).

Pinned in ``tests/test_oracle.py`` against the reference docstring
printouts (``coherent(4, 0.5)``, ``coherent_dm(4, 0.5)``, ``displace(4, 0.5)``,
``position(3)``, ``momentum(3)``) and the Poisson photon statistics of a
coherent state.
"""

from __future__ import annotations

import numpy as np
import scipy.linalg as sla

C128 = np.complex128


def eye(*dims: int) -> np.ndarray:
    return np.eye(int(np.prod(dims)), dtype=C128)


def _embed(op: np.ndarray, dims: tuple[int, ...], k: int) -> np.ndarray:
    out = np.ones((1, 1), dtype=C128)
    for i, d in enumerate(dims):
        out = np.kron(out, op if i == k else np.eye(d, dtype=C128))
    return out


def destroy(*dims: int):
    """``a`` with ``a[m-1, m] = sqrt(m)``; tuple of embedded ops for several dims."""
    singles = [np.diag(np.sqrt(np.arange(1, d, dtype=np.float64)), k=1).astype(C128) for d in dims]
    if len(dims) == 1:
        return singles[0]
    return tuple(_embed(s, dims, k) for k, s in enumerate(singles))


def create(*dims: int):
    a = destroy(*dims)
    if isinstance(a, tuple):
        return tuple(x.conj().T for x in a)
    return a.conj().T


def number(*dims: int):
    singles = [np.diag(np.arange(d, dtype=np.float64)).astype(C128) for d in dims]
    if len(dims) == 1:
        return singles[0]
    return tuple(_embed(s, dims, k) for k, s in enumerate(singles))


def displace(dim: int, alpha: complex) -> np.ndarray:
    a = destroy(dim)
    return sla.expm(alpha * a.conj().T - np.conj(alpha) * a)


def position(dim: int, hbar: float = 0.5) -> np.ndarray:
    a = destroy(dim)
    return np.sqrt(hbar / 2) * (a.conj().T + a)


def momentum(dim: int, hbar: float = 0.5) -> np.ndarray:
    a = destroy(dim)
    return np.sqrt(hbar / 2) * 1j * (a.conj().T - a)


def sigmax() -> np.ndarray:
    return np.array([[0, 1], [1, 0]], dtype=C128)


def sigmay() -> np.ndarray:
    return np.array([[0, -1j], [1j, 0]], dtype=C128)


def sigmaz() -> np.ndarray:
    return np.array([[1, 0], [0, -1]], dtype=C128)


def sigmam() -> np.ndarray:
    return np.array([[0, 0], [1, 0]], dtype=C128)


def fock(dim: int, k: int) -> np.ndarray:
    v = np.zeros((dim, 1), dtype=C128)
    v[k, 0] = 1.0
    return v


def fock_dm(dim: int, k: int) -> np.ndarray:
    v = fock(dim, k)
    return v @ v.conj().T


def coherent(dim, alpha) -> np.ndarray:
    """Ket ``D(alpha)|0>`` (multi-mode: tensor product, first mode most significant)."""
    if isinstance(dim, (tuple, list)):
        out = np.ones((1, 1), dtype=C128)
        for d, a in zip(dim, alpha):
            out = np.kron(out, coherent(int(d), a))
        return out
    return displace(int(dim), complex(alpha)) @ fock(int(dim), 0)


def coherent_dm(dim, alpha) -> np.ndarray:
    v = coherent(dim, alpha)
    return v @ v.conj().T


def todm(x: np.ndarray) -> np.ndarray:
    """``dynamiqs/utils/general.py:821-854``: ket -> |x><x|, density matrix unchanged."""
    if x.shape[-1] == 1:
        return x @ np.swapaxes(x.conj(), -1, -2)
    return x


def tensor(*ops: np.ndarray) -> np.ndarray:
    out = np.ones((1, 1), dtype=C128)
    for op in ops:
        out = np.kron(out, op)
    return out


def tracemm(x: np.ndarray, y: np.ndarray) -> complex:
    """``dynamiqs/utils/general.py:241-276``: Tr(x @ y) as sum(x * y.T)."""
    return (x * np.swapaxes(y, -1, -2)).sum((-2, -1))


def expect(O: np.ndarray, x: np.ndarray) -> complex:
    """``dynamiqs/utils/general.py:467-474`` for a single operator."""
    if x.shape[-1] == 1:
        return (np.swapaxes(x.conj(), -1, -2) @ O @ x)[..., 0, 0]
    return tracemm(O, x)
