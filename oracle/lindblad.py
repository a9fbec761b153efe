"""Lindblad vector fields in complex128 (oracle).

Restates ``dynamiqs/integrators/core/diffrax_integrator.py``:

* ``vector_field_unvec_standard`` :312-317 -> :func:`rhs_standard`
* ``vector_field_unvec_hermitian`` :319-323 -> :func:`rhs_hermitian`
  (``Hnh = -iH - 1/2 sum L^dag L``, ``tmp = Hnh rho + 1/2 sum L rho L^dag``,
  ``drho = tmp + tmp^dag``; the default form, ``assume_hermitian=True``)
* ``vector_field_vec`` :325-327 with ``slindbladian``
  (``dynamiqs/utils/vectorization.py:25-58, 100-177, 213-259``, column stacking)
  -> :func:`slindbladian`
* ``dissipator`` / ``lindbladian`` (``dynamiqs/utils/general.py:560-654``).

Pinned in ``tests/test_oracle.py`` by the reference doctest KATs
(``dissipator(destroy(4), fock_dm(4, 2)) = diag(0, 2, -2, 0)`` and
``lindbladian(a^dag a, [a, a^dag a], fock_dm(4, 1)) = diag(1, -1, 0, 0)``),
trace preservation, Hermiticity, agreement of the three forms, and a brute-force
element-wise sum over indices on tiny inputs.
"""

from __future__ import annotations

import numpy as np


def dag(x: np.ndarray) -> np.ndarray:
    return np.swapaxes(x.conj(), -1, -2)


def rhs_standard(H: np.ndarray, Ls, rho: np.ndarray) -> np.ndarray:
    half_LdL = 0.5 * sum((dag(L) @ L for L in Ls), np.zeros_like(H))
    nojump = (-1j * H - half_LdL) @ rho + rho @ (1j * H - half_LdL)
    jump = sum((L @ rho @ dag(L) for L in Ls), np.zeros_like(rho))
    return nojump + jump


def hnh(H: np.ndarray, Ls) -> np.ndarray:
    """Non-Hermitian generator ``-iH - 1/2 sum L^dag L`` of the Hermitian form."""
    return -1j * H - 0.5 * sum((dag(L) @ L for L in Ls), np.zeros_like(H))


def rhs_hermitian(H: np.ndarray, Ls, rho: np.ndarray) -> np.ndarray:
    tmp = hnh(H, Ls) @ rho + 0.5 * sum((L @ rho @ dag(L) for L in Ls), np.zeros_like(rho))
    return tmp + dag(tmp)


def vectorize(x: np.ndarray) -> np.ndarray:
    """Column stacking: ``(..., n, n) -> (..., n^2, 1)``."""
    return np.swapaxes(x, -1, -2).reshape(*x.shape[:-2], -1, 1)


def unvectorize(x: np.ndarray) -> np.ndarray:
    n = int(round(np.sqrt(x.shape[-2])))
    return np.swapaxes(x.reshape(*x.shape[:-2], n, n), -1, -2)


def spre(x: np.ndarray) -> np.ndarray:
    n = x.shape[-1]
    return np.kron(np.eye(n), x)


def spost(x: np.ndarray) -> np.ndarray:
    n = x.shape[-1]
    return np.kron(x.T, np.eye(n))


def sprepost(x: np.ndarray, y: np.ndarray) -> np.ndarray:
    return np.kron(y.T, x)


def slindbladian(H: np.ndarray, Ls) -> np.ndarray:
    n = H.shape[-1]
    out = -1j * spre(H) + 1j * spost(H)
    for L in Ls:
        LdL = dag(L) @ L
        out = out + sprepost(L, dag(L)) - 0.5 * spre(LdL) - 0.5 * spost(LdL)
    assert out.shape == (n * n, n * n)
    return out


def dissipator(L: np.ndarray, rho: np.ndarray) -> np.ndarray:
    LdL = dag(L) @ L
    return L @ rho @ dag(L) - 0.5 * LdL @ rho - 0.5 * rho @ LdL


def lindbladian(H: np.ndarray, Ls, rho: np.ndarray) -> np.ndarray:
    return -1j * H @ rho + 1j * rho @ H + sum((dissipator(L, rho) for L in Ls), np.zeros_like(rho))


def adjoint_lindbladian(H: np.ndarray, Ls, O: np.ndarray) -> np.ndarray:
    """Heisenberg-picture generator ``L^dagger[O] = i[H,O] + sum(L^dag O L - 1/2{L^dag L, O})``,
    the adjoint of :func:`lindbladian` for the Hilbert-Schmidt product
    ``<A, B> = Tr(A^dag B)``."""
    out = 1j * (H @ O - O @ H)
    for L in Ls:
        LdL = dag(L) @ L
        out = out + dag(L) @ O @ L - 0.5 * (LdL @ O + O @ LdL)
    return out
