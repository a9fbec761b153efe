"""``sepropagator`` / ``mepropagator`` in complex128 (oracle).

Restates what ``dynamiqs/integrators/apis/sepropagator.py`` and
``mepropagator.py`` compute with an adaptive method: the propagator is the solution
for the identity as initial state (``_mepropagator`` :286-290 evolves
``eye(n^2)``), i.e. column by column for a basis of initial states:

* ``sepropagator``: column j of U(t) is ``oracle.sesolve`` from |j>;
* ``mepropagator``: column j n + i is ``vectorize`` (column stacking,
  ``dynamiqs/utils/vectorization.py:25-58``) of ``oracle.mesolve`` in the standard
  form (valid for the non-Hermitian |i><j|) from |i><j|.

Pinned in ``tests/test_oracle.py``: for constant generators both equal
``scipy.linalg.expm`` of ``-i H t`` and of the superoperator ``slindbladian`` times
t (``oracle.lindblad``, itself pinned against the reference's KATs and the other
forms); a piecewise-constant H gives the product of segment exponentials.
"""

from __future__ import annotations

import numpy as np

from . import lindblad
from . import mesolve as om
from . import sesolve as ose


def sepropagator(H, tsave, **kw) -> np.ndarray:
    """(ntsave, n, n) propagators."""
    n = np.asarray(H.const if isinstance(H, om.TimeOp) else H).shape[-1]
    cols = [ose.sesolve(H, np.eye(n)[:, j : j + 1], tsave, **kw)['states'][..., 0] for j in range(n)]
    return np.stack(cols, axis=-1)


def mepropagator(H, Ls, tsave, **kw) -> np.ndarray:
    """(ntsave, n^2, n^2) propagators, column-stacking convention."""
    n = np.asarray(H.const if isinstance(H, om.TimeOp) else H).shape[-1]
    out = np.zeros((len(tsave), n * n, n * n), dtype=np.complex128)
    for j in range(n):
        for i in range(n):
            e = np.zeros((n, n), dtype=np.complex128)
            e[i, j] = 1.0
            st = om.mesolve(H, Ls, e, tsave, form='standard', **kw)['states']
            out[:, :, j * n + i] = lindblad.vectorize(st)[..., 0]
    return out
