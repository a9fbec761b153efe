"""Dormand-Prince 5(4) tableau and dense output for the oracle integrator
(test infrastructure, like the rest of ``oracle/``).

The reference's ``Dopri5`` method (``dynamiqs/method.py:402-435``) runs
``dx.Dopri5()`` (``dynamiqs/integrators/core/diffrax_integrator.py:361-372``)
with the same PID controller, error norm and saving as ``Tsit5``. diffrax is not
vendored, so the published method is restated:

* Tableau: J. R. Dormand and P. J. Prince, "A family of embedded Runge-Kutta
  formulae", J. Comput. Appl. Math. 6 (1980) 19-26 (RK5(4)7M: 7 stages, FSAL,
  ``b = a[6]``; ``E = b - bhat``).
* Embedded weights: ``bhat = (1951/21600, 0, 22642/50085, 451/720,
  -12231/42400, 649/6300, 1/60)``, the alternative fourth-order solution that
  Shampine (1986, below) pairs with these stages. The 1980 paper's own
  ``(5179/57600, ..., 1/40)`` is not used. With Shampine's weights, the
  reference's printed Dopri5 run (``docs/documentation/basics/workflow.md:88``,
  Rabi oscillation at ``atol = 1e-8``) takes 48 accepted and 8 rejected steps, as
  printed. The 1980 weights take 51-52 accepted steps.
* Dense output: the quartic through ``y0``, ``y1``, ``f0``, ``f1`` and the
  midpoint value ``y0 + h sum_i cmid_i f_i`` with Shampine's coefficients (L. F.
  Shampine, "Some practical Runge-Kutta formulas", Math. Comp. 46 (1986)
  135-150), which is the interpolation diffrax uses for Dopri5. Solving the five
  interpolation conditions for the quartic gives it in the form
  ``y0 + h sum_i b_i(theta) f_i`` with ``b_i(theta) = sum_p R[i, p] theta^(p+1)``
  (the same form as Tsit5's continuous extension); the theta^4 column equals the
  dense-output constants of Hairer's DOPRI5 code.

Pinned in ``tests/test_oracle.py``: row sums, the 17 order conditions to order 5
for ``b`` and to order 4 for ``bhat``, the order-4 conditions of the midpoint
weights at theta = 1/2 and of ``b_i(theta)`` at several theta, ``b_i(1) = b_i``,
and empirical 5th-order convergence.
"""

from __future__ import annotations

import numpy as np

C = np.array([0.0, 1 / 5, 3 / 10, 4 / 5, 8 / 9, 1.0, 1.0])
A = np.zeros((7, 7))
A[1, 0] = 1 / 5
A[2, :2] = [3 / 40, 9 / 40]
A[3, :3] = [44 / 45, -56 / 15, 32 / 9]
A[4, :4] = [19372 / 6561, -25360 / 2187, 64448 / 6561, -212 / 729]
A[5, :5] = [9017 / 3168, -355 / 33, 46732 / 5247, 49 / 176, -5103 / 18656]
A[6, :6] = [35 / 384, 0.0, 500 / 1113, 125 / 192, -2187 / 6784, 11 / 84]
B = A[6].copy()
BHAT = np.array([1951 / 21600, 0.0, 22642 / 50085, 451 / 720, -12231 / 42400, 649 / 6300, 1 / 60])
E = B - BHAT
# Shampine's midpoint weights: y(t + h/2) ~ y0 + h sum_i CMID_i f_i
CMID = np.array([6025192743 / 30085553152, 0.0, 51252292925 / 65400821598, -2691868925 / 45128329728,
                 187940372067 / 1594534317056, -1776094331 / 19743644256, 11237099 / 235043384]) / 2


def _dense_output_weights() -> np.ndarray:
    """R (7 x 4): the quartic Y(theta) = (p(theta) - y0) / h = sum_p a_p theta^(p+1)
    with Y'(0) = f_1, Y(1) = sum b_i f_i, Y'(1) = f_7 and Y(1/2) = sum cmid_i f_i,
    solved per stage coefficient."""
    a1 = np.eye(7)[0]
    rhs = np.stack([B - a1, np.eye(7)[6] - a1, CMID - a1 / 2])  # (3, 7)
    M = np.array([[1.0, 1.0, 1.0], [2.0, 3.0, 4.0], [1 / 4, 1 / 8, 1 / 16]])
    a234 = np.linalg.solve(M, rhs)  # (3, 7): rows a2, a3, a4
    return np.concatenate([a1[:, None], a234.T], axis=1)


R = _dense_output_weights()
