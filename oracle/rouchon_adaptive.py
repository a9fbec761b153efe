"""Adaptive Rouchon 1-2 and 2-3 pairs with the PID step controller, complex128 (oracle).

Test infrastructure: only ``tests/``, ``smoke()`` and the bench's CPU baseline use
this package.

Restates ``MESolveAdaptiveRouchonIntegrator`` (``dynamiqs/integrators/core/
rouchon_integrator.py:455-549``) with its two subclasses (:585-606), selected by
``Rouchon2()`` / ``Rouchon3()`` without ``dt`` (:620-651):

* One step from ``rho`` at ``t`` over ``h`` builds the no-jump propagators of an
  embedded Runge-Kutta pair on dU/dt = G(tau) U, U(t) = I, G = -i H - 1/2 sum
  L^dag L (``make_nojump_propagators`` :475-510, ``G`` :310-312): the high-order
  solution and the low-order one ``y1_high - error``.

  - order 2: explicit midpoint with its embedded Euler solution, U_high(1) =
    I + h G(t + h/2) (I + h/2 G(t)). The high map is ``KrausHeun2`` with U_high;
    the low map is ``KrausEuler``, whose own U (:250-253) is I + h G(t + h/2) and
    whose jump operators are taken at t + h/2.
  - order 3: Bogacki-Shampine (Bosh3) with its embedded second-order solution
    (P. Bogacki and L. F. Shampine, Appl. Math. Lett. 2 (1989) 321-325: stages at
    t, t + h/2, t + 3h/4, t + h; weights 2/9, 1/3, 4/9, 0 and 7/24, 1/4, 1/3, 1/8
    over k1..k4, k4 = h G(t + h) U_high(1)). The high map is ``KrausHeun3`` on the
    Bosh3 cubic Hermite dense output through (I, h G(t)) and (U_high(1), k4); the
    low map is ``KrausHeun2``, which needs only U_low(1) (its nodes are 0 and 1,
    where the midpoint dense output of :602-606 takes its end values).

* Both maps are applied to the same ``rho``, each after its own Cholesky
  normalisation when ``normalize`` is set; the step returns ``rho_high`` and the
  error ``(rho_high - rho_low) / 2`` (:513-539).
* Controller: diffrax's ``PIDController`` as built by the reference
  (``diffrax_integrator.py:59-72``) with ``step_ts = tsave`` (:544-549) and the
  solver's order (2 or 3) as error order, i.e. the integral controller of
  ``oracle.tsit5`` with the exponent ``1/order``. Steps land on every save time
  and on the breakpoints of piecewise-constant terms (``jump_ts``), so every saved
  state is a step end; a piecewise-constant term takes its value at the step start
  for the whole step.
* Initial step 0.01: the reference passes no ``dt0`` for adaptive methods
  (``diffrax_integrator.py:74-76``); diffrax's initial-step selection estimates
  one only for ODE terms and returns 0.01 for other terms, and the Rouchon term is
  not an ODE term (its vector field is empty, ``rouchon_integrator.py:23-38``).
  This value is taken from that reading, not from a run of the reference:
  parity unpinned for the first trial step (later steps and the result do not
  depend on it beyond the tolerance).

Pinned in ``tests/test_oracle.py``: for constant operators the propagators and
maps equal the fixed-step ones of ``oracle.rouchon`` (pinned there); the
embedded Bosh3 propagator against its closed form I + hG + h^2/2 G^2 + 3h^3/16 G^3
+ h^4/48 G^4; the error estimate scaling as h^order (the local error of the low
map) and the high map's local error as h^(order+1); trace preservation of every
accepted step; convergence to ``expm`` of the Lindbladian and, for a driven qubit,
to a tight Tsit5 solve as the tolerance tightens.
"""

from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from .lindblad import dag
from .rouchon import HEUN2_A, HEUN2_B, HEUN2_C, HEUN3_A, HEUN3_B, HEUN3_C, hermite3, kraus_rk_step, kraus_step_dense
from .tsit5 import SNAP, MaxStepsError, rms_norm

DT0 = 0.01


def _as_hf(H):
    """``H`` as a function of (tau, tseg): constant arrays are wrapped."""
    if callable(H):
        return H
    H = np.asarray(H, dtype=np.complex128)
    return lambda tau, tseg: H


def _generator(Hf, LdL, tseg):
    return lambda tau: -1j * Hf(tau, tseg) - 0.5 * LdL


def propagators(Hf, Ls, t: float, h: float, order: int, tseg: float | None = None) -> dict:
    """The no-jump propagators of one trial step: ``{'high': {c: U(c)}, 'low': U_low(1)}``
    (order 2: ``'low'`` is the KrausEuler M0 = I + h G(t + h/2))."""
    tseg = t if tseg is None else tseg
    n = Ls[0].shape[-1] if Ls else Hf(t, tseg).shape[-1]
    eye = np.eye(n)
    LdL = sum((dag(L) @ L for L in Ls), np.zeros((n, n), dtype=np.complex128))
    G = _generator(Hf, LdL, tseg)
    if order == 2:
        Gm = G(t + 0.5 * h)
        return {'high': {0.0: eye, 1.0: eye + h * Gm @ (eye + 0.5 * h * G(t))}, 'low': eye + h * Gm}
    if order == 3:
        k1 = G(t)
        k2 = G(t + 0.5 * h) @ (eye + 0.5 * h * k1)
        k3 = G(t + 0.75 * h) @ (eye + 0.75 * h * k2)
        U1 = eye + h * (2.0 / 9.0 * k1 + 1.0 / 3.0 * k2 + 4.0 / 9.0 * k3)
        k4 = G(t + h) @ U1
        low = eye + h * (7.0 / 24.0 * k1 + 0.25 * k2 + 1.0 / 3.0 * k3 + 0.125 * k4)
        d0, d1 = h * k1, h * k4
        high = {0.0: eye, HEUN3_C[1]: hermite3(eye, U1, d0, d1, HEUN3_C[1]),
                HEUN3_C[2]: hermite3(eye, U1, d0, d1, HEUN3_C[2]), 1.0: U1}
        return {'high': high, 'low': low}
    raise ValueError(f'adaptive Rouchon order must be 2 or 3, got {order}')


def pair_step(H, Ls, rho: np.ndarray, h: float, order: int, normalize: bool = True, t: float = 0.0,
              tseg: float | None = None):
    """One trial step of the pair from ``rho`` at ``t``: ``(rho_high, error)``.
    ``H``: constant array or a function ``H(tau, tseg)``."""
    Hf = _as_hf(H)
    tseg = t if tseg is None else tseg
    U = propagators(Hf, Ls, t, h, order, tseg)
    if order == 2:
        high = kraus_rk_step(U['high'], HEUN2_C, HEUN2_A, HEUN2_B, Ls, rho, h, normalize)
        low = kraus_step_dense(Hf(t + 0.5 * h, tseg), Ls, rho, h, normalize)
    else:
        high = kraus_rk_step(U['high'], HEUN3_C, HEUN3_A, HEUN3_B, Ls, rho, h, normalize)
        n = rho.shape[-1]
        low = kraus_rk_step({0.0: np.eye(n), 1.0: U['low']}, HEUN2_C, HEUN2_A, HEUN2_B, Ls, rho, h, normalize)
    return high, 0.5 * (high - low)


@dataclass
class ARStats:
    nsteps: int = 0
    naccepted: int = 0
    nrejected: int = 0
    dts: list = field(default_factory=list)


def integrate(H, Ls, rho0: np.ndarray, tsave, *, order: int, t0: float | None = None, rtol: float = 1e-6,
              atol: float = 1e-6, safety: float = 0.9, factormin: float = 0.2, factormax: float = 5.0,
              max_steps: int = 100_000, normalize: bool = True, dt0: float = DT0, jump_ts=None):
    """Adaptive Rouchon solve from ``t0`` (default ``tsave[0]``) to ``tsave[-1]``.
    ``H``: constant array or ``H(tau, tseg)`` (``tseg``: start of the step, for
    piecewise-constant terms); ``jump_ts``: their breakpoints.
    Returns ``(states [ntsave, n, n], stats)``."""
    tsave = np.asarray(tsave, dtype=np.float64)
    t = float(tsave[0] if t0 is None else t0)
    t_end = float(tsave[-1])
    rho = np.array(rho0, dtype=np.complex128)
    Ls = [np.asarray(L, dtype=np.complex128) for L in Ls]
    jumps = np.sort(np.asarray([] if jump_ts is None else jump_ts, dtype=np.float64))
    stats = ARStats()
    saved = []
    isave = 0
    while isave < len(tsave) and tsave[isave] <= t:
        saved.append(rho.copy())
        isave += 1
    dt = dt0
    while t < t_end:
        if stats.nsteps >= max_steps:
            raise MaxStepsError('The maximum number of method steps was reached.')
        target = t_end
        if isave < len(tsave) and tsave[isave] < target:
            target = float(tsave[isave])
        nj = jumps[(jumps > t) & (jumps < target)]
        if nj.size:
            target = float(nj[0])
        h = dt
        t_new = t + h
        on_target = False
        if t + h >= target - SNAP * max(1.0, abs(target)):
            h = target - t
            t_new = target
            on_target = True
        high, err = pair_step(H, Ls, rho, h, order, normalize, t=t)
        scale = atol + rtol * np.maximum(np.abs(rho), np.abs(high))
        norm = rms_norm(err / scale)
        finite = bool(np.isfinite(norm))
        keep = finite and norm < 1.0
        if not finite:
            factor = factormin
        elif norm == 0.0:
            factor = factormax
        else:
            factor = safety * norm ** (-1.0 / order)
        factor = min(max(factor, 1.0 if keep else factormin), factormax)
        stats.nsteps += 1
        if keep:
            stats.naccepted += 1
            stats.dts.append(h)
            t = target if on_target else t_new
            rho = high
            while isave < len(tsave) and tsave[isave] <= t:
                saved.append(rho.copy())
                isave += 1
        else:
            stats.nrejected += 1
        dt = factor * h
    while isave < len(tsave):
        saved.append(rho.copy())
        isave += 1
    return np.stack(saved), stats
