"""Tsit5 with a PID step controller and dense output, in float64/complex128 (oracle).

The reference delegates the arithmetic to ``dx.diffeqsolve(..., dx.Tsit5(),
stepsize_controller=dx.PIDController(rtol, atol, safety, factormin, factormax,
jump_ts))`` (``dynamiqs/integrators/core/diffrax_integrator.py:59-81, 101-147``;
defaults at ``dynamiqs/method.py:121-127``). diffrax (``diffrax>=0.6.1``,
``pyproject.toml:30``) is a third-party dependency that is not vendored in the
reference, so this module restates its published algorithm:

* Tableau and continuous extension: Ch. Tsitouras, "Runge-Kutta pairs of order
  5(4) satisfying only the first column simplifying assumption", Comput. Math.
  Appl. 62 (2011) 770-775 (7 stages, FSAL, ``b = a[6]``, error weights ``e``).
* Error norm: RMS over all n*n complex entries of
  ``err / (atol + rtol * max(|y0|, |y1|))`` with ``err = sum_i e_i k_i`` and
  ``k_i = dt f_i``.
* Integral controller: ``factor = clip(safety * norm**(-1/5), fmin_eff, fmax)``,
  ``fmin_eff = 1`` after an accepted step and ``factormin`` after a rejected
  one; a step is accepted when ``norm < 1``; ``dt_next = factor * dt``.
* Initial step: Hairer-Norsett-Wanner, Solving ODEs I, II.4 (the ``HINIT``
  routine), scale ``atol + rtol |y0|``, with ``h1 = (0.01 / max(d1, d2))^(1/4)``:
  the exponent is one over the order of the embedded solution (4), while the
  controller uses one over the order of the pair (5). The exponent is fitted to
  the reference's printed step counts, since diffrax's source is not in the
  reference. With ``1/4``, this module and ``oracle.dopri5`` take exactly the
  five counts that the reference prints (below). With ``1/5`` (HINIT's own
  value) every Tsit5 example takes one accepted step fewer, because its first
  step is about 2.5 times longer. Any first step 0.2-0.5 times the ``1/5`` value
  fits the Tsit5 counts; ``1/4`` gives 0.35-0.4 on these systems.
* ``jump_ts``: a step never crosses a discontinuity time; all stages of a step
  use the generator of the segment in which the step starts, and the FSAL stage
  is recomputed after landing on a jump. A step that would stop within
  ``SNAP * max(1, |target|)`` before the next discontinuity or the end time ends
  on it: floating-point breakpoints (``linspace``) are not equally spaced, and a
  step proposed equal to the previous segment length can stop an ulp short of
  the next breakpoint, which would leave a step of ~1e-15 to it.
* Saving: states at ``tsave`` by the continuous extension of the step that
  contains them; ``tsave[0] == t0`` saves the initial state.

Pinned in ``tests/test_oracle.py``: row sums ``c_i = sum_j a_ij``, the 17
order conditions up to order 5 for ``b`` and the order-4 conditions for
``b - e``, the interpolant's order-4 conditions at several theta and
``b_i(1) = b_i``; empirical 5th-order convergence; the closed-form lossy
oscillator; the exact step counts that the reference prints for five runs of its
own: ``README.md:75`` (mesolve, 40 steps), ``dynamiqs/integrators/apis/mesolve.py:225``
(9), ``dynamiqs/integrators/apis/sesolve.py:164`` (11),
``dynamiqs/integrators/apis/floquet.py:139`` (a time-dependent propagator, 10
accepted and 1 rejected) and ``docs/documentation/basics/workflow.md:88`` (Dopri5
at ``atol = 1e-8``: 48 accepted and 8 rejected). The step sequence is not pinned
beyond those counts, and neither is the bookkeeping at jumps.
"""

from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

# --- Tsitouras 5(4) tableau ---------------------------------------------------
C = np.array([0.0, 0.161, 0.327, 0.9, 0.9800255409045097, 1.0, 1.0])
A = np.zeros((7, 7))
A[1, 0] = 0.161
A[2, :2] = [-0.008480655492356989, 0.335480655492357]
A[3, :3] = [2.897153057105493, -6.359448489975075, 4.3622954328695815]
A[4, :4] = [5.325864828439257, -11.748883564062828, 7.4955393428898365, -0.09249506636175525]
A[5, :5] = [5.86145544294642, -12.92096931784711, 8.159367898576159, -0.071584973281401, -0.028269050394068383]
A[6, :6] = [0.09646076681806523, 0.01, 0.4798896504144996, 1.379008574103742, -3.290069515436081, 2.324710524099774]
B = A[6].copy()  # FSAL: solution weights are the last stage row (b7 = 0)
# error weights e = b - bhat
E = np.array(
    [
        -0.00178001105222577714,
        -0.0008164344596567469,
        0.007880878010261995,
        -0.1447110071732629,
        0.5823571654525552,
        -0.45808210592918697,
        0.015151515151515152,
    ]
)
# continuous extension b_i(theta) = sum_p R[i, p] theta^(p+1), p = 0..3
R = np.array(
    [
        [1.0, -2.763706197274826, 2.9132554618219126, -1.0530884977290216],
        [0.0, 0.13169999999999998, -0.2234, 0.1017],
        [0.0, 3.9302962368947516, -5.941033872131505, 2.490627285651253],
        [0.0, -12.411077166933676, 30.33818863028232, -16.548102889244902],
        [0.0, 37.50931341651104, -88.1789048947664, 47.37952196281928],
        [0.0, -27.896526289197286, 65.09189467479366, -34.87065786149661],
        [0.0, 1.5, -4.0, 2.5],
    ]
)
SNAP = 1e-12
ORDER = 5  # controller exponent 1/ORDER
INIT_ORDER = 4  # initial-step exponent 1/INIT_ORDER (order of the embedded solution)


@dataclass(frozen=True)
class Tableau:
    """A 7-stage FSAL pair of order 5(4): stage coefficients ``A`` (the last row
    holds the solution weights, ``b_7 = 0``), error weights ``E = b - bhat`` and
    continuous-extension weights ``R`` (``b_i(theta) = sum_p R[i, p] theta^(p+1)``)."""

    C: np.ndarray
    A: np.ndarray
    E: np.ndarray
    R: np.ndarray


TSIT5 = Tableau(C, A, E, R)


def _dopri5() -> Tableau:
    from . import dopri5

    return Tableau(dopri5.C, dopri5.A, dopri5.E, dopri5.R)


def tableau(name: str) -> Tableau:
    """``'tsit5'`` or ``'dopri5'`` (the reference's ``Tsit5`` and ``Dopri5``)."""
    if name == 'tsit5':
        return TSIT5
    if name == 'dopri5':
        return _dopri5()
    raise ValueError(name)


def interp_weights(theta: float, tab: Tableau = TSIT5) -> np.ndarray:
    th = np.array([theta, theta**2, theta**3, theta**4])
    return tab.R @ th


class MaxStepsError(RuntimeError):
    pass


def rms_norm(x: np.ndarray) -> float:
    return float(np.sqrt(np.mean(np.abs(x) ** 2))) if x.size else 0.0


@dataclass
class Tsit5Stats:
    nsteps: int = 0
    naccepted: int = 0
    nrejected: int = 0
    nrhs: int = 0
    dts: list = field(default_factory=list)
    ends: list = field(default_factory=list)  # end time of every accepted step


def initial_step(f, t0: float, y0: np.ndarray, f0: np.ndarray, rtol: float, atol: float, stats=None) -> float:
    scale = atol + np.abs(y0) * rtol
    d0 = rms_norm(y0 / scale)
    d1 = rms_norm(f0 / scale)
    if d0 < 1e-5 or d1 < 1e-5:
        h0 = 1e-6
    else:
        h0 = 0.01 * d0 / d1
    y1 = y0 + h0 * f0
    f1 = f(t0 + h0, y1, t0)
    if stats is not None:
        stats.nrhs += 1
    d2 = rms_norm((f1 - f0) / scale) / h0
    max_d = max(d1, d2)
    if max_d <= 1e-15:
        h1 = max(1e-6, h0 * 1e-3)
    else:
        h1 = (0.01 / max_d) ** (1.0 / INIT_ORDER)
    return min(100 * h0, h1)


def integrate(
    f,
    y0: np.ndarray,
    tsave: np.ndarray,
    *,
    t0: float | None = None,
    rtol: float = 1e-6,
    atol: float = 1e-6,
    safety: float = 0.9,
    factormin: float = 0.2,
    factormax: float = 5.0,
    max_steps: int = 100_000,
    jump_ts=None,
    save=None,
    tab: Tableau = TSIT5,
    dtype=np.complex128,
):
    """Integrate ``dy/dt = f(t, y, tseg)`` from ``t0`` to ``tsave[-1]`` with the
    pair ``tab`` (Tsit5 by default, or ``tableau('dopri5')``).

    ``dtype=np.complex64`` runs the same algorithm in single precision (state,
    stages and tableau in complex64 / float32, the error norm in float64): the
    complex64 CPU restatement that is timed beside the GPU, not a parity reference.

    ``f(t, y, tseg)`` evaluates the vector field at time ``t``; ``tseg`` is the
    start of the current step, used by piecewise-constant generators to pick the
    segment (so that stages of a step ending on a jump stay on its left side).
    ``save(y)`` maps a state to what is stored at each ``tsave``.

    Returns ``(saved_list, ylast, stats)``.
    """
    tsave = np.asarray(tsave, dtype=np.float64)
    t = float(tsave[0] if t0 is None else t0)
    t_end = float(tsave[-1])
    save = save or (lambda y: y.copy())
    jumps = np.sort(np.asarray([] if jump_ts is None else jump_ts, dtype=np.float64))
    jumps = jumps[(jumps > t) & (jumps < t_end)]

    stats = Tsit5Stats()
    y = np.array(y0, dtype=dtype)
    real = np.float32 if np.dtype(dtype) == np.complex64 else np.float64
    tA, tE = tab.A.astype(real), tab.E.astype(real)
    saved = []
    isave = 0
    while isave < len(tsave) and tsave[isave] <= t:
        saved.append(save(y))
        isave += 1

    k1 = f(t, y, t)
    stats.nrhs += 1
    dt = initial_step(f, t, y, k1, rtol, atol, stats)
    need_k1 = False

    while t < t_end:
        if stats.nsteps >= max_steps:
            raise MaxStepsError('The maximum number of method steps was reached.')
        if need_k1:
            k1 = f(t, y, t)
            stats.nrhs += 1
            need_k1 = False
        # clip the step to the end time or the next discontinuity; a step that would
        # stop within SNAP * max(1, |target|) short of it ends on it (no sliver step)
        h = dt
        nj = jumps[jumps > t]
        target, target_jump = (float(nj[0]), True) if nj.size else (t_end, False)
        made_jump = False
        t_new = t + h
        if t + h >= target - SNAP * max(1.0, abs(target)):
            h = target - t
            t_new = target
            made_jump = target_jump

        hs = real(h)
        ks = [hs * k1]
        for s in range(1, 7):
            ys = y + sum(tA[s, j] * ks[j] for j in range(s))
            ks.append(hs * f(t + tab.C[s] * h, ys, t))
            stats.nrhs += 1
        y_new = y + sum(tA[6, j] * ks[j] for j in range(6))
        err = sum(tE[j] * ks[j] for j in range(7))
        scale = atol + rtol * np.maximum(np.abs(y), np.abs(y_new)).astype(np.float64)
        norm = rms_norm(err.astype(np.complex128) / scale)
        keep = norm < 1.0
        if norm == 0.0:
            factor = factormax
        else:
            factor = safety * norm ** (-1.0 / ORDER)
        factor = min(max(factor, 1.0 if keep else factormin), factormax)
        stats.nsteps += 1
        if keep:
            stats.naccepted += 1
            stats.dts.append(h)
            stats.ends.append(t_new)
            while isave < len(tsave) and tsave[isave] <= t_new:
                theta = (tsave[isave] - t) / h
                w = interp_weights(theta, tab).astype(real)
                saved.append(save(y + sum(w[j] * ks[j] for j in range(7))))
                isave += 1
            t = t_new
            y = y_new
            k1 = ks[6] / hs
            need_k1 = made_jump
        else:
            stats.nrejected += 1
        dt = factor * h
    while isave < len(tsave):  # tsave[-1] == t_end always saved above
        saved.append(save(y))
        isave += 1
    return saved, y, stats


def integrate_steps(f, y0: np.ndarray, tsave: np.ndarray, step_ends, *, t0: float | None = None, save=None,
                    tab: Tableau = TSIT5):
    """Tsit5 over a prescribed sequence of accepted steps (``step_ends``: the end
    time of each step, as an adaptive solve recorded them): the same stages,
    solution weights and continuous extension as ``integrate`` without the
    controller, so that a solve in another precision can be replayed in float64
    on the identical step sequence. ``f(t, y, tseg)`` as in ``integrate``."""
    tsave = np.asarray(tsave, dtype=np.float64)
    t = float(tsave[0] if t0 is None else t0)
    save = save or (lambda x: x.copy())
    y = np.array(y0, dtype=np.complex128)
    saved = []
    isave = 0
    while isave < len(tsave) and tsave[isave] <= t:
        saved.append(save(y))
        isave += 1
    for t_new in np.asarray(step_ends, dtype=np.float64):
        h = t_new - t
        ks = [h * f(t, y, t)]
        for s in range(1, 7):
            ys = y + sum(tab.A[s, j] * ks[j] for j in range(s))
            ks.append(h * f(t + tab.C[s] * h, ys, t))
        while isave < len(tsave) and tsave[isave] <= t_new:
            w = interp_weights((tsave[isave] - t) / h, tab)
            saved.append(save(y + sum(w[j] * ks[j] for j in range(7))))
            isave += 1
        y = y + sum(tab.A[6, j] * ks[j] for j in range(6))
        t = t_new
    while isave < len(tsave):
        saved.append(save(y))
        isave += 1
    return saved, y
