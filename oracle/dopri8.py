"""Dormand-Prince 8(7) with the PID step controller, in float64/complex128 (oracle).

Test infrastructure: only ``tests/``, ``smoke()`` and the bench's CPU baseline use
this package.

The reference's ``Dopri8`` method (``dynamiqs/method.py``) runs
``dx.diffeqsolve(..., dx.Dopri8(), stepsize_controller=dx.PIDController(...))``
(``dynamiqs/integrators/core/diffrax_integrator.py:59-81, 101-147``). diffrax is not
vendored in the reference, so this module restates the published method:

* Tableau: P. J. Prince and J. R. Dormand, "High order embedded Runge-Kutta
  formulae", J. Comput. Appl. Math. 7 (1981) 67-75, method RK8(7)13M: 13 stages,
  the 8th-order solution propagated (local extrapolation), the 7th-order one
  for the error estimate.
* Error norm, integral controller and initial step as ``oracle.tsit5``: the
  controller exponent is ``1/8`` (the order of the pair). The initial-step exponent
  is ``1/7``, the order of the embedded solution, by the rule fitted to the 5(4)
  pairs. No printed count pins it for this pair (parity unpinned).
* Saving: steps land on the save times (and on ``jump_ts`` and ``t1``), so every
  saved state is a step end. diffrax evaluates its own continuous extension at
  the save times instead; that interpolant is not restated here, so saved values
  agree with the reference to within the tolerance, not step for step.

Pinned in ``tests/test_oracle.py``: the 200 order conditions up to order 8 for
``b`` and the 85 up to order 7 for ``bhat`` (rooted trees, exact rationals), row
sums ``c_i = sum_j a_ij``, empirical 8th-order convergence, and the closed-form
lossy oscillator.
"""

from __future__ import annotations

from fractions import Fraction as Fr

import numpy as np

from .tsit5 import SNAP, MaxStepsError, Tsit5Stats, rms_norm

ORDER = 8
NSTAGES = 13

C_FR = [Fr(0), Fr(1, 18), Fr(1, 12), Fr(1, 8), Fr(5, 16), Fr(3, 8), Fr(59, 400), Fr(93, 200),
        Fr(5490023248, 9719169821), Fr(13, 20), Fr(1201146811, 1299019798), Fr(1), Fr(1)]
A_FR = [[Fr(0)] * NSTAGES for _ in range(NSTAGES)]
for _i, _row in {
    1: {0: Fr(1, 18)},
    2: {0: Fr(1, 48), 1: Fr(1, 16)},
    3: {0: Fr(1, 32), 2: Fr(3, 32)},
    4: {0: Fr(5, 16), 2: Fr(-75, 64), 3: Fr(75, 64)},
    5: {0: Fr(3, 80), 3: Fr(3, 16), 4: Fr(3, 20)},
    6: {0: Fr(29443841, 614563906), 3: Fr(77736538, 692538347), 4: Fr(-28693883, 1125000000),
        5: Fr(23124283, 1800000000)},
    7: {0: Fr(16016141, 946692911), 3: Fr(61564180, 158732637), 4: Fr(22789713, 633445777),
        5: Fr(545815736, 2771057229), 6: Fr(-180193667, 1043307555)},
    8: {0: Fr(39632708, 573591083), 3: Fr(-433636366, 683701615), 4: Fr(-421739975, 2616292301),
        5: Fr(100302831, 723423059), 6: Fr(790204164, 839813087), 7: Fr(800635310, 3783071287)},
    9: {0: Fr(246121993, 1340847787), 3: Fr(-37695042795, 15268766246), 4: Fr(-309121744, 1061227803),
        5: Fr(-12992083, 490766935), 6: Fr(6005943493, 2108947869), 7: Fr(393006217, 1396673457),
        8: Fr(123872331, 1001029789)},
    10: {0: Fr(-1028468189, 846180014), 3: Fr(8478235783, 508512852), 4: Fr(1311729495, 1432422823),
         5: Fr(-10304129995, 1701304382), 6: Fr(-48777925059, 3047939560), 7: Fr(15336726248, 1032824649),
         8: Fr(-45442868181, 3398467696), 9: Fr(3065993473, 597172653)},
    11: {0: Fr(185892177, 718116043), 3: Fr(-3185094517, 667107341), 4: Fr(-477755414, 1098053517),
         5: Fr(-703635378, 230739211), 6: Fr(5731566787, 1027545527), 7: Fr(5232866602, 850066563),
         8: Fr(-4093664535, 808688257), 9: Fr(3962137247, 1805957418), 10: Fr(65686358, 487910083)},
    12: {0: Fr(403863854, 491063109), 3: Fr(-5068492393, 434740067), 4: Fr(-411421997, 543043805),
         5: Fr(652783627, 914296604), 6: Fr(11173962825, 925320556), 7: Fr(-13158990841, 6184727034),
         8: Fr(3936647629, 1978049680), 9: Fr(-160528059, 685178525), 10: Fr(248638103, 1413531060)},
}.items():
    for _j, _v in _row.items():
        A_FR[_i][_j] = _v
B_FR = [Fr(14005451, 335480064), Fr(0), Fr(0), Fr(0), Fr(0), Fr(-59238493, 1068277825), Fr(181606767, 758867731),
        Fr(561292985, 797845732), Fr(-1041891430, 1371343529), Fr(760417239, 1151165299),
        Fr(118820643, 751138087), Fr(-528747749, 2220607170), Fr(1, 4)]
BHAT_FR = [Fr(13451932, 455176623), Fr(0), Fr(0), Fr(0), Fr(0), Fr(-808719846, 976000145),
           Fr(1757004468, 5645159321), Fr(656045339, 265891186), Fr(-3867574721, 1518517206),
           Fr(465885868, 322736535), Fr(53011238, 667516719), Fr(2, 45), Fr(0)]

C = np.array([float(x) for x in C_FR])
A = np.array([[float(x) for x in r] for r in A_FR])
B = np.array([float(x) for x in B_FR])
E = np.array([float(b - bh) for b, bh in zip(B_FR, BHAT_FR)])


def initial_step(f, t0: float, y0: np.ndarray, f0: np.ndarray, rtol: float, atol: float, stats=None) -> float:
    """Hairer-Norsett-Wanner II.4 with the exponent 1/7 (order of the embedded solution)."""
    scale = atol + np.abs(y0) * rtol
    d0 = rms_norm(y0 / scale)
    d1 = rms_norm(f0 / scale)
    h0 = 1e-6 if (d0 < 1e-5 or d1 < 1e-5) else 0.01 * d0 / d1
    f1 = f(t0 + h0, y0 + h0 * f0, t0)
    if stats is not None:
        stats.nrhs += 1
    d2 = rms_norm((f1 - f0) / scale) / h0
    max_d = max(d1, d2)
    h1 = max(1e-6, h0 * 1e-3) if max_d <= 1e-15 else (0.01 / max_d) ** (1.0 / (ORDER - 1))
    return min(100 * h0, h1)


def integrate(f, y0: np.ndarray, tsave: np.ndarray, *, t0: float | None = None, rtol: float = 1e-6,
              atol: float = 1e-6, safety: float = 0.9, factormin: float = 0.2, factormax: float = 5.0,
              max_steps: int = 100_000, jump_ts=None, save=None):
    """Integrate ``dy/dt = f(t, y, tseg)`` (``tseg``: start of the step) from ``t0``
    to ``tsave[-1]``; steps end on every save time, discontinuity and ``t1``.
    Returns ``(saved_list, ylast, stats)``."""
    tsave = np.asarray(tsave, dtype=np.float64)
    t = float(tsave[0] if t0 is None else t0)
    t_end = float(tsave[-1])
    save = save or (lambda y: y.copy())
    jumps = np.sort(np.asarray([] if jump_ts is None else jump_ts, dtype=np.float64))
    stats = Tsit5Stats()
    y = np.array(y0, dtype=np.complex128)
    saved = []
    isave = 0
    while isave < len(tsave) and tsave[isave] <= t:
        saved.append(save(y))
        isave += 1
    k0 = f(t, y, t)
    stats.nrhs += 1
    dt = initial_step(f, t, y, k0, rtol, atol, stats)
    need_k0 = False
    while t < t_end:
        if stats.nsteps >= max_steps:
            raise MaxStepsError('The maximum number of method steps was reached.')
        if need_k0:
            k0 = f(t, y, t)
            stats.nrhs += 1
        need_k0 = True
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
        ks = [k0]
        for s in range(1, NSTAGES):
            ys = y + h * sum(A[s, j] * ks[j] for j in range(s) if A[s, j] != 0.0)
            ks.append(f(t + C[s] * h, ys, t))
            stats.nrhs += 1
        y_new = y + h * sum(B[j] * ks[j] for j in range(NSTAGES) if B[j] != 0.0)
        err = h * sum(E[j] * ks[j] for j in range(NSTAGES) if E[j] != 0.0)
        scale = atol + rtol * np.maximum(np.abs(y), np.abs(y_new))
        norm = rms_norm(err / scale)
        keep = norm < 1.0
        factor = factormax if norm == 0.0 else safety * norm ** (-1.0 / ORDER)
        factor = min(max(factor, 1.0 if keep else factormin), factormax)
        stats.nsteps += 1
        if keep:
            stats.naccepted += 1
            stats.dts.append(h)
            t = target if on_target else t_new
            stats.ends.append(t)
            y = y_new
            while isave < len(tsave) and tsave[isave] <= t:
                saved.append(save(y))
                isave += 1
        else:
            stats.nrejected += 1
            need_k0 = False
        dt = factor * h
    while isave < len(tsave):
        saved.append(save(y))
        isave += 1
    return saved, y, stats
