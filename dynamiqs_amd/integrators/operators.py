"""Decomposition of (batched, time-dependent) operators into what the kernels consume.

A time-qarray of batch shape ``b`` becomes a constant dense part of shape
``(*b, n, n)``, a list of piecewise-constant terms ``(times, values (*b, nv),
op (n, n))`` (``dynamiqs/time_qarray.py:497-569, 572-666, 836-956``) and a list of
modulated terms ``(f, op (n, n))`` with ``f(t)`` broadcastable to ``b``
(``:669-760``). The kernels read a modulated term through a piecewise-cubic table
of ``f`` (``spline_table``) at every stage time.
"""

from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from ..time_qarray import (
    CallableTimeQArray,
    ConstantTimeQArray,
    ModulatedTimeQArray,
    PWCTimeQArray,
    SummedTimeQArray,
    TimeQArray,
)


@dataclass
class OpTerms:
    batch: tuple
    n: int
    const: np.ndarray | None  # (*batch, n, n) complex128 or None
    pwc: list = field(default_factory=list)  # [(times, values (*batch, nv), op (n, n))]
    dia: object = None  # (offsets, diags (*batch, D, n)) when every term is DIA and constant
    mod: list = field(default_factory=list)  # [(f, op (n, n))], f(t) broadcastable to batch

    @property
    def is_constant(self) -> bool:
        return not self.pwc and not self.mod

    def mod_at(self, idx):
        """The modulated terms of batch element ``idx`` as scalar functions of t."""
        out = []
        for f, o in self.mod:
            if np.ndim(f(0.0)) == 0:
                out.append((f, o, f))  # one function for the whole batch (vectorised sampling possible)
                continue

            def g(t, f=f, idx=idx):
                return complex(np.broadcast_to(np.asarray(f(t)), self.batch)[idx])

            out.append((g, o, f))
        return out

    def const_at(self, idx) -> np.ndarray:
        if self.const is None:
            return np.zeros((self.n, self.n), dtype=np.complex128)
        return self.const[idx]

    def pwc_at(self, idx):
        return [(t, v[idx], o) for (t, v, o) in self.pwc]


def decompose(x: TimeQArray) -> OpTerms:
    """Split a time-qarray into constant + PWC terms broadcast to its batch shape."""
    shape = tuple(x.shape)
    batch, n = shape[:-2], shape[-1]
    terms = x.timeqarrays if isinstance(x, SummedTimeQArray) else [x]
    const = None
    pwc = []
    mod = []
    dia_parts = []
    for t in terms:
        if isinstance(t, ConstantTimeQArray):
            q = t.qarray
            m = np.broadcast_to(q.to_numpy().astype(np.complex128), (*batch, n, n))
            const = m.copy() if const is None else const + m
            dia_parts.append(q)
        elif isinstance(t, PWCTimeQArray):
            op = t.qarray.to_numpy().astype(np.complex128)
            if op.ndim != 2:
                raise NotImplementedError('batched operators inside a PWC timeqarray are not supported')
            vals = np.broadcast_to(np.asarray(t.values, dtype=np.complex128), (*batch, t.values.shape[-1]))
            pwc.append((np.asarray(t.times, dtype=np.float64), vals, op))
        elif isinstance(t, ModulatedTimeQArray):
            if np.size(t.discontinuity_ts):
                raise NotImplementedError(
                    'modulated operators with `discontinuity_ts` are not supported by the MI355X solver (the '
                    'kernels tabulate f as a smooth piecewise cubic): use `dq.pwc` for the discontinuous part.'
                )
            op = t.qarray.to_numpy().astype(np.complex128)
            if op.ndim != 2:
                raise NotImplementedError('batched operators inside a modulated timeqarray are not supported')
            mod.append((t.f, op))
        elif isinstance(t, CallableTimeQArray):
            raise NotImplementedError(
                'CallableTimeQArray operators are not supported by the MI355X solver (constant, piecewise-constant'
                ' and modulated operators, `dq.constant` / `dq.pwc` / `dq.modulated`, run on the GPU).'
            )
        else:
            raise TypeError(f'unknown timeqarray type {type(t).__name__}')
    dia = None
    if not pwc and not mod and dia_parts and all(q.layout.name == 'dia' for q in dia_parts):
        tot = dia_parts[0]
        for q in dia_parts[1:]:
            tot = tot + q
        dia = (tot.offsets, np.broadcast_to(tot.diags, (*batch, *tot.diags.shape[-2:])))
    return OpTerms(batch=batch, n=n, const=const, pwc=pwc, dia=dia, mod=mod)


def _sample(f, t: np.ndarray) -> np.ndarray:
    """f at every time of ``t`` as a complex128 array of shape ``(len(t), *shape(f(t0)))``:
    one vectorised call when ``f`` maps an array of times elementwise, else a loop."""
    probe = np.asarray(f(float(t[0])))
    if probe.ndim == 0:
        try:
            v = np.asarray(f(t), dtype=np.complex128)
            if v.shape == t.shape:
                idx = [0, len(t) // 2, len(t) - 1]
                if np.allclose(v[idx], [complex(f(float(t[i]))) for i in idx], rtol=1e-14, atol=1e-300):
                    return v
        except Exception:  # noqa: BLE001 - f need not accept arrays; fall back to scalar calls
            pass
    return np.stack([np.asarray(f(float(x)), dtype=np.complex128) for x in t])


def spline_table(f, ta: float, tb: float, *, tol: float = 1e-12, n0: int = 256, nmax: int = 1 << 20):
    """Piecewise-cubic table of ``f`` on [ta, tb] for the kernels (``dqh_me_inst``
    fields ``mod_t0``, ``mod_dt``, ``mod_nint``, ``mod_coef``): a not-a-knot cubic
    spline on a uniform grid, refined by doubling until it matches ``f`` at every
    interval midpoint to ``tol * max(1, max |f|)`` (or ``nmax`` intervals).

    Returns ``(t0, dt, coef)`` with ``coef`` of shape ``(nint, 4, *shape(f))``:
    on [t0 + i dt, t0 + (i+1) dt], ``f = ((c0 x + c1) x + c2) x + c3``,
    ``x = t - t0 - i dt``."""
    from scipy.interpolate import CubicSpline

    ta, tb = float(ta), float(tb)
    if not tb > ta:
        tb = ta + 1.0
    tol = max(tol, float(getattr(f, 'table_tol', 0.0)))  # f known to limited precision
    n = n0
    while True:
        t = np.linspace(ta, tb, n + 1)
        y = _sample(f, t)
        cr = CubicSpline(t, y.real, axis=0)
        ci = CubicSpline(t, y.imag, axis=0)
        tm = 0.5 * (t[:-1] + t[1:])
        err = np.max(np.abs(cr(tm) + 1j * ci(tm) - _sample(f, tm)))
        ok = err <= tol * max(1.0, float(np.max(np.abs(y))))
        if ok or n >= nmax:
            if not ok and nmax > n0:
                raise ValueError(
                    f'the time-dependent coefficient could not be tabulated to {tol:.1e} with {n} cubic intervals on '
                    f'[{ta}, {tb}] (error {err:.1e}): it is not smooth there. Pass its jumps as `dq.pwc` terms, or '
                    'split the time interval at them.'
                )
            coef = np.moveaxis(cr.c + 1j * ci.c, 0, 1)  # (n, 4, *b)
            return ta, (tb - ta) / n, np.ascontiguousarray(coef)
        n *= 2


def union_breakpoints(pwc_lists) -> np.ndarray:
    ts = [np.asarray(t, dtype=np.float64) for lst in pwc_lists for (t, _, _) in lst]
    if not ts:
        return np.zeros(0)
    return np.unique(np.concatenate(ts))


def resample_pwc(times: np.ndarray, values: np.ndarray, grid: np.ndarray) -> np.ndarray:
    """Coefficient of a PWC term on each interval ``[grid[i], grid[i+1])`` of a finer grid."""
    times = np.asarray(times, dtype=np.float64)
    t = np.asarray(grid[:-1], dtype=np.float64)
    idx = np.searchsorted(times, t, side='right') - 1
    inside = (t >= times[0]) & (t < times[-1])
    idx = np.clip(idx, 0, values.shape[-1] - 1)
    out = np.asarray(values, dtype=np.complex128)[..., idx]
    return np.where(inside, out, 0.0)


def _reject_discontinuities(c) -> None:
    if np.size(c.discontinuity_ts):
        raise NotImplementedError(
            'timecallable operators with `discontinuity_ts` are not supported by the MI355X solver (the kernels '
            'tabulate their coefficients as smooth piecewise cubics): use `dq.pwc` for the discontinuous part.'
        )


def resolve_callables(x, ta: float, tb: float, max_terms: int = 4, seed: int = 0):
    """A ``timecallable`` operator H(t) = f(t) (``CallableTimeQArray``,
    ``dynamiqs/time_qarray.py:763-833``), alone or inside a sum, rewritten as a
    constant part plus at most ``max_terms`` modulated terms g_r(t) O_r, which the
    kernels run through their piecewise-cubic tables.

    The operator span of f over [ta, tb] is found from the samples f(t_i) at 65
    uniform and 16 random times: with s_i = [Re f(t_i), Im f(t_i)] and mean m, the
    right singular vectors of (s_i - m) above 1e-6 of the largest scale give
    orthonormal directions B_r (smaller ones are the rounding of f's own arithmetic,
    below what the complex64 kernels resolve); O_r = B_r as a matrix (Hermitian when
    f is), the constant part is m and g_r(t) = (s(t) - m) . B_r, tabulated to 10x
    the largest dropped direction. The span is checked at 32 other random times
    (residual <= 1e-5 relative). A callable whose samples need more
    than ``max_terms`` directions, or leave the span, raises NotImplementedError;
    so does a batched one. Other operators pass through unchanged."""
    from ..qarrays import QArray, asqarray
    from ..time_qarray import CallableTimeQArray, ConstantTimeQArray, ModulatedTimeQArray, SummedTimeQArray

    if isinstance(x, SummedTimeQArray):
        # the callables of a sum are resolved together (their joint span), within
        # the term budget the sum's modulated terms leave
        calls = [t for t in x.timeqarrays if isinstance(t, CallableTimeQArray)]
        if not calls:
            return x
        for c in calls:
            _reject_discontinuities(c)
        others = [t for t in x.timeqarrays if not isinstance(t, CallableTimeQArray)]
        budget = max_terms - sum(isinstance(t, ModulatedTimeQArray) for t in others)

        def fsum(t, fs=tuple(c.f for c in calls)):
            out = asqarray(fs[0](t))
            for f in fs[1:]:
                out = out + asqarray(f(t))
            return out

        r = resolve_callables(CallableTimeQArray(fsum), ta, tb, max(budget, 0), seed)
        return SummedTimeQArray(others + (list(r.timeqarrays) if isinstance(r, SummedTimeQArray) else [r]))
    if not isinstance(x, CallableTimeQArray):
        return x
    _reject_discontinuities(x)
    ta, tb = float(ta), float(tb)
    if not tb > ta:
        tb = ta + 1.0
    probe = asqarray(x.f(ta))
    dims = probe.dims
    if probe.ndim != 2:
        raise NotImplementedError('batched CallableTimeQArray operators are not supported by the MI355X solver')
    n = probe.shape[-1]
    cache = {}
    # directions below 1e-6 of the largest scale are rounding of f's own arithmetic
    # (single-precision qarrays by default): below what the complex64 kernels resolve
    rank_tol, span_tol = 1e-6, 1e-5

    def s_of(t: float) -> np.ndarray:
        t = float(t)
        if t not in cache:
            m = asqarray(x.f(t)).to_numpy().astype(np.complex128)
            cache[t] = np.concatenate([m.real.ravel(), m.imag.ravel()])
        return cache[t]

    rng = np.random.default_rng(seed)
    ts = np.concatenate([np.linspace(ta, tb, 65), rng.uniform(ta, tb, 16)])
    S = np.stack([s_of(t) for t in ts])
    mean = S.mean(axis=0)
    _, sv, Vh = np.linalg.svd(S - mean, full_matrices=False)
    scale = max(float(sv[0]) if sv.size else 0.0, float(np.linalg.norm(mean)), 1e-300)
    r = int(np.sum(sv > rank_tol * scale))
    noise = float(sv[r]) / scale if r < sv.size else 0.0  # the largest dropped direction
    if r > max_terms:
        raise NotImplementedError(
            f'this timecallable operator spans {r} independent operators over [{ta}, {tb}]; the MI355X solver '
            f'runs at most {max_terms} time-dependent terms (use dq.modulated / dq.pwc terms instead)'
        )
    B = Vh[:r]
    check = rng.uniform(ta, tb, 32)
    for t in check:
        v = s_of(t) - mean
        res = np.linalg.norm(v - B.T @ (B @ v))
        if res > span_tol * max(np.linalg.norm(s_of(t)), 1e-300):
            raise NotImplementedError(
                'this timecallable operator is not spanned by a few fixed operators over the solve interval; '
                'the MI355X solver runs it only in that form (use dq.modulated / dq.pwc terms instead)'
            )

    def as_op(v: np.ndarray) -> np.ndarray:
        return (v[: n * n] + 1j * v[n * n :]).reshape(n, n)

    # a mean inside the span is absorbed by the terms (f(t) L becomes one modulated
    # term without a constant part, the form a time-dependent jump operator takes)
    center = mean
    if r > 0 and np.linalg.norm(mean - B.T @ (B @ mean)) <= span_tol * scale:
        center = np.zeros_like(mean)
    terms = [] if not center.any() else [ConstantTimeQArray(QArray(as_op(center), dims=dims))]
    for k in range(r):
        def g(t, b=B[k]):
            t = np.asarray(t, dtype=np.float64)
            if t.ndim == 0:
                return float((s_of(t) - center) @ b)
            return np.array([(s_of(u) - center) @ b for u in t])

        g.table_tol = max(1e-12, 10 * noise)
        terms.append(ModulatedTimeQArray(g, QArray(as_op(B[k]), dims=dims)))
    if r == 0:
        return ConstantTimeQArray(QArray(as_op(mean), dims=dims))
    return terms[0] if len(terms) == 1 else SummedTimeQArray(terms)
