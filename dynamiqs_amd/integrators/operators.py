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
    n = n0
    while True:
        t = np.linspace(ta, tb, n + 1)
        y = _sample(f, t)
        cr = CubicSpline(t, y.real, axis=0)
        ci = CubicSpline(t, y.imag, axis=0)
        tm = 0.5 * (t[:-1] + t[1:])
        err = np.max(np.abs(cr(tm) + 1j * ci(tm) - _sample(f, tm)))
        if err <= tol * max(1.0, float(np.max(np.abs(y)))) or n >= nmax:
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
