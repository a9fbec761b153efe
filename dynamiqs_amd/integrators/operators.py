"""Decomposition of (batched, time-dependent) operators into what the kernels consume.

A time-qarray of batch shape ``b`` becomes a constant dense part of shape
``(*b, n, n)`` plus a list of piecewise-constant terms ``(times, values (*b, nv),
op (n, n))`` (``dynamiqs/time_qarray.py:497-569, 572-666, 836-956``).
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

    @property
    def is_constant(self) -> bool:
        return not self.pwc

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
        elif isinstance(t, (ModulatedTimeQArray, CallableTimeQArray)):
            raise NotImplementedError(
                f'{type(t).__name__} operators are not supported by the MI355X solver (only constant and'
                ' piecewise-constant operators, `dq.constant` / `dq.pwc`, run on the GPU).'
            )
        else:
            raise TypeError(f'unknown timeqarray type {type(t).__name__}')
    dia = None
    if not pwc and dia_parts and all(q.layout.name == 'dia' for q in dia_parts):
        tot = dia_parts[0]
        for q in dia_parts[1:]:
            tot = tot + q
        dia = (tot.offsets, np.broadcast_to(tot.diags, (*batch, *tot.diags.shape[-2:])))
    return OpTerms(batch=batch, n=n, const=const, pwc=pwc, dia=dia)


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
