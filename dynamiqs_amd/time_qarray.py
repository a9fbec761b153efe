"""Time-dependent operators (host side).

Mirrors ``dynamiqs/time_qarray.py``: ``TimeQArray.__call__ = prefactor * operator``
(:450-451), :class:`ConstantTimeQArray` (:497-569), :class:`PWCTimeQArray`
(:572-666, prefactor ``values[..., searchsorted(times, t, 'right') - 1]`` and zero
outside ``[times[0], times[-1])``, :649-659), :class:`SummedTimeQArray`
(:836-956), :class:`ModulatedTimeQArray` (:669-760) and
:class:`CallableTimeQArray` (:763-834). The solvers take modulated terms through
piecewise-cubic coefficient tables evaluated on the device, and rewrite
``timecallable`` operators as constant + modulated terms over the solve interval
(``dynamiqs_amd.integrators.operators``).
"""

from __future__ import annotations

import numpy as np

from .qarrays import QArray, asqarray, isqarraylike


class TimeQArray:
    _is_timeqarray = True  # QArray arithmetic defers to the time-dependent operand

    def __call__(self, t: float) -> QArray:
        raise NotImplementedError

    @property
    def discontinuity_ts(self) -> np.ndarray:
        return np.zeros(0)

    def __add__(self, y):
        if isqarraylike(y):
            y = ConstantTimeQArray(asqarray(y))
        if isinstance(y, SummedTimeQArray):
            return SummedTimeQArray([self, *y.timeqarrays])
        return SummedTimeQArray([self, y])

    __radd__ = __add__

    def __sub__(self, y):
        return self + (-1) * y

    def __rsub__(self, y):
        return (-1) * self + y

    def __neg__(self):
        return self * (-1)


class ConstantTimeQArray(TimeQArray):
    def __init__(self, qarray: QArray):
        self.qarray = asqarray(qarray)

    @property
    def shape(self):
        return self.qarray.shape

    @property
    def ndim(self):
        return self.qarray.ndim

    @property
    def dims(self):
        return self.qarray.dims

    @property
    def layout(self):
        return self.qarray.layout

    def __call__(self, t: float) -> QArray:
        return self.qarray

    def __mul__(self, y):
        return ConstantTimeQArray(self.qarray * y)

    __rmul__ = __mul__


class PWCTimeQArray(TimeQArray):
    def __init__(self, times, values, qarray: QArray):
        self.times = np.asarray(times, dtype=np.float64)
        self.values = np.asarray(values)
        self.qarray = asqarray(qarray)
        if self.times.ndim != 1 or self.values.shape[-1] != len(self.times) - 1:
            raise ValueError(
                'For a PWC timeqarray, `values` must have shape (..., len(times) - 1), but has shape'
                f' {self.values.shape} with len(times)={len(self.times)}.'
            )

    @property
    def shape(self):
        return (*self.values.shape[:-1], *self.qarray.shape[-2:])

    @property
    def ndim(self):
        return len(self.shape)

    @property
    def dims(self):
        return self.qarray.dims

    @property
    def layout(self):
        return self.qarray.layout

    @property
    def discontinuity_ts(self) -> np.ndarray:
        return self.times

    def prefactor(self, t: float):
        if t < self.times[0] or t >= self.times[-1]:
            return np.zeros(self.values.shape[:-1], dtype=self.values.dtype)
        idx = int(np.searchsorted(self.times, t, side='right')) - 1
        return self.values[..., idx]

    def __call__(self, t: float) -> QArray:
        return self.qarray * np.asarray(self.prefactor(t))[..., None, None]

    def __mul__(self, y):
        return PWCTimeQArray(self.times, self.values, self.qarray * y)

    __rmul__ = __mul__


class ModulatedTimeQArray(TimeQArray):
    def __init__(self, f, qarray: QArray):
        self.f = f
        self.qarray = asqarray(qarray)

    @property
    def shape(self):
        return (*np.shape(self.f(0.0)), *self.qarray.shape[-2:])

    @property
    def ndim(self):
        return len(self.shape)

    @property
    def dims(self):
        return self.qarray.dims

    def __call__(self, t: float) -> QArray:
        return self.qarray * np.asarray(self.f(t))[..., None, None]

    _disc_ts = np.zeros(0)

    @property
    def discontinuity_ts(self) -> np.ndarray:
        return self._disc_ts

    def __mul__(self, y):
        out = ModulatedTimeQArray(self.f, self.qarray * y)
        out._disc_ts = self._disc_ts
        return out

    __rmul__ = __mul__


class CallableTimeQArray(TimeQArray):
    def __init__(self, f, discontinuity_ts=None):
        self.f = f
        self._disc_ts = (np.zeros(0) if discontinuity_ts is None
                         else np.sort(np.asarray(discontinuity_ts, dtype=float).reshape(-1)))

    @property
    def discontinuity_ts(self) -> np.ndarray:
        return self._disc_ts

    @property
    def shape(self):
        return asqarray(self.f(0.0)).shape

    @property
    def ndim(self):
        return len(self.shape)

    @property
    def dims(self):
        return asqarray(self.f(0.0)).dims

    def __call__(self, t: float) -> QArray:
        return asqarray(self.f(t))

    def __mul__(self, y):
        f = self.f
        return CallableTimeQArray(lambda t: asqarray(f(t)) * y, self._disc_ts)

    __rmul__ = __mul__


class SummedTimeQArray(TimeQArray):
    def __init__(self, timeqarrays):
        self.timeqarrays = list(timeqarrays)
        dims = {t.dims for t in self.timeqarrays}
        if len(dims) > 1:
            raise ValueError(f'All terms of a SummedTimeArray must have the same Hilbert space dimensions, got {dims}')

    @property
    def shape(self):
        return np.broadcast_shapes(*[t.shape for t in self.timeqarrays])

    @property
    def ndim(self):
        return len(self.shape)

    @property
    def dims(self):
        return self.timeqarrays[0].dims

    @property
    def discontinuity_ts(self) -> np.ndarray:
        ts = [t.discontinuity_ts for t in self.timeqarrays]
        return np.unique(np.concatenate(ts)) if ts else np.zeros(0)

    def __call__(self, t: float) -> QArray:
        out = None
        for x in self.timeqarrays:
            out = x(t) if out is None else out + x(t)
        return out

    def __mul__(self, y):
        return SummedTimeQArray([t * y for t in self.timeqarrays])

    __rmul__ = __mul__

    def __add__(self, y):
        if isqarraylike(y):
            y = ConstantTimeQArray(asqarray(y))
        if isinstance(y, SummedTimeQArray):
            return SummedTimeQArray(self.timeqarrays + y.timeqarrays)
        return SummedTimeQArray([*self.timeqarrays, y])

    __radd__ = __add__


def constant(qarray) -> ConstantTimeQArray:
    return ConstantTimeQArray(asqarray(qarray))


def pwc(times, values, qarray) -> PWCTimeQArray:
    return PWCTimeQArray(times, values, asqarray(qarray))


def modulated(f, qarray, *, discontinuity_ts=None) -> ModulatedTimeQArray:
    """``O(t) = f(t) O`` (``dynamiqs/time_qarray.py:124-181``)."""
    if not callable(f):
        raise TypeError(f'Argument `f` must be a function, but has type {type(f).__name__}.')
    out = ModulatedTimeQArray(f, asqarray(qarray))
    if discontinuity_ts is not None:
        out._disc_ts = np.sort(np.asarray(discontinuity_ts, dtype=float).reshape(-1))
    return out


def timecallable(f, *, discontinuity_ts=None) -> CallableTimeQArray:
    """``O(t) = f(t)`` (``dynamiqs/time_qarray.py:184-237``)."""
    if not callable(f):
        raise TypeError(f'Argument `f` must be a function, but has type {type(f).__name__}.')
    return CallableTimeQArray(f, discontinuity_ts)


def astimeqarray(x) -> TimeQArray:
    """``dynamiqs/integrators/_utils.py:26-38``."""
    if isinstance(x, TimeQArray):
        return x
    try:
        return ConstantTimeQArray(asqarray(x))
    except (TypeError, ValueError) as e:
        raise TypeError(f'Argument must be a qarray-like or a timeqarray, but has type {type(x).__name__}.') from e
