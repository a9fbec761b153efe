"""Result containers, mirroring ``dynamiqs/result.py:29-194`` and the step infos of
``dynamiqs/integrators/core/diffrax_integrator.py:25-50``."""

from __future__ import annotations

import numpy as np

from .qarrays import QArray


def _memory_str(x) -> str:
    mem = x.dtype.itemsize * int(np.prod(x.shape))
    if mem < 1024**2:
        return f'{mem / 1024:.1f} Kb'
    elif mem < 1024**3:
        return f'{mem / 1024**2:.1f} Mb'
    return f'{mem / 1024**3:.1f} Gb'


def _array_str(x) -> str | None:
    if x is None:
        return None
    type_name = 'QArray' if isinstance(x, QArray) else 'Array'
    return f'{type_name} {x.dtype} {tuple(x.shape)} | {_memory_str(x)}'


class FixedStepInfos:
    def __init__(self, nsteps):
        self.nsteps = np.asarray(nsteps)

    def __str__(self) -> str:
        if self.nsteps.ndim >= 1:
            return f'{int(self.nsteps.mean())} steps | infos shape {self.nsteps.shape}'
        return f'{self.nsteps} steps'

    __repr__ = __str__


class AdaptiveStepInfos:
    def __init__(self, nsteps, naccepted, nrejected):
        self.nsteps = np.asarray(nsteps)
        self.naccepted = np.asarray(naccepted)
        self.nrejected = np.asarray(nrejected)

    def __str__(self) -> str:
        if self.nsteps.ndim >= 1:
            return (
                f'avg. {self.nsteps.mean():.1f} steps ({self.naccepted.mean():.1f}'
                f' accepted, {self.nrejected.mean():.1f} rejected) | infos shape'
                f' {self.nsteps.shape}'
            )
        return f'{self.nsteps} steps ({self.naccepted} accepted, {self.nrejected} rejected)'

    __repr__ = __str__


class Result:
    def __init__(self, tsave, method, gradient, options, ysave, Esave, extra, infos):
        self.tsave = tsave
        self.method = method
        self.gradient = gradient
        self.options = options
        self._ysave = ysave
        self._Esave = Esave
        self._extra = extra
        self.infos = infos

    @property
    def extra(self):
        return self._extra

    def block_until_ready(self) -> Result:
        return self

    def _str_parts(self) -> dict:
        return {
            'Method': type(self.method).__name__,
            'Gradient': type(self.gradient).__name__ if self.gradient is not None else None,
            'Infos': self.infos if self.infos is not None else None,
            'Extra': str(self.extra) if self.extra is not None else None,
        }

    def __str__(self) -> str:
        parts = {k: v for k, v in self._str_parts().items() if v is not None}
        padding = max(len(k) for k in parts) + 1
        parts_str = '\n'.join(f'{k:<{padding}}: {v}' for k, v in parts.items())
        return f'==== {self.__class__.__name__} ====\n' + parts_str

    __repr__ = __str__


class SolveResult(Result):
    @property
    def states(self) -> QArray:
        return self._ysave

    @property
    def final_state(self) -> QArray:
        return self.states[..., -1, :, :]

    @property
    def expects(self):
        return self._Esave

    def _str_parts(self) -> dict:
        d = super()._str_parts()
        d.update({'States': _array_str(self.states), 'Expects': _array_str(self.expects)})
        return d


class MESolveResult(SolveResult):
    """Result of :func:`dynamiqs_amd.mesolve`.

    ``states`` (..., nsave, n, n), ``final_state``, ``expects`` (..., nE, ntsave)
    or None, ``extra``, ``infos``, ``tsave``, ``method``, ``gradient``,
    ``options`` (``dynamiqs/result.py:193-194``).
    """


class SESolveResult(SolveResult):
    """Result of :func:`dynamiqs_amd.sesolve`.

    ``states`` (..., nsave, n, 1), ``final_state``, ``expects`` (..., nE, ntsave)
    or None, ``infos``, ``tsave``, ``method``, ``gradient``, ``options``
    (``dynamiqs/result.py:189-190``).
    """


class PropagatorResult(Result):
    """``propagators`` (..., nsave, N, N) with nsave = ntsave, or 1 without
    ``save_propagators``; ``final_propagator`` (``dynamiqs/result.py:149-162``)."""

    @property
    def propagators(self) -> QArray:
        return self._ysave

    @property
    def final_propagator(self) -> QArray:
        return self.propagators[..., -1, :, :]

    def _str_parts(self) -> dict:
        d = super()._str_parts()
        d['Propagators'] = _array_str(self.propagators)
        return d


class SEPropagatorResult(PropagatorResult):
    """Result of :func:`dynamiqs_amd.sepropagator`: N = n (``dynamiqs/result.py:213-214``)."""


class MEPropagatorResult(PropagatorResult):
    """Result of :func:`dynamiqs_amd.mepropagator`: N = n^2, columns in the
    column-stacking order of ``dq.vectorize`` (``dynamiqs/result.py:217-218``)."""
