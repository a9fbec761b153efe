"""``mesolve`` driver in complex128 (oracle).

Restates the single-instance path of ``dynamiqs/integrators/apis/mesolve.py``
(``_mesolve`` :339-387) for ``Tsit5`` and ``Rouchon1``:

* ``rho0`` is converted with ``todm`` (``diffrax_integrator.py:338-345``);
* the vector field is the Hermitian form by default (``:319-323``), the standard
  form with ``assume_hermitian=False`` (``:312-317``) and the vectorised form with
  ``vectorized=True`` (``:325-327``);
* saving (``save_mixin.py:47-68``): states at every ``tsave`` (or the final one
  only with ``save_states=False``), ``Esave[k, i] = Tr(E_k rho(tsave[i]))``
  through ``tracemm`` (``utils/general.py:241-276``);
* infos: ``(nsteps, naccepted, nrejected)`` for Tsit5, ``nsteps`` for Rouchon1
  (``diffrax_integrator.py:149-157``).

Time dependence (``time_qarray.py``): an operator is a constant array or a
:class:`TimeOp` = constant part + sum of modulated terms ``f(t) * op`` evaluated at
the exact stage time + sum of piecewise-constant terms
``values[idx] * op`` with ``idx = searchsorted(times, t, 'right') - 1`` and zero
outside ``[times[0], times[-1])`` (``PWCTimeQArray._prefactor`` :649-659); the
discontinuity times feed the controller's ``jump_ts`` (``:627-628``,
``diffrax_integrator.py:64-71``).
"""

from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from . import lindblad, rouchon, tsit5
from .operators import todm, tracemm


@dataclass
class PWCTerm:
    times: np.ndarray  # (nv + 1,)
    values: np.ndarray  # (nv,) complex
    op: np.ndarray  # (n, n)

    def coeff(self, t: float) -> complex:
        times = self.times
        if t < times[0] or t >= times[-1]:
            return 0.0
        idx = int(np.searchsorted(times, t, side='right')) - 1
        return complex(self.values[idx])


@dataclass
class ModTerm:
    """``f(t) * op`` (``ModulatedTimeQArray``, ``time_qarray.py:669-760``): a smooth
    term, evaluated at the exact stage time."""

    f: object
    op: np.ndarray  # (n, n)

    def coeff(self, t: float) -> complex:
        return complex(self.f(t))


@dataclass
class TimeOp:
    const: np.ndarray
    terms: list = field(default_factory=list)

    def __call__(self, t: float, t_stage: float | None = None) -> np.ndarray:
        """The operator with its piecewise-constant terms at ``t`` (the segment time)
        and its modulated terms at ``t_stage`` (default ``t``)."""
        out = np.array(self.const, dtype=np.complex128)
        for term in self.terms:
            c = term.coeff(t if isinstance(term, PWCTerm) or t_stage is None else t_stage)
            if c != 0:
                out = out + c * term.op
        return out

    @property
    def jump_ts(self) -> np.ndarray:
        pwc = [t for t in self.terms if isinstance(t, PWCTerm)]
        if not pwc:
            return np.zeros(0)
        return np.unique(np.concatenate([np.asarray(t.times, dtype=np.float64) for t in pwc]))


def as_timeop(x) -> TimeOp:
    return x if isinstance(x, TimeOp) else TimeOp(np.asarray(x, dtype=np.complex128))


def _segment_time(t_eval: float, tseg: float, jumps: np.ndarray) -> float:
    # stages of a step that ends on a discontinuity are evaluated on its left side:
    # use the step start to select the segment whenever no jump lies in [tseg, t_eval]
    if jumps.size and np.any((jumps > tseg) & (jumps <= t_eval)):
        return tseg
    return t_eval


def mesolve(
    H,
    Ls,
    rho0: np.ndarray,
    tsave,
    *,
    exp_ops=None,
    method: str = 'tsit5',
    rtol: float = 1e-6,
    atol: float = 1e-6,
    safety: float = 0.9,
    factormin: float = 0.2,
    factormax: float = 5.0,
    max_steps: int = 100_000,
    dt: float | None = None,
    normalize: bool = True,
    t0: float | None = None,
    save_states: bool = True,
    form: str = 'hermitian',
) -> dict:
    H = as_timeop(H)
    Ls = [as_timeop(L) for L in Ls]
    rho0 = todm(np.asarray(rho0, dtype=np.complex128))
    tsave = np.asarray(tsave, dtype=np.float64)
    Es = [np.asarray(E, dtype=np.complex128) for E in (exp_ops or [])]
    jumps = np.unique(np.concatenate([H.jump_ts, *[L.jump_ts for L in Ls]]))

    def save(y):
        return (y.copy() if save_states else None, np.array([tracemm(E, y) for E in Es], dtype=np.complex128))

    if method in ('tsit5', 'dopri5', 'dopri8'):
        if form == 'hermitian':
            rhs = lindblad.rhs_hermitian
        elif form == 'standard':
            rhs = lindblad.rhs_standard
        elif form == 'vectorized':
            rhs = lambda Hm, Lm, y: lindblad.unvectorize(lindblad.slindbladian(Hm, Lm) @ lindblad.vectorize(y))
        else:
            raise ValueError(form)

        def f(t, y, tseg):
            ts = _segment_time(t, tseg, jumps)
            return rhs(H(ts, t), [L(ts, t) for L in Ls], y)

        common = dict(t0=t0, rtol=rtol, atol=atol, safety=safety, factormin=factormin, factormax=factormax,
                      max_steps=max_steps, jump_ts=jumps, save=save)
        if method == 'dopri8':
            from . import dopri8

            saved, ylast, st = dopri8.integrate(f, rho0, tsave, **common)
        else:
            saved, ylast, st = tsit5.integrate(f, rho0, tsave, tab=tsit5.tableau(method), **common)
        infos = dict(nsteps=st.nsteps, naccepted=st.naccepted, nrejected=st.nrejected, nrhs=st.nrhs)
    elif method in ('rouchon1', 'rouchon2', 'rouchon3', 'euler'):
        if dt is None:
            raise ValueError('fixed-step methods need dt')
        if method == 'euler':
            # diffrax Euler: y + h f(t, y), the vector field at the step start

            def step(t, h, rho):
                return rho + h * lindblad.rhs_standard(H(t), [L(t) for L in Ls], rho)
        else:
            kstep = {'rouchon1': rouchon.kraus_step_dense, 'rouchon2': rouchon.rouchon2_step,
                     'rouchon3': rouchon.rouchon3_step}[method]

            def step(t, h, rho):
                tm = t + 0.5 * h
                return kstep(H(tm), [L(tm) for L in Ls], rho, h, normalize)

        saved, ylast, nsteps = rouchon.integrate(step, rho0, tsave, dt, t0=t0, save=save)
        infos = dict(nsteps=nsteps)
    else:
        raise ValueError(method)

    states = np.stack([s[0] for s in saved]) if save_states else ylast[None]
    expects = np.stack([s[1] for s in saved], axis=-1) if Es else None
    return dict(states=states, final_state=ylast, expects=expects, infos=infos)
