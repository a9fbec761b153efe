"""``sesolve`` driver in complex128 (oracle).

Restates the single-instance path of ``dynamiqs/integrators/apis/sesolve.py``
(``_sesolve`` :263-304) for ``Tsit5``, ``Dopri5`` and ``Euler`` (fixed steps, the
vector field at the step start, linear interpolation to ``tsave``):

* the vector field is ``-1j * H(t) @ psi`` (``SEDiffraxIntegrator.terms``,
  ``dynamiqs/integrators/core/diffrax_integrator.py:225-228``), with the time
  dependence of :class:`oracle.mesolve.TimeOp` (piecewise-constant terms on the
  segment of the step start, modulated terms at the exact stage time) and the
  discontinuity times of the piecewise-constant terms as the controller's
  ``jump_ts``;
* the controller, initial step and dense output are those of ``oracle.tsit5``;
* saving (``save_mixin.py:47-68``): kets at every ``tsave`` (or the final one only
  with ``save_states=False``) and ``Esave[k, i] = <psi|E_k|psi>`` (``expect``,
  ``dynamiqs/utils/general.py:420-470``, ``dag(x) @ O @ x`` for a ket);
* infos: ``(nsteps, naccepted, nrejected)``.

Pinned in ``tests/test_oracle.py`` against closed forms that a dropped factor of
``-1j``, a transposed ``H`` or a wrong segment rule would fail: the reference's
``Cavity`` system (``tests/systems/closed_system.py:46-104``: ``delta a^dag a``
rotates a coherent state, every Fock amplitude picks up ``exp(-1j delta k t)``),
its ``TDQubit`` (:107-176: ``H = eps cos(omega t) sigma_x`` from ``|0>`` gives
``cos(theta/2)|0> - 1j sin(theta/2)|1>`` with ``theta = 2 eps/omega sin(omega t)``),
a two-segment piecewise-constant Hamiltonian against the product of matrix
exponentials (``scipy.linalg.expm``), and the oracle ``mesolve`` without jump
operators on ``|psi><psi|``.
"""

from __future__ import annotations

import numpy as np

from . import tsit5
from .mesolve import _segment_time, as_timeop


def sesolve(
    H,
    psi0: np.ndarray,
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
    t0: float | None = None,
    save_states: bool = True,
    dt: float | None = None,
) -> dict:
    H = as_timeop(H)
    psi0 = np.asarray(psi0, dtype=np.complex128).reshape(-1, 1)
    tsave = np.asarray(tsave, dtype=np.float64)
    Es = [np.asarray(E, dtype=np.complex128) for E in (exp_ops or [])]
    jumps = H.jump_ts

    def save(y):
        ex = np.array([(y.conj().T @ E @ y)[0, 0] for E in Es], dtype=np.complex128)
        return (y.copy() if save_states else None, ex)

    def f(t, y, tseg):
        ts = _segment_time(t, tseg, jumps)
        return -1j * (H(ts, t) @ y)

    if method == 'euler':
        # diffrax Euler: psi + dt f(t, psi) at the step start, fixed steps, linear
        # interpolation to tsave (the fixed-step loop of oracle.rouchon)
        from . import rouchon

        if dt is None:
            raise ValueError('Euler needs dt')
        saved, ylast, nsteps = rouchon.integrate(lambda t, h, y: y + h * f(t, y, t), psi0, tsave, dt, t0=t0,
                                                 save=save)
        states = np.stack([s[0] for s in saved]) if save_states else ylast[None]
        expects = np.stack([s[1] for s in saved], axis=-1) if Es else None
        return dict(states=states, final_state=ylast, expects=expects, infos=dict(nsteps=nsteps))
    if method not in ('tsit5', 'dopri5'):
        raise ValueError(method)
    saved, ylast, st = tsit5.integrate(
        f,
        psi0,
        tsave,
        t0=t0,
        rtol=rtol,
        atol=atol,
        safety=safety,
        factormin=factormin,
        factormax=factormax,
        max_steps=max_steps,
        jump_ts=jumps,
        save=save,
        tab=tsit5.tableau(method),
    )
    infos = dict(nsteps=st.nsteps, naccepted=st.naccepted, nrejected=st.nrejected, nrhs=st.nrhs, ends=list(st.ends))
    states = np.stack([s[0] for s in saved]) if save_states else ylast[None]
    expects = np.stack([s[1] for s in saved], axis=-1) if Es else None
    return dict(states=states, final_state=ylast, expects=expects, infos=infos)
