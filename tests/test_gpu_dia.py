"""GPU parity of the sparse-DIA Rouchon1 path (``dqh_mesolve_rouchon1_dia``)
against the complex128 oracle. Tolerance: 1e-5 in Frobenius norm for every
saved state (FP64 state with complex64 increments)."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TOL = 1e-5


def _two_mode(na, nb_):
    from oracle import operators as op

    a, b = op.destroy(na, nb_)
    ad, bd = a.conj().T, b.conj().T
    H = 5 * ad @ a + 6 * bd @ b + 0.1 * (ad @ b + a @ bd) - 0.01 * ad @ a @ bd @ b
    return H, [np.sqrt(0.02) * a, np.sqrt(0.05) * b], (a, b)


@pytest.mark.parametrize('na,nb_,nsteps', [(8, 9, 300), (12, 12, 200)])
def test_dia_rouchon1_matches_oracle(na, nb_, nsteps):
    from dynamiqs_amd.integrators import dia
    from oracle import mesolve as om
    from oracle import operators as op

    H, Ls, _ = _two_mode(na, nb_)
    dt = 1e-3
    T = nsteps * dt
    tsave = np.linspace(0, T, 6)
    alphas = [0.5, 1.2, 2.0]
    rho0 = np.stack([op.coherent_dm((na, nb_), (al, 0.0)) for al in alphas])
    plan = dia.build_plan(H, Ls, rho0, tsave, dt)
    plan.run()
    out = plan.results()
    for b in range(len(alphas)):
        ref = om.mesolve(H, Ls, rho0[b], tsave, method='rouchon1', dt=dt)
        err = np.linalg.norm(out['states'][b] - ref['states'], axis=(-2, -1)).max()
        assert err <= TOL, (b, err)
        # FP32 increments on an FP64 state: trace drift stays far below complex64 rounding
        assert abs(np.trace(out["final_state"][b]) - 1) < 1e-7


def test_dia_rouchon1_config3_shape_short_run():
    """Full 32 x 32 two-mode size (n = 1024, half-bandwidth 62) for 5 steps vs the
    banded FP64 oracle, plus trace preservation."""
    from dynamiqs_amd.integrators import dia
    from oracle import operators as op
    from oracle import rouchon

    H, Ls, _ = _two_mode(32, 32)
    dt = 1e-3
    nsteps = 5
    tsave = np.linspace(0, nsteps * dt, 2)
    rho0 = np.stack([op.coherent_dm((32, 32), (al, 0.0)) for al in (0.5, 2.0)])
    plan = dia.build_plan(H, Ls, rho0, tsave, dt)
    plan.run()
    out = plan.results()
    M0, Ms, S = rouchon.kraus_ops(H, Ls, dt)
    Tab = rouchon.banded_factor(S, 62)
    for b in range(2):
        r = rho0[b]
        for _ in range(nsteps):
            r = rouchon.kraus_step_banded(H, Ls, r, dt, 62, Tab=Tab)
        err = np.linalg.norm(out['final_state'][b] - r)
        assert err <= TOL, (b, err)
        # FP32 increments on an FP64 state: trace drift stays far below complex64 rounding
        assert abs(np.trace(out["final_state"][b]) - 1) < 1e-7
