"""GPU parity of the Dopri5 method (``DQH_FLAG_DOPRI5``: the Tsit5 kernels with the
Dormand-Prince pair and its quartic dense output) against the complex128 Dopri5
oracle, in every kernel form: the LDS-resident body (n <= 64), the global-memory
body, the grid-wide solve and the DIA sweeps. Bar: 1e-5 Frobenius at every save
time against the oracle's adaptive run, and against the oracle replaying the GPU's
own accepted steps where the step log is available."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TOL = 1e-5


def _rand_instance(rng, n, K):
    from dynamiqs_amd.integrators.dense import Instance

    H = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
    H = (H + H.conj().T) / (2 * np.sqrt(n))
    Ls = [0.3 * (rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))) / np.sqrt(n) for _ in range(K)]
    psi = rng.normal(size=(n, 1)) + 1j * rng.normal(size=(n, 1))
    psi /= np.linalg.norm(psi)
    return Instance(H=H, Ls=Ls, rho0=psi @ psi.conj().T)


def _replay(inst, tsave, ends):
    from oracle import lindblad, tsit5

    f = lambda t, y, tseg: lindblad.rhs_hermitian(inst.H, inst.Ls, y)  # noqa: E731
    saved, _ = tsit5.integrate_steps(f, inst.rho0, tsave, ends, tab=tsit5.tableau('dopri5'))
    return np.stack(saved)


@pytest.mark.parametrize('n,K,form', [(20, 2, 'resident'), (64, 1, 'resident'), (40, 3, 'global'), (100, 2, 'auto')])
def test_dopri5_dense_matches_fp64_oracle_and_replay(n, K, form):
    from dynamiqs_amd.integrators import dense
    from oracle import mesolve as om

    rng = np.random.default_rng(n + K)
    insts = [_rand_instance(rng, n, K) for _ in range(2)]
    tsave = np.linspace(0.0, 2.0, 9)
    plan = dense.build_plan('tsit5', insts, tsave, rtol=1e-6, atol=1e-6, form=form, grid_tsit5=False,
                            step_log_cap=4000, tableau='dopri5')
    plan.run()
    out = plan.results()
    assert np.all(out['status'] == 0)
    for b, inst in enumerate(insts):
        ref = om.mesolve(inst.H, inst.Ls, inst.rho0, tsave, method='dopri5', rtol=1e-6, atol=1e-6)
        err = np.linalg.norm(out['states'][b] - ref['states'], axis=(-2, -1)).max()
        assert err <= TOL, (n, form, b, err)
        nref = int(ref['infos']['nsteps'])
        assert abs(int(out['nsteps'][b]) - nref) <= max(2, nref // 20), (out['nsteps'][b], nref)
        rep = _replay(inst, tsave, out['step_ends'][b])
        assert np.linalg.norm(out['states'][b] - rep, axis=(-2, -1)).max() <= 2e-6
    # a different pair from Tsit5: its own step sequence
    tplan = dense.build_plan('tsit5', insts, tsave, rtol=1e-6, atol=1e-6, form=form, grid_tsit5=False)
    tplan.run()
    assert not np.array_equal(tplan.results()['nsteps'], out['nsteps']) or not np.allclose(
        tplan.results()['states'], out['states'], atol=0)


def test_dopri5_grid_matches_fp64_replay():
    from dynamiqs_amd.integrators import dense

    rng = np.random.default_rng(7)
    insts = [_rand_instance(rng, 144, 2)]
    tsave = np.linspace(0.0, 1.5, 6)
    plan = dense.build_plan('tsit5', insts, tsave, rtol=1e-6, atol=1e-6, grid_tsit5=True, step_log_cap=4000,
                            tableau='dopri5')
    assert plan.grid is not None
    plan.run()
    out = plan.results()
    assert np.all(out['status'] == 0)
    rep = _replay(insts[0], tsave, out['step_ends'][0])
    assert np.linalg.norm(out['states'][0] - rep, axis=(-2, -1)).max() <= 2e-6


def test_dopri5_dia_matches_fp64_oracle():
    from dynamiqs_amd.integrators import dia
    from oracle import mesolve as om
    from oracle import operators as op

    na = nb = 12
    a, b = op.destroy(na, nb)
    H = 5 * a.conj().T @ a + 6 * b.conj().T @ b + 0.1 * (a.conj().T @ b + a @ b.conj().T)
    Ls = [np.sqrt(0.02) * a, np.sqrt(0.05) * b]
    rho0 = op.coherent_dm((na, nb), (1.0, 0.0))[None]
    tsave = np.linspace(0.0, 0.3, 5)
    plan = dia.tsit5_plan(H, Ls, rho0, tsave, tableau='dopri5')
    plan.run()
    out = plan.results()
    assert np.all(out['status'] == 0)
    ref = om.mesolve(H, Ls, rho0[0], tsave, method='dopri5', rtol=1e-6, atol=1e-6)
    assert np.linalg.norm(out['states'][0] - ref['states'], axis=(-2, -1)).max() <= TOL
    nref = int(ref['infos']['nsteps'])
    assert abs(int(out['nsteps'][0]) - nref) <= max(2, nref // 20)


def test_mesolve_dopri5_api():
    """``dq.mesolve(..., method=dq.method.Dopri5(...))`` on the first configuration's
    system against the oracle's Dopri5, the result's method name and infos."""
    import dynamiqs_amd as dq
    from oracle import mesolve as om
    from oracle import operators as op

    n = 16
    a = dq.destroy(n)
    tsave = np.linspace(0.0, 2 * np.pi, 41)
    res = dq.mesolve(a.dag() @ a, [np.sqrt(0.1) * a], dq.coherent(n, 1.0), tsave, exp_ops=[dq.number(n)],
                     method=dq.method.Dopri5(rtol=1e-6, atol=1e-8))
    ao = op.destroy(n)
    ref = om.mesolve(ao.conj().T @ ao, [np.sqrt(0.1) * ao], op.coherent(n, 1.0), tsave, method='dopri5',
                     rtol=1e-6, atol=1e-8, exp_ops=[op.number(n)])
    st = res.states.to_numpy()
    assert np.linalg.norm(st - ref['states'], axis=(-2, -1)).max() <= TOL
    np.testing.assert_allclose(res.expects[0], ref['expects'][0], atol=TOL)
    assert abs(int(res.infos.nsteps) - ref['infos']['nsteps']) <= 2
    assert str(res).startswith('==== MESolveResult ====\nMethod  : Dopri5\n')
