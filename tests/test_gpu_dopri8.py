"""GPU parity of the Dopri8 method (``DQH_FLAG_DOPRI8``: the Dormand-Prince 8(7)
pair on the dense global-memory body, steps landing on the save times) against
the complex128 Dopri8 oracle (``oracle.dopri8``, the same algorithm), against the
superoperator exponential, and through ``dq.mesolve``. Bar: 1e-5 Frobenius at
every save time."""

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


@pytest.mark.parametrize('n,K', [(6, 1), (20, 2), (64, 3), (100, 2)])
def test_dopri8_dense_matches_fp64_oracle_and_expm(n, K):
    import scipy.linalg as sla

    from dynamiqs_amd.integrators import dense
    from oracle import lindblad
    from oracle import mesolve as om

    rng = np.random.default_rng(80 + n + K)
    insts = [_rand_instance(rng, n, K) for _ in range(2)]
    tsave = np.linspace(0.0, 2.0, 9)
    plan = dense.build_plan('tsit5', insts, tsave, rtol=1e-6, atol=1e-6, tableau='dopri8', step_log_cap=4000)
    assert plan.uniform_np == 0 and not plan.fp64_diag and plan.grid is None
    assert np.all(plan.descs['flags'] & 128)
    plan.run()
    out = plan.results()
    assert np.all(out['status'] == 0)
    for b, inst in enumerate(insts):
        ref = om.mesolve(inst.H, inst.Ls, inst.rho0, tsave, method='dopri8', rtol=1e-6, atol=1e-6)
        err = np.linalg.norm(out['states'][b] - ref['states'], axis=(-2, -1)).max()
        nref = int(ref['infos']['nsteps'])
        print(f'dopri8 n={n} K={K} b={b}: err vs oracle {err:.2e}, steps {int(out["nsteps"][b])} (oracle {nref})')
        assert err <= TOL, (n, b, err)
        # the 8(7) error estimate cancels heavily: on FP32 stages its rounding floor
        # (about 1e-7 of |f| h) keeps the GPU's proposals below the oracle's, so the
        # GPU takes at least as many steps, never fewer than the save intervals
        assert int(out['naccepted'][b]) >= len(tsave) - 1 and int(out['nsteps'][b]) >= nref - 2
        # every save time after t0 is the end of an accepted step
        ends = np.asarray(out['step_ends'][b])
        assert all(np.min(np.abs(ends - t)) < 1e-12 for t in tsave[1:])
        if n <= 20:  # the superoperator exponential is n^2 x n^2
            Lsup = lindblad.slindbladian(inst.H, inst.Ls)
            ex = np.stack([lindblad.unvectorize(sla.expm(t * Lsup) @ lindblad.vectorize(inst.rho0)) for t in tsave])
            assert np.linalg.norm(out['states'][b] - ex, axis=(-2, -1)).max() <= 2e-5


def test_dopri8_through_mesolve_pwc_and_closed_form():
    """dq.mesolve with Dopri8: the lossy cavity against its closed form (coherent
    state alpha0 e^{-(i + kappa/2) t}), and a PWC drive against the oracle."""
    import dynamiqs_amd as dq
    from oracle import mesolve as om
    from oracle import operators as op

    n, kappa = 16, 0.1
    a = dq.destroy(n)
    tsave = np.linspace(0, 2 * np.pi, 11)
    res = dq.mesolve(a.dag() @ a, [np.sqrt(kappa) * a], dq.coherent(n, 1.0), tsave,
                     method=dq.method.Dopri8(rtol=1e-8, atol=1e-10))
    st = res.states.to_numpy()
    for i, t in enumerate(tsave):
        al = np.exp(-1j * t - 0.5 * kappa * t)
        assert np.linalg.norm(st[i] - op.coherent_dm(n, al)) < 1e-5
    assert int(res.infos.nsteps) >= len(tsave) - 1
    # PWC Hamiltonian
    times = np.array([0.0, 0.7, 1.3, 2.0])
    vals = np.array([0.5, -1.0, 0.8])
    H = dq.pwc(times, vals, a + a.dag())
    r = dq.mesolve(H, [np.sqrt(kappa) * a], dq.fock(n, 0), np.linspace(0, 2, 5), method=dq.method.Dopri8())
    A = op.destroy(n)
    Hop = om.TimeOp(np.zeros((n, n), dtype=complex), [om.PWCTerm(times, vals, A + A.conj().T)])
    ref = om.mesolve(Hop, [np.sqrt(kappa) * A], op.fock(n, 0), np.linspace(0, 2, 5), method='dopri8')
    err = np.linalg.norm(r.states.to_numpy() - ref['states'], axis=(-2, -1)).max()
    print(f'dopri8 PWC drive: err vs oracle {err:.2e}')
    assert err <= TOL
