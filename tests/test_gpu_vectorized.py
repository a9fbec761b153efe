"""``Options(vectorized=True)``: the Liouvillian superoperator integrated as a GEMV
on the sesolve kernel (``vector_field_vec``, diffrax_integrator.py:325-327), against
the superoperator exponential for a constant generator, against the oracle's
vectorized form with piecewise-constant and modulated Hamiltonian terms, and
against the default (Hermitian, unvectorized) GPU path. Tolerance 1e-5 (Frobenius,
per saved state) at rtol = 1e-8."""

import numpy as np
import pytest
import scipy.linalg as sla

pytestmark = pytest.mark.gpu

TOL = 1e-5


def _random_problem(n, K, seed):
    from oracle import operators as op

    rng = np.random.default_rng(seed)
    A = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
    H = (A + A.conj().T) / 2
    Ls = [0.3 * (rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))) / np.sqrt(n) for _ in range(K)]
    psi = rng.normal(size=(n, 1)) + 1j * rng.normal(size=(n, 1))
    psi /= np.linalg.norm(psi)
    return H, Ls, psi, op


def test_vectorized_constant_matches_superoperator_expm():
    import dynamiqs_amd as dq
    from oracle import lindblad

    H, Ls, psi, _ = _random_problem(6, 2, 0)
    tsave = np.linspace(0.0, 2.0, 9)
    res = dq.mesolve(H, Ls, psi, tsave, exp_ops=[H], method=dq.method.Tsit5(rtol=1e-8, atol=1e-10),
                     options=dq.Options(vectorized=True))
    S = lindblad.slindbladian(H, Ls)
    r0 = lindblad.vectorize(psi @ psi.conj().T)
    want = np.stack([lindblad.unvectorize(sla.expm(S * t) @ r0) for t in tsave])
    got = res.states.to_numpy()
    assert got.shape == (9, 6, 6)
    err = np.linalg.norm(got - want, axis=(-2, -1)).max()
    assert err <= TOL, err
    ex = np.array([np.trace(H @ w) for w in want])
    assert np.abs(res.expects[0] - ex).max() <= TOL
    assert int(res.infos.nsteps) > 0


def test_vectorized_time_dependent_matches_oracle_and_default_path():
    import dynamiqs_amd as dq
    from oracle import mesolve as om

    H, Ls, psi, op = _random_problem(5, 1, 1)
    rng = np.random.default_rng(2)
    X = np.diag(np.arange(5.0)) + 0.2
    times = np.linspace(0.0, 1.5, 7)
    vals = rng.normal(size=6)
    f = lambda t: np.cos(2.0 * t)  # noqa: E731
    tsave = np.linspace(0.0, 1.5, 11)
    Hq = dq.constant(H) + dq.pwc(times, vals, X) + dq.modulated(f, 0.5 * H)
    m = dq.method.Tsit5(rtol=1e-9, atol=1e-11)
    vec = dq.mesolve(Hq, Ls, psi, tsave, method=m, options=dq.Options(vectorized=True)).states.to_numpy()
    Hop = om.TimeOp(H, [om.PWCTerm(times, vals, X), om.ModTerm(f, 0.5 * H)])
    ref = om.mesolve(Hop, Ls, psi, tsave, rtol=1e-10, atol=1e-12, form='vectorized')['states']
    err = np.linalg.norm(vec - ref, axis=(-2, -1)).max()
    assert err <= TOL, err
    default = dq.mesolve(Hq, Ls, psi, tsave, method=m).states.to_numpy()
    assert np.linalg.norm(vec - default, axis=(-2, -1)).max() <= TOL


def test_vectorized_batching_and_limits():
    import dynamiqs_amd as dq

    H, Ls, psi, _ = _random_problem(4, 1, 3)
    Hb = np.stack([H, 2 * H, 3 * H])
    tsave = np.linspace(0.0, 1.0, 5)
    m = dq.method.Dopri5(rtol=1e-8, atol=1e-10)
    vec = dq.mesolve(Hb, Ls, psi, tsave, method=m, options=dq.Options(vectorized=True, save_states=False))
    assert vec.states.shape == (3, 1, 4, 4)
    for b in range(3):
        solo = dq.mesolve(Hb[b], Ls, psi, tsave, method=m, options=dq.Options(vectorized=True))
        np.testing.assert_array_equal(vec.states.to_numpy()[b, 0], solo.states.to_numpy()[-1])
    a = dq.destroy(33)
    with pytest.raises(NotImplementedError, match='vectorized'):
        dq.mesolve(a.dag() @ a, [a], dq.fock(33, 1), tsave, options=dq.Options(vectorized=True))
