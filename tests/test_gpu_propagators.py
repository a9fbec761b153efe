"""GPU parity of ``dq.sepropagator`` / ``dq.mepropagator`` (batches of basis-state
solves on ``dqh_sesolve_tsit5`` and the dense Tsit5 kernels) against matrix
exponentials of the generator (``scipy.linalg.expm`` of -iHt and of the oracle's
``slindbladian`` times t, the closed form for constant and piecewise-constant
generators), against ``dq.mesolve`` applied to a state, and the reference's shapes."""

import numpy as np
import pytest
import scipy.linalg as sla

pytestmark = pytest.mark.gpu

TIGHT = dict(rtol=1e-8, atol=1e-8)


def _herm(rng, n):
    x = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
    return (x + x.conj().T) / (2 * np.sqrt(n))


def test_sepropagator_constant_and_batched():
    import dynamiqs_amd as dq

    rng = np.random.default_rng(1)
    n = 5
    Hs = np.stack([_herm(rng, n) for _ in range(2)])
    ts = np.linspace(0.0, 2.0, 6)
    res = dq.sepropagator(Hs, ts, method=dq.method.Tsit5(**TIGHT))
    assert res.propagators.shape == (2, 6, n, n) and res.final_propagator.shape == (2, n, n)
    U = res.propagators.to_numpy()
    err = max(np.max(np.abs(U[b, k] - sla.expm(-1j * Hs[b] * t))) for b in range(2) for k, t in enumerate(ts))
    print('sepropagator max |U - expm| =', err)
    assert err <= 1e-5
    fin = dq.sepropagator(Hs[0], ts, options=dq.Options(save_propagators=False))
    assert fin.propagators.shape == (1, n, n)
    np.testing.assert_allclose(fin.final_propagator.to_numpy(), sla.expm(-1j * Hs[0] * ts[-1]), atol=1e-4)
    assert 'SEPropagatorResult' in str(res)


def test_sepropagator_pwc():
    """H0 + v(t) O with two segments: the product of segment exponentials."""
    import dynamiqs_amd as dq

    rng = np.random.default_rng(2)
    n = 4
    H0, O = _herm(rng, n), _herm(rng, n)
    H = dq.asqarray(H0) + dq.pwc(np.array([0.0, 0.5, 2.0]), np.array([1.0, -2.0]), dq.asqarray(O))
    ts = np.array([0.0, 0.5, 0.8])
    U = dq.sepropagator(H, ts, method=dq.method.Tsit5(**TIGHT)).propagators.to_numpy()
    exact = sla.expm(-1j * (H0 - 2 * O) * 0.3) @ sla.expm(-1j * (H0 + O) * 0.5)
    err = np.max(np.abs(U[2] - exact))
    print('PWC sepropagator max |U - exact| =', err)
    assert err <= 1e-5
    np.testing.assert_allclose(U[1], sla.expm(-1j * (H0 + O) * 0.5), atol=1e-5)


def test_mepropagator_reference_example_against_expm():
    """The reference docstring example (mepropagator.py:143-165): n = 16, H = a^dag a,
    jump_ops = [a], 11 save times on [0, 1] -> propagators (11, 256, 256), here equal
    to expm(t slindbladian) to the method's tolerance."""
    import dynamiqs_amd as dq
    from oracle import lindblad
    from oracle import operators as op

    n = 16
    a = dq.destroy(n)
    ts = np.linspace(0.0, 1.0, 11)
    res = dq.mepropagator(a.dag() @ a, [a], ts, method=dq.method.Tsit5(**TIGHT))
    assert res.propagators.shape == (11, n * n, n * n)
    ao = op.destroy(n)
    S = lindblad.slindbladian(ao.conj().T @ ao, [ao])
    P = res.propagators.to_numpy()
    err = max(np.max(np.abs(P[k] - sla.expm(S * t))) for k, t in enumerate(ts))
    print('mepropagator n=16 max |P - expm| =', err)
    assert err <= 1e-5
    assert 'MEPropagatorResult' in str(res)


def test_mepropagator_batched_random_and_consistent_with_mesolve():
    """H batch (2,) x L batch (3,): shapes, each element vs expm(t S), and
    P(t) vec(rho0) = vec(mesolve(rho0)(t)) on a random density matrix."""
    import dynamiqs_amd as dq
    from oracle import lindblad

    rng = np.random.default_rng(3)
    n = 4
    Hs = np.stack([_herm(rng, n) for _ in range(2)])
    Ls = np.stack([0.5 * (rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))) / np.sqrt(n) for _ in range(3)])
    L2 = 0.3 * np.diag(np.arange(n, dtype=float)).astype(complex)
    ts = np.linspace(0.0, 1.5, 4)
    res = dq.mepropagator(Hs, [Ls, L2], ts, method=dq.method.Tsit5(**TIGHT))
    assert res.propagators.shape == (2, 3, 4, n * n, n * n)
    assert np.shape(res.infos.nsteps) == (2, 3)
    P = res.propagators.to_numpy()
    err = 0.0
    for i in range(2):
        for j in range(3):
            S = lindblad.slindbladian(Hs[i], [Ls[j], L2])
            err = max(err, max(np.max(np.abs(P[i, j, k] - sla.expm(S * t))) for k, t in enumerate(ts)))
    print('batched mepropagator max |P - expm| =', err)
    assert err <= 1e-5
    psi = rng.normal(size=(n, 1)) + 1j * rng.normal(size=(n, 1))
    rho = psi @ psi.conj().T / np.vdot(psi, psi)
    me = dq.mesolve(Hs[1], [Ls[2], L2], rho, ts, method=dq.method.Tsit5(**TIGHT)).states.to_numpy()
    via_p = lindblad.unvectorize(P[1, 2] @ lindblad.vectorize(rho))
    np.testing.assert_allclose(via_p, me, atol=1e-5)
    fin = dq.mepropagator(Hs, [Ls, L2], ts, options=dq.Options(save_propagators=False))
    assert fin.propagators.shape == (2, 3, 1, n * n, n * n)


def test_expm_default_method_mepropagator_reference_example():
    """With method=None a constant Liouvillian takes Expm, as in the reference
    (mepropagator.py:269-270): its docstring example (n = 16, H = a^dag a, [a],
    11 save times on [0, 1]) prints 'Method : Expm', 'Infos : 10 steps' and
    propagators (11, 256, 256); here equal to expm(t slindbladian) within 2e-5
    (complex64 GEMMs, Taylor degree 8 plus squarings)."""
    import time

    import dynamiqs_amd as dq
    from oracle import lindblad
    from oracle import operators as op

    n = 16
    a = dq.destroy(n)
    ts = np.linspace(0.0, 1.0, 11)
    dq.mepropagator(a.dag() @ a, [a], ts)
    t0 = time.perf_counter()
    res = dq.mepropagator(a.dag() @ a, [a], ts)
    print('Expm mepropagator n=16 wall ms', 1e3 * (time.perf_counter() - t0))
    assert type(res.method).__name__ == 'Expm'
    assert int(res.infos.nsteps) == 10
    assert res.propagators.shape == (11, n * n, n * n)
    ao = op.destroy(n)
    S = lindblad.slindbladian(ao.conj().T @ ao, [ao])
    P = res.propagators.to_numpy()
    err = max(np.max(np.abs(P[k] - sla.expm(S * t))) for k, t in enumerate(ts))
    print('Expm mepropagator n=16 max |P - expm| =', err)
    assert err <= 2e-5
    s = str(res)
    assert 'Expm' in s and '10 steps' in s and 'Propagators' in s


def test_expm_sepropagator_batched_and_pwc():
    """Expm on a batch of constant Hamiltonians and on a PWC Hamiltonian whose
    breakpoints fall between the save times (the product of segment exponentials)."""
    import dynamiqs_amd as dq

    rng = np.random.default_rng(7)
    n = 6
    Hs = np.stack([3.0 * _herm(rng, n) for _ in range(3)])
    ts = np.linspace(0.0, 2.0, 5)
    res = dq.sepropagator(Hs, ts)
    assert type(res.method).__name__ == 'Expm' and res.propagators.shape == (3, 5, n, n)
    U = res.propagators.to_numpy()
    err = max(np.max(np.abs(U[b, k] - sla.expm(-1j * Hs[b] * t))) for b in range(3) for k, t in enumerate(ts))
    print('Expm sepropagator batched max err', err)
    assert err <= 2e-5
    H0, O = _herm(rng, n), _herm(rng, n)
    H = dq.asqarray(H0) + dq.pwc(np.array([0.1, 0.7, 1.3]), np.array([2.0, -1.0]), dq.asqarray(O))
    ts = np.array([0.0, 1.0, 2.0])
    res = dq.sepropagator(H, ts)
    assert int(res.infos.nsteps) == 5  # intervals between 0, 0.1, 0.7, 1.0, 1.3, 2.0
    E = lambda M, h: sla.expm(-1j * M * h)  # noqa: E731
    U1 = E(H0 - O, 0.3) @ E(H0 + 2 * O, 0.6) @ E(H0, 0.1)
    U2 = E(H0, 0.7) @ E(H0 - O, 0.3) @ U1
    got = res.propagators.to_numpy()
    np.testing.assert_allclose(got[1], U1, atol=2e-5)
    np.testing.assert_allclose(got[2], U2, atol=2e-5)


def test_expm_mepropagator_batched_jump_operators():
    """Cartesian batching over H (2,) and a jump operator (3,) with Expm, against
    expm(t S) per element; flat batching broadcasts."""
    import dynamiqs_amd as dq
    from oracle import lindblad

    rng = np.random.default_rng(8)
    n = 3
    Hs = np.stack([_herm(rng, n) for _ in range(2)])
    Ls = np.stack([0.4 * (rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))) for _ in range(3)])
    ts = np.linspace(0.0, 1.0, 3)
    res = dq.mepropagator(Hs, [Ls], ts, method=dq.method.Expm())
    assert res.propagators.shape == (2, 3, 3, n * n, n * n)
    P = res.propagators.to_numpy()
    for i in range(2):
        for j in range(3):
            S = lindblad.slindbladian(Hs[i], [Ls[j]])
            for k, t in enumerate(ts):
                np.testing.assert_allclose(P[i, j, k], sla.expm(S * t), atol=2e-5)
    flat = dq.mepropagator(Hs[:, None].repeat(3, axis=1), [Ls], ts, options=dq.Options(cartesian_batching=False))
    assert flat.propagators.shape == (2, 3, 3, n * n, n * n)
    np.testing.assert_allclose(flat.propagators.to_numpy(), P, atol=1e-6)


def test_expm_sesolve_and_mesolve_against_matrix_exponentials():
    """method=Expm() for sesolve and mesolve (the propagator chain applied to the
    initial states on the GPU): cartesian batches over H and the initial state, a
    PWC drive, expectation values, FixedStepInfos with the interval count."""
    import dynamiqs_amd as dq
    from oracle import lindblad

    rng = np.random.default_rng(12)
    n = 5
    Hs = np.stack([_herm(rng, n) for _ in range(2)])
    psis = np.stack([rng.normal(size=(n, 1)) + 1j * rng.normal(size=(n, 1)) for _ in range(3)])
    psis /= np.linalg.norm(psis, axis=1, keepdims=True)
    E = np.diag(np.arange(n, dtype=float)).astype(complex)
    ts = np.linspace(0.0, 1.5, 4)
    res = dq.sesolve(Hs, psis, ts, exp_ops=[E], method=dq.method.Expm())
    assert res.states.shape == (2, 3, 4, n, 1) and res.expects.shape == (2, 3, 1, 4)
    st = res.states.to_numpy()[..., 0]
    for i in range(2):
        for j in range(3):
            for k, t in enumerate(ts):
                ex = sla.expm(-1j * Hs[i] * t) @ psis[j, :, 0]
                np.testing.assert_allclose(st[i, j, k], ex, atol=2e-5)
                np.testing.assert_allclose(res.expects[i, j, 0, k], np.vdot(ex, E @ ex), atol=1e-4)
    assert int(np.asarray(res.infos.nsteps).max()) == 3
    # mesolve: PWC H, one jump operator, two initial states
    a = dq.destroy(n)
    X = (a + a.dag()).to_numpy()
    H0 = (a.dag() @ a).to_numpy()
    H = dq.asqarray(H0) + dq.pwc(np.array([0.0, 0.6, 2.0]), np.array([0.5, -0.5]), dq.asqarray(X))
    L = 0.3 * a.to_numpy()
    rho = np.stack([dq.coherent_dm(n, al).to_numpy() for al in (0.2, 0.8)])
    ts = np.array([0.0, 0.6, 1.0])
    rm = dq.mesolve(H, [L], rho, ts, exp_ops=[dq.asqarray(H0)], method=dq.method.Expm())
    S1 = lindblad.slindbladian(H0 + 0.5 * X, [L])
    S2 = lindblad.slindbladian(H0 - 0.5 * X, [L])
    P06, P10 = sla.expm(0.6 * S1), sla.expm(0.4 * S2) @ sla.expm(0.6 * S1)
    got = rm.states.to_numpy()
    for b in range(2):
        v = lindblad.vectorize(rho[b])
        np.testing.assert_allclose(got[b, 1], lindblad.unvectorize(P06 @ v), atol=2e-5)
        ref = lindblad.unvectorize(P10 @ v)
        np.testing.assert_allclose(got[b, 2], ref, atol=2e-5)
        np.testing.assert_allclose(rm.expects[b, 0, 2], np.trace(H0 @ ref), atol=1e-4)
    assert str(rm).split('\n')[1] == 'Method  : Expm'


def test_euler_propagators_against_oracle():
    """Euler(dt) propagators: sepropagator on the sesolve kernel's fixed-step mode,
    mepropagator on the dense Kraus kernel with the Euler sets, against the oracle's
    column-by-column Euler solves (linear interpolation between steps included)."""
    import dynamiqs_amd as dq
    from oracle import propagators as op_

    rng = np.random.default_rng(14)
    n = 4
    H = _herm(rng, n)
    L = 0.4 * (rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))) / np.sqrt(n)
    ts = np.array([0.0, 0.25, 0.5])
    m = dq.method.Euler(dt=0.01)
    U = dq.sepropagator(H, ts, method=m)
    ref = op_.sepropagator(H, ts, method='euler', dt=0.01)
    np.testing.assert_allclose(U.propagators.to_numpy(), ref, atol=1e-5)
    assert int(U.infos.nsteps) == 50
    P = dq.mepropagator(H, [L], ts, method=m)
    refP = op_.mepropagator(H, [L], ts, method='euler', dt=0.01)
    err = np.max(np.abs(P.propagators.to_numpy() - refP))
    print('Euler mepropagator max err vs oracle', err)
    assert err <= 1e-5


def test_single_save_time_edge_cases():
    """tsave with one time (nothing to integrate): sesolve returns psi0, the Expm
    propagators the identity with zero steps, mesolve with Expm rho0."""
    import dynamiqs_amd as dq

    H = dq.sigmax()
    ts = np.array([0.0])
    r = dq.sesolve(H, dq.fock(2, 0), ts)
    assert r.states.shape == (1, 2, 1)
    np.testing.assert_allclose(r.states.to_numpy()[0, :, 0], [1, 0], atol=1e-7)
    U = dq.sepropagator(H, ts)
    assert type(U.method).__name__ == 'Expm' and int(U.infos.nsteps) == 0
    np.testing.assert_allclose(U.propagators.to_numpy()[0], np.eye(2), atol=0)
    rm = dq.mesolve(H, [0.1 * dq.sigmam()], dq.fock(2, 0), ts, method=dq.method.Expm())
    np.testing.assert_allclose(rm.states.to_numpy()[0], np.diag([1.0, 0.0]), atol=1e-7)
    # t0 before the first save time
    U2 = dq.sepropagator(H, np.array([0.5]), options=dq.Options(t0=0.0))
    np.testing.assert_allclose(U2.propagators.to_numpy()[0], sla.expm(-1j * 0.5 * H.to_numpy()), atol=2e-6)


def test_time_dependent_batches_and_modulated_propagator():
    """sesolve over a batch of PWC control vectors (3) and kets (2): each element
    equals its own single solve bit for bit; mepropagator (Tsit5) of a modulated
    Hamiltonian against the oracle's column-by-column propagator."""
    import dynamiqs_amd as dq
    from oracle import mesolve as om
    from oracle import propagators as op_

    rng = np.random.default_rng(33)
    n = 5
    H0, O = _herm(rng, n), _herm(rng, n)
    times = np.array([0.0, 0.3, 0.7, 1.0])
    vals = rng.normal(size=(3, 3))
    psis = np.stack([rng.normal(size=(n, 1)) + 1j * rng.normal(size=(n, 1)) for _ in range(2)])
    psis /= np.linalg.norm(psis, axis=1, keepdims=True)
    ts = np.linspace(0.0, 1.0, 4)
    H = dq.asqarray(H0) + dq.pwc(times, vals, dq.asqarray(O))
    res = dq.sesolve(H, psis, ts)
    assert res.states.shape == (3, 2, 4, n, 1)
    for i in range(3):
        for j in range(2):
            one = dq.sesolve(dq.asqarray(H0) + dq.pwc(times, vals[i], dq.asqarray(O)), psis[j], ts)
            np.testing.assert_array_equal(res.states.to_numpy()[i, j], one.states.to_numpy())
    f = lambda t: 0.8 * np.sin(2.0 * t)  # noqa: E731
    L = 0.3 * np.diag(np.ones(n - 1), 1).astype(complex)
    P = dq.mepropagator(dq.asqarray(H0) + dq.modulated(f, dq.asqarray(O)), [L], ts,
                        method=dq.method.Tsit5(rtol=1e-8, atol=1e-8)).propagators.to_numpy()
    ref = op_.mepropagator(om.TimeOp(H0, [om.ModTerm(f, O)]), [L], ts, rtol=1e-10, atol=1e-10)
    err = np.max(np.abs(P - ref))
    print('modulated mepropagator max err vs oracle', err)
    assert err < 1e-5
