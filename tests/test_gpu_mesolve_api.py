"""GPU checks of the drop-in ``mesolve``: the output shapes of cartesian and flat
batching (``tests/mesolve/test_batching.py:20-87`` of the reference, same shape
cases), every batch element equal to its own single solve, and the result
object (infos shapes, ``final_state``, ``save_states=False``)."""

import numpy as np
import pytest

import dynamiqs_amd as dq

pytestmark = pytest.mark.gpu


def _rand_op(rng, n, batch, hermitian=True):
    x = rng.normal(size=(*batch, n, n)) + 1j * rng.normal(size=(*batch, n, n))
    if hermitian:
        x = x + np.swapaxes(x.conj(), -1, -2)
    return x


def _rand_ket(rng, n, batch):
    x = rng.normal(size=(*batch, n, 1)) + 1j * rng.normal(size=(*batch, n, 1))
    return x / np.linalg.norm(x, axis=(-2, -1), keepdims=True)


def _args(n, nH, nLs, npsi0, nEs, seed=42):
    rng = np.random.default_rng(seed)
    H = _rand_op(rng, n, nH)
    Ls = [_rand_op(rng, n, nL, hermitian=False) for nL in nLs]
    psi0 = _rand_ket(rng, n, npsi0)
    Es = _rand_op(rng, n, (nEs,), hermitian=False)
    return H, Ls, psi0, list(Es)


@pytest.mark.parametrize('nH,npsi0,nL1,nL2', [((), (), (), ()), ((3,), (5,), (), (9,)), ((3, 4), (), (7, 8), ()),
                                              ((3,), (5,), (7, 8), (9,))])
def test_cartesian_batching_shapes(nH, npsi0, nL1, nL2):
    n, nEs, ntsave = 2, 10, 11
    H, Ls, psi0, Es = _args(n, nH, [nL1, nL2], npsi0, nEs)
    tsave = np.linspace(0, 0.01, ntsave)
    res = dq.mesolve(H, Ls, psi0, tsave, exp_ops=Es)
    assert res.states.shape == (*nH, *nL1, *nL2, *npsi0, ntsave, n, n)
    assert res.expects.shape == (*nH, *nL1, *nL2, *npsi0, nEs, ntsave)
    assert np.shape(res.infos.nsteps) == (*nH, *nL1, *nL2, *npsi0)


@pytest.mark.parametrize('nL1', [(), (5, 1, 4)])
@pytest.mark.parametrize('npsi0', [(), (1,), (4,), (3, 1), (3, 4), (5, 1, 4)])
def test_flat_batching_shapes(nL1, npsi0):
    n, nH, nEs, ntsave = 2, (3, 4), 6, 11
    H, Ls, psi0, Es = _args(n, nH, [nL1, ()], npsi0, nEs)
    tsave = np.linspace(0, 0.01, ntsave)
    res = dq.mesolve(H, Ls, psi0, tsave, exp_ops=Es, options=dq.Options(cartesian_batching=False))
    shape = np.broadcast_shapes(nH, nL1, npsi0)
    assert res.states.shape == (*shape, ntsave, n, n)
    assert res.expects.shape == (*shape, nEs, ntsave)


def test_batch_elements_equal_single_solves():
    """Cartesian batch (H: 2, L: 3, rho0: 2) against 12 single mesolve calls."""
    n = 4
    H, Ls, psi0, _ = _args(n, (2,), [(3,)], (2,), 1, seed=3)
    tsave = np.linspace(0, 0.5, 4)
    res = dq.mesolve(H, Ls, psi0, tsave)
    got = res.states.to_numpy()
    for i in range(2):
        for j in range(3):
            for k in range(2):
                one = dq.mesolve(H[i], [Ls[0][j]], psi0[k], tsave)
                np.testing.assert_allclose(got[i, j, k], one.states.to_numpy(), atol=1e-6)


def test_result_fields_and_save_states_false():
    n = 6
    a = dq.destroy(n)
    H = a.dag() @ a
    tsave = np.linspace(0, 1, 5)
    res = dq.mesolve(H, [0.3 * a], dq.coherent(n, 0.8), tsave, exp_ops=[H],
                     options=dq.Options(save_states=False))
    # the reference keeps only the final state, nsave = 1 (mesolve.py:184-187)
    assert res.states.shape == (1, n, n)
    full = dq.mesolve(H, [0.3 * a], dq.coherent(n, 0.8), tsave, exp_ops=[H])
    assert full.states.shape == (5, n, n)
    np.testing.assert_allclose(np.asarray(full.final_state.to_numpy()), full.states.to_numpy()[-1], atol=0)
    np.testing.assert_allclose(res.states.to_numpy()[0], full.states.to_numpy()[-1], atol=1e-6)
    np.testing.assert_allclose(res.final_state.to_numpy(), full.states.to_numpy()[-1], atol=1e-6)
    np.testing.assert_allclose(res.expects, full.expects, atol=1e-6)
    s = str(full)
    assert '|' in s or 'MESolveResult' in s


def _expm_solution(H, Ls, rho0, tsave):
    """Exact propagation rho(t) = exp(t L) rho0 with the superoperator in
    complex128 (linear in rho0, so valid for non-Hermitian states)."""
    from scipy.linalg import expm

    from oracle import lindblad

    S = lindblad.slindbladian(H, Ls)
    v0 = lindblad.vectorize(rho0)
    return np.stack([lindblad.unvectorize(expm(t * S) @ v0) for t in tsave])


def test_non_hermitian_rho0_tsit5_and_rouchon1():
    """assume_hermitian=False (Tsit5) and Rouchon1 (which assumes nothing) on a
    non-Hermitian initial state, against exact propagation: the host splits rho0
    into Hermitian parts A + iB, solves both on the GPU and recombines states and
    expectation values."""
    rng = np.random.default_rng(8)
    n = 6
    a = np.diag(np.sqrt(np.arange(1, n)), k=1).astype(complex)
    H = a.conj().T @ a + 0.3 * (a + a.conj().T)
    Ls = [0.4 * a]
    x = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
    rho0 = x / np.linalg.norm(x)
    E = a.conj().T @ a
    tsave = np.linspace(0.0, 1.0, 6)
    ref = _expm_solution(H, Ls, rho0, tsave)
    ref_e = np.einsum('ij,tji->t', E, ref)
    res = dq.mesolve(H, Ls, rho0, tsave, exp_ops=[E], options=dq.Options(assume_hermitian=False))
    np.testing.assert_allclose(res.states.to_numpy(), ref, atol=2e-5)
    np.testing.assert_allclose(res.expects[0], ref_e, atol=1e-4)
    # Rouchon1 (first order): against exact propagation at its own O(dt) accuracy
    # and against the complex128 Rouchon1 oracle at FP32 accuracy
    from oracle import mesolve as omesolve

    dt = 1e-3
    r1 = dq.mesolve(H, Ls, rho0, tsave, exp_ops=[E], method=dq.method.Rouchon1(dt=dt))
    ro = omesolve.mesolve(H, Ls, rho0, tsave, method='rouchon1', dt=dt, exp_ops=[E])
    np.testing.assert_allclose(r1.states.to_numpy(), ro['states'], atol=1e-5)
    np.testing.assert_allclose(r1.states.to_numpy(), ref, atol=5e-3)
    # a batch mixing a Hermitian and a non-Hermitian state keeps each element's own answer
    herm = 0.5 * (rho0 + rho0.conj().T)
    rb = dq.mesolve(H, Ls, np.stack([herm, rho0]), tsave, options=dq.Options(assume_hermitian=False))
    np.testing.assert_allclose(rb.states.to_numpy()[1], res.states.to_numpy(), atol=1e-6)
    np.testing.assert_allclose(rb.states.to_numpy()[0], _expm_solution(H, Ls, herm, tsave), atol=2e-5)


def test_save_extra_every_save_time_with_batching_and_no_states():
    """Options(save_extra=fn) sees the state at every save time, also with
    save_states=False (save_mixin.py:47-54, 62-66). Its outputs have shape
    (*batch, ntsave, ...) and equal fn applied to the states of the same solve with
    save_states=True."""
    n = 5
    a = dq.destroy(n)
    H = np.stack([(a.dag() @ a).to_numpy() * w for w in (1.0, 1.5, 2.0)])
    Ls = [0.2 * a]
    num = (a.dag() @ a).to_numpy()
    psi = dq.coherent(n, 0.7)
    tsave = np.linspace(0.0, 1.0, 7)

    def fn(rho):
        r = rho.to_numpy()
        return {'pop': np.real(np.trace(num @ r)), 'corner': rho[0, 1]}

    full = dq.mesolve(H, Ls, psi, tsave)
    res = dq.mesolve(H, Ls, psi, tsave, options=dq.Options(save_extra=fn, save_states=False))
    assert res.states.shape == (3, 1, n, n)
    assert res.extra['pop'].shape == (3, len(tsave))
    st = full.states.to_numpy()
    np.testing.assert_allclose(res.extra['pop'], np.real(np.einsum('ij,btji->bt', num, st)), atol=1e-6)
    np.testing.assert_allclose(np.asarray(res.extra['corner'].to_numpy()), st[:, :, 0, 1], atol=1e-6)
    np.testing.assert_allclose(res.states.to_numpy()[:, 0], st[:, -1], atol=1e-6)


def test_rouchon2_fixed_step_matches_oracle():
    """dq.method.Rouchon2(dt) on the dense Kraus kernel (1 + 2K + K^2 normalised
    Kraus operators folded on the host) against the complex128 oracle's
    restatement of the reference's Kraus-Heun map, with and without normalisation,
    on a batch of initial states; trace preserved with normalisation."""
    from oracle import mesolve as omesolve

    n = 8
    a = dq.destroy(n)
    H = a.dag() @ a + 0.3 * (a + a.dag())
    Ls = [0.5 * a, 0.2 * a.dag() @ a]
    rho0 = np.stack([dq.coherent_dm(n, al).to_numpy() for al in (0.3, 0.9)])
    tsave = np.linspace(0.0, 1.0, 6)
    for normalize in (True, False):
        res = dq.mesolve(H, Ls, rho0, tsave, exp_ops=[a.dag() @ a],
                         method=dq.method.Rouchon2(dt=1e-2, normalize=normalize))
        st = res.states.to_numpy()
        for b in range(2):
            ref = omesolve.mesolve(H.to_numpy(), [L.to_numpy() for L in Ls], rho0[b], tsave, method='rouchon2',
                                   dt=1e-2, normalize=normalize, exp_ops=[(a.dag() @ a).to_numpy()])
            assert np.linalg.norm(st[b] - ref['states'], axis=(-2, -1)).max() <= 1e-5
            np.testing.assert_allclose(res.expects[b, 0], ref['expects'][0], atol=1e-5)
        if normalize:
            assert np.abs(np.trace(st[:, -1], axis1=-2, axis2=-1) - 1).max() < 1e-5
    assert str(res).split('\n')[1] == 'Method  : Rouchon2'


def test_rouchon3_fixed_step_matches_oracle():
    """dq.method.Rouchon3(dt) on the dense Kraus kernel (1 + 2K + K^2 + K^3
    normalised Kraus operators folded on the host) against the oracle's stage-form
    restatement of KrausHeun3 with the Bosh3 no-jump propagator, with and without
    normalisation, on a batch of initial states; closer to expm than Rouchon2 at
    the same dt."""
    import scipy.linalg as sla

    from oracle import lindblad
    from oracle import mesolve as omesolve

    n = 8
    a = dq.destroy(n)
    H = a.dag() @ a + 0.3 * (a + a.dag())
    Ls = [0.5 * a, 0.2 * a.dag() @ a]
    rho0 = np.stack([dq.coherent_dm(n, al).to_numpy() for al in (0.3, 0.9)])
    tsave = np.linspace(0.0, 1.0, 6)
    for normalize in (True, False):
        res = dq.mesolve(H, Ls, rho0, tsave, exp_ops=[a.dag() @ a],
                         method=dq.method.Rouchon3(dt=2e-2, normalize=normalize))
        st = res.states.to_numpy()
        for b in range(2):
            ref = omesolve.mesolve(H.to_numpy(), [L.to_numpy() for L in Ls], rho0[b], tsave, method='rouchon3',
                                   dt=2e-2, normalize=normalize, exp_ops=[(a.dag() @ a).to_numpy()])
            err = np.linalg.norm(st[b] - ref['states'], axis=(-2, -1)).max()
            print('rouchon3 normalize', normalize, 'instance', b, 'err vs oracle', err)
            assert err <= 1e-5
            np.testing.assert_allclose(res.expects[b, 0], ref['expects'][0], atol=1e-5)
        if normalize:
            assert np.abs(np.trace(st[:, -1], axis1=-2, axis2=-1) - 1).max() < 1e-5
    S = lindblad.slindbladian(H.to_numpy(), [L.to_numpy() for L in Ls])
    exact = lindblad.unvectorize(sla.expm(S) @ lindblad.vectorize(rho0[1]))
    r3 = dq.mesolve(H, Ls, rho0[1], tsave, method=dq.method.Rouchon3(dt=2e-2)).states.to_numpy()[-1]
    r2 = dq.mesolve(H, Ls, rho0[1], tsave, method=dq.method.Rouchon2(dt=2e-2)).states.to_numpy()[-1]
    assert np.linalg.norm(r3 - exact) < np.linalg.norm(r2 - exact)
    assert str(res).split('\n')[1] == 'Method  : Rouchon3'


def test_euler_fixed_step_matches_oracle():
    """dq.method.Euler(dt) for mesolve on the dense Kraus kernel (sets D = [dt G,
    sqrt(dt) L_k], C = [I, sqrt(dt) L_k^dag / 2], the operators at the step start)
    against the oracle's Euler, with a PWC drive; trace preserved."""
    from oracle import mesolve as omesolve

    n = 6
    a = dq.destroy(n)
    H0 = (a.dag() @ a).to_numpy()
    X = (a + a.dag()).to_numpy()
    times, vals = np.array([0.0, 0.33, 0.8]), np.array([0.4, -0.7])
    H = dq.asqarray(H0) + dq.pwc(times, vals, dq.asqarray(X))
    Ls = [0.4 * a]
    rho0 = dq.coherent_dm(n, 0.6).to_numpy()
    tsave = np.linspace(0.0, 1.0, 5)
    res = dq.mesolve(H, Ls, rho0, tsave, method=dq.method.Euler(dt=1e-3))
    Hop = omesolve.TimeOp(H0, [omesolve.PWCTerm(times, vals, X)])
    ref = omesolve.mesolve(Hop, [L.to_numpy() for L in Ls], rho0, tsave, method='euler', dt=1e-3)
    err = np.linalg.norm(res.states.to_numpy() - ref['states'], axis=(-2, -1)).max()
    print('euler mesolve err vs oracle', err)
    assert err <= 1e-5
    assert abs(np.trace(res.states.to_numpy()[-1]) - 1) < 1e-5
    assert int(res.infos.nsteps) == 1000
