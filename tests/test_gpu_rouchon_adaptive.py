"""GPU parity of the adaptive Rouchon pairs (``Rouchon2()`` / ``Rouchon3()`` without
dt, ``dqh_mesolve_rouchon_adaptive_dense``, FP64 on the device) against the
complex128 oracle running the same algorithm (``oracle.rouchon_adaptive``): the
same step sequence and the same states to 1e-9; the reference's own test systems
(``tests/mesolve/test_adaptive_rouchon.py``: dense OCavity and OTDQubit) through
``dq.mesolve`` against their closed forms; and the FP64 MFMA product tile against
numpy."""

import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TOL = 1e-9


def _problem(n=5):
    from oracle import operators as op

    a = op.destroy(n)
    H = a.conj().T @ a + 0.2 * (a + a.conj().T)
    Ls = [0.5 * a, 0.3 * a.conj().T @ a]
    return H, Ls, op.coherent_dm(n, 0.7)


def _rand(rng, n, K, scale=1.0):
    from dynamiqs_amd.integrators.dense import Instance

    H = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
    H = scale * (H + H.conj().T) / (2 * np.sqrt(n))
    Ls = [0.3 * (rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))) / np.sqrt(n) for _ in range(K)]
    psi = rng.normal(size=(n, 1)) + 1j * rng.normal(size=(n, 1))
    psi /= np.linalg.norm(psi)
    return Instance(H=H, Ls=Ls, rho0=psi @ psi.conj().T)


@pytest.mark.parametrize('npd', [16, 32, 48, 64])
def test_fp64_mfma_product_tile(npd):
    import torch

    from dynamiqs_amd import _native

    rng = np.random.default_rng(npd)
    A = rng.normal(size=(3, npd, npd)) + 1j * rng.normal(size=(3, npd, npd))
    B = rng.normal(size=(3, npd, npd)) + 1j * rng.normal(size=(3, npd, npd))
    dA = torch.from_numpy(A.view(np.float64)).cuda()
    dB = torch.from_numpy(B.view(np.float64)).cuda()
    dC = torch.empty_like(dA)
    lib = _native.lib()
    for opA in (0, 1):
        for opB in (0, 1):
            s = torch.cuda.current_stream()
            rc = lib.dqh_zgemm_dense(ctypes.c_void_p(dA.data_ptr()), ctypes.c_void_p(dB.data_ptr()),
                                     ctypes.c_void_p(dC.data_ptr()), npd, opA, opB, 3, ctypes.c_void_p(s.cuda_stream))
            _native.check(rc, 'dqh_zgemm_dense')
            torch.cuda.synchronize()
            C = dC.cpu().numpy().view(np.complex128)
            a = np.conj(np.swapaxes(A, 1, 2)) if opA else A
            b = np.conj(np.swapaxes(B, 1, 2)) if opB else B
            want = a @ b
            assert np.max(np.abs(C - want)) < 1e-12 * np.max(np.abs(want)), (opA, opB)


@pytest.mark.parametrize('order', [2, 3])
@pytest.mark.parametrize('normalize', [True, False])
def test_adaptive_rouchon_matches_oracle_step_for_step(order, normalize):
    from dynamiqs_amd.integrators import rouchon_adaptive as ra
    from dynamiqs_amd.integrators.dense import Instance
    from oracle import operators as op
    from oracle import rouchon_adaptive as ora

    H, Ls, rho0 = _problem()
    tsave = np.linspace(0.0, 1.0, 6)
    Es = [op.destroy(5).conj().T @ op.destroy(5)]
    print(f'adaptive Rouchon order {order} normalize {normalize} n=5', flush=True)
    res = ra.solve([Instance(H=H, Ls=Ls, rho0=rho0)], tsave, order=order, exp_ops=Es, normalize=normalize)
    want, st = ora.integrate(H, Ls, rho0, tsave, order=order, normalize=normalize)
    assert res['status'][0] == 0
    assert (res['nsteps'][0], res['naccepted'][0], res['nrejected'][0]) == (st.nsteps, st.naccepted, st.nrejected)
    assert np.max(np.abs(res['states'][0] - want)) < TOL
    np.testing.assert_allclose(res['final_state'][0], want[-1], atol=TOL)
    ex = np.array([np.trace(Es[0] @ r) for r in want])
    np.testing.assert_allclose(res['expects'][0, 0], ex, atol=TOL)
    if normalize:
        assert np.max(np.abs(np.trace(res['states'][0], axis1=1, axis2=2) - 1)) < 1e-12


@pytest.mark.parametrize('order', [2, 3])
def test_adaptive_rouchon_heterogeneous_batch(order):
    """One launch of instances of different operators and jump-operator counts per
    padded size; each against its own oracle solve."""
    from dynamiqs_amd.integrators import rouchon_adaptive as ra
    from oracle import rouchon_adaptive as ora

    rng = np.random.default_rng(30 + order)
    tsave = np.linspace(0.0, 1.5, 4)
    for n, K in ((12, 1), (30, 2), (48, 0), (64, 3)):
        insts = [_rand(rng, n, K, scale=s) for s in (0.5, 2.0)]
        print(f'adaptive Rouchon order {order} batch n={n} K={K}', flush=True)
        res = ra.solve(insts, tsave, order=order, rtol=1e-6, atol=1e-6)
        for b, inst in enumerate(insts):
            want, st = ora.integrate(inst.H, inst.Ls, inst.rho0, tsave, order=order)
            assert res['status'][b] == 0
            assert res['naccepted'][b] == st.naccepted, (n, K, b)
            assert np.max(np.abs(res['states'][b] - want)) < TOL, (n, K, b)


def test_adaptive_rouchon_pwc_and_modulated_terms():
    """Piecewise-constant and modulated Hamiltonian terms (G at every stage time, PWC
    value of the step start, steps landing on the breakpoints) against the oracle."""
    from dynamiqs_amd.integrators import rouchon_adaptive as ra
    from dynamiqs_amd.integrators.dense import Instance
    from oracle import operators as op
    from oracle import rouchon_adaptive as ora

    H, Ls, rho0 = _problem()
    a = op.destroy(5)
    X, Y = a + a.conj().T, 1j * (a - a.conj().T)
    bps = np.array([0.0, 0.33, 0.71, 1.0])
    vals = np.array([0.4, -0.8, 0.2])
    f = lambda t: 0.3 * np.cos(3.0 * t)  # noqa: E731

    def Hf(tau, tseg):
        i = int(np.searchsorted(bps, tseg, side='right')) - 1
        return H + (vals[i] if 0 <= i < len(vals) else 0.0) * X + f(tau) * Y

    tsave = np.linspace(0.0, 1.2, 5)
    inst = Instance(H=H, Ls=Ls, rho0=rho0, pwc=[(bps, vals, X)], mod=[(f, Y, None)])
    for order in (2, 3):
        print(f'adaptive Rouchon order {order} PWC + modulated', flush=True)
        res = ra.solve([inst], tsave, order=order)
        want, st = ora.integrate(Hf, Ls, rho0, tsave, order=order, jump_ts=bps)
        assert res['status'][0] == 0
        assert res['naccepted'][0] == st.naccepted
        # the modulated coefficient goes through the kernels' cubic tables (1e-12 of |f|)
        assert np.max(np.abs(res['states'][0] - want)) < 1e-8


@pytest.mark.parametrize('method', ['Rouchon2', 'Rouchon3'])
def test_reference_systems_through_mesolve(method, kats):
    """``tests/mesolve/test_adaptive_rouchon.py`` runs both pairs on dense OCavity and
    OTDQubit: closed forms at the reference's tolerances."""
    import dynamiqs_amd as dq
    from oracle import operators as op

    m = getattr(dq.method, method)()
    k = kats['ocavity']
    a = dq.destroy(k['n'])
    tsave = np.linspace(0, k['t_end'], k['ntsave'])
    psi0 = dq.coherent(k['n'], k['alpha0'])
    print(f'{method} OCavity through dq.mesolve', flush=True)
    r = dq.mesolve(k['delta'] * a.dag() @ a, [np.sqrt(k['kappa']) * a], psi0, tsave, method=m,
                   exp_ops=[dq.position(k['n']), dq.momentum(k['n'])])
    st = r.states.to_numpy()
    for i, t in enumerate(tsave):
        al = k['alpha0'] * np.exp(-1j * k['delta'] * t - 0.5 * k['kappa'] * t)
        assert np.linalg.norm(st[i] - op.coherent_dm(k['n'], al)) <= k['ysave_atol']
        np.testing.assert_allclose(r.expects[:, i], [al.real, al.imag], rtol=k['esave_rtol'], atol=k['esave_atol'])
    assert int(r.infos.naccepted) > 0 and int(r.infos.nsteps) == int(r.infos.naccepted) + int(r.infos.nrejected)

    q = kats['otdqubit']
    eps, w, g = q['eps'], q['omega'], q['gamma']
    sx = dq.sigmax()
    H = dq.timecallable(lambda t: eps * np.cos(w * t) * sx)
    tq = np.linspace(0, q['t_end'], q['ntsave'])
    print(f'{method} OTDQubit through dq.mesolve', flush=True)
    r = dq.mesolve(H, [np.sqrt(g) * sx], dq.fock(2, 0), tq, method=m)
    st = r.states.to_numpy()
    for i, t in enumerate(tq):
        th = 2 * eps / w * np.sin(w * t)
        eta = np.exp(-2 * g * t)
        want = np.array([[0.5 * (1 + eta * np.cos(th)), 0.5j * eta * np.sin(th)],
                         [-0.5j * eta * np.sin(th), 0.5 * (1 - eta * np.cos(th))]])
        assert np.linalg.norm(st[i] - want) < 1e-4


def test_adaptive_rouchon_max_steps():
    import dynamiqs_amd as dq

    a = dq.destroy(4)
    with pytest.raises(RuntimeError, match='maximum number of method steps'):
        dq.mesolve(a.dag() @ a, [a], dq.fock(4, 2), np.linspace(0.0, 5.0, 3), method=dq.method.Rouchon2(max_steps=3))
