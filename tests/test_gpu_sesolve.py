"""GPU parity of ``dq.sesolve`` (``dqh_sesolve_tsit5``) against the complex128
oracle (``oracle/sesolve.py``): the oracle's replay of the GPU's own accepted steps
(the FP32 arithmetic alone), closed forms of the reference's own test systems
(``tests/systems/closed_system.py``: Cavity, TDQubit), products of matrix
exponentials for piecewise-constant Hamiltonians, batching shapes and per-instance
independence, saving options, Tsit5 and Dopri5."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

REPLAY_TOL = 2e-6


def _replay(Hop, psi0, tsave, ends, method='tsit5'):
    from oracle import tsit5
    from oracle.mesolve import _segment_time

    jumps = Hop.jump_ts

    def f(t, y, tseg):
        return -1j * (Hop(_segment_time(t, tseg, jumps), t) @ y)

    saved, _ = tsit5.integrate_steps(f, np.asarray(psi0, dtype=complex).reshape(-1, 1), tsave, ends,
                                     tab=tsit5.tableau(method))
    return np.stack(saved)[..., 0]


def _run_plan(insts, tsave, Hops, psis, method='tsit5', exp_ops=None, **kw):
    """Solve through the plan with a step log; return results and the replay error
    of every instance."""
    from dynamiqs_amd.integrators.sesolve import build_plan

    plan = build_plan(insts, tsave, exp_ops=exp_ops, tableau=method, step_log_cap=100_000, **kw)
    plan.run()
    res = plan.results()
    assert np.all(res['status'] == 0), res['status']
    errs = []
    for b in range(len(insts)):
        ref = _replay(Hops[b], psis[b], tsave, res['step_ends'][b], method)
        errs.append(np.max(np.abs(res['states'][b] - ref)))
    return res, np.asarray(errs)


def _herm(rng, n, scale=1.0):
    x = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
    return scale * (x + x.conj().T) / (2 * np.sqrt(n))


def _ket(rng, n):
    v = rng.normal(size=n) + 1j * rng.normal(size=n)
    return v / np.linalg.norm(v)


def test_cavity_closed_form_and_reference_tolerances():
    """The reference's Cavity system (closed_system.py:46-104) with the reference
    test's tolerances (integrator_tester.py:18-45) and the exact Fock-basis
    rotation; dense and DIA layouts give the same result."""
    import dynamiqs_amd as dq

    n, delta, a0 = 8, 2 * np.pi, 0.5
    ts = np.linspace(0.0, 0.3, 11)
    for layout in (dq.dense, dq.dia):
        H = delta * dq.number(n, layout=layout)
        Es = [dq.position(n, layout=layout), dq.momentum(n, layout=layout)]
        res = dq.sesolve(H, dq.coherent(n, a0), ts, exp_ops=Es)
        assert res.states.shape == (11, n, 1) and res.expects.shape == (2, 11)
        psi0 = dq.coherent(n, a0).to_numpy()[:, 0]
        exact = np.stack([np.exp(-1j * delta * np.arange(n) * t) * psi0 for t in ts])
        got = res.states.to_numpy()[..., 0]
        print(layout, 'max |psi - exact| =', np.max(np.abs(got - exact)))
        assert np.max(np.abs(got - exact)) <= 1e-5
        alpha = a0 * np.exp(-1j * delta * ts)
        np.testing.assert_allclose(res.expects[0], alpha.real, rtol=1e-3, atol=1e-4)
        np.testing.assert_allclose(res.expects[1], alpha.imag, rtol=1e-3, atol=1e-4)
        assert 'SESolveResult' in str(res)


@pytest.mark.parametrize('method', ['tsit5', 'dopri5'])
def test_tdqubit_closed_form(method):
    """The reference's TDQubit (closed_system.py:107-176) as a modulated term:
    psi(t) = cos(theta/2)|0> - 1j sin(theta/2)|1>, theta = 2 eps/omega sin(omega t)."""
    import dynamiqs_amd as dq

    eps, w = 3.0, 10.0
    ts = np.linspace(0.0, 1.0, 11)
    H = dq.modulated(lambda t: eps * np.cos(w * t), dq.sigmax())
    m = dq.method.Tsit5() if method == 'tsit5' else dq.method.Dopri5()
    res = dq.sesolve(H, dq.fock(2, 0), ts, exp_ops=[dq.sigmax(), dq.sigmay(), dq.sigmaz()], method=m)
    th = 2 * eps / w * np.sin(w * ts)
    exact = np.stack([np.cos(th / 2), -1j * np.sin(th / 2)], axis=-1)
    err = np.max(np.abs(res.states.to_numpy()[..., 0] - exact))
    print(method, 'max |psi - exact| =', err)
    assert err <= 1e-4  # rtol = atol = 1e-6 over 40 oscillation-driven steps
    np.testing.assert_allclose(res.expects.real, np.stack([0 * th, -np.sin(th), np.cos(th)]), atol=1e-4)


@pytest.mark.parametrize('method', ['tsit5', 'dopri5'])
def test_replay_constant_pwc_modulated(method):
    """Random Hermitian H0 plus a PWC term and a modulated term (n = 37, padded to
    48), three kets: every instance matches the oracle's replay of its own steps,
    and the expectation values match <psi|E|psi> of the replayed states."""
    from oracle import mesolve as om

    from dynamiqs_amd.integrators.sesolve import SEInstance

    rng = np.random.default_rng(21)
    n = 37
    H0 = _herm(rng, n, 2.0)
    O = np.diag(np.linspace(-1, 1, n)).astype(complex)
    M = np.diag(np.ones(n - 1), 1).astype(complex)
    M = M + M.conj().T
    times, vals = np.array([0.0, 0.3, 0.55, 1.2]), np.array([0.8, -1.1 + 0.2j, 0.4])
    f = lambda t: 0.7 * np.sin(3.0 * t)  # noqa: E731
    tsave = np.linspace(0.0, 1.5, 9)
    E = np.diag(np.arange(n, dtype=float)).astype(complex)
    psis = [_ket(rng, n) for _ in range(3)]
    Hop = om.TimeOp(H0, [om.PWCTerm(times, vals, O), om.ModTerm(f, M)])
    insts = [SEInstance(H=H0, psi0=p, pwc=[(times, vals, O)], mod=[(f, M, ('f', None))], key_H=()) for p in psis]
    res, errs = _run_plan(insts, tsave, [Hop] * 3, psis, method, exp_ops=[E])
    print(method, 'replay errors', errs, 'steps', res['naccepted'])
    assert np.all(errs <= REPLAY_TOL)
    for b in range(3):
        ref = _replay(Hop, psis[b], tsave, res['step_ends'][b], method)
        ex = np.einsum('ti,ij,tj->t', ref.conj(), E, ref)
        np.testing.assert_allclose(res['expects'][b, 0], ex, atol=REPLAY_TOL * n * 4)
        # the steps land on every breakpoint inside the interval
        ends = res['step_ends'][b]
        for tb in times[1:]:
            assert np.any(ends == tb), tb


def test_pwc_against_matrix_exponentials():
    """A two-segment PWC Hamiltonian (zero coefficient after the last breakpoint):
    the GPU solve within 1e-5 of the product of segment propagators."""
    import scipy.linalg as sla

    import dynamiqs_amd as dq

    rng = np.random.default_rng(5)
    n = 20
    H0, O = _herm(rng, n), _herm(rng, n)
    v = np.array([0.7, -1.3])
    H = dq.asqarray(H0) + dq.pwc(np.array([0.0, 0.4, 1.1]), v, dq.asqarray(O))
    psi0 = _ket(rng, n)
    ts = np.array([0.0, 0.4, 0.9, 1.1, 1.5])
    res = dq.sesolve(H, dq.asqarray(psi0[:, None]), ts, method=dq.method.Tsit5(rtol=1e-8, atol=1e-8))
    U = lambda k, h: sla.expm(-1j * (H0 + v[k] * O) * h)  # noqa: E731
    p04 = U(0, 0.4) @ psi0
    p11 = U(1, 0.7) @ p04
    exact = np.stack([psi0, p04, U(1, 0.5) @ p04, p11, sla.expm(-1j * H0 * 0.4) @ p11])
    err = np.max(np.abs(res.states.to_numpy()[..., 0] - exact))
    print('max |psi - exact| =', err)
    assert err <= 1e-5


def test_batching_shapes_and_instance_independence():
    """Cartesian (...H, ...psi0) and flat broadcast batching; every instance equals
    the single solve of its own (H, psi0) bit for bit (one workgroup each, same
    code path); save_states=False returns the final states of the same solve."""
    import dynamiqs_amd as dq

    rng = np.random.default_rng(8)
    n = 10
    Hs = np.stack([_herm(rng, n) for _ in range(2)])
    psis = np.stack([_ket(rng, n)[:, None] for _ in range(3)])
    ts = np.linspace(0.0, 1.0, 5)
    res = dq.sesolve(Hs, psis, ts)
    assert res.states.shape == (2, 3, 5, n, 1)
    assert res.infos.nsteps.shape == (2, 3)
    for i in range(2):
        for j in range(3):
            one = dq.sesolve(Hs[i], psis[j], ts)
            np.testing.assert_array_equal(res.states.to_numpy()[i, j], one.states.to_numpy())
    flat = dq.sesolve(Hs[:, None].repeat(3, axis=1), psis, ts, options=dq.Options(cartesian_batching=False))
    assert flat.states.shape == (2, 3, 5, n, 1)
    np.testing.assert_array_equal(flat.states.to_numpy(), res.states.to_numpy())
    fin = dq.sesolve(Hs, psis, ts, options=dq.Options(save_states=False))
    assert fin.states.shape == (2, 3, 1, n, 1)  # nsave = 1 (sesolve.py:127-130)
    np.testing.assert_array_equal(fin.states.to_numpy()[:, :, 0], res.states.to_numpy()[:, :, -1])
    np.testing.assert_array_equal(fin.final_state.to_numpy(), res.states.to_numpy()[:, :, -1])


def test_large_dimension_replay():
    """n = 200 (padded to 208, one row per thread): replay within 1e-5 (FP32
    products over 200-term rows) and the norm conserved to the same level."""
    from oracle import mesolve as om

    from dynamiqs_amd.integrators.sesolve import SEInstance

    rng = np.random.default_rng(13)
    n = 200
    H0 = _herm(rng, n, 3.0)
    psi = _ket(rng, n)
    tsave = np.linspace(0.0, 1.0, 6)
    res, errs = _run_plan([SEInstance(H=H0, psi0=psi, key_H=())], tsave, [om.TimeOp(H0)], [psi])
    print('n=200 replay error', errs, 'steps', res['naccepted'])
    assert np.all(errs <= 1e-5)
    norms = np.linalg.norm(res['states'][0], axis=-1)
    assert np.max(np.abs(norms - 1.0)) <= 1e-5


def test_save_extra_and_t0():
    """Options(save_extra=fn) sees the ket at every save time; Options(t0=...)
    starts the integration before tsave[0]."""
    import dynamiqs_amd as dq

    H = 2 * np.pi * dq.number(6)
    psi0 = dq.coherent(6, 0.4)
    ts = np.linspace(0.5, 1.0, 4)
    res = dq.sesolve(H, psi0, ts, options=dq.Options(save_extra=lambda s: s.to_numpy()[0, 0], t0=0.0))
    assert res.extra.shape == (4,)
    np.testing.assert_allclose(res.extra, res.states.to_numpy()[:, 0, 0])
    k = np.arange(6)
    exact = np.stack([np.exp(-1j * 2 * np.pi * k * t) * psi0.to_numpy()[:, 0] for t in ts])
    assert np.max(np.abs(res.states.to_numpy()[..., 0] - exact)) <= 1e-5


def test_euler_fixed_step_against_oracle():
    """dq.method.Euler(dt) on the sesolve kernel (fixed steps, the operators at the
    step start, linear interpolation to the save times) against the oracle's Euler,
    with a PWC and a modulated term; FixedStepInfos counts the steps."""
    import dynamiqs_amd as dq
    from oracle import mesolve as om
    from oracle import sesolve as ose

    rng = np.random.default_rng(41)
    n = 12
    H0, O, M = _herm(rng, n), _herm(rng, n), _herm(rng, n)
    times, vals = np.array([0.0, 0.37, 1.0]), np.array([0.5, -0.8])
    f = lambda t: 0.6 * np.cos(2.0 * t)  # noqa: E731
    psi0 = _ket(rng, n)
    ts = np.linspace(0.0, 1.2, 7)
    H = dq.asqarray(H0) + dq.pwc(times, vals, dq.asqarray(O)) + dq.modulated(f, dq.asqarray(M))
    res = dq.sesolve(H, dq.asqarray(psi0[:, None]), ts, exp_ops=[dq.asqarray(O)], method=dq.method.Euler(dt=1e-3))
    Hop = om.TimeOp(H0, [om.PWCTerm(times, vals, O), om.ModTerm(f, M)])
    ref = ose.sesolve(Hop, psi0, ts, exp_ops=[O], method='euler', dt=1e-3)
    err = np.max(np.abs(res.states.to_numpy()[..., 0] - ref['states'][..., 0]))
    print('euler sesolve max err vs oracle', err, 'steps', res.infos.nsteps)
    assert err <= 1e-5
    np.testing.assert_allclose(res.expects[0], ref['expects'][0], atol=1e-5)
    assert int(res.infos.nsteps) == 1200 and 'steps' in str(res.infos)


def test_large_dimension_variant_replay():
    """n = 600 (padded to 608) runs on the four-wave kernel whose LDS holds up to
    np = 1024 (112 KB): replay of its own steps within 1e-5 and the norm conserved;
    DIA operators of that size are accepted through their dense form."""
    import dynamiqs_amd as dq
    from oracle import mesolve as om

    from dynamiqs_amd.integrators.sesolve import SEInstance

    rng = np.random.default_rng(17)
    n = 600
    a = np.diag(np.sqrt(np.arange(1, n)), 1).astype(complex)
    H0 = a.conj().T @ a * 0.01 + 0.3 * (a + a.conj().T) / np.sqrt(n)
    psi = _ket(rng, n)
    tsave = np.linspace(0.0, 1.0, 4)
    res, errs = _run_plan([SEInstance(H=H0, psi0=psi, key_H=())], tsave, [om.TimeOp(H0)], [psi])
    print('n=600 replay error', errs, 'steps', res['naccepted'])
    assert np.all(errs <= 1e-5)
    assert np.max(np.abs(np.linalg.norm(res['states'][0], axis=-1) - 1.0)) <= 1e-5
    r = dq.sesolve(dq.asqarray(H0, layout=dq.dia), dq.asqarray(psi[:, None]), tsave)
    np.testing.assert_allclose(r.states.to_numpy()[-1, :, 0], res['states'][0, -1], atol=1e-5)
