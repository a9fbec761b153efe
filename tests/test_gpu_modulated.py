"""GPU parity of modulated Hamiltonians H(t) = H0 + sum_j f_j(t) O_j
(``dq.modulated``, ``time_qarray.py:669-760``): the kernels evaluate each f_j at
every stage time from a piecewise-cubic table built on the host (refined until it
matches f to 1e-12 relative). Checked against the complex128 oracle, which calls
f exactly: its replay of the GPU's own accepted steps (2e-6, the arithmetic
alone) and a tight FP64 solve (the GPU within 3x of the FP64 oracle's own error at
the same tolerance, whose steps differ); the reference's OTDQubit closed form; PWC
and modulated terms together; batched f; resident and global kernel forms; Tsit5
and Dopri5."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TOL = 1e-5


def _oracle_op(H0, mods, pwcs=()):
    from oracle import mesolve as om

    terms = [om.ModTerm(f, o) for f, o in mods] + [om.PWCTerm(t, v, o) for t, v, o in pwcs]
    return om.TimeOp(np.asarray(H0, dtype=complex), terms)


def _replay(Hop, Ls, rho0, tsave, ends, method='tsit5'):
    from oracle import lindblad, tsit5
    from oracle.mesolve import _segment_time

    jumps = Hop.jump_ts

    def f(t, y, tseg):
        ts = _segment_time(t, tseg, jumps)
        Lt = [L(ts, t) if callable(L) else L for L in Ls]  # TimeOp jump operators at the stage time
        return lindblad.rhs_hermitian(Hop(ts, t), Lt, y)

    saved, _ = tsit5.integrate_steps(f, rho0, tsave, ends, tab=tsit5.tableau(method))
    return np.stack(saved)


def _random(rng, n, K):
    H0 = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
    H0 = (H0 + H0.conj().T) / (2 * np.sqrt(n))
    O1 = np.diag(np.linspace(-1, 1, n)).astype(complex)
    O2 = np.diag(np.ones(n - 1), 1).astype(complex)
    O2 = O2 + O2.conj().T
    Ls = [0.3 * (rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))) / np.sqrt(n) for _ in range(K)]
    psi = rng.normal(size=(n, 1)) + 1j * rng.normal(size=(n, 1))
    psi /= np.linalg.norm(psi)
    return H0, O1, O2, Ls, psi @ psi.conj().T


def test_otdqubit_closed_form(kats):
    """The reference's OTDQubit (tests/systems/open_system.py:134-229):
    H = eps cos(w t) sigma_x, L = sqrt(gamma) sigma_x, rho0 = |0><0|."""
    import dynamiqs_amd as dq

    k = kats['otdqubit']
    eps, w, g = k['eps'], k['omega'], k['gamma']
    tsave = np.linspace(0, k['t_end'], k['ntsave'])
    H = dq.modulated(lambda t: eps * np.cos(w * t), dq.sigmax())
    res = dq.mesolve(H, [np.sqrt(g) * dq.sigmax()], dq.fock(2, 0), tsave,
                     method=dq.method.Tsit5(rtol=1e-8, atol=1e-8))
    st = res.states.to_numpy()
    for i, t in enumerate(tsave):
        th = 2 * eps / w * np.sin(w * t)
        eta = np.exp(-2 * g * t)
        want = np.array([[0.5 * (1 + eta * np.cos(th)), 0.5j * eta * np.sin(th)],
                         [-0.5j * eta * np.sin(th), 0.5 * (1 - eta * np.cos(th))]])
        assert np.linalg.norm(st[i] - want) < 2e-6, (i, np.linalg.norm(st[i] - want))


@pytest.mark.parametrize('n,form', [(20, 'resident'), (80, 'global')])
@pytest.mark.parametrize('method', ['tsit5', 'dopri5'])
def test_modulated_and_pwc_drive_matches_oracle_and_replay(n, form, method):
    from dynamiqs_amd.integrators import dense
    from oracle import mesolve as om

    rng = np.random.default_rng(n)
    H0, O1, O2, Ls, rho0 = _random(rng, n, 2)
    f1 = lambda t: 0.8 * np.cos(3.0 * t)  # noqa: E731
    f2 = lambda t: 0.5 * np.exp(-t) * np.sin(7.0 * t) + 0.2j * t  # noqa: E731 - complex coefficient
    ptimes = np.array([0.0, 0.4, 1.1, 1.7])
    pvals = np.array([0.3, -0.5, 0.7], dtype=complex)
    tsave = np.linspace(0.0, 2.0, 11)
    inst = dense.Instance(H=H0, Ls=Ls, rho0=rho0, pwc=[(ptimes, pvals, O2)], mod=[(f1, O1, 'f1'), (f2, O2, 'f2')])
    plan = dense.build_plan('tsit5', [inst], tsave, rtol=1e-6, atol=1e-6, form=form, step_log_cap=20000,
                            tableau=method)
    plan.run()
    out = plan.results()
    assert np.all(out['status'] == 0)
    Hop = _oracle_op(H0, [(f1, O1), (f2, O2)], [(ptimes, pvals, O2)])
    # the arithmetic: the oracle on the GPU's own accepted steps
    rep = _replay(Hop, Ls, rho0, tsave, out['step_ends'][0], method)
    err_rep = np.linalg.norm(out['states'][0] - rep, axis=(-2, -1)).max()
    assert err_rep <= 2e-6, (n, form, method, err_rep)
    # the solution: against a tight FP64 solve, as close as the FP64 oracle at the same tolerance
    ref = om.mesolve(Hop, Ls, rho0, tsave, method=method, rtol=1e-6, atol=1e-6)
    tight = om.mesolve(Hop, Ls, rho0, tsave, method=method, rtol=1e-10, atol=1e-10)
    e_gpu = np.linalg.norm(out['states'][0] - tight['states'], axis=(-2, -1)).max()
    e_ora = np.linalg.norm(ref['states'] - tight['states'], axis=(-2, -1)).max()
    assert e_gpu <= 3 * e_ora + 2e-6, (n, form, method, e_gpu, e_ora)
    nref = int(ref['infos']['nsteps'])
    assert abs(int(out['nsteps'][0]) - nref) <= max(2, nref // 20), (out['nsteps'][0], nref)


def test_mesolve_batched_modulation():
    """A modulation whose f returns one value per batch element (shape (3,)):
    each element gets its own table and equals its own unbatched solve."""
    import dynamiqs_amd as dq

    n = 6
    a = dq.destroy(n)
    ws = np.array([1.0, 2.0, 3.5])
    H = a.dag() @ a + dq.modulated(lambda t: 0.4 * np.cos(ws * t), a + a.dag())
    tsave = np.linspace(0.0, 2.0, 5)
    res = dq.mesolve(H, [0.2 * a], dq.coherent(n, 0.5), tsave)
    assert res.states.shape == (3, 5, n, n)
    for b, w in enumerate(ws):
        one = dq.mesolve(a.dag() @ a + dq.modulated(lambda t, w=w: 0.4 * np.cos(w * t), a + a.dag()), [0.2 * a],
                         dq.coherent(n, 0.5), tsave)
        np.testing.assert_allclose(res.states.to_numpy()[b], one.states.to_numpy(), atol=1e-6)


def test_timecallable_hamiltonians(kats):
    """timecallable Hamiltonians, rewritten on the host as a constant part plus
    modulated terms spanning their samples: the reference's OTDQubit and TDQubit
    written as dq.timecallable (open_system.py:134-229, closed_system.py:107-176)
    against their closed forms, and a two-direction callable against the oracle
    with the exact f at every stage time."""
    import dynamiqs_amd as dq
    from oracle import mesolve as om

    k = kats['otdqubit']
    eps, w, g = k['eps'], k['omega'], k['gamma']
    tsave = np.linspace(0, k['t_end'], k['ntsave'])
    H = dq.timecallable(lambda t: eps * np.cos(w * t) * dq.sigmax())
    res = dq.mesolve(H, [np.sqrt(g) * dq.sigmax()], dq.fock(2, 0), tsave, method=dq.method.Tsit5(rtol=1e-8, atol=1e-8))
    st = res.states.to_numpy()
    for i, t in enumerate(tsave):
        th = 2 * eps / w * np.sin(w * t)
        eta = np.exp(-2 * g * t)
        want = np.array([[0.5 * (1 + eta * np.cos(th)), 0.5j * eta * np.sin(th)],
                         [-0.5j * eta * np.sin(th), 0.5 * (1 - eta * np.cos(th))]])
        assert np.linalg.norm(st[i] - want) < 2e-6, (i, np.linalg.norm(st[i] - want))
    ts = np.linspace(0.0, 1.0, 11)
    rs = dq.sesolve(dq.timecallable(lambda t: 3.0 * np.cos(10.0 * t) * dq.sigmax()), dq.fock(2, 0), ts,
                    method=dq.method.Tsit5(rtol=1e-8, atol=1e-8))
    th = 0.6 * np.sin(10.0 * ts)
    exact = np.stack([np.cos(th / 2), -1j * np.sin(th / 2)], axis=-1)
    assert np.max(np.abs(rs.states.to_numpy()[..., 0] - exact)) < 2e-6
    n = 6
    a = dq.destroy(n)
    A, X, Y = (a.dag() @ a).to_numpy(), (a + a.dag()).to_numpy(), (a @ a + a.dag() @ a.dag()).to_numpy()
    f = lambda t: dq.asqarray(A + np.sin(2 * t) * X + t**2 * Y)  # noqa: E731
    L = 0.3 * a.to_numpy()
    rho0 = dq.coherent_dm(n, 0.5).to_numpy()
    ts = np.linspace(0.0, 1.0, 6)
    got = dq.mesolve(dq.timecallable(f), [L], rho0, ts, method=dq.method.Tsit5(rtol=1e-8, atol=1e-8)).states.to_numpy()
    Hop = om.TimeOp(A, [om.ModTerm(lambda t: np.sin(2 * t), X), om.ModTerm(lambda t: t**2, Y)])
    ref = om.mesolve(Hop, [L], rho0, ts, rtol=1e-10, atol=1e-10)['states']
    err = np.linalg.norm(got - ref, axis=(-2, -1)).max()
    print('two-direction timecallable vs oracle', err)
    assert err < 1e-5


@pytest.mark.parametrize('n', [6, 40])
def test_modulated_jump_operators_match_oracle(n):
    """Jump operators f(t) L (dq.modulated): the kernels rebuild f L and
    conj(f) L^dag / 2 at every stage time and carry -1/2 |f|^2 L^dag L as a
    modulated term of the effective Hamiltonian. Against the oracle with the exact
    f at every stage time: a decaying rate, a complex phase (no effect on the
    dynamics, a check of the conj), together with a modulated H term, batched over
    two initial states; the replay of the GPU's steps within 2e-6."""
    import dynamiqs_amd as dq
    from oracle import mesolve as om

    from dynamiqs_amd.integrators import dense
    from dynamiqs_amd.integrators.mesolve import enumerate_instances

    rng = np.random.default_rng(n)
    H0, O1, _, _, rho0 = _random(rng, n, 0)
    a = np.diag(np.sqrt(np.arange(1, n)), 1).astype(complex)
    fr = lambda t: 0.6 * np.exp(-0.5 * t)  # noqa: E731 - decaying amplitude
    fp = lambda t: 0.3 * np.exp(2j * t)  # noqa: E731 - pure phase on a second channel
    fh = lambda t: 0.4 * np.sin(3 * t)  # noqa: E731
    H = dq.asqarray(H0) + dq.modulated(fh, dq.asqarray(O1))
    Ls = [dq.modulated(fr, dq.asqarray(a)), dq.modulated(fp, dq.asqarray(a.conj().T @ a))]
    psi = rng.normal(size=(n, 1)) + 1j * rng.normal(size=(n, 1))
    rho1 = psi @ psi.conj().T / np.vdot(psi, psi)
    rhos = np.stack([rho0, rho1])
    tsave = np.linspace(0.0, 2.0, 5)
    res = dq.mesolve(H, Ls, rhos, tsave, method=dq.method.Tsit5(rtol=1e-8, atol=1e-8))
    Hop = om.TimeOp(H0, [om.ModTerm(fh, O1)])
    Lops = [om.TimeOp(np.zeros((n, n), complex), [om.ModTerm(fr, a)]),
            om.TimeOp(np.zeros((n, n), complex), [om.ModTerm(fp, a.conj().T @ a)])]
    for b in range(2):
        ref = om.mesolve(Hop, Lops, rhos[b], tsave, rtol=1e-10, atol=1e-10)['states']
        err = np.linalg.norm(res.states.to_numpy()[b] - ref, axis=(-2, -1)).max()
        print('n', n, 'instance', b, 'err vs oracle', err)
        assert err < 1e-5
    # replay of the GPU's own steps
    from dynamiqs_amd.time_qarray import astimeqarray

    _, insts = enumerate_instances(astimeqarray(H), [astimeqarray(L) for L in Ls], dq.asqarray(rho0), True)
    plan = dense.build_plan('tsit5', insts, tsave, step_log_cap=100_000)
    plan.run()
    r = plan.results()
    assert r['status'][0] == 0
    ref = _replay(Hop, Lops, rho0, tsave, r['step_ends'][0])
    err = np.linalg.norm(r['states'][0] - ref, axis=(-2, -1)).max()
    print('replay err', err)
    assert err < 2e-6


def test_timecallable_jump_operator():
    """A timecallable jump operator sqrt(kappa(t)) a becomes one modulated term
    f(t) a (its mean absorbed) and runs as a modulated jump operator; against the
    oracle with the exact operator at every stage time."""
    import dynamiqs_amd as dq
    from oracle import mesolve as om

    n = 8
    a = dq.destroy(n)
    an = a.to_numpy().astype(complex)
    kappa = lambda t: 0.4 * np.exp(-0.7 * t) + 0.05  # noqa: E731
    L = dq.timecallable(lambda t: np.sqrt(kappa(t)) * a)
    H = a.dag() @ a
    rho0 = dq.coherent_dm(n, 1.0).to_numpy()
    ts = np.linspace(0.0, 3.0, 7)
    got = dq.mesolve(H, [L], rho0, ts, method=dq.method.Tsit5(rtol=1e-8, atol=1e-8)).states.to_numpy()
    Lop = om.TimeOp(np.zeros((n, n), complex), [om.ModTerm(lambda t: np.sqrt(kappa(t)), an)])
    ref = om.mesolve(H.to_numpy(), [Lop], rho0, ts, rtol=1e-10, atol=1e-10)['states']
    err = np.linalg.norm(got - ref, axis=(-2, -1)).max()
    print('timecallable jump operator err vs oracle', err)
    assert err < 1e-5


@pytest.mark.parametrize('method', ['rouchon1', 'rouchon2', 'euler'])
def test_fixed_step_methods_with_modulated_operators(method):
    """Rouchon1 / Rouchon2 / Euler with a modulated Hamiltonian term and a modulated
    jump operator: one Kraus set per step, the operators at the step midpoint
    (Rouchon) or start (Euler), against the oracle's fixed-step solve."""
    import dynamiqs_amd as dq
    from oracle import mesolve as om

    n = 6
    a = np.diag(np.sqrt(np.arange(1, n)), 1).astype(complex)
    H0 = a.conj().T @ a
    X = a + a.conj().T
    fh = lambda t: 0.5 * np.cos(4.0 * t)  # noqa: E731
    fl = lambda t: 0.5 * np.exp(-0.3 * t) * np.exp(1j * t)  # noqa: E731
    rho0 = np.zeros((n, n), complex)
    rho0[1, 1] = 1.0
    ts = np.linspace(0.0, 1.0, 5)
    m = {'rouchon1': dq.method.Rouchon1(dt=2e-3), 'rouchon2': dq.method.Rouchon2(dt=2e-3),
         'euler': dq.method.Euler(dt=2e-3)}[method]
    res = dq.mesolve(dq.asqarray(H0) + dq.modulated(fh, dq.asqarray(X)), [dq.modulated(fl, dq.asqarray(a))], rho0,
                     ts, method=m)
    Hop = om.TimeOp(H0, [om.ModTerm(fh, X)])
    Lop = om.TimeOp(np.zeros((n, n), complex), [om.ModTerm(fl, a)])
    ref = om.mesolve(Hop, [Lop], rho0, ts, method=method, dt=2e-3)['states']
    err = np.linalg.norm(res.states.to_numpy() - ref, axis=(-2, -1)).max()
    print(method, 'err vs oracle', err)
    assert err < 1e-5


def test_pwc_jump_operators_against_oracle():
    """``dq.pwc(times, values, L)`` as a jump operator (``time_qarray.py:572-666``): the
    kernels scale L by the coefficient of the step's PWC segment (zero outside the
    breakpoints) and carry -1/2 |c|^2 L^dag L as a PWC Hamiltonian-side term. Alone,
    beside a constant jump operator, with a batched coefficient next to a PWC
    Hamiltonian term and a modulated jump operator, and on the fixed-step Rouchon1
    kernel: each against the complex128 oracle at the same tolerances."""
    import dynamiqs_amd as dq
    from oracle import mesolve as om
    from oracle import operators as op

    n = 8
    a = op.destroy(n)
    zero = np.zeros((n, n), dtype=complex)
    H0 = a.conj().T @ a + 0.3 * (a + a.conj().T)
    times = np.array([0.0, 0.4, 1.1, 2.0])
    vals = np.array([0.5, 1.2 - 0.4j, 0.0])
    tsave = np.linspace(0.0, 2.5, 6)
    rho0 = op.coherent_dm(n, 0.8)
    Lc = np.sqrt(0.1) * a.conj().T @ a
    m = dq.method.Tsit5(rtol=1e-7, atol=1e-9)
    pw = dq.pwc(times, vals, dq.asqarray(a))
    pw_or = om.TimeOp(zero, [om.PWCTerm(times, vals, a)])
    for name, Ls_dq, Ls_or in [('alone', [pw], [pw_or]), ('with a constant one', [pw, dq.asqarray(Lc)], [pw_or, Lc])]:
        print(f'PWC jump operator {name}', flush=True)
        r = dq.mesolve(dq.asqarray(H0), Ls_dq, dq.asqarray(rho0), tsave, method=m)
        ref = om.mesolve(H0, Ls_or, rho0, tsave, rtol=1e-7, atol=1e-9)
        err = np.linalg.norm(r.states.to_numpy() - ref['states'], axis=(-2, -1)).max()
        assert err <= TOL, (name, err)
        assert abs(int(r.infos.nsteps) - ref['infos']['nsteps']) <= 2, (name, r.infos.nsteps, ref['infos']['nsteps'])

    # batched PWC coefficient, a PWC Hamiltonian term and a modulated jump operator
    vals2 = np.array([[0.5, 1.2 - 0.4j, 0.0], [0.2j, 0.7, 1.5]])
    X = a + a.conj().T
    htimes, hvals = np.array([0.3, 1.7]), np.array([0.6])
    g = lambda t: 0.4 * np.cos(2.0 * t)  # noqa: E731
    Hdq = dq.asqarray(H0) + dq.pwc(htimes, hvals, dq.asqarray(X))
    print('batched PWC jump operator with a PWC H term and a modulated jump operator', flush=True)
    r = dq.mesolve(Hdq, [dq.pwc(times, vals2, dq.asqarray(a)), dq.modulated(g, dq.asqarray(a.conj().T @ a))],
                   dq.asqarray(rho0), tsave, method=m)
    st = r.states.to_numpy()
    assert st.shape == (2, 6, n, n)
    Hor = om.TimeOp(H0, [om.PWCTerm(htimes, hvals, X)])
    for b in range(2):
        Ls_or = [om.TimeOp(zero, [om.PWCTerm(times, vals2[b], a)]), om.TimeOp(zero, [om.ModTerm(g, a.conj().T @ a)])]
        ref = om.mesolve(Hor, Ls_or, rho0, tsave, rtol=1e-7, atol=1e-9)
        err = np.linalg.norm(st[b] - ref['states'], axis=(-2, -1)).max()
        assert err <= TOL, (b, err)

    print('PWC jump operator on Rouchon1', flush=True)
    r = dq.mesolve(dq.asqarray(H0), [pw], dq.asqarray(rho0), tsave, method=dq.method.Rouchon1(dt=0.01))
    ref = om.mesolve(H0, [pw_or], rho0, tsave, method='rouchon1', dt=0.01)
    err = np.linalg.norm(r.states.to_numpy() - ref['states'], axis=(-2, -1)).max()
    assert err <= TOL, err


def test_jump_operator_sums_against_oracle():
    """Jump operators that sum a constant and a modulated term, and a constant and a
    PWC term (``time_qarray.py:836-956`` as jump operators): the kernels rebuild the
    D/C pairs of the sum at every stage (DQH_FLAG_LPAIR) and carry the products of
    -1/2 L^dag L as generator terms. Tsit5 and Rouchon1 against the complex128 oracle
    at the same tolerances / step."""
    import dynamiqs_amd as dq
    from oracle import mesolve as om
    from oracle import operators as op

    n = 8
    a = op.destroy(n)
    ad = a.conj().T
    H0 = ad @ a + 0.3 * (a + ad)
    g = lambda t: 0.5 * np.cos(2.0 * t)  # noqa: E731
    times, vals = np.array([0.0, 0.5, 1.4]), np.array([0.8, 0.3 - 0.5j])
    rho0 = op.coherent_dm(n, 0.8)
    tsave = np.linspace(0.0, 2.0, 5)
    A, N = dq.asqarray(a), dq.asqarray(ad @ a)
    Ls_dq = [np.sqrt(0.2) * A + dq.modulated(g, N), 0.4 * N + dq.pwc(times, vals, A)]
    Ls_or = [om.TimeOp(np.sqrt(0.2) * a, [om.ModTerm(g, ad @ a)]), om.TimeOp(0.4 * ad @ a, [om.PWCTerm(times, vals, a)])]
    # rtol 1e-6 / atol 1e-8: at 1e-7 / 1e-9 the error estimate of the FP32 stages (four D/C
    # pairs here) nears its rounding floor and the step counts drift by a few steps
    print('sums of jump-operator terms, Tsit5', flush=True)
    r = dq.mesolve(dq.asqarray(H0), Ls_dq, dq.asqarray(rho0), tsave, method=dq.method.Tsit5(rtol=1e-6, atol=1e-8))
    ref = om.mesolve(H0, Ls_or, rho0, tsave, rtol=1e-6, atol=1e-8)
    err = np.linalg.norm(r.states.to_numpy() - ref['states'], axis=(-2, -1)).max()
    assert err <= TOL, err
    assert abs(int(r.infos.nsteps) - ref['infos']['nsteps']) <= 2, (r.infos.nsteps, ref['infos']['nsteps'])
    print('sums of jump-operator terms, Rouchon1', flush=True)
    r = dq.mesolve(dq.asqarray(H0), Ls_dq, dq.asqarray(rho0), tsave, method=dq.method.Rouchon1(dt=0.01))
    ref = om.mesolve(H0, Ls_or, rho0, tsave, method='rouchon1', dt=0.01)
    err = np.linalg.norm(r.states.to_numpy() - ref['states'], axis=(-2, -1)).max()
    assert err <= TOL, err
