"""GPU parity of the dense HIP path (C-ABI ``libdqh.so``) against the complex128
oracle. Tolerances: complex64 MFMA arithmetic with FP64 state accumulation; every
saved state must be within 1e-5 of the FP64 oracle in Frobenius norm and in
trace distance unless a test states otherwise."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TOL = 1e-5


def trace_distance(a, b):
    return 0.5 * np.abs(np.linalg.eigvalsh(0.5 * ((a - b) + (a - b).conj().T))).sum()


def _rand_instance(rng, n, K, hermitian_scale=1.0, lscale=0.3):
    H = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
    H = hermitian_scale * (H + H.conj().T) / (2 * np.sqrt(n))
    Ls = [lscale * (rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))) / np.sqrt(n) for _ in range(K)]
    psi = rng.normal(size=(n, 1)) + 1j * rng.normal(size=(n, 1))
    psi /= np.linalg.norm(psi)
    return H, Ls, psi @ psi.conj().T


def test_mfma_lane_layout_single_tile():
    """16x16x16 complex product through one MFMA tile (checks the builtin's lane layout)."""
    from dynamiqs_amd.integrators import dense
    from oracle import lindblad

    rng = np.random.default_rng(0)
    n = 16
    # K = 0, Hnh = -iH: rhs = -i[H, rho]; choose H real diagonal-free to exercise all lanes
    H = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
    H = H + H.conj().T
    rho = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
    rho = rho + rho.conj().T
    inst = dense.Instance(H=H, Ls=[], rho0=rho)
    plan = dense.build_plan('rhs', [inst], np.array([0.0, 1.0]))
    plan.run()
    out = plan.results()['final_state'][0]
    ref = lindblad.rhs_hermitian(H, [], rho)
    assert np.abs(out - ref).max() < 1e-4 * np.abs(ref).max()


@pytest.mark.parametrize('n,K', [(4, 1), (16, 1), (20, 2), (37, 3), (64, 2), (100, 1)])
def test_rhs_matches_oracle(n, K):
    from dynamiqs_amd.integrators import dense
    from oracle import lindblad

    rng = np.random.default_rng(n * 10 + K)
    insts, refs = [], []
    for _ in range(3):
        H, Ls, rho = _rand_instance(rng, n, K)
        insts.append(dense.Instance(H=H, Ls=Ls, rho0=rho))
        refs.append(lindblad.rhs_hermitian(H, Ls, rho))
    plan = dense.build_plan('rhs', insts, np.array([0.0, 1.0]))
    plan.run()
    out = plan.results()['final_state']
    for b in range(3):
        scale = np.abs(refs[b]).max()
        assert np.abs(out[b] - refs[b]).max() < 2e-6 * np.sqrt(n) * scale + 1e-7


def _c1():
    from oracle import operators as op

    n = 16
    a = op.destroy(n)
    return 1.0 * a.conj().T @ a, [np.sqrt(0.1) * a], op.coherent(n, 1.0), np.linspace(0, 2 * np.pi, 101)


def test_tsit5_config1_parity_vs_fp64_oracle():
    import dynamiqs_amd as dq
    from oracle import mesolve as om
    from oracle import operators as op

    H, Ls, psi0, tsave = _c1()
    ref = om.mesolve(H, Ls, psi0, tsave, rtol=1e-6, atol=1e-8)
    a = dq.destroy(16, layout=dq.dense)
    res = dq.mesolve(
        1.0 * a.dag() @ a,
        [np.sqrt(0.1) * a],
        dq.coherent(16, 1.0),
        tsave,
        exp_ops=[dq.number(16)],
        method=dq.method.Tsit5(rtol=1e-6, atol=1e-8),
    )
    st = res.states.to_numpy()
    assert st.shape == (101, 16, 16) and st.dtype == np.complex64
    errs = np.linalg.norm(st - ref['states'], axis=(-2, -1))
    tds = [trace_distance(st[i].astype(np.complex128), ref['states'][i]) for i in range(101)]
    assert errs.max() <= TOL, errs.max()
    assert max(tds) <= TOL, max(tds)
    # same controller: step counts agree with the FP64 oracle up to rounding-induced drift
    assert abs(int(res.infos.nsteps) - ref['infos']['nsteps']) <= 2
    # expectation values fused at save points
    nbar = np.array([op.expect(op.number(16), s) for s in ref['states']])
    np.testing.assert_allclose(res.expects[0], nbar, atol=TOL)
    text = str(res)
    assert text.startswith('==== MESolveResult ====\nMethod  : Tsit5\nInfos   : ')
    assert 'States  : QArray complex64 (101, 16, 16) | 202.0 Kb' in text
    assert 'Expects : Array complex64 (1, 101) | 0.8 Kb' in text


def test_tsit5_batched_random_instances_alone_and_in_batch():
    """Batch of uneven random problems; each instance equals its solo solve (batch invariance)."""
    from dynamiqs_amd.integrators import dense
    from oracle import mesolve as om

    rng = np.random.default_rng(7)
    n, K = 24, 2
    tsave = np.linspace(0, 2.0, 9)
    insts = []
    for _ in range(5):
        H, Ls, rho = _rand_instance(rng, n, K)
        insts.append(dense.Instance(H=H, Ls=Ls, rho0=rho))
    plan = dense.build_plan('tsit5', insts, tsave, rtol=1e-6, atol=1e-8)
    plan.run()
    out = plan.results()
    for b, inst in enumerate(insts):
        ref = om.mesolve(inst.H, inst.Ls, inst.rho0, tsave, rtol=1e-6, atol=1e-8)
        err = np.linalg.norm(out['states'][b] - ref['states'], axis=(-2, -1)).max()
        assert err <= TOL, (b, err)
        solo = dense.build_plan('tsit5', [inst], tsave, rtol=1e-6, atol=1e-8)
        solo.run()
        so = solo.results()
        np.testing.assert_array_equal(so['states'][0], out['states'][b])
        assert so['nsteps'][0] == out['nsteps'][b]


def test_rouchon1_dense_parity():
    import dynamiqs_amd as dq
    from oracle import mesolve as om
    from oracle import operators as op

    n = 8
    delta = kappa = 2 * np.pi
    a = op.destroy(n)
    tsave = np.linspace(0, 0.3, 11)
    ref = om.mesolve(delta * a.conj().T @ a, [np.sqrt(kappa) * a], op.coherent(n, 0.5), tsave, method='rouchon1', dt=1e-3)
    A = dq.destroy(n)
    res = dq.mesolve(delta * A.dag() @ A, [np.sqrt(kappa) * A], dq.coherent(n, 0.5), tsave, method=dq.method.Rouchon1(dt=1e-3))
    st = res.states.to_numpy()
    errs = np.linalg.norm(st - ref['states'], axis=(-2, -1))
    assert errs.max() <= TOL, errs.max()
    assert int(res.infos.nsteps) == ref['infos']['nsteps'] == 300
    assert abs(np.trace(st[-1]) - 1) < 1e-6


def test_tsit5_pwc_parity():
    import dynamiqs_amd as dq
    from oracle import mesolve as om
    from oracle import operators as op

    n = 12
    rng = np.random.default_rng(3)
    times = np.linspace(0, 2.0, 11)
    u = rng.normal(0, 0.5, size=(3, 10))
    a = op.destroy(n)
    tsave = np.linspace(0, 2.0, 5)
    A = dq.destroy(n, layout=dq.dense)
    H = dq.pwc(times, u, A + A.dag()) + A.dag() @ A
    res = dq.mesolve(H, [np.sqrt(0.05) * A], dq.fock(n, 0), tsave, method=dq.method.Tsit5(rtol=1e-7, atol=1e-9))
    st = res.states.to_numpy()
    assert st.shape == (3, 5, n, n)
    for b in range(3):
        Hop = om.TimeOp(a.conj().T @ a, [om.PWCTerm(times, u[b], a + a.conj().T)])
        ref = om.mesolve(Hop, [np.sqrt(0.05) * a], op.fock(n, 0), tsave, rtol=1e-7, atol=1e-9)
        err = np.linalg.norm(st[b] - ref['states'], axis=(-2, -1)).max()
        assert err <= TOL, (b, err)


def test_printed_step_counts_through_mesolve_and_sesolve(kats):
    """The reference prints these infos for its README example (``README.md:75``),
    its mesolve docstring (``dynamiqs/integrators/apis/mesolve.py:225``), its
    sesolve docstring (``dynamiqs/integrators/apis/sesolve.py:164``) and its
    workflow tutorial (Dopri5, ``docs/documentation/basics/workflow.md:88``). The
    kernels take exactly those step counts."""
    import dynamiqs_amd as dq

    k = kats['readme_lossy_oscillator']
    a = dq.destroy(k['n'])
    print('README lossy oscillator through dq.mesolve', flush=True)
    r = dq.mesolve(k['omega'] * a.dag() @ a, [np.sqrt(k['kappa']) * a], dq.coherent(k['n'], k['alpha0']),
                   np.linspace(0, k['T'], k['ntsave']))
    assert str(r.infos) == f"{k['nsteps']} steps ({k['naccepted']} accepted, {k['nrejected']} rejected)"
    d = kats['mesolve_docstring']
    print('mesolve docstring example', flush=True)
    r = dq.mesolve(a.dag() @ a, [a], dq.coherent(d['n'], d['alpha0']), np.linspace(0, d['T'], d['ntsave']))
    assert str(r.infos) == f"{d['nsteps']} steps ({d['naccepted']} accepted, {d['nrejected']} rejected)"
    s = kats['sesolve_docstring']
    print('sesolve docstring example', flush=True)
    r = dq.sesolve(a.dag() @ a, dq.coherent(s['n'], s['alpha0']), np.linspace(0, s['T'], s['ntsave']))
    assert str(r.infos) == f"{s['nsteps']} steps ({s['naccepted']} accepted, {s['nrejected']} rejected)"
    # the workflow tutorial's Rabi run with Dopri5 at atol 1e-8 (docs/documentation/basics/workflow.md:88)
    w = kats['workflow_rabi_dopri5']
    print('workflow tutorial, Dopri5', flush=True)
    r = dq.sesolve(w['delta'] * dq.sigmaz() + w['Omega'] * dq.sigmax(), dq.fock(2, 0),
                   np.linspace(0.0, w['sim_time'], w['num_save']),
                   method=dq.method.Dopri5(rtol=w['rtol'], atol=w['atol']))
    assert str(r.infos) == f"{w['nsteps']} steps ({w['naccepted']} accepted, {w['nrejected']} rejected)"


def test_max_steps_raises_reference_error():
    import dynamiqs_amd as dq

    a = dq.destroy(4)
    with pytest.raises(RuntimeError, match='maximum number of method steps'):
        dq.mesolve(a.dag() @ a, [a], dq.fock(4, 2), np.linspace(0, 10, 3), method=dq.method.Tsit5(max_steps=3))


def test_tsit5_kerr_matches_fp64_replay_of_its_own_steps():
    """Batched Kerr oscillators (n = 64, 2 jump operators, t in [0, 1]): the GPU
    states against the complex128 oracle replaying the GPU's logged accepted steps
    (same stages in FP64), which isolates the complex64 arithmetic from the step
    sequence. Bar: 1e-5 Frobenius at every save time."""
    import sys

    from dynamiqs_amd.integrators import dense

    sys.path.insert(0, __import__('os').path.dirname(__import__('os').path.dirname(__file__)))
    import bench

    insts, _ = bench.kerr_instances(4, 64)
    tsave = np.linspace(0.0, 1.0, 21)
    plan = dense.build_plan('tsit5', insts, tsave, rtol=1e-6, atol=1e-6, step_log_cap=5000)
    plan.run()
    out = plan.results()
    assert np.all(out['status'] == 0)
    from oracle import tsit5

    for b in (0, 3):
        assert len(out['step_ends'][b]) == out['naccepted'][b]
        H, Ls = insts[b].H, insts[b].Ls
        Hnh = -1j * H - 0.5 * sum(L.conj().T @ L for L in Ls)

        def f(t, y, tseg, Hnh=Hnh, Ls=Ls):
            tmp = Hnh @ y + 0.5 * sum(L @ y @ L.conj().T for L in Ls)
            return tmp + tmp.conj().T

        saved, _ = tsit5.integrate_steps(f, insts[b].rho0, tsave, out['step_ends'][b])
        err = np.linalg.norm(out['states'][b] - np.stack(saved), axis=(-2, -1)).max()
        assert err <= TOL, (b, err)


@pytest.mark.parametrize('fp64_diag', [True, False])
def test_tsit5_kerr_fp64_diagonal_against_fp64_replay(fp64_diag):
    """The driven Kerr batch (n = 64, eps2 up to 4, t in [0, 2]) with and without the
    generator's diagonal in FP64 (hdiag on the single-size resident kernel), each
    against the complex128 oracle replaying its own accepted steps: both within the
    1e-5 Frobenius bar at every save time."""
    import os
    import sys

    from dynamiqs_amd.integrators import dense
    from oracle import tsit5

    sys.path.insert(0, os.path.dirname(os.path.dirname(__file__)))
    import bench

    insts, eps2 = bench.kerr_instances(32, 64)
    pick = [int(np.argmax(eps2)), int(np.argmin(eps2))]
    insts = [insts[i] for i in pick]
    tsave = np.linspace(0.0, 2.0, 41)
    plan = dense.build_plan('tsit5', insts, tsave, rtol=1e-6, atol=1e-6, step_log_cap=20000, fp64_diag=fp64_diag)
    assert plan.fp64_diag == fp64_diag and plan.uniform_np == 64
    plan.run()
    out = plan.results()
    assert np.all(out['status'] == 0)
    for b, inst in enumerate(insts):
        Hnh = -1j * inst.H - 0.5 * sum(L.conj().T @ L for L in inst.Ls)

        def f(t, y, tseg, Hnh=Hnh, Ls=inst.Ls):
            tmp = Hnh @ y + 0.5 * sum(L @ y @ L.conj().T for L in Ls)
            return tmp + tmp.conj().T

        saved, _ = tsit5.integrate_steps(f, inst.rho0, tsave, out['step_ends'][b])
        err = np.linalg.norm(out['states'][b] - np.stack(saved), axis=(-2, -1)).max()
        print(f'fp64_diag={fp64_diag} eps2={eps2[pick[b]]:.3f} steps={len(out["step_ends"][b])} replay err {err:.3e}')
        assert err <= TOL, (b, err)
