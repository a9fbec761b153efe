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


@pytest.mark.parametrize('normalisation', ['banded', 'blocks'])
@pytest.mark.parametrize('na,nb_,nsteps', [(8, 9, 300), (12, 12, 200)])
def test_dia_rouchon1_matches_oracle(na, nb_, nsteps, normalisation):
    from dynamiqs_amd.integrators import dia
    from oracle import mesolve as om
    from oracle import operators as op

    H, Ls, _ = _two_mode(na, nb_)
    dt = 1e-3
    T = nsteps * dt
    tsave = np.linspace(0, T, 6)
    alphas = [0.5, 1.2, 2.0]
    rho0 = np.stack([op.coherent_dm((na, nb_), (al, 0.0)) for al in alphas])
    plan = dia.build_plan(H, Ls, rho0, tsave, dt, normalisation=normalisation)
    assert plan.normalisation == normalisation
    plan.run()
    out = plan.results()
    for b in range(len(alphas)):
        ref = om.mesolve(H, Ls, rho0[b], tsave, method='rouchon1', dt=dt)
        err = np.linalg.norm(out['states'][b] - ref['states'], axis=(-2, -1)).max()
        assert err <= TOL, (b, err)
        # FP32 increments on an FP64 state: trace drift stays far below complex64 rounding
        assert abs(np.trace(out["final_state"][b]) - 1) < 1e-7


@pytest.mark.parametrize('variant', ['global', 'split', 'h16'])
def test_dia_rouchon1_sweep_forms_agree(variant):
    """Every Kraus sweep form (one-pass global, two-pass split, LDS-tiled halo 16)
    reproduces the oracle's Rouchon1 at n = 144 (padded to 192)."""
    from dynamiqs_amd.integrators import dia
    from oracle import mesolve as om
    from oracle import operators as op

    H, Ls, _ = _two_mode(12, 12)
    dt = 1e-3
    tsave = np.linspace(0, 0.1, 3)
    rho0 = np.stack([op.coherent_dm((12, 12), (al, 0.0)) for al in (0.8, 1.6)])
    v = {'global': dia.SWEEP_GLOBAL, 'split': dia.SWEEP_SPLIT, 'h16': dia.SWEEP_TILED_H16}[variant]
    plan = dia.build_plan(H, Ls, rho0, tsave, dt, variant=v)
    plan.run()
    out = plan.results()
    for b in range(2):
        ref = om.mesolve(H, Ls, rho0[b], tsave, method='rouchon1', dt=dt)
        err = np.linalg.norm(out['states'][b] - ref['states'], axis=(-2, -1)).max()
        assert err <= TOL, (variant, b, err)


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
    outs = []
    for v, norm in ((dia.SWEEP_SPLIT, 'banded'), (dia.SWEEP_GLOBAL, 'banded'), (dia.SWEEP_GLOBAL, 'blocks')):
        plan = dia.build_plan(H, Ls, rho0, tsave, dt, variant=v, normalisation=norm)
        plan.run()
        outs.append(plan.results())
    out = outs[0]
    assert np.abs(outs[0]['final_state'] - outs[1]['final_state']).max() <= 1e-6
    assert np.abs(outs[0]['final_state'] - outs[2]['final_state']).max() <= 1e-6
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


def test_mesolve_routes_dia_rouchon1_with_expects_and_close_saves():
    """``dq.mesolve`` with DIA operators and Rouchon1 at n = 144 runs the DIA kernels;
    states and ``Tr(E rho)`` match the oracle, including two save times inside one
    step (each interpolated with its own weight)."""
    import importlib

    import dynamiqs_amd as dq
    from oracle import mesolve as om

    ms = importlib.import_module('dynamiqs_amd.integrators.mesolve')
    from oracle import operators as op

    a, b = dq.destroy(12, 12)
    H = 5 * a.dag() @ a + 6 * b.dag() @ b + 0.1 * (a.dag() @ b + a @ b.dag()) - 0.01 * a.dag() @ a @ b.dag() @ b
    Ls = [np.sqrt(0.02) * a, np.sqrt(0.05) * b]
    assert ms.dia_route(dq.method.Rouchon1(dt=1e-3), dq.constant(H), [dq.constant(L) for L in Ls], 144) is not None
    alphas = (0.5, 1.5)
    kets = np.stack([op.coherent((12, 12), (al, 0.0)) for al in alphas])
    E = (a.dag() @ a).to_numpy()
    dt = 1e-3
    tsave = np.array([0.0, 0.0502, 0.0504, 0.1, 0.2])
    res = dq.mesolve(H, Ls, kets, tsave, exp_ops=[a.dag() @ a], method=dq.method.Rouchon1(dt=dt))
    assert res.states.shape == (2, len(tsave), 144, 144)
    Hn, Ln = H.to_numpy(), [L.to_numpy() for L in Ls]
    for i in range(len(alphas)):
        rho0 = kets[i] @ kets[i].conj().T
        ref = om.mesolve(Hn, Ln, rho0, tsave, exp_ops=[E], method='rouchon1', dt=dt)
        err = np.linalg.norm(res.states.to_numpy()[i] - ref['states'], axis=(-2, -1)).max()
        assert err <= TOL, (i, err)
        eerr = np.abs(res.expects[i, 0] - ref['expects'][0]).max()
        assert eerr <= 1e-4, (i, eerr)
    assert int(np.asarray(res.infos.nsteps).ravel()[0]) == 200


def _random_dm(rng, n, B):
    out = []
    for _ in range(B):
        x = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
        r = x @ x.conj().T
        out.append(r / np.trace(r).real)
    return np.stack(out)


@pytest.mark.parametrize(
    'dims,variant',
    [((12, 12), 'auto'), ((12, 12), 'global'), ((32, 32), 'tiled'), ((32, 32), 'global'), ((6, 6, 6), 'auto')],
)
def test_dia_rhs_matches_oracle(dims, variant):
    """``dqh_lindblad_rhs_dia`` against the FP64 Hermitian-form RHS of the oracle on
    random density matrices: n = 144 (padded to 192), the full 1024 two-mode size
    (tiled and global sweeps) and a 3-mode 216 space whose offsets (36) exceed the
    LDS halo, so only the global sweep applies. Bar: 1e-5 relative Frobenius."""
    from dynamiqs_amd.integrators import dia
    from oracle import lindblad
    from oracle import operators as op

    rng = np.random.default_rng(7)
    modes = op.destroy(*dims)
    n = int(np.prod(dims))
    H = sum((w + 1) * m.conj().T @ m for w, m in enumerate(modes))
    H = H + 0.1 * (modes[0].conj().T @ modes[1] + modes[0] @ modes[1].conj().T)
    H = H - 0.01 * modes[0].conj().T @ modes[0] @ modes[1].conj().T @ modes[1]
    Ls = [np.sqrt(0.02 * (k + 1)) * m for k, m in enumerate(modes)]
    rho = _random_dm(rng, n, 2)
    v = {'auto': dia.SWEEP_AUTO, 'global': dia.SWEEP_GLOBAL, 'tiled': dia.SWEEP_TILED}[variant]
    plan = dia.rhs_plan(H, Ls, rho, variant=v)
    plan.run()
    out = plan.result()
    for b in range(2):
        ref = lindblad.rhs_hermitian(H, Ls, rho[b])
        err = np.linalg.norm(out[b] - ref) / max(1.0, np.linalg.norm(ref))
        assert err <= 1e-5, (dims, variant, b, err)
        # the vector field of a density matrix is traceless and Hermitian
        assert abs(np.trace(out[b])) <= 1e-4 * max(1.0, np.linalg.norm(ref))


def test_dia_rhs_tiled_rejects_wide_offsets():
    """Offsets beyond the LDS halo (a 3-mode 6 x 6 x 6 space: offset 36) make the
    tiled sweep refuse the problem on the host before any launch."""
    from dynamiqs_amd.integrators import dia
    from oracle import operators as op

    modes = op.destroy(6, 6, 6)
    H = sum(m.conj().T @ m for m in modes)
    plan = dia.rhs_plan(H, [modes[0]], _random_dm(np.random.default_rng(0), 216, 1), variant=dia.SWEEP_TILED)
    with pytest.raises(RuntimeError, match='tiled sweep'):
        plan.run()


def _sparse_lindblad(H, Ls):
    """FP64 vector field through scipy.sparse products (the oracle's Hermitian
    form, rhs_hermitian, with sparse operators so n = 1024 runs in milliseconds)."""
    import scipy.sparse as sp

    Hnh = sp.csr_matrix(-1j * H - 0.5 * sum(L.conj().T @ L for L in Ls))
    Hd = Hnh.conj().T.tocsr()
    Lsp = [sp.csr_matrix(L) for L in Ls]
    Ld = [L.conj().T.tocsr() for L in Lsp]

    def f(t, y, tseg):
        out = Hnh @ y + (Hd.T @ y.T).T
        for L, D in zip(Lsp, Ld):
            out = out + (D.T @ (L @ y).T).T
        return out

    return f


@pytest.mark.parametrize('variant', ['tiled', 'global'])
def test_dia_tsit5_matches_oracle(variant):
    """Adaptive Tsit5 on the DIA sweeps (n = 144, padded to 192) against the oracle's
    FP64 Tsit5 on the same problem and tolerances: states within 1e-5, identical
    step counts up to a few steps (FP32 stages vs FP64 error estimate)."""
    from dynamiqs_amd.integrators import dia
    from oracle import mesolve as om
    from oracle import operators as op

    H, Ls, _ = _two_mode(12, 12)
    alphas = (0.5, 1.5)
    rho0 = np.stack([op.coherent_dm((12, 12), (al, 0.0)) for al in alphas])
    tsave = np.linspace(0.0, 0.3, 7)
    v = {'tiled': dia.SWEEP_TILED, 'global': dia.SWEEP_GLOBAL}[variant]
    plan = dia.tsit5_plan(H, Ls, rho0, tsave, variant=v)
    plan.run()
    out = plan.results()
    assert np.all(out['status'] == 0)
    for b in range(len(alphas)):
        ref = om.mesolve(H, Ls, rho0[b], tsave, method='tsit5', rtol=1e-6, atol=1e-6)
        err = np.linalg.norm(out['states'][b] - ref['states'], axis=(-2, -1)).max()
        assert err <= TOL, (variant, b, err)
        nref = int(ref['infos']['nsteps'])
        assert abs(int(out['nsteps'][b]) - nref) <= max(2, nref // 20), (out['nsteps'][b], nref)
        assert abs(np.trace(out['final_state'][b]) - 1) < 1e-6


def test_dia_tsit5_graph_replays_equal_plain_launches():
    """The captured hipGraph (replayed many times, 4 steps each) and plain launches
    of the same kernels give bit-identical states and counters."""
    from dynamiqs_amd.integrators import dia
    from oracle import operators as op

    H, Ls, _ = _two_mode(12, 12)
    rho0 = np.stack([op.coherent_dm((12, 12), (al, 0.0)) for al in (0.7, 1.9)])
    tsave = np.linspace(0.0, 0.2, 5)
    outs = []
    for gs in (4, 0):
        plan = dia.tsit5_plan(H, Ls, rho0, tsave, graph_steps=gs)
        plan.run()
        outs.append(plan.results())
    assert np.array_equal(outs[0]['states'], outs[1]['states'])
    assert np.array_equal(outs[0]['nsteps'], outs[1]['nsteps'])
    assert np.array_equal(outs[0]['final_state'], outs[1]['final_state'])


def test_dia_tsit5_full_size_short_run():
    """The 32 x 32 two-mode size (n = 1024) for t in [0, 0.05] against the oracle's
    Tsit5 driven by sparse FP64 products, plus the per-instance step counters."""
    from dynamiqs_amd.integrators import dia
    from oracle import operators as op
    from oracle import tsit5

    H, Ls, _ = _two_mode(32, 32)
    rho0 = np.stack([op.coherent_dm((32, 32), (al, 0.0)) for al in (0.5, 2.0)])
    tsave = np.linspace(0.0, 0.05, 3)
    plan = dia.tsit5_plan(H, Ls, rho0, tsave)
    plan.run()
    out = plan.results()
    assert np.all(out['status'] == 0)
    f = _sparse_lindblad(H, Ls)
    for b in range(2):
        saved, _, st = tsit5.integrate(f, rho0[b], tsave, rtol=1e-6, atol=1e-6)
        err = np.linalg.norm(out['states'][b] - np.stack(saved), axis=(-2, -1)).max()
        assert err <= TOL, (b, err)
        assert out['nrhs'][b] == 2 + 6 * out['nsteps'][b]


def test_mesolve_routes_dia_tsit5():
    """``dq.mesolve`` with DIA operators and Tsit5 at n = 144 runs the DIA Tsit5
    kernels and returns adaptive-step infos and expectation values."""
    import dynamiqs_amd as dq
    from oracle import mesolve as om

    a, b = dq.destroy(12, 12)
    H = 5 * a.dag() @ a + 6 * b.dag() @ b + 0.1 * (a.dag() @ b + a @ b.dag())
    Ls = [np.sqrt(0.02) * a, np.sqrt(0.05) * b]
    psi = dq.coherent((12, 12), (1.0, 0.0))
    tsave = np.linspace(0.0, 0.2, 5)
    res = dq.mesolve(H, Ls, psi, tsave, exp_ops=[a.dag() @ a])
    rho0 = psi.to_numpy() @ psi.to_numpy().conj().T
    ref = om.mesolve(H.to_numpy(), [L.to_numpy() for L in Ls], rho0, tsave, exp_ops=[(a.dag() @ a).to_numpy()])
    assert np.linalg.norm(res.states.to_numpy() - ref['states'], axis=(-2, -1)).max() <= TOL
    assert np.abs(res.expects[0] - ref['expects'][0]).max() <= 1e-4
    assert int(res.infos.nsteps) > 0


def test_autotuner_times_and_caches_dia_and_dense_forms(tmp_path, monkeypatch):
    """The autotuner times every admissible form on the GPU, keeps the fastest and
    answers the next request for the same key from its cache."""
    import importlib

    monkeypatch.setenv('DQH_AUTOTUNE_CACHE', str(tmp_path / 'tune.json'))
    from dynamiqs_amd import autotune
    from dynamiqs_amd.integrators import dia

    importlib.reload(autotune)
    H, Ls, _ = _two_mode(12, 12)
    v = autotune.dia_sweep(H, Ls, 2)
    assert v in (dia.SWEEP_GLOBAL, dia.SWEEP_TILED_H16, dia.SWEEP_TILED_H32)
    entries = [e for k, e in autotune._memo.items() if k.startswith('rhs|dia|')]
    assert len(entries) == 1 and len(entries[0]['us']) == 3
    assert autotune.dia_sweep(H, Ls, 2) == v
    g = autotune.dense_global(32, 2, 64)
    assert isinstance(g, bool)
    assert any(k.startswith('dense|') for k in autotune._memo)
    importlib.reload(autotune)


def test_dia_block_solve_panel_widths_agree():
    """The banded solves with 32- and 64-column panels per workgroup compute every
    column with the same operations: bit-identical states."""
    from dynamiqs_amd.integrators import dia
    from oracle import operators as op

    H, Ls, _ = _two_mode(16, 16)
    rho0 = np.stack([op.coherent_dm((16, 16), (al, 0.0)) for al in (0.6, 1.4)])
    tsave = np.linspace(0, 0.02, 3)
    outs = []
    for pw in (32, 64):
        plan = dia.build_plan(H, Ls, rho0, tsave, 1e-3, variant=dia.SWEEP_SPLIT, solve_panel=pw, normalisation='banded')
        plan.run()
        outs.append(plan.results())
    assert np.array_equal(outs[0]['states'], outs[1]['states'])


@pytest.mark.parametrize('normalisation', ['banded', 'blocks'])
def test_dia_rouchon1_long_run_is_stable_and_matches_dense(normalisation):
    """5000 steps (T = 5, dt = 1e-3) of the 16 x 16 two-mode problem (n = 256): the
    state stays trace-preserving and Hermitian and agrees with the dense Rouchon1
    kernel (a different kernel for the same scheme) at every save. Rounding seeds
    an anti-Hermitian part in the state; the block solves apply the exact map to
    it, which contracts it (a Hermitian-only form of the normalisation amplified it
    by ~2 % per step here)."""
    from dynamiqs_amd.integrators import dense, dia
    from oracle import operators as op

    H, Ls, _ = _two_mode(16, 16)
    v = np.kron(op.coherent(16, 1.5).ravel(), np.eye(16)[:, 0])
    rho0 = np.outer(v, v.conj())
    tsave = np.linspace(0.0, 5.0, 11)
    plan = dia.build_plan(H, Ls, rho0[None], tsave, 1e-3, normalisation=normalisation)
    plan.run()
    st = plan.results()['states'][0]
    tr = np.abs(np.trace(st, axis1=-2, axis2=-1) - 1)
    assert tr.max() < 1e-6, tr
    assert np.abs(st - np.conj(np.swapaxes(st, -1, -2))).max() < 1e-6
    dplan = dense.build_plan('rouchon1', [dense.Instance(H=H, Ls=Ls, rho0=rho0)], tsave, dt=1e-3)
    dplan.run()
    ref = dplan.results()['states'][0]
    err = np.linalg.norm(st - ref, axis=(-2, -1)).max()
    assert err < 1e-5, err


@pytest.mark.parametrize('method', ['tsit5', 'rouchon1'])
def test_batched_dia_operators_above_the_dense_limit(method):
    """Batched constant DIA operators at n = 1100, beyond the dense kernels' 1024: one
    DIA plan per distinct operator set. Every element of the batch equals the same
    problem solved alone through the shared-operator DIA path (pinned against the
    oracle by the tests above), with cartesian and flat batching; shapes and infos
    follow the reference's batching rules (mesolve.py:307-336)."""
    import dynamiqs_amd as dq

    n = 1100
    a = dq.destroy(n)
    eps = np.array([0.1, 0.3])
    kap = np.array([0.05, 0.2])
    H = a.dag() @ a + eps[:, None, None] * (a + a.dag())
    L = np.sqrt(kap)[:, None, None] * a
    psi = dq.coherent(n, 1.0)
    tsave = np.linspace(0.0, 0.3, 4)
    m = dq.method.Tsit5() if method == 'tsit5' else dq.method.Rouchon1(dt=0.01)
    E = [dq.number(n)]
    print(f'{method}: batched DIA n={n}, cartesian', flush=True)
    r = dq.mesolve(H, [L], psi, tsave, exp_ops=E, method=m)
    assert r.states.shape == (2, 2, 4, n, n) and r.expects.shape == (2, 2, 1, 4)
    st = r.states.to_numpy()
    for i in range(2):
        for j in range(2):
            print(f'{method}: alone eps={eps[i]} kappa={kap[j]}', flush=True)
            ri = dq.mesolve(a.dag() @ a + eps[i] * (a + a.dag()), [np.sqrt(kap[j]) * a], psi, tsave, exp_ops=E, method=m)
            assert np.max(np.abs(st[i, j] - ri.states.to_numpy())) <= 1e-7, (i, j)
            np.testing.assert_allclose(r.expects[i, j], ri.expects, atol=1e-6)
            if method == 'tsit5':
                assert int(r.infos.nsteps[i, j]) == int(ri.infos.nsteps)
    print(f'{method}: batched DIA n={n}, flat', flush=True)
    rf = dq.mesolve(H, [L], psi, tsave, method=m, options=dq.Options(cartesian_batching=False))
    assert rf.states.shape == (2, 4, n, n)
    for i in range(2):
        assert np.max(np.abs(rf.states.to_numpy()[i] - st[i, i])) <= 1e-7, i


@pytest.mark.parametrize('tableau', ['tsit5', 'dopri5'])
def test_dia_tsit5_time_dependent_terms_match_oracle(tableau):
    """PWC and modulated Hamiltonian terms on the DIA Tsit5 sweeps: the kernels rebuild
    Hnh's diagonals per instance before every stage (PWC value of the step start's
    segment, modulated coefficient at the stage time), land steps on the breakpoints
    and evaluate k_0 anew there. Against the complex128 oracle at the same
    tolerances, n = 96 (padded to 128), two initial states."""
    from dynamiqs_amd.integrators import dia
    from oracle import mesolve as om
    from oracle import operators as op

    n = 96
    a = op.destroy(n)
    H0 = a.conj().T @ a
    X, Y = a + a.conj().T, 1j * (a.conj().T - a)
    times, vals = np.array([0.0, 0.3, 0.8]), np.array([0.5, -0.7])
    g = lambda t: 0.4 * np.cos(3.0 * t)  # noqa: E731
    Ls = [np.sqrt(0.1) * a]
    rho0 = np.stack([op.coherent_dm(n, 1.0), op.coherent_dm(n, 0.5j)])
    tsave = np.linspace(0.0, 1.2, 5)
    print(f'DIA {tableau} with PWC + modulated terms, n={n}', flush=True)
    plan = dia.tsit5_plan(H0, Ls, rho0, tsave, rtol=1e-7, atol=1e-9, tableau=tableau,
                          pwc=[(times, vals, X)], mod=[(g, Y)])
    plan.run()
    res = plan.results()
    Hop = om.TimeOp(H0, [om.PWCTerm(times, vals, X), om.ModTerm(g, Y)])
    for b in range(2):
        ref = om.mesolve(Hop, Ls, rho0[b], tsave, rtol=1e-7, atol=1e-9, method=tableau)
        assert res['status'][b] == 0
        err = np.linalg.norm(res['states'][b] - ref['states'], axis=(-2, -1)).max()
        assert err <= 1e-5, (b, err)
        assert abs(int(res['nsteps'][b]) - ref['infos']['nsteps']) <= 2, (res['nsteps'][b], ref['infos']['nsteps'])


def test_mesolve_routes_time_dependent_dia_above_the_dense_limit():
    """``dq.mesolve`` with a PWC or modulated DIA Hamiltonian at n = 1100 runs on the
    DIA Tsit5 sweeps. A PWC term held at one value over the whole interval, and a
    constant modulation, reproduce the constant-Hamiltonian solve."""
    import dynamiqs_amd as dq

    n = 1100
    a = dq.destroy(n)
    tsave = np.linspace(0.0, 0.4, 3)
    psi = dq.coherent(n, 1.0)
    c = 0.3
    Ls = [np.sqrt(0.1) * a]
    print(f'constant H, n={n}', flush=True)
    r0 = dq.mesolve(a.dag() @ a + c * (a + a.dag()), Ls, psi, tsave)
    for name, H in [('pwc', a.dag() @ a + dq.pwc(np.array([0.0, 0.4]), np.array([c]), a + a.dag())),
                    ('modulated', a.dag() @ a + dq.modulated(lambda t: c + 0.0 * t, a + a.dag()))]:
        print(f'{name} H, n={n}', flush=True)
        r = dq.mesolve(H, Ls, psi, tsave)
        assert np.max(np.abs(r.states.to_numpy() - r0.states.to_numpy())) <= 1e-5, name
        assert abs(int(r.infos.nsteps) - int(r0.infos.nsteps)) <= 1, name


def test_batched_time_dependent_dia_above_the_dense_limit():
    """A batch of PWC control vectors on a DIA Hamiltonian at n = 1100 runs one DIA
    Tsit5 plan per batch element; each element equals its own unbatched solve (the
    time-dependent DIA path pinned against the oracle above)."""
    import dynamiqs_amd as dq

    n = 1100
    a = dq.destroy(n)
    tsave = np.linspace(0.0, 0.3, 3)
    psi = dq.coherent(n, 1.0)
    times = np.array([0.0, 0.1, 0.3])
    vals = np.array([[0.2, -0.4], [0.5, 0.1]])
    Ls = [np.sqrt(0.1) * a]
    print(f'batched PWC DIA Hamiltonian, n={n}', flush=True)
    r = dq.mesolve(a.dag() @ a + dq.pwc(times, vals, a + a.dag()), Ls, psi, tsave)
    assert r.states.shape == (2, 3, n, n)
    st = r.states.to_numpy()
    for b in range(2):
        print(f'element {b} alone', flush=True)
        rb = dq.mesolve(a.dag() @ a + dq.pwc(times, vals[b], a + a.dag()), Ls, psi, tsave)
        assert np.max(np.abs(st[b] - rb.states.to_numpy())) <= 1e-7, b
        assert int(r.infos.nsteps[b]) == int(rb.infos.nsteps)
