"""CPU checks of the drop-in ``mesolve`` error contract: everything the reference
raises before integrating (``dynamiqs/integrators/apis/mesolve.py:280-291, 350-414``,
``dynamiqs/_checks.py``, ``dynamiqs/integrators/_utils.py:84-90``) raises here the
same exception type with the same message, without touching the GPU."""

import re

import numpy as np
import pytest

import dynamiqs_amd as dq


def _args(n=3):
    a = dq.destroy(n)
    return a.dag() @ a, [a], dq.fock(n, 0), np.linspace(0.0, 1.0, 3)


def test_unsupported_method_type_error():
    H, Ls, psi, ts = _args()
    with pytest.raises(TypeError, match=re.escape('Method of type `Kvaerno5` is not supported (supported method types:')):
        dq.mesolve(H, Ls, psi, ts, method=dq.method.Kvaerno5())
    with pytest.raises(TypeError, match=re.escape('Method of type `Dopri8` is not supported (supported method types:')):
        dq.sesolve(H, psi, ts, method=dq.method.Dopri8())


def test_shape_errors():
    H, Ls, psi, ts = _args()
    with pytest.raises(ValueError, match=re.escape('Argument `H` must have shape (...H, n, n), but has shape H.shape=(3, 2)')):
        dq.mesolve(np.ones((3, 2)), Ls, psi, ts)
    with pytest.raises(ValueError, match=re.escape('Argument `jump_ops[0]` must have shape (...L0, n, n)')):
        dq.mesolve(H, [np.ones((3, 2))], psi, ts)
    with pytest.raises(ValueError, match=re.escape('Argument `rho0` must have shape (...rho0, n, 1) or (...rho0, n, n)')):
        dq.mesolve(H, Ls, np.ones((3, 2)), ts)
    with pytest.raises(ValueError, match=re.escape('Argument `exp_ops[0]` must have shape (n, n)')):
        dq.mesolve(H, Ls, psi, ts, exp_ops=[np.ones((2, 3, 3))])


def test_tsave_errors():
    H, Ls, psi, _ = _args()
    with pytest.raises(ValueError, match='must be a 1D array, but is a 2D array'):
        dq.mesolve(H, Ls, psi, np.zeros((2, 2)))
    with pytest.raises(ValueError, match='must contain at least one element'):
        dq.mesolve(H, Ls, psi, np.zeros(0))
    with pytest.raises(ValueError, match='must be sorted in strictly ascending order'):
        dq.mesolve(H, Ls, psi, np.array([0.0, 1.0, 0.5]))


def test_rho0_layout_error():
    H, Ls, _, ts = _args()
    rho_dia = dq.asqarray(np.eye(3) / 3, layout=dq.dia)
    with pytest.raises(ValueError, match=re.escape('Argument rho0 must have layout `dense` but has layout rho0.layout=')):
        dq.mesolve(H, Ls, rho_dia, ts)


def test_option_whitelist():
    H, Ls, psi, ts = _args()
    # every mesolve option is accepted by the check; an option of another solver is not
    opts = dq.Options(save_states=False, cartesian_batching=False, t0=0.0, vectorized=True, assume_hermitian=False)
    from dynamiqs_amd.options import check_options

    check_options(opts, 'mesolve')
    with pytest.raises(ValueError, match=re.escape('Option `vectorized` was set to `True` but is not used by the solver `dq.sesolve()`')):
        check_options(dq.Options(vectorized=True), 'sesolve')


def test_empty_jump_ops_with_ket_warns():
    H, _, psi, ts = _args()
    with pytest.warns(UserWarning, match='consider using `dq.sesolve\\(\\)`'):
        try:
            dq.mesolve(H, [], psi, ts)
        except RuntimeError:
            pass  # no GPU here: the warning is raised before the solve


def test_unsupported_time_dependence_is_reported():
    """What the kernels cannot run raises before any device work: callable
    operators that span more than four time-dependent directions, modulated
    operators with discontinuity times, jump operators that sum modulated and
    piecewise-constant terms; a non-callable f is a TypeError (time_qarray.py:161-165)."""
    H, Ls, psi, ts = _args()
    rng = np.random.default_rng(0)
    ops = [(lambda x: x + x.conj().T)(rng.normal(size=(3, 3)) + 1j * rng.normal(size=(3, 3))) for _ in range(5)]
    five = dq.timecallable(lambda t: dq.asqarray(sum(np.cos((6 * k + 7) * t) * o for k, o in enumerate(ops[:3]))))
    with pytest.raises(NotImplementedError, match='timecallable operator'):
        dq.mesolve(five + dq.timecallable(lambda t: dq.asqarray(np.sin(29 * t) * ops[3] + t * ops[4])), Ls, psi, ts)
    with pytest.raises(NotImplementedError, match='discontinuity_ts'):
        dq.mesolve(dq.modulated(np.cos, H, discontinuity_ts=[0.5]), Ls, psi, ts)
    with pytest.raises(NotImplementedError, match='time-dependent jump operators'):
        pw = dq.pwc(np.array([0.0, 1.0]), np.array([1.0]), Ls[0])
        dq.mesolve(H, [pw + dq.modulated(np.cos, Ls[0])], psi, ts)

    with pytest.raises(TypeError, match=re.escape('Argument `f` must be a function, but has type float.')):
        dq.modulated(1.0, H)
    # timecallable takes the reference's keyword-only discontinuity_ts; the kernels
    # cannot tabulate a jump, so it is reported
    step = dq.timecallable(lambda t: dq.asqarray((1.0 if t < 0.5 else 2.0) * H), discontinuity_ts=[0.5])
    assert list(step.discontinuity_ts) == [0.5]
    with pytest.raises(NotImplementedError, match='discontinuity_ts'):
        dq.mesolve(step, Ls, psi, ts)
    # a jump passed without discontinuity_ts is caught by the coefficient table
    # (it cannot meet its tolerance) instead of being silently smoothed
    from dynamiqs_amd.integrators.operators import spline_table

    with pytest.raises(ValueError, match='not smooth'):
        spline_table(lambda t: np.where(np.asarray(t) < 0.37, 1.0, 2.0) + 0j, 0.0, 1.0, nmax=1 << 12)


def test_spline_table_accuracy_and_layout():
    """The piecewise-cubic table of a modulated coefficient: Horner evaluation in
    the kernels' layout reproduces f to ~1e-12 at arbitrary times, for a scalar f
    (vectorised sampling) and a batched f (per-element columns)."""
    from dynamiqs_amd.integrators.operators import spline_table

    def horner(t0, dt, coef, t):
        i = np.clip(np.floor((t - t0) / dt).astype(int), 0, len(coef) - 1)
        x = t - (t0 + i * dt)
        c = coef[i]
        return ((c[:, 0] * x + c[:, 1]) * x + c[:, 2]) * x + c[:, 3]

    f = lambda t: np.cos(2 * np.pi * t) + 0.3j * np.sin(5 * t)  # noqa: E731
    t0, dt, coef = spline_table(f, 0.0, 10.0)
    assert coef.shape[1] == 4 and np.isclose(t0 + dt * len(coef), 10.0)
    t = np.random.default_rng(1).uniform(0.0, 10.0, 2000)
    assert np.max(np.abs(horner(t0, dt, coef, t) - f(t))) < 1e-11
    g = lambda t: np.array([np.cos(t), np.exp(-t) * 1j])  # noqa: E731
    t0, dt, coef = spline_table(g, 0.5, 3.0)
    assert coef.shape[1:] == (4, 2)
    t = np.linspace(0.5, 3.0, 77)
    for k in range(2):
        ref = np.array([g(x)[k] for x in t])
        assert np.max(np.abs(horner(t0, dt, coef[..., k], t) - ref)) < 1e-11


def test_non_hermitian_rho0_error_with_assume_hermitian():
    # dynamiqs/integrators/core/diffrax_integrator.py:339-343 + dynamiqs/_checks.py:102-108
    H, Ls, _, ts = _args()
    rho = np.zeros((3, 3), dtype=complex)
    rho[0, 1] = 1.0
    with pytest.raises(ValueError, match=re.escape('Argument y0 is not hermitian.')):
        dq.mesolve(H, Ls, rho, ts)


def test_hermitian_parts_split_and_recombine():
    """rho0 = A + iB with A, B Hermitian; a Hermitian state is not split; the
    recombination inverts the split on a stack of states and expectation values."""
    from dynamiqs_amd.integrators.mesolve import combine_parts, hermitian_parts

    rng = np.random.default_rng(5)
    n = 5
    x = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
    h = x + x.conj().T
    rhos = [x, h, 1j * h, x.T]
    parts, index = hermitian_parts(rhos, assume_hermitian=False)
    assert index.tolist() == [[0, 1], [2, -1], [3, 4], [5, 6]]
    for p in parts:
        np.testing.assert_array_equal(p, p.conj().T)
    # the parts of a Hermitian state are the state itself
    np.testing.assert_array_equal(parts[2], h)
    # recombination of the parts themselves, and of a (parts, 2, n, n) stack
    np.testing.assert_allclose(combine_parts(np.stack(parts), index), np.stack(rhos), atol=1e-15)
    stack = np.stack([np.stack([p, 2 * p]) for p in parts])
    np.testing.assert_allclose(combine_parts(stack, index), np.stack([np.stack([r, 2 * r]) for r in rhos]), atol=1e-14)
    # with assume_hermitian every state is replaced by its Hermitian part, never split
    parts, index = hermitian_parts(rhos, assume_hermitian=True)
    assert index[:, 1].tolist() == [-1] * 4
    np.testing.assert_allclose(parts[0], 0.5 * (x + x.conj().T))


def test_save_extra_outputs_stacking():
    """save_extra outputs are stacked leaf by leaf with shape (*batch, ntsave, *leaf)
    (dicts, named tuples, lists, QArray leaves), as the reference's vmapped scan
    stacks them (dynamiqs/integrators/core/save_mixin.py:47-54)."""
    from collections import namedtuple

    from dynamiqs_amd.integrators.mesolve import save_extra_outputs

    P = namedtuple('P', 'a b')
    st = np.random.default_rng(0).normal(size=(6, 4, 3, 3)) + 0j

    def fn(r):
        x = r.to_numpy()
        return {'tr': r.trace(), 'pair': P(x[0, 0], [x[1, 1]]), 'q': r @ r}

    out = save_extra_outputs(fn, st, (3,), (2, 3))
    assert out['tr'].shape == (2, 3, 4)
    assert out['pair'].a.shape == (2, 3, 4) and out['pair'].b[0].shape == (2, 3, 4)
    assert isinstance(out['q'], dq.QArray) and out['q'].shape == (2, 3, 4, 3, 3)
    np.testing.assert_allclose(out['tr'][1, 2, 3], np.trace(st[5, 3]))
    np.testing.assert_allclose(out['q'].to_numpy()[0, 1, 2], st[1, 2] @ st[1, 2])
    # no batch: the time axis leads
    one = save_extra_outputs(fn, st[:1], (3,), ())
    assert one['tr'].shape == (4,) and one['q'].shape == (4, 3, 3)


def test_rouchon_method_support():
    """Rouchon2 and Rouchon3 run fixed-step with dt and as the adaptive pairs
    without; adaptive solves outside the kernel's scope (n > 64, time-dependent
    jump operators) raise NotImplementedError before any device work."""
    a = dq.destroy(80)
    psi, ts = dq.fock(80, 0), np.linspace(0.0, 1.0, 3)
    for m in (dq.method.Rouchon2(), dq.method.Rouchon3()):
        with pytest.raises(NotImplementedError, match='n <= 64'):
            dq.mesolve(a.dag() @ a, [a], psi, ts, method=m)
    b = dq.destroy(3)
    Lmod = dq.modulated(lambda t: np.cos(t), b)
    with pytest.raises(NotImplementedError, match='constant jump operators'):
        dq.mesolve(b.dag() @ b, [Lmod], dq.fock(3, 1), ts, method=dq.method.Rouchon2())
    m = dq.method.Rouchon2(dt=0.01, normalize=False)
    assert (m.dt, m.normalize, m.rtol, m.max_steps) == (0.01, False, 1e-6, 100_000)


def test_sesolve_argument_errors():
    """``dq.sesolve`` raises what the reference raises before integrating
    (dynamiqs/integrators/apis/sesolve.py:222-225, 282, 307-318): shapes, save
    times, options of another solver, unsupported method types; callable
    batched callable Hamiltonians raise NotImplementedError before any device work."""
    a = dq.destroy(3)
    H, psi, ts = a.dag() @ a, dq.fock(3, 0), np.linspace(0.0, 1.0, 3)
    with pytest.raises(ValueError, match=re.escape('Argument `H` must have shape (...H, n, n), but has shape H.shape=(3, 2)')):
        dq.sesolve(np.ones((3, 2)), psi, ts)
    with pytest.raises(ValueError, match=re.escape('Argument `psi0` must have shape (...psi0, n, 1)')):
        dq.sesolve(H, dq.fock_dm(3, 0), ts)
    with pytest.raises(ValueError, match=re.escape('Argument `exp_ops[0]` must have shape (n, n)')):
        dq.sesolve(H, psi, ts, exp_ops=[np.ones((2, 3, 3))])
    with pytest.raises(ValueError, match='must be sorted in strictly ascending order'):
        dq.sesolve(H, psi, np.array([0.0, 1.0, 0.5]))
    with pytest.raises(ValueError, match=re.escape('Option `vectorized` was set to `True` but is not used by the solver `dq.sesolve()`')):
        dq.sesolve(H, psi, ts, options=dq.Options(vectorized=True))
    with pytest.raises(TypeError, match=re.escape('Method of type `Rouchon1` is not supported (supported method types: `Tsit5`, `Dopri5`, `Euler`, `Expm`)')):
        dq.sesolve(H, psi, ts, method=dq.method.Rouchon1(dt=0.1))
    with pytest.raises(NotImplementedError, match='batched CallableTimeQArray'):
        dq.sesolve(dq.timecallable(lambda t: dq.asqarray(np.stack([H.to_numpy(), t * H.to_numpy()]))), psi, ts)


def test_sesolve_batching_shapes():
    """Cartesian batching concatenates (...H, ...psi0); flat batching broadcasts
    them (sesolve.py:234-260); each instance gets its H element and its ket."""
    from dynamiqs_amd.integrators.sesolve import enumerate_instances

    n = 3
    Hs = dq.asqarray(np.stack([k * np.eye(n) for k in range(4)]).reshape(2, 2, n, n))
    psis = dq.asqarray(np.stack([dq.fock(n, k).to_numpy() for k in range(2)]))
    shape, insts = enumerate_instances(dq.constant(Hs), psis, True)
    assert shape == (2, 2, 2) and len(insts) == 8
    np.testing.assert_array_equal(insts[5].H, 2 * np.eye(n))  # index (1, 0, 1)
    np.testing.assert_array_equal(insts[5].psi0, [0, 1, 0])
    shape, insts = enumerate_instances(dq.constant(Hs), psis, False)
    assert shape == (2, 2) and len(insts) == 4
    np.testing.assert_array_equal(insts[3].H, 3 * np.eye(n))
    np.testing.assert_array_equal(insts[3].psi0, [0, 1, 0])
    np.testing.assert_array_equal(insts[2].psi0, [1, 0, 0])


def test_sesolve_constant_batch_matches_instance_order():
    """The array packing of a constant batch (one staged block of Hamiltonians,
    index columns) lists the same (H, psi0) pairs in the same order as the
    per-instance enumeration, for cartesian and broadcast batching."""
    from dynamiqs_amd.integrators.operators import decompose
    from dynamiqs_amd.integrators.sesolve import constant_batch, enumerate_instances

    rng = np.random.default_rng(2)
    n = 3
    Hs = dq.asqarray(rng.normal(size=(2, 1, n, n)) + 0j)
    psis = dq.asqarray(rng.normal(size=(3, n, 1)) + 1j * rng.normal(size=(3, n, 1)))
    for cart in (True, False):
        H = dq.constant(Hs)
        shape, insts = enumerate_instances(H, psis, cart)
        shape2, Hflat, hi, pflat, pi = constant_batch(decompose(H), psis, cart)
        assert shape == shape2 and len(insts) == len(hi) == len(pi)
        for b, inst in enumerate(insts):
            np.testing.assert_array_equal(Hflat[hi[b]], inst.H)
            np.testing.assert_array_equal(pflat[pi[b]], inst.psi0)


def test_batched_scalar_multiplication():
    """A qarray times a scalar or a batch of scalars of shape (..., 1, 1) broadcasts
    (dynamiqs/_utils.py:48-57, qarray.py:496-504), as in the batching tutorial's
    omega[:, None, None] * sigma_z + epsilon[:, None, None, None] * sigma_x; any
    other array raises NotImplementedError."""
    omega = np.linspace(0.0, 10.0, 100)[:, None, None]
    epsilon = np.linspace(0.0, 1.0, 30)[:, None, None, None]
    H = omega * dq.sigmaz() + epsilon * dq.sigmax()
    assert H.shape == (30, 100, 2, 2)
    np.testing.assert_allclose(H.to_numpy()[3, 7], omega[7, 0, 0] * np.diag([1, -1]) + epsilon[3, 0, 0, 0] * np.array([[0, 1], [1, 0]]))
    assert (np.ones((4, 1, 1)) * dq.destroy(3, layout=dq.dia)).shape == (4, 3, 3)
    with pytest.raises(NotImplementedError, match='Element-wise multiplication of two qarrays'):
        dq.sigmaz() * np.ones(3)
    np.testing.assert_array_equal(dq.sigmaz().elmul(dq.sigmaz()).to_numpy(), np.eye(2))


def test_superoperator_assembly_from_hermitian_basis():
    """mepropagator's host assembly: the images of the n^2 Hermitian basis matrices
    under a random linear map S (given as an n^2 x n^2 matrix in the column-stacking
    convention) give back S exactly, for two save times and two batch elements."""
    from dynamiqs_amd.integrators.propagators import hermitian_basis, superoperator_from_basis

    rng = np.random.default_rng(9)
    n = 4
    basis = hermitian_basis(n)
    assert basis.shape == (n * n, n, n)
    for b in basis:
        np.testing.assert_array_equal(b, b.conj().T)
    # the basis spans the matrices: rank n^2 as vectors
    assert np.linalg.matrix_rank(basis.reshape(n * n, -1)) == n * n
    S = rng.normal(size=(2, 2, n * n, n * n)) + 1j * rng.normal(size=(2, 2, n * n, n * n))
    vec = lambda x: np.swapaxes(x, -1, -2).reshape(*x.shape[:-2], -1)  # noqa: E731
    unvec = lambda v: np.swapaxes(v.reshape(*v.shape[:-1], n, n), -1, -2)  # noqa: E731
    phi = np.stack([unvec(np.einsum('btrc,c->btr', S, vec(bm))) for bm in basis], axis=1)  # (B, n^2, ns, n, n)
    np.testing.assert_allclose(superoperator_from_basis(phi), S, atol=1e-12)


def test_propagator_argument_errors():
    """sepropagator / mepropagator checks (sepropagator.py:261-263,
    mepropagator.py:309-321, options.py:80-123) and the method registry."""
    a = dq.destroy(3)
    H, ts = a.dag() @ a, np.linspace(0.0, 1.0, 3)
    with pytest.raises(ValueError, match=re.escape('Argument `H` must have shape (...H, n, n)')):
        dq.sepropagator(np.ones((3, 2)), ts)
    with pytest.raises(ValueError, match=re.escape('Option `cartesian_batching` was set to `False` but is not used by the solver `dq.sepropagator()`')):
        dq.sepropagator(H, ts, options=dq.Options(cartesian_batching=False))
    with pytest.raises(ValueError, match=re.escape('Option `save_states` was set to `False` but is not used by the solver `dq.mepropagator()`')):
        dq.mepropagator(H, [a], ts, options=dq.Options(save_states=False))
    with pytest.raises(TypeError, match=re.escape('Method of type `Dopri8` is not supported (supported method types: `Expm`, `Tsit5`, `Dopri5`, `Euler`)')):
        dq.mepropagator(H, [a], ts, method=dq.method.Dopri8())
    with pytest.raises(TypeError, match='Method `Expm` requires a constant or piecewise constant Hamiltonian.'):
        dq.sepropagator(dq.modulated(np.cos, H), ts, method=dq.method.Expm())
    with pytest.raises(ValueError, match=re.escape('Argument `jump_ops[0]` must have shape (...L0, n, n)')):
        dq.mepropagator(H, [np.ones((3, 2))], ts)
    with pytest.warns(UserWarning, match='consider using `dq.sepropagator\\(\\)`'):
        try:
            dq.mepropagator(H, [], ts)
        except RuntimeError:
            pass  # no GPU here: the warning comes first


def test_expm_rejects_large_liouvillians_before_building_them():
    """n = 65 gives a 4225 x 4225 Liouvillian, above what the Expm kernels take: the
    call raises before any generator is built or any device memory is touched."""
    import time

    a = dq.destroy(65)
    ts = np.linspace(0.0, 1.0, 11)
    t0 = time.perf_counter()
    with pytest.raises(NotImplementedError, match='Expm'):
        dq.mepropagator(a.dag() @ a, [a], ts, method=dq.method.Expm())
    with pytest.raises(NotImplementedError, match='Expm'):
        dq.mesolve(a.dag() @ a, [a], dq.fock(65, 1), ts, method=dq.method.Expm())
    assert time.perf_counter() - t0 < 10.0


def test_expm_interval_bookkeeping_and_liouvillian():
    """Expm method host side: the interval boundaries are t0, the save times and the
    discontinuities clipped to [t0, t1] without zero-length intervals
    (expm_integrator.py:63-69); each save time maps to the number of intervals
    before it; the column-stacking Liouvillian equals the oracle's; the PWC operator
    at an interval's left end follows PWCTimeQArray (zero outside [t0, t_end))."""
    from oracle import lindblad

    from dynamiqs_amd.integrators.operators import decompose
    from dynamiqs_amd.integrators.propagators import _op_at, expm_times, ispwc, slindbladian

    times, nb = expm_times(0.0, np.array([0.0, 0.5, 1.0]), np.array([-1.0, 0.25, 0.5, 3.0]))
    np.testing.assert_array_equal(times, [0.0, 0.25, 0.5, 1.0])
    np.testing.assert_array_equal(nb, [0, 2, 3])
    times, nb = expm_times(-0.5, np.array([0.0, 1.0]), np.zeros(0))
    np.testing.assert_array_equal(times, [-0.5, 0.0, 1.0])
    np.testing.assert_array_equal(nb, [1, 2])
    rng = np.random.default_rng(6)
    n = 3
    H = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
    H = H + H.conj().T
    L = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
    np.testing.assert_allclose(slindbladian(H, [L]), lindblad.slindbladian(H, [L]), atol=1e-13)
    O = np.diag([1.0, 2.0, 3.0]).astype(complex)
    Ht = dq.asqarray(H) + dq.pwc(np.array([0.0, 1.0, 2.0]), np.array([5.0, -1.0]), dq.asqarray(O))
    terms = decompose(Ht)
    np.testing.assert_allclose(_op_at(terms, (), 0.0), H + 5 * O)
    np.testing.assert_allclose(_op_at(terms, (), 1.0), H - O)
    np.testing.assert_allclose(_op_at(terms, (), 2.0), H)
    assert ispwc(Ht) and not ispwc(dq.modulated(np.cos, dq.asqarray(O)))


def test_timecallable_low_rank_rewrite():
    """A timecallable H(t) is rewritten as a constant part plus at most four
    modulated terms spanning its samples: the rewrite reproduces f at times off the
    sampling grid to rounding, a constant callable becomes a constant operator, and
    the reference's TDQubit callable (closed_system.py:121-123) needs one term."""
    from dynamiqs_amd.integrators.operators import resolve_callables
    from dynamiqs_amd.time_qarray import ConstantTimeQArray, ModulatedTimeQArray

    H = dq.timecallable(lambda t: 3.0 * np.cos(10.0 * t) * dq.sigmax())
    R = resolve_callables(H, 0.0, 1.0)
    # f(t) O with the mean of f absorbed: one modulated term, no constant part
    assert isinstance(R, ModulatedTimeQArray)
    np.testing.assert_allclose(R(0.37).to_numpy(), H(0.37).to_numpy(), atol=1e-12)
    a = dq.destroy(4)
    f = lambda t: dq.asqarray(a.dag() @ a) + np.sin(2 * t) * (a + a.dag()) + t**2 * (a @ a + a.dag() @ a.dag())  # noqa: E731
    R = resolve_callables(dq.timecallable(f) + dq.asqarray(0.5 * np.eye(4)), -1.0, 2.0)
    for t in np.random.default_rng(3).uniform(-1.0, 2.0, 20):
        np.testing.assert_allclose(R(t).to_numpy(), f(t).to_numpy() + 0.5 * np.eye(4), atol=1e-12)
    assert isinstance(resolve_callables(dq.timecallable(lambda t: dq.sigmaz()), 0.0, 1.0), ConstantTimeQArray)


def test_dia_batched_operator_groups():
    """Batched DIA operators above the dense size run one DIA plan per distinct
    operator set: the grouping follows the reference's cartesian and flat batching
    rules (mesolve.py:307-336) and covers every output position once."""
    from dynamiqs_amd.integrators.mesolve import _dia_groups
    from dynamiqs_amd.integrators.operators import decompose
    from dynamiqs_amd.time_qarray import astimeqarray

    n = 6
    a = dq.destroy(n)
    eps, kap = np.array([0.1, 0.3]), np.array([0.05, 0.2, 0.4])
    H = a.dag() @ a + eps[:, None, None] * (a + a.dag())
    L = np.sqrt(kap)[:, None, None] * a
    rho = dq.asqarray(np.stack([dq.coherent(n, 0.5).to_numpy(), dq.coherent(n, 1.0).to_numpy()]))
    Ht, Lts = decompose(astimeqarray(H)), [decompose(astimeqarray(L))]
    ad = a.to_numpy()
    out, groups = _dia_groups(Ht, Lts, rho, True)
    assert out == (2, 3, 2) and len(groups) == 6
    assert sorted(p for _, m, _ in groups for p, _ in m) == list(range(12))
    for (Hd, Ld), members, terms in groups:
        assert terms == {}
        for pos, ir in members:
            i, j, k = np.unravel_index(pos, out)
            assert ir == (k,)
            np.testing.assert_allclose(Hd, ad.conj().T @ ad + eps[i] * (ad + ad.conj().T), atol=1e-6)
            np.testing.assert_allclose(Ld[0], np.sqrt(kap[j]) * ad, atol=1e-6)
    H2 = a.dag() @ a + np.array([0.1, 0.3, 0.5])[:, None, None] * (a + a.dag())
    out, groups = _dia_groups(decompose(astimeqarray(H2)), Lts, dq.coherent(n, 1.0), False)
    assert out == (3,) and len(groups) == 3
    assert all(len(m) == 1 and m[0][1] == () for _, m, _ in groups)


def test_jump_operator_sums_expand_to_the_lindbladian():
    """A jump operator that sums constant, modulated or piecewise-constant terms runs as
    D/C pairs (conj(c_j) for C, cross pairs j < l) plus generator terms. The vector field
    the kernels form from them, tmp = Hnh(t) rho + sum_k a_k D_k rho conj(b_k) C_k and
    tmp + tmp^dag, equals the oracle's Lindbladian with L(t) = sum_j c_j(t) L_j at
    several times (inside and outside the PWC range)."""
    from dynamiqs_amd.integrators import dense
    from oracle import lindblad

    rng = np.random.default_rng(3)
    n = 5

    def rnd():
        return rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))

    H = rnd()
    H = H + H.conj().T
    L0, L1, L2, M = rnd(), rnd(), rnd(), rnd()
    g = lambda t: 0.7 * np.exp(1j * t) + 0.2  # noqa: E731
    times, vals = np.array([0.0, 0.4, 1.0]), np.array([0.5 - 0.2j, 1.3])
    inst = dense.Instance(
        H=H, Ls=[L0, M], rho0=np.eye(n) / n,
        Lsum=[[(('const', 1.0), L0), (('mod', g, None), L1)],
              [(('const', 0.3j), M), (('pwc', times, vals), L2)]],
    )
    ex = dense.expand_jump_sums(inst)
    assert len(ex.Ls) == 6 and ex.expanded

    def coef(mod, pw, t):
        c = 1.0 + 0j
        if mod is not None:
            c *= mod[0](t)
        if pw is not None:
            c *= dense.pwc_value(pw[0], pw[1], t)
        return c

    rho = rnd()
    rho = rho @ rho.conj().T
    for t in (0.1, 0.55, 1.3):
        Hnh = ex.Hnh + sum(gm(t) * (-1j * o) for gm, o, _ in ex.mod) + sum(
            dense.pwc_value(tt, vv, t) * (-1j * o) for tt, vv, o in ex.pwc)
        tmp = Hnh @ rho
        for k in range(len(ex.Ls)):
            a = coef(ex.Lmod[k], ex.Lpwc[k], t)
            b = coef(ex.Lbmod[k], ex.Lbpwc[k], t)
            tmp = tmp + a * ex.Ls[k] @ rho @ (np.conj(b) * ex.Cs[k])
        got = tmp + tmp.conj().T
        Lt = [L0 + g(t) * L1, 0.3j * M + dense.pwc_value(times, vals, t) * L2]
        want = lindblad.rhs_standard(H, Lt, rho)
        np.testing.assert_allclose(got, want, atol=1e-10 * np.abs(want).max())
