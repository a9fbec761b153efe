"""CPU checks of the descriptor packing (``dynamiqs_amd.integrators.dense.build_plan``)
with the device replaced by host memory: every pointer a kernel would dereference
lies inside an allocation of the plan, staged operators land at their offsets, and
the pre-launch guard rejects a descriptor that does not."""

import numpy as np
import pytest
import torch

from dynamiqs_amd.integrators import dense


@pytest.fixture
def host_device(monkeypatch):
    monkeypatch.setattr(dense, '_device', lambda: torch.device('cpu'))
    orig = torch.empty

    def empty(*a, **k):
        k.pop('pin_memory', None)
        return orig(*a, **k)

    monkeypatch.setattr(torch, 'empty', empty)


def _inst(rng, n, K, pwc=False):
    H = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
    H = H + H.conj().T
    rho = np.eye(n) / n
    p = [(np.linspace(0, 1, 5), rng.normal(size=4), H)] if pwc else []
    return dense.Instance(H=H, Ls=[0.1 * H] * K, rho0=rho, pwc=p)


@pytest.mark.parametrize('kind,K,pwc', [('rhs', 0, False), ('tsit5', 2, False), ('tsit5', 1, True), ('rouchon1', 1, False)])
def test_descriptors_point_into_plan_allocations(host_device, kind, K, pwc):
    rng = np.random.default_rng(0)
    n = 20
    insts = [_inst(rng, n, K, pwc) for _ in range(3)]
    plan = dense.build_plan(kind, insts, np.array([0.0, 0.5, 1.0]), dt=0.1 if kind == 'rouchon1' else None, form='resident')
    tensors = [*plan.arena.tensors, plan.states, plan.final, plan.expects, plan.stats]
    dense.check_descriptors(plan.descs, tensors)
    if kind in ('rhs', 'tsit5'):
        # the staged Hnh of instance 1 holds -iH - 1/2 sum L^dag L
        H, Ls = insts[1].H, insts[1].Ls
        want = -1j * H - 0.5 * sum((L.conj().T @ L for L in Ls), np.zeros_like(H))
        base = plan.arena.tensors[0].data_ptr()
        off = int(plan.descs[1]['Hnh']) - base
        host = plan.arena.tensors[0].numpy()
        got = host[off : off + 32 * 32 * 8].view(np.complex64).reshape(32, 32)[:n, :n].astype(np.complex128)
        if kind == 'tsit5':
            # the single-size resident kernel takes the diagonal in FP64 (hdiag, exact)
            # and Hnh holds the rest
            assert plan.fp64_diag and plan.uniform_np == 32
            offd = int(plan.descs[1]['hdiag']) - base
            dg = host[offd : offd + 32 * 16].view(np.complex128)
            assert not np.any(np.diag(got)) and not np.any(dg[n:])
            np.testing.assert_array_equal(dg[:n], np.diag(want))
            got = got + np.diag(dg[:n])
        np.testing.assert_allclose(got, want, rtol=1e-6, atol=1e-5)


def test_guard_rejects_unpatched_descriptor(host_device):
    rng = np.random.default_rng(1)
    plan = dense.build_plan('rhs', [_inst(rng, 8, 1)], np.array([0.0, 1.0]), form='global')
    assert plan.descs['flags'][0] & 4  # the global kernel form is requested in the descriptor
    descs = plan.descs.copy()
    descs['Hnh'] = 0x101  # a raw staging offset, never a device address
    with pytest.raises(RuntimeError, match='outside the plan allocations'):
        dense.check_descriptors(descs, [*plan.arena.tensors, plan.states, plan.final, plan.expects, plan.stats])


def test_pwc_values_packed_as_resample_pwc(host_device):
    """The cached per-grid sampling of build_plan packs the same PWC coefficients as
    operators.resample_pwc on the union grid (two terms on different time grids,
    shared by several instances, one with its own grid)."""
    from dynamiqs_amd.integrators.operators import resample_pwc

    rng = np.random.default_rng(3)
    n = 8
    t1 = np.linspace(0.0, 1.0, 6)
    t2 = np.array([0.1, 0.35, 0.8, 1.2])
    t3 = np.linspace(0.0, 1.2, 4)
    H = np.diag(np.arange(n, dtype=float)).astype(complex)
    X = np.eye(n, k=1) + np.eye(n, k=-1)
    Y = 1j * np.eye(n, k=1) - 1j * np.eye(n, k=-1)
    insts = []
    for b in range(4):
        tt = t3 if b == 3 else t1
        pwc = [(tt, rng.normal(size=len(tt) - 1) + 1j * rng.normal(size=len(tt) - 1), X),
               (t2, rng.normal(size=3), Y)]
        insts.append(dense.Instance(H=H, Ls=[], rho0=np.eye(n) / n, pwc=pwc))
    plan = dense.build_plan('tsit5', insts, np.array([0.0, 0.5, 1.0]), form='resident', grid_tsit5=False)
    host = plan.arena.tensors[0].numpy()
    base = plan.arena.tensors[0].data_ptr()
    for b, inst in enumerate(insts):
        grid = np.unique(np.concatenate([t for (t, _, _) in inst.pwc]))
        want = np.stack([resample_pwc(t, np.asarray(v)[None], grid)[0] for (t, v, _) in inst.pwc])
        d = plan.descs[b]
        assert int(d['nseg']) == len(grid)
        off = int(d['pwc_vals']) - base
        got = host[off : off + want.size * 16].view(np.complex128).reshape(want.shape)
        np.testing.assert_array_equal(got, want)


def test_shared_kraus_sets_built_once_and_guarded(host_device, monkeypatch):
    """Rouchon1 with a modulated Hamiltonian term (one Kraus set per step): instances
    that share key_H / key_L (a batched rho0) build their sets once and point at the
    same device copy; the size guard counts D and C over the distinct sets of the
    batch."""
    rng = np.random.default_rng(3)
    n = 6
    H = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
    H = H + H.conj().T
    L = 0.2 * np.diag(np.sqrt(np.arange(1, n)), k=1)
    g = lambda t: np.cos(3.0 * t)  # noqa: E731
    calls = []
    orig = dense._kraus_sets

    def counting(*a, **k):
        calls.append(1)
        return orig(*a, **k)

    monkeypatch.setattr(dense, '_kraus_sets', counting)
    insts = [dense.Instance(H=H, Ls=[L], rho0=np.eye(n) / n * (1 + 0.1 * b), mod=[(g, H, ('g',))], key_H=('h',),
                            key_L=('l',), key_rho=(b,)) for b in range(3)]
    tsave = np.linspace(0.0, 0.2, 3)
    plan = dense.build_plan('rouchon1', insts, tsave, dt=0.01)
    nsteps = 20
    assert len(calls) == nsteps  # one set per step, built for the first instance only
    assert len(set(plan.descs['Dops'])) == 1 and len(set(plan.descs['Cops'])) == 1
    assert len(set(plan.descs['step_set'])) == 3  # each instance keeps its own index array
    assert np.all(plan.descs['max_steps'] == nsteps)
    dense.check_descriptors(plan.descs, [*plan.arena.tensors, plan.states, plan.final, plan.expects, plan.stats])
    # distinct keys: the guard sums the D and C sets of every instance
    monkeypatch.setattr(dense, 'KRAUS_SET_BUDGET', 2 * 2 * nsteps * 2 * 16 * 16 * 8 + 1)
    distinct = [dense.Instance(H=H, Ls=[L], rho0=np.eye(n) / n, mod=[(g, H, ('g',))], key_H=(b,), key_L=('l',),
                               key_rho=()) for b in range(3)]
    with pytest.raises(NotImplementedError, match='exceed'):
        dense.build_plan('rouchon1', distinct, tsave, dt=0.01)


def test_fp64_diagonal_split_leaves_caller_generator_intact(host_device):
    """An explicit generator part (Instance.Hnh, as the adjoint's backward solve
    passes) at a dimension that needs no padding: the plan stages its off-diagonal
    part and its diagonal in FP64 without writing into the caller's array."""
    rng = np.random.default_rng(4)
    n = 16
    H = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
    Hnh = -1j * (H + H.conj().T) - np.diag(rng.uniform(0.5, 2.0, size=n))
    keep = Hnh.copy()
    inst = dense.Instance(H=np.zeros((n, n), dtype=complex), Ls=[0.1 * np.eye(n, k=1)], rho0=np.eye(n) / n, Hnh=Hnh)
    plan = dense.build_plan('tsit5', [inst], np.array([0.0, 1.0]), form='resident')
    np.testing.assert_array_equal(Hnh, keep)
    assert plan.fp64_diag and plan.uniform_np == 16
    host = plan.arena.tensors[0].numpy()
    base = plan.arena.tensors[0].data_ptr()
    off = int(plan.descs[0]['Hnh']) - base
    offd = int(plan.descs[0]['hdiag']) - base
    got = host[off : off + n * n * 8].view(np.complex64).reshape(n, n)
    dg = host[offd : offd + n * 16].view(np.complex128)
    np.testing.assert_array_equal(dg, np.diag(keep))
    np.testing.assert_array_equal(got, (keep - np.diag(np.diag(keep))).astype(np.complex64))


def test_grid_plan_carries_no_fp64_diagonal(host_device, monkeypatch):
    """The grid-wide Tsit5 does not take the FP64 diagonal: a plan forced onto it
    stages the full generator part and leaves hdiag at zero."""
    monkeypatch.setattr(torch, 'zeros', lambda *a, **k: torch.empty(*a, **{kk: v for kk, v in k.items() if kk != 'pin_memory'}).zero_())
    rng = np.random.default_rng(5)
    insts = [_inst(rng, 20, 1) for _ in range(2)]
    plan = dense.build_plan('tsit5', insts, np.array([0.0, 0.5, 1.0]), form='resident', grid_tsit5=True)
    assert plan.grid is not None and not plan.fp64_diag
    assert np.all(plan.descs['hdiag'] == 0)
