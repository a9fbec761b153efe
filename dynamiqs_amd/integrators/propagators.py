"""Drop-in ``sepropagator`` and ``mepropagator`` running on the MI355X.

Mirrors ``dynamiqs/integrators/apis/sepropagator.py`` and
``dynamiqs/integrators/apis/mepropagator.py``: argument conversion and checks
(sepropagator.py:183-193, 261-263; mepropagator.py:212-229, 309-321), batching over
``H`` (and the jump operators), the method registry and ``SEPropagatorResult`` /
``MEPropagatorResult``.

A propagator is the solution for a basis of initial states, so each one is a batch of
independent solves on the existing kernels:

* ``sepropagator``: the n kets |j> through ``dqh_sesolve_tsit5``; column j of U(t) is
  psi_j(t).
* ``mepropagator``: n^2 Hermitian basis matrices through the dense Tsit5 kernels
  (which evolve Hermitian states): |i><i|, X_ij = |i><j| + |j><i| and
  Y_ij = -i(|i><j| - |j><i|) for i < j. By linearity the image of |i><j| is
  (X_ij(t) + i Y_ij(t)) / 2 and of |j><i| is (X_ij(t) - i Y_ij(t)) / 2; column
  j n + i of the superoperator is the column-stacked image of |i><j| (``vectorize``,
  ``dynamiqs/utils/vectorization.py:25-58``).

Every basis solve takes its own adaptive step (the reference steps the whole N x N
matrix with one controller); both are within the method's tolerance of the exact
propagator. ``Tsit5`` and ``Dopri5`` are supported, and ``Expm``, the reference's default for
constant and piecewise-constant generators (sepropagator.py:224-225,
mepropagator.py:269-270): as ``ExpmIntegrator.run`` (expm_integrator.py:60-90) the
generator A (-iH, or the Liouvillian ``slindbladian`` in the column-stacking
convention) is taken at the left end of every interval between t0, the save times
and the discontinuities, exponentiated and chained, U_k = expm(dt_k A_k) U_{k-1}.
The exponentials and products run in ``dqh_expm_propagators`` (Taylor degree 8 on
dt A / 2^s with ||dt A / 2^s||_1 <= 1/2, then s squarings, all as batched complex64
GEMMs on MFMA). With ``method=None`` the generator type picks ``Expm`` or ``Tsit5``
as in the reference.
"""

from __future__ import annotations

import warnings
from dataclasses import replace

import numpy as np

from .._checks import check_shape, check_times
from ..gradient import Gradient, reject_gradient
from .. import _native
from ..method import Dopri5, Euler, Expm, Method, Tsit5
from ..options import Options, check_options
from ..qarrays import QArray
from ..result import AdaptiveStepInfos, FixedStepInfos, MEPropagatorResult, SEPropagatorResult
from ..time_qarray import ConstantTimeQArray, PWCTimeQArray, SummedTimeQArray, astimeqarray
from . import dense
from .mesolve import assert_method_supported, max_steps_error, save_extra_outputs
from .mesolve import enumerate_instances as me_instances
from .operators import decompose, resolve_callables
from .sesolve import build_plan, build_plan_constant, constant_batch, pad16
from .sesolve import enumerate_instances as se_instances

SUPPORTED_METHODS = (Expm, Tsit5, Dopri5, Euler)
TAYLOR_DEGREE = 8  # remainder 0.5^9 / 9! e^0.5 < 1e-8 at ||X||_1 <= THETA
THETA = 0.5


def ispwc(x) -> bool:
    """Constant or piecewise constant, or a sum of such (``integrators/_utils.py:41-48``)."""
    if isinstance(x, (ConstantTimeQArray, PWCTimeQArray)):
        return True
    if isinstance(x, SummedTimeQArray):
        return all(ispwc(t) for t in x.timeqarrays)
    return False


def expm_times(t0: float, tsave: np.ndarray, disc: np.ndarray):
    """Interval boundaries of the Expm method: t0, the save times and the
    discontinuities inside [t0, tsave[-1]] (expm_integrator.py:63-69), without the
    zero-length intervals; and for each save time the number of intervals before it."""
    disc = np.clip(np.asarray(disc, dtype=np.float64), t0, tsave[-1])
    times = np.unique(np.concatenate([[t0], tsave, disc]))
    return times, np.searchsorted(times, tsave)


# Largest padded generator dimension dqh_expm_propagators takes (N = n^2 for the
# Liouvillian: n <= 64), and the memory one Expm call may use for its generators,
# propagators and work matrices (host complex128 copy included).
EXPM_MAX_N = 4096
EXPM_MAX_BYTES = 16 << 30


def check_expm_size(nchain: int, nint: int, N: int) -> None:
    """Raise before any generator is built when an Expm call would exceed what
    ``dqh_expm_propagators`` accepts or a sane memory footprint."""
    Np = pad16(N)
    if Np > EXPM_MAX_N:
        raise NotImplementedError(
            f'Method `Expm` exponentiates {N} x {N} generators here (at most {EXPM_MAX_N}: n <= 64 for the '
            'Liouvillian of mesolve / mepropagator); use `Tsit5` for larger systems.'
        )
    nbytes = nchain * nint * (N * N * 16 + 4 * Np * Np * 8)
    if nbytes > EXPM_MAX_BYTES:
        raise NotImplementedError(
            f'Method `Expm` would hold {nbytes / 2**30:.1f} GiB of generators and propagators for {nchain} chains of '
            f'{nint} intervals (limit {EXPM_MAX_BYTES / 2**30:.0f} GiB); use fewer save times or `Tsit5`.'
        )


def expm_chain(X: np.ndarray, Y0: np.ndarray | None = None) -> np.ndarray:
    """Chained exponentials on the GPU: ``X`` (nchain, nint, N, N) complex holds
    dt_k A_k; returns U (nchain, nint, N, N) complex64 with
    U[b, k] = expm(X[b, k]) ... expm(X[b, 0]) (``dqh_expm_propagators``), or, with
    ``Y0`` (nchain, N, M) a block of initial states per chain as columns, the states
    U[b, k] Y0[b] (nchain, nint, N, M) (``dqh_cgemm_batched`` on the resident U)."""
    import torch

    from .dense import _device

    nchain, nint, N = X.shape[0], X.shape[1], X.shape[-1]
    Np = pad16(N)
    norm = float(np.max(np.sum(np.abs(X), axis=-2), initial=0.0))  # largest 1-norm
    s = int(max(0, np.ceil(np.log2(norm / THETA)))) if norm > THETA else 0
    Xs = np.zeros((nchain, nint, Np, Np), dtype=np.complex64)
    Xs[..., :N, :N] = X / 2.0**s
    dev = _device()
    lib = _native.lib()
    x = torch.from_numpy(Xs.view(np.float32)).to(dev)
    work = torch.empty(int(lib.dqh_expm_work_bytes(nchain, nint, Np)), dtype=torch.uint8, device=dev)
    U = torch.empty_like(x)
    stream = torch.cuda.current_stream()
    c = _native.ctypes.c_void_p
    _native.check(
        lib.dqh_expm_propagators(c(x.data_ptr()), c(work.data_ptr()), c(U.data_ptr()), nchain, nint, Np,
                                 TAYLOR_DEGREE, s, c(stream.cuda_stream)),
        'dqh_expm_propagators',
    )
    if Y0 is not None:
        M = Y0.shape[-1]
        blocks = []
        Y = torch.empty_like(x)
        for c0 in range(0, M, Np):  # at most Np states per product
            c1 = min(M, c0 + Np)
            Yp = np.zeros((nchain, Np, Np), dtype=np.complex64)
            Yp[:, :N, : c1 - c0] = Y0[..., c0:c1]
            y0 = torch.from_numpy(Yp.view(np.float32)).to(dev)
            _native.check(
                lib.dqh_cgemm_batched(c(U.data_ptr()), Np * Np, c(y0.data_ptr()), Np * Np, nint, c(Y.data_ptr()), Np * Np,
                                      Np, nchain * nint, c(stream.cuda_stream)),
                'dqh_cgemm_batched',
            )
            torch.cuda.synchronize()
            blocks.append(Y.cpu().numpy().view(np.complex64).reshape(nchain, nint, Np, Np)[..., :N, : c1 - c0])
        return np.concatenate(blocks, axis=-1)
    torch.cuda.synchronize()
    out = U.cpu().numpy().view(np.complex64).reshape(nchain, nint, Np, Np)
    return out[..., :N, :N]


def _op_at(terms, idx, t: float) -> np.ndarray:
    """Constant part plus the PWC terms of batch element ``idx`` at time t
    (``PWCTimeQArray``: zero outside [times[0], times[-1]))."""
    out = terms.const_at(idx).copy()
    for times, vals, op in terms.pwc_at(idx):
        if times[0] <= t < times[-1]:
            out = out + vals[int(np.searchsorted(times, t, side='right')) - 1] * op
    return out


def _expm_result(ops_at, build, chains, N, dims, out_shape, tsave, t0, disc, method, gradient, options, cls):
    """Propagators by ``expm_chain``: ``ops_at(chain, t)`` gives the operators at t
    (a tuple of n x n arrays) and ``build(ops)`` the N x N generator."""
    times, nbefore = expm_times(t0, tsave, disc)
    nint = len(times) - 1
    ns = len(tsave)
    U = np.empty((len(chains), ns, N, N), dtype=np.complex64)
    if nint > 0:
        check_expm_size(len(chains), nint, N)
        dts = np.diff(times)
        X = np.empty((len(chains), nint, N, N), dtype=np.complex128)
        for b, c in enumerate(chains):
            cache = {}  # one generator per distinct set of operators (a constant one: once)
            for k in range(nint):
                ops = ops_at(c, times[k])
                key = b''.join(o.tobytes() for o in ops)
                if key not in cache:
                    cache[key] = build(ops)
                X[b, k] = dts[k] * cache[key]
        Uc = expm_chain(X)
    eye = np.eye(N, dtype=np.complex64)
    for i, nb in enumerate(nbefore):
        U[:, i] = eye if nb == 0 else Uc[:, nb - 1]
    infos = FixedStepInfos(np.full(out_shape, nint, dtype=np.int64) if out_shape else np.int64(nint))
    extra = None
    if options.save_extra is not None:
        extra = save_extra_outputs(options.save_extra, U, dims, out_shape)
    U = U.reshape(*out_shape, ns, N, N)
    if not options.save_propagators:
        U = U[..., -1:, :, :]
    return cls(tsave, method, gradient, options, QArray(U, dims=dims), None, extra, infos)


def slindbladian(H: np.ndarray, Ls) -> np.ndarray:
    """The Liouvillian as an n^2 x n^2 matrix in the column-stacking convention
    (``dynamiqs/utils/vectorization.py:100-259``: spre(x) = I (x) x,
    spost(x) = x^T (x) I, sprepost(x, y) = y^T (x) x)."""
    n = H.shape[-1]
    eye = np.eye(n)
    out = -1j * np.kron(eye, H) + 1j * np.kron(H.T, eye)
    for L in Ls:
        LdL = L.conj().T @ L
        out = out + np.kron(L.conj(), L) - 0.5 * np.kron(eye, LdL) - 0.5 * np.kron(LdL.T, eye)
    return out


def _method_kwargs(method, options, save_all: bool) -> dict:
    if isinstance(method, Euler):
        return dict(t0=None if options.t0 is None else float(options.t0), save_states=save_all,
                    euler_dt=float(method.dt))
    return dict(
        t0=None if options.t0 is None else float(options.t0),
        rtol=method.rtol,
        atol=method.atol,
        safety=method.safety_factor,
        factormin=method.min_factor,
        factormax=method.max_factor,
        max_steps=method.max_steps,
        save_states=save_all,
        tableau='dopri5' if isinstance(method, Dopri5) else 'tsit5',
    )


def _check_status(status) -> None:
    if np.any(status == -3):
        raise max_steps_error()
    if np.any(status == -4):
        raise RuntimeError('The solver stopped on a non-finite error estimate or a vanishing step size.')
    if np.any(status == -6):
        raise RuntimeError('an instance descriptor was rejected by the kernel (unsupported size or term count).')


def _infos(res: dict, nbasis: int, out_shape: tuple, fixed: bool = False):
    """Step counts per propagator: those of its slowest basis solve."""

    def per(x):
        x = np.asarray(x).reshape(-1, nbasis).max(axis=1).reshape(out_shape)
        return x if out_shape else x[()]

    if fixed:
        return FixedStepInfos(per(res['nsteps']))
    return AdaptiveStepInfos(per(res['nsteps']), per(res['naccepted']), per(res['nrejected']))


def _finish(U: np.ndarray, dims, out_shape, tsave, method, gradient, options, res, nbasis, cls):
    """U: (B, ns, N, N) -> result with the reference's shapes and saving rules."""
    N = U.shape[-1]
    extra = None
    if options.save_extra is not None:
        extra = save_extra_outputs(options.save_extra, U, dims, out_shape)
    U = U.reshape(*out_shape, -1, N, N)
    if not options.save_propagators:  # the final propagator only, nsave = 1 (save_mixin.py:26-41)
        U = U[..., -1:, :, :]
    infos = _infos(res, nbasis, out_shape, fixed=isinstance(method, Euler))
    return cls(tsave, method, gradient, options, QArray(U, dims=dims), None, extra, infos)


def sepropagator(
    H,
    tsave,
    *,
    method: Method | None = None,
    gradient: Gradient | None = None,
    options: Options = Options(),  # noqa: B008
) -> SEPropagatorResult:
    r"""Propagator U(t) of the Schrodinger equation on the MI355X
    (``dynamiqs/integrators/apis/sepropagator.py:26-193``): ``propagators`` has shape
    ``(...H, nsave, n, n)``."""
    H = astimeqarray(H)
    tsave = np.asarray(tsave, dtype=np.float64)
    check_shape(H, 'H', '(..., n, n)', subs={'...': '...H'})
    tsave = check_times(tsave, 'tsave')
    check_options(options, 'sepropagator')
    options = options.initialise()
    if method is None:  # default method (sepropagator.py:224-225)
        method = Expm() if ispwc(H) else Tsit5()
    assert_method_supported(method, SUPPORTED_METHODS)
    method.assert_supports_gradient(gradient)
    reject_gradient(gradient, '`dq.sepropagator()`')

    n = H.shape[-1]
    if isinstance(method, Expm):
        if not ispwc(H):
            raise TypeError('Method `Expm` requires a constant or piecewise constant Hamiltonian.')
        Ht = decompose(H)
        chains = list(np.ndindex(*Ht.batch)) if Ht.batch else [()]
        t0 = float(tsave[0] if options.t0 is None else options.t0)
        return _expm_result(lambda c, t: (_op_at(Ht, c, t),), lambda ops: -1j * ops[0], chains, n, H.dims,
                            tuple(Ht.batch), tsave, t0, H.discontinuity_ts, method, gradient, options,
                            SEPropagatorResult)
    H = resolve_callables(H, float(tsave[0] if options.t0 is None else options.t0), float(tsave[-1]))
    basis = QArray(np.eye(n, dtype=np.complex128)[:, :, None], dims=H.dims)  # (n, n, 1): the kets |j>
    kw = _method_kwargs(method, options, options.save_propagators or options.save_extra is not None)
    Ht = decompose(H)
    if Ht.is_constant:
        shape, Hs, h_index, psis, p_index = constant_batch(Ht, basis, True)
        plan = build_plan_constant(Hs, h_index, psis, p_index, tsave, **kw)
    else:
        shape, insts = se_instances(H, basis, True)
        plan = build_plan(insts, tsave, **kw)
    plan.run()
    res = plan.results()
    _check_status(res['status'])
    out_shape = tuple(shape[:-1])
    st = res['states']  # (B n, ns, n): column j of U is psi_j
    B = int(np.prod(out_shape, dtype=np.int64))
    U = st.reshape(B, n, st.shape[1], n).transpose(0, 2, 3, 1)  # (B, ns, p, j)
    return _finish(np.ascontiguousarray(U), H.dims, out_shape, tsave, method, gradient, options, res, n,
                   SEPropagatorResult)


def hermitian_basis(n: int) -> np.ndarray:
    """The n^2 Hermitian basis matrices of ``mepropagator``: |i><i| (i < n), then
    X_ij = |i><j| + |j><i| and Y_ij = -i(|i><j| - |j><i|) for the pairs i < j in
    ``np.triu_indices(n, 1)`` order."""
    iu, ju = np.triu_indices(n, 1)
    K = len(iu)
    out = np.zeros((n + 2 * K, n, n), dtype=np.complex128)
    out[np.arange(n), np.arange(n), np.arange(n)] = 1.0
    k = np.arange(K)
    out[n + k, iu, ju] = 1.0
    out[n + k, ju, iu] = 1.0
    out[n + K + k, iu, ju] = -1j
    out[n + K + k, ju, iu] = 1j
    return out


def superoperator_from_basis(phi: np.ndarray) -> np.ndarray:
    """Column-stacked superoperators from the images of ``hermitian_basis``:
    ``phi`` (B, n^2, ns, n, n) -> (B, ns, n^2, n^2), column j n + i = vec(image of
    |i><j|), vec(x)[q n + p] = x[p, q]."""
    B, m, ns, n, _ = phi.shape
    iu, ju = np.triu_indices(n, 1)
    K = len(iu)
    M = np.zeros((B, ns, n, n, n, n), dtype=np.result_type(phi.dtype, np.complex64))  # [b, t, i, j, p, q]
    d = np.arange(n)
    M[:, :, d, d] = np.moveaxis(phi[:, :n], 1, 2)
    X = np.moveaxis(phi[:, n : n + K], 1, 2)
    Y = np.moveaxis(phi[:, n + K :], 1, 2)
    M[:, :, iu, ju] = 0.5 * (X + 1j * Y)
    M[:, :, ju, iu] = 0.5 * (X - 1j * Y)
    return M.transpose(0, 1, 5, 4, 3, 2).reshape(B, ns, n * n, n * n)


def mepropagator(
    H,
    jump_ops,
    tsave,
    *,
    method: Method | None = None,
    gradient: Gradient | None = None,
    options: Options = Options(),  # noqa: B008
) -> MEPropagatorResult:
    r"""Propagator of the Lindblad master equation on the MI355X
    (``dynamiqs/integrators/apis/mepropagator.py:37-229``): ``propagators`` has shape
    ``(...H, ...L0, ..., nsave, n^2, n^2)`` (cartesian batching) in the column-stacking
    convention of ``dq.vectorize``."""
    H = astimeqarray(H)
    Ls = [astimeqarray(L) for L in jump_ops]
    tsave = np.asarray(tsave, dtype=np.float64)
    check_shape(H, 'H', '(..., n, n)', subs={'...': '...H'})
    for i, L in enumerate(Ls):
        check_shape(L, f'jump_ops[{i}]', '(..., n, n)', subs={'...': f'...L{i}'})
    if len(Ls) == 0:
        warnings.warn(
            'Argument `jump_ops` is an empty list, consider using `dq.sepropagator()`'
            ' to compute propagators for the Schrödinger equation.',
            stacklevel=2,
        )
    tsave = check_times(tsave, 'tsave')
    check_options(options, 'mepropagator')
    options = options.initialise()
    if method is None:  # default method (mepropagator.py:269-270)
        method = Expm() if all(ispwc(x) for x in [H, *Ls]) else Tsit5()
    assert_method_supported(method, SUPPORTED_METHODS)
    method.assert_supports_gradient(gradient)
    reject_gradient(gradient, '`dq.mepropagator()`')

    n = H.shape[-1]
    if isinstance(method, Expm):
        if not ispwc(H):
            raise TypeError('Method `Expm` requires a constant or piecewise constant Hamiltonian.')
        if not all(ispwc(L) for L in Ls):
            raise TypeError('Method `Expm` requires constant or piecewise constant jump operators.')
        Ht = decompose(H)
        Lts = [decompose(L) for L in Ls]
        shapes = [Ht.batch, *[Lt.batch for Lt in Lts]]
        if options.cartesian_batching:
            out_shape = tuple(d for sh in shapes for d in sh)
        else:
            out_shape = tuple(np.broadcast_shapes(*shapes))
        chains = []
        for idx in np.ndindex(*out_shape) if out_shape else [()]:
            if options.cartesian_batching:
                parts, o = [], 0
                for sh in shapes:
                    parts.append(tuple(idx[o : o + len(sh)]))
                    o += len(sh)
            else:
                parts = [tuple(i if d > 1 else 0 for i, d in zip(idx[len(idx) - len(sh) :] if sh else (), sh))
                         for sh in shapes]
            chains.append(parts)
        disc = np.unique(np.concatenate([H.discontinuity_ts, *[L.discontinuity_ts for L in Ls]]))
        t0 = float(tsave[0] if options.t0 is None else options.t0)

        def ops_at(parts, t):
            return (_op_at(Ht, parts[0], t), *[_op_at(Lt, p, t) for Lt, p in zip(Lts, parts[1:])])

        return _expm_result(ops_at, lambda ops: slindbladian(ops[0], ops[1:]), chains, n * n, H.dims, out_shape,
                            tsave, t0, disc, method, gradient, options, MEPropagatorResult)
    t_first = float(tsave[0] if options.t0 is None else options.t0)
    H = resolve_callables(H, t_first, float(tsave[-1]))
    Ls = [resolve_callables(L, t_first, float(tsave[-1])) for L in Ls]
    out_shape, insts = me_instances(H, Ls, QArray(np.zeros((n, n), dtype=np.complex128), dims=H.dims),
                                    options.cartesian_batching)
    basis = hermitian_basis(n)
    solve = [replace(inst, rho0=b, key_rho=None) for inst in insts for b in basis]
    kw = _method_kwargs(method, options, options.save_propagators or options.save_extra is not None)
    if isinstance(method, Euler):  # the dense Kraus kernel with the Euler sets (no normalisation)
        plan = dense.build_plan('rouchon1', solve, tsave, t0=kw['t0'], dt=kw['euler_dt'], normalize=False,
                                save_states=kw['save_states'], rouchon_order=0)
    else:
        plan = dense.build_plan('tsit5', solve, tsave, **kw)
    plan.run()
    res = plan.results()
    _check_status(res['status'])
    st = res['states']  # (B n^2, ns, n, n)
    phi = st.reshape(len(insts), n * n, *st.shape[1:])
    U = superoperator_from_basis(phi)
    return _finish(U, H.dims, tuple(out_shape), tsave, method, gradient, options, res, n * n, MEPropagatorResult)


def _batch_parts(shapes, cartesian: bool):
    """(out_shape, [per-instance tuple of per-argument batch indices]) with the
    reference's cartesian / broadcast batching rules."""
    if cartesian:
        out_shape = tuple(d for sh in shapes for d in sh)
    else:
        out_shape = tuple(np.broadcast_shapes(*shapes))
    parts_all = []
    for idx in np.ndindex(*out_shape) if out_shape else [()]:
        if cartesian:
            parts, o = [], 0
            for sh in shapes:
                parts.append(tuple(idx[o : o + len(sh)]))
                o += len(sh)
        else:
            parts = [tuple(i if d > 1 else 0 for i, d in zip(idx[len(idx) - len(sh) :] if sh else (), sh)) for sh in shapes]
        parts_all.append(tuple(parts))
    return out_shape, parts_all


def expm_solve(H, Ls, y0: QArray, tsave, Es, method, gradient, options, me: bool):
    """``sesolve`` / ``mesolve`` with ``Expm`` (expm_integrator.py:60-90 with a state
    as carry): the instances that share their operators form one chain whose
    propagators are applied on the GPU to a block of initial states (psi0, or
    vec(rho0) in the column-stacking convention, as columns). The expectation
    values <psi|E|psi> / Tr(E rho) are formed on the host from the saved states."""
    from ..result import MESolveResult, SESolveResult

    if not ispwc(H):
        raise TypeError('Method `Expm` requires a constant or piecewise constant Hamiltonian.')
    if me and not all(ispwc(L) for L in Ls):
        raise TypeError('Method `Expm` requires constant or piecewise constant jump operators.')
    Ht = decompose(H)
    Lts = [decompose(L) for L in Ls]
    n = H.shape[-1]
    y = y0.to_numpy().astype(np.complex128)
    yb = y.shape[:-2]
    y = y.reshape(-1, *y.shape[-2:])
    if me:
        if y.shape[-1] == 1:
            y = y @ np.conj(np.swapaxes(y, -1, -2))
        cols = np.swapaxes(y, -1, -2).reshape(len(y), n * n)  # vec(rho0), column stacking
        N = n * n
    else:
        cols = y[..., 0]
        N = n
    shapes = [Ht.batch, *[Lt.batch for Lt in Lts], yb]
    out_shape, parts_all = _batch_parts(shapes, options.cartesian_batching)
    chain_keys = sorted({p[:-1] for p in parts_all})
    chain_of = {k: i for i, k in enumerate(chain_keys)}
    members = [[] for _ in chain_keys]  # per chain: (instance, y index)
    for b, p in enumerate(parts_all):
        yi = int(np.ravel_multi_index(p[-1], yb)) if yb else 0
        members[chain_of[p[:-1]]].append((b, yi))
    M = max(len(m) for m in members)
    Y0 = np.zeros((len(chain_keys), N, M), dtype=np.complex128)
    for ci, mem in enumerate(members):
        for j, (_, yi) in enumerate(mem):
            Y0[ci, :, j] = cols[yi]
    disc = np.unique(np.concatenate([H.discontinuity_ts, *[L.discontinuity_ts for L in Ls]]))
    t0 = float(tsave[0] if options.t0 is None else options.t0)
    times, nbefore = expm_times(t0, tsave, disc)
    nint = len(times) - 1
    B, ns = len(parts_all), len(tsave)
    out = np.empty((B, ns, N), dtype=np.complex64)
    if nint > 0:
        check_expm_size(len(chain_keys), nint, N)
        dts = np.diff(times)
        X = np.empty((len(chain_keys), nint, N, N), dtype=np.complex128)
        for ci, key in enumerate(chain_keys):
            cache = {}
            for k in range(nint):
                ops = (_op_at(Ht, key[0], times[k]), *[_op_at(Lt, p, times[k]) for Lt, p in zip(Lts, key[1:])])
                kb = b''.join(o.tobytes() for o in ops)
                if kb not in cache:
                    cache[kb] = slindbladian(ops[0], ops[1:]) if me else -1j * ops[0]
                X[ci, k] = dts[k] * cache[kb]
        Y = expm_chain(X, Y0)
    for ci, mem in enumerate(members):
        for j, (b, _) in enumerate(mem):
            for i, nb in enumerate(nbefore):
                out[b, i] = Y0[ci, :, j] if nb == 0 else Y[ci, nb - 1, :, j]
    if me:
        states = np.swapaxes(out.reshape(B, ns, n, n), -1, -2)  # unvec
        Esave = None if Es is None else np.stack([np.einsum('ij,btji->bt', E, states) for E in Es], axis=1)
    else:
        states = out[..., None]
        Esave = None if Es is None else np.stack(
            [np.einsum('bti,ij,btj->bt', states[..., 0].conj(), E, states[..., 0]) for E in Es], axis=1)
    extra = None
    if options.save_extra is not None:
        extra = save_extra_outputs(options.save_extra, states, y0.dims, out_shape)
    states = states.reshape(*out_shape, ns, *states.shape[-2:])
    if not options.save_states:
        states = states[..., -1:, :, :]
    if Esave is not None:
        Esave = Esave.astype(np.complex64).reshape(*out_shape, len(Es), ns)
    infos = FixedStepInfos(np.full(out_shape, nint, dtype=np.int64) if out_shape else np.int64(nint))
    cls = MESolveResult if me else SESolveResult
    return cls(tsave, method, gradient, options, QArray(states, dims=y0.dims), Esave, extra, infos)
