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
propagator. ``Tsit5`` and ``Dopri5`` are supported. The reference's default for
constant and piecewise-constant generators, ``Expm`` (sepropagator.py:224-225,
mepropagator.py:269-270), is not built: ``method=None`` runs ``Tsit5``.
"""

from __future__ import annotations

import warnings
from dataclasses import replace

import numpy as np

from .._checks import check_shape, check_times
from ..gradient import Gradient
from ..method import Dopri5, Method, Tsit5
from ..options import Options, check_options
from ..qarrays import QArray
from ..result import AdaptiveStepInfos, MEPropagatorResult, SEPropagatorResult
from ..time_qarray import astimeqarray
from . import dense
from .mesolve import assert_method_supported, max_steps_error, save_extra_outputs
from .mesolve import enumerate_instances as me_instances
from .operators import decompose
from .sesolve import build_plan, build_plan_constant, constant_batch
from .sesolve import enumerate_instances as se_instances

SUPPORTED_METHODS = (Tsit5, Dopri5)


def _method_kwargs(method, options, save_all: bool) -> dict:
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


def _infos(res: dict, nbasis: int, out_shape: tuple) -> AdaptiveStepInfos:
    """Step counts per propagator: those of its slowest basis solve."""

    def per(x):
        x = np.asarray(x).reshape(-1, nbasis).max(axis=1).reshape(out_shape)
        return x if out_shape else x[()]

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
    return cls(tsave, method, gradient, options, QArray(U, dims=dims), None, extra, _infos(res, nbasis, out_shape))


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
    method = Tsit5() if method is None else method
    assert_method_supported(method, SUPPORTED_METHODS)
    method.assert_supports_gradient(gradient)

    n = H.shape[-1]
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
    method = Tsit5() if method is None else method
    assert_method_supported(method, SUPPORTED_METHODS)
    method.assert_supports_gradient(gradient)

    n = H.shape[-1]
    out_shape, insts = me_instances(H, Ls, QArray(np.zeros((n, n), dtype=np.complex128), dims=H.dims),
                                    options.cartesian_batching)
    basis = hermitian_basis(n)
    solve = [replace(inst, rho0=b, key_rho=None) for inst in insts for b in basis]
    kw = _method_kwargs(method, options, options.save_propagators or options.save_extra is not None)
    plan = dense.build_plan('tsit5', solve, tsave, **kw)
    plan.run()
    res = plan.results()
    _check_status(res['status'])
    st = res['states']  # (B n^2, ns, n, n)
    phi = st.reshape(len(insts), n * n, *st.shape[1:])
    U = superoperator_from_basis(phi)
    return _finish(U, H.dims, tuple(out_shape), tsave, method, gradient, options, res, n * n, MEPropagatorResult)
