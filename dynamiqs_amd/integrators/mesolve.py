"""Drop-in ``mesolve`` running on the MI355X.

Mirrors ``dynamiqs/integrators/apis/mesolve.py``: argument conversion (:280-285),
checks (:288-291, :390-414), cartesian / flat batching (:307-336,
``integrators/_utils.py:93-187``) and the method registry (:350-366). The
integrator seam (``integrator_constructor(...).run()``, :372-384) is served by
the HIP kernels through ``dynamiqs_amd.integrators.dense``: every batch element
becomes one instance, solved by one persistent workgroup with its own adaptive
step.
"""

from __future__ import annotations

import warnings

from dataclasses import replace

import numpy as np

from .._checks import check_hermitian, check_qarray_is_dense, check_shape, check_times
from ..gradient import Gradient, reject_gradient
from ..method import Dopri5, Dopri8, Euler, Expm, Method, Rouchon1, Rouchon2, Rouchon3, Tsit5, _DEAdaptiveStep
from ..options import Options, check_options
from ..qarrays import QArray, asqarray
from ..result import AdaptiveStepInfos, FixedStepInfos, MESolveResult
from ..time_qarray import TimeQArray, astimeqarray
from . import dense
from . import dia as dia_path
from .operators import decompose, resolve_callables

SUPPORTED_METHODS = (Tsit5, Dopri5, Dopri8, Euler, Rouchon1, Rouchon2, Rouchon3, Expm)
ADAPTIVE = (Tsit5, Dopri5, Dopri8)


def _tableau(method) -> str:
    return 'dopri8' if isinstance(method, Dopri8) else ('dopri5' if isinstance(method, Dopri5) else 'tsit5')


def assert_method_supported(method: Method, supported_methods) -> None:
    """``integrators/_utils.py:84-90``."""
    if not isinstance(method, tuple(supported_methods)):
        supported_str = ', '.join(f'`{x.__name__}`' for x in supported_methods)
        raise TypeError(
            f'Method of type `{type(method).__name__}` is not supported (supported method types: {supported_str}).'
        )


def max_steps_error() -> RuntimeError:
    """Message of ``integrators/_utils.py:66-72``."""
    default_max_steps = _DEAdaptiveStep.max_steps
    return RuntimeError(
        'The maximum number of method steps has been reached (the default'
        f' value is `max_steps={default_max_steps:_}`). Try increasing'
        ' `max_steps` with the `method` argument, e.g.'
        ' `method=dq.method.Tsit5(max_steps=1_000_000)`.'
    )


def _fixed_step_order(method) -> int:
    """The dense Kraus kernel's map for a fixed-step method: 0 Euler, 1-3 Rouchon."""
    if isinstance(method, Euler):
        return 0
    return 3 if isinstance(method, Rouchon3) else (2 if isinstance(method, Rouchon2) else 1)


def _check_mesolve_args(H: TimeQArray, Ls, rho0: QArray, exp_ops) -> None:
    check_shape(H, 'H', '(..., n, n)', subs={'...': '...H'})
    for i, L in enumerate(Ls):
        check_shape(L, f'jump_ops[{i}]', '(..., n, n)', subs={'...': f'...L{i}'})
    if len(Ls) == 0 and rho0.isket():
        warnings.warn(
            'Argument `jump_ops` is an empty list and argument `rho0` is a ket,'
            ' consider using `dq.sesolve()` to solve the Schrödinger equation.',
            stacklevel=3,
        )
    check_shape(rho0, 'rho0', '(..., n, 1)', '(..., n, n)', subs={'...': '...rho0'})
    check_qarray_is_dense(rho0, 'rho0')
    if exp_ops is not None:
        for i, E in enumerate(exp_ops):
            check_shape(E, f'exp_ops[{i}]', '(n, n)')


def _broadcast_index(idx: tuple, shape: tuple) -> tuple:
    sub = idx[len(idx) - len(shape) :] if shape else ()
    return tuple(i if d > 1 else 0 for i, d in zip(sub, shape))


# Constant un-batched DIA operators run on the DIA sweeps from this dimension on
# (smaller problems fit one workgroup of the dense kernel, which is faster there)
DIA_MIN_N = 128


def dia_route(method, H: TimeQArray, Ls, n: int):
    """Operators for the DIA kernels, or None when the dense path applies: Tsit5, or
    Rouchon1 with normalisation, constant DIA operators, n >= DIA_MIN_N; batched
    operators only above the dense kernels' size limit (``('batched', H, Ls)``)."""
    if n < DIA_MIN_N:
        return None
    if isinstance(method, Dopri8):  # the 13-stage pair runs on the dense kernels only
        return None
    if not (isinstance(method, ADAPTIVE) or (isinstance(method, Rouchon1) and method.normalize)):
        return None
    Ht = decompose(H)
    Lts = [decompose(L) for L in Ls]
    if any(Lt.dia is None for Lt in Lts):
        return None
    if Ht.dia is None:
        # PWC / modulated Hamiltonian terms on the DIA Tsit5 sweeps (the per-stage
        # generator rebuild of dqh_mesolve_tsit5_dia) where the dense kernels cannot take
        # the size: un-batched operators, constant jump operators
        if n > dense.MAX_NP and isinstance(method, ADAPTIVE) and (Ht.pwc or Ht.mod):
            return 'timedep', Ht, Lts
        return None
    if Ht.batch == () and all(Lt.batch == () for Lt in Lts):
        return Ht.const_at(()), [Lt.const_at(()) for Lt in Lts]
    # batched constant DIA operators: one DIA plan per distinct operator set, where
    # the dense kernels (one launch for the whole batch) cannot take the size
    if n > dense.MAX_NP:
        return 'batched', Ht, Lts
    return None


def enumerate_instances(H: TimeQArray, Ls, rho0: QArray, cartesian: bool):
    """Return (out_shape, [Instance]) following the reference's batching rules."""
    Ht = decompose(H)
    Lts = [decompose(L) for L in Ls]
    for i, Lt in enumerate(Lts):
        # a time-dependent jump operator is one modulated or PWC term, or a sum of
        # constant and modulated terms, or of constant and PWC terms (expanded into D/C
        # pairs by the dense plan)
        if Lt.mod and Lt.pwc:
            raise NotImplementedError(
                'time-dependent jump operators are supported by the MI355X solver as sums of constant and modulated'
                ' terms, or of constant and piecewise-constant terms, not both kinds'
            )
    r = rho0.to_numpy().astype(np.complex128)
    rb = r.shape[:-2]
    shapes = [Ht.batch, *[Lt.batch for Lt in Lts], rb]
    if cartesian:
        out_shape = tuple(d for s in shapes for d in s)
    else:
        out_shape = tuple(np.broadcast_shapes(*shapes))
    insts = []
    for idx in np.ndindex(*out_shape) if out_shape else [()]:
        if cartesian:
            parts, o = [], 0
            for s in shapes:
                parts.append(tuple(idx[o : o + len(s)]))
                o += len(s)
        else:
            parts = [_broadcast_index(idx, s) for s in shapes]
        iH, iLs, ir = parts[0], parts[1:-1], parts[-1]
        rho = r[ir]
        if rho.shape[-1] == 1:
            rho = rho @ rho.conj().T
        Lmods, Lpwcs, Lsums, Lops = [], [], [], []
        for Lt, i in zip(Lts, iLs):
            nterms = len(Lt.mod) + len(Lt.pwc) + (Lt.const is not None and bool(np.any(Lt.const)))
            if (Lt.mod or Lt.pwc) and nterms > 1:
                terms = []
                if Lt.const is not None and np.any(Lt.const):
                    terms.append((('const', 1.0), Lt.const_at(i)))
                for g, o, f in Lt.mod_at(i):
                    terms.append((('mod', g, (id(f), i if np.ndim(f(0.0)) else None)), o))
                for times, vals, o in Lt.pwc_at(i):
                    terms.append((('pwc', times, vals), o))
                Lmods.append(None)
                Lpwcs.append(None)
                Lsums.append(terms)
                Lops.append(terms[0][1])
                continue
            Lsums.append(None)
            if Lt.mod:
                g, _, f = Lt.mod_at(i)[0]
                Lmods.append((g, (id(f), i if np.ndim(f(0.0)) else None)))
                Lpwcs.append(None)
                Lops.append(Lt.mod[0][1])
            elif Lt.pwc:
                times, vals, o = Lt.pwc_at(i)[0]
                Lmods.append(None)
                Lpwcs.append((times, vals))
                Lops.append(o)
            else:
                Lmods.append(None)
                Lpwcs.append(None)
                Lops.append(Lt.const_at(i))
        insts.append(
            dense.Instance(
                H=Ht.const_at(iH),
                Ls=Lops,
                Lmod=Lmods if any(m is not None for m in Lmods) else [],
                Lpwc=Lpwcs if any(m is not None for m in Lpwcs) else [],
                Lsum=Lsums if any(m is not None for m in Lsums) else [],
                rho0=rho,
                pwc=Ht.pwc_at(iH),
                mod=[(g, o, (id(f), iH if np.ndim(f(0.0)) else None)) for (g, o, f) in Ht.mod_at(iH)],
                key_H=iH,
                key_L=tuple(iLs),
                key_rho=ir,
            )
        )
    return out_shape, insts


def mesolve(
    H,
    jump_ops,
    rho0,
    tsave,
    *,
    exp_ops=None,
    method: Method = Tsit5(),  # noqa: B008
    gradient: Gradient | None = None,
    options: Options = Options(),  # noqa: B008
) -> MESolveResult:
    r"""Solve the Lindblad master equation on the MI355X.

    Same signature, batching rules, errors and result as ``dq.mesolve``
    (``dynamiqs/integrators/apis/mesolve.py:62-304``). Methods: ``Tsit5`` and
    ``Dopri5`` (adaptive; with ``Options(vectorized=True)`` the Liouvillian as a
    superoperator, n <= 32), ``Dopri8`` (adaptive 8(7) pair on the dense kernels,
    n <= 1024, steps landing on the save times), ``Euler``, ``Rouchon1``, ``Rouchon2`` /
    ``Rouchon3`` (fixed step with ``dt``; without ``dt`` the adaptive 1-2 / 2-3 pairs in
    FP64, n <= 64, constant jump operators), ``Expm``. Operators: constant (dense or DIA; constant DIA
    operators from n = 128 on run on the DIA kernels), piecewise-constant
    (``dq.pwc``), modulated and ``timecallable`` Hamiltonians, constant jump
    operators, modulated or PWC terms and their sums with constant terms. ``gradient`` is validated as in the
    reference and then rejected (no autodiff here; see ``dynamiqs_amd.adjoint``).
    """
    # === convert arguments
    H = astimeqarray(H)
    Ls = [astimeqarray(L) for L in jump_ops]
    rho0 = asqarray(rho0)
    tsave = np.asarray(tsave, dtype=np.float64)
    if exp_ops is not None:
        exp_ops = [asqarray(E) for E in exp_ops] if len(exp_ops) > 0 else None

    # === check arguments
    _check_mesolve_args(H, Ls, rho0, exp_ops)
    tsave = check_times(tsave, 'tsave')
    check_options(options, 'mesolve')
    options = options.initialise()
    assert_method_supported(method, SUPPORTED_METHODS)
    method.assert_supports_gradient(gradient)
    reject_gradient(gradient, '`dq.mesolve()`')
    # Rouchon2() / Rouchon3() without dt: the adaptive pairs (rouchon_integrator.py:620-651)
    adaptive_rouchon = isinstance(method, (Rouchon2, Rouchon3)) and method.dt is None
    if isinstance(method, Expm):  # propagators on MFMA applied to the initial states
        from .propagators import expm_solve

        Es_np = [E.to_numpy() for E in exp_ops] if exp_ops is not None else None
        return expm_solve(H, Ls, rho0, tsave, Es_np, method, gradient, options, me=True)
    # timecallable operators become constant + modulated terms over the solve interval
    t_first = float(tsave[0] if options.t0 is None else options.t0)
    H = resolve_callables(H, t_first, float(tsave[-1]))
    Ls = [resolve_callables(L, t_first, float(tsave[-1])) for L in Ls]

    if options.vectorized and isinstance(method, ADAPTIVE):
        if isinstance(method, Dopri8):
            raise NotImplementedError('Option `vectorized=True` is supported with Tsit5 and Dopri5.')
        return _mesolve_vectorized(H, Ls, rho0, tsave, exp_ops, method, gradient, options)

    route = dia_route(method, H, Ls, H.shape[-1])
    if route is not None:
        return _mesolve_dia(route, rho0, tsave, exp_ops, method, gradient, options)

    out_shape, insts = enumerate_instances(H, Ls, rho0, options.cartesian_batching)
    if options.assume_hermitian and isinstance(method, ADAPTIVE):
        for inst in insts:
            check_hermitian(inst.rho0, 'y0')
    parts, pindex = hermitian_parts([inst.rho0 for inst in insts], options.assume_hermitian and isinstance(method, ADAPTIVE))
    if len(parts) > len(insts):
        solve = []
        for inst, (ia, ib) in zip(insts, pindex):
            solve.append(replace(inst, rho0=parts[ia], key_rho=None))
            if ib >= 0:
                solve.append(replace(inst, rho0=parts[ib], key_rho=None))
        insts = solve
    else:
        insts = [replace(inst, rho0=a) for inst, a in zip(insts, parts)]
    Es = [E.to_numpy() for E in exp_ops] if exp_ops is not None else None
    t0 = None if options.t0 is None else float(options.t0)
    # save_extra sees the state at every save time, whatever save_states is
    save_all = options.save_states or options.save_extra is not None

    if adaptive_rouchon:
        from . import rouchon_adaptive

        plan = rouchon_adaptive.build_plan(
            insts,
            tsave,
            order=3 if isinstance(method, Rouchon3) else 2,
            exp_ops=Es,
            t0=t0,
            rtol=method.rtol,
            atol=method.atol,
            safety=method.safety_factor,
            factormin=method.min_factor,
            factormax=method.max_factor,
            max_steps=method.max_steps,
            normalize=method.normalize,
            save_states=save_all,
        )
    elif isinstance(method, ADAPTIVE):
        plan = dense.build_plan(
            'tsit5',
            insts,
            tsave,
            exp_ops=Es,
            t0=t0,
            rtol=method.rtol,
            atol=method.atol,
            safety=method.safety_factor,
            factormin=method.min_factor,
            factormax=method.max_factor,
            max_steps=method.max_steps,
            save_states=save_all,
            tableau=_tableau(method),
        )
    else:
        plan = dense.build_plan(
            'rouchon1',
            insts,
            tsave,
            exp_ops=Es,
            t0=t0,
            dt=method.dt,
            normalize=getattr(method, 'normalize', False),
            save_states=save_all,
            rouchon_order=_fixed_step_order(method),
        )
    plan.run()
    res = plan.results()
    if adaptive_rouchon:  # FP64 on the device; results in the dtype of the other paths
        res['states'] = res['states'].astype(np.complex64)
        if 'expects' in res:
            res['expects'] = res['expects'].astype(np.complex64)
    if np.any(res['status'] == -6):
        raise RuntimeError('an instance was rejected by the MI355X kernel (unsupported shape or options)')
    if np.any(res['status'] == -3):
        raise max_steps_error()
    if np.any(res['status'] == -4):
        raise RuntimeError('The solver stopped on a non-finite error estimate or a vanishing step size.')

    n = plan.n
    states = combine_parts(res['states'], pindex)
    extra = None
    if options.save_extra is not None:
        extra = save_extra_outputs(options.save_extra, states, rho0.dims, out_shape)
    states = states.reshape(*out_shape, -1, n, n)
    if not options.save_states:  # the final state only, nsave = 1 (save_mixin.py:62-66, mesolve.py:184-187)
        states = states[..., -1:, :, :]
    ysave = QArray(states, dims=rho0.dims)
    Esave = None
    if Es is not None:
        Esave = combine_parts(res['expects'], pindex).reshape(*out_shape, len(Es), len(tsave))

    def shaped(x):
        x = np.asarray(x)[pindex[:, 0]].reshape(out_shape)
        return x if out_shape else x[()]

    if isinstance(method, ADAPTIVE) or adaptive_rouchon:
        infos = AdaptiveStepInfos(shaped(res['nsteps']), shaped(res['naccepted']), shaped(res['nrejected']))
    else:
        infos = FixedStepInfos(shaped(res['nsteps']))
    return MESolveResult(tsave, method, gradient, options, ysave, Esave, extra, infos)


# n^2 <= this: the vectorized form runs on the sesolve kernel variant of that size
VECTORIZED_MAX_N2 = 1024


def _mesolve_vectorized(H, Ls, rho0, tsave, exp_ops, method, gradient, options) -> MESolveResult:
    """``Options(vectorized=True)``: the reference's ``vector_field_vec``
    (diffrax_integrator.py:325-327), d vec(rho)/dt = slindbladian(H, Ls) vec(rho) in
    the column-stacking convention (utils/vectorization.py:213-259), integrated as a
    GEMV on the ``dqh_sesolve_tsit5`` kernel with the n^2 x n^2 superoperator as
    generator (its PWC and modulated Hamiltonian terms as superoperator terms). The
    adaptive controller sees the same n^2 entries as the matrix form. Expectation
    values are taken from the saved states on the host."""
    from .propagators import slindbladian
    from .sesolve import SEInstance
    from .sesolve import build_plan as se_build_plan

    out_shape, insts = enumerate_instances(H, Ls, rho0, options.cartesian_batching)
    n = H.shape[-1]
    if n * n > VECTORIZED_MAX_N2:
        raise NotImplementedError(
            f'Option `vectorized=True` integrates the {n * n} x {n * n} Liouvillian as a matrix-vector product on the '
            f'MI355X, which supports n^2 <= {VECTORIZED_MAX_N2} (n <= 32); use the default unvectorized form.'
        )
    if any(inst.Lmod or inst.Lpwc or inst.Lsum for inst in insts):
        raise NotImplementedError('Option `vectorized=True` is supported with constant jump operators only.')
    if options.assume_hermitian:
        for inst in insts:
            check_hermitian(inst.rho0, 'y0')
    sup = lambda O: 1j * slindbladian(O, [])  # noqa: E731  (G = -i (1j S) = S for a term O: -i[O, .])
    se_insts = []
    for inst in insts:
        se_insts.append(SEInstance(
            H=1j * slindbladian(inst.H, inst.Ls),
            psi0=np.swapaxes(inst.rho0, -1, -2).reshape(-1),  # vec(rho), column stacking
            pwc=[(t, v, sup(O)) for (t, v, O) in inst.pwc],
            mod=[(g, sup(O), key) for (g, O, key) in inst.mod],
            key_H=(inst.key_H, inst.key_L),
        ))
    plan = se_build_plan(
        se_insts, tsave, t0=None if options.t0 is None else float(options.t0), rtol=method.rtol, atol=method.atol,
        safety=method.safety_factor, factormin=method.min_factor, factormax=method.max_factor,
        max_steps=method.max_steps, save_states=True, tableau='dopri5' if isinstance(method, Dopri5) else 'tsit5',
    )
    plan.run()
    res = plan.results()
    if np.any(res['status'] == -3):
        raise max_steps_error()
    if np.any(res['status'] == -4):
        raise RuntimeError('The solver stopped on a non-finite error estimate or a vanishing step size.')
    B = len(insts)
    states = np.swapaxes(res['states'].reshape(B, len(tsave), n, n), -1, -2)  # unvectorize
    extra = None
    if options.save_extra is not None:
        extra = save_extra_outputs(options.save_extra, states, rho0.dims, out_shape)
    Esave = None
    if exp_ops is not None:
        Es = np.stack([E.to_numpy() for E in exp_ops]).astype(np.complex128)
        Esave = np.einsum('eij,btji->bet', Es, states.astype(np.complex128)).astype(np.complex64)
        Esave = Esave.reshape(*out_shape, len(Es), len(tsave))
    states = states.reshape(*out_shape, len(tsave), n, n)
    if not options.save_states:
        states = states[..., -1:, :, :]

    def shaped(x):
        x = np.asarray(x).reshape(out_shape)
        return x if out_shape else x[()]

    infos = AdaptiveStepInfos(shaped(res['nsteps']), shaped(res['naccepted']), shaped(res['nrejected']))
    return MESolveResult(tsave, method, gradient, options, QArray(states, dims=rho0.dims), Esave, extra, infos)


def _stack_trees(trees: list):
    """Stack a list of equally structured outputs of ``save_extra`` leaf by leaf
    (dicts, lists, tuples and named tuples, ``QArray``, array-likes; ``None``
    stays ``None``) along a new first axis."""
    t0 = trees[0]
    if t0 is None:
        return None
    if isinstance(t0, dict):
        return {k: _stack_trees([t[k] for t in trees]) for k in t0}
    if isinstance(t0, tuple) and hasattr(t0, '_fields'):
        return type(t0)(*[_stack_trees([t[i] for t in trees]) for i in range(len(t0))])
    if isinstance(t0, (list, tuple)):
        return type(t0)(_stack_trees([t[i] for t in trees]) for i in range(len(t0)))
    if isinstance(t0, QArray):
        return QArray(np.stack([t.to_numpy() for t in trees]), dims=t0.dims)
    return np.stack([np.asarray(t) for t in trees])


def _reshape_tree(tree, lead: tuple):
    """Split the first axis of every leaf into ``lead``."""
    if tree is None:
        return None
    if isinstance(tree, dict):
        return {k: _reshape_tree(v, lead) for k, v in tree.items()}
    if isinstance(tree, tuple) and hasattr(tree, '_fields'):
        return type(tree)(*[_reshape_tree(v, lead) for v in tree])
    if isinstance(tree, (list, tuple)):
        return type(tree)(_reshape_tree(v, lead) for v in tree)
    if isinstance(tree, QArray):
        x = tree.to_numpy()
        return QArray(x.reshape(*lead, *x.shape[1:]), dims=tree.dims)
    return tree.reshape(*lead, *tree.shape[1:])


def save_extra_outputs(fn, states: np.ndarray, dims, out_shape: tuple):
    """``Options(save_extra=fn)`` as a host callback: ``fn`` is applied to the state
    (a ``QArray``) at every save time, as ``SolveSaveMixin.save`` does
    (``dynamiqs/integrators/core/save_mixin.py:47-54``). The outputs are stacked
    along time and then over the batch, so each leaf has shape
    ``(*batch, ntsave, *leaf)`` as in the reference's vmapped scan. ``states`` is
    ``(nbatch, ntsave, n, n)``."""
    nb, nt = states.shape[:2]
    per_inst = [_stack_trees([fn(QArray(states[b, t], dims=dims)) for t in range(nt)]) for b in range(nb)]
    return _reshape_tree(_stack_trees(per_inst), tuple(out_shape))


def hermitian_parts(rhos, assume_hermitian: bool):
    """Hermitian initial states for the kernels, which evolve Hermitian states only
    (the vector field in the form tmp + tmp^dag, as ``vector_field_unvec_hermitian``
    at ``dynamiqs/integrators/core/diffrax_integrator.py:319-323``).

    With ``assume_hermitian`` (Tsit5 with ``Options(assume_hermitian=True)``, whose
    states are checked) every state is Hermitian up to the checked tolerance and is
    replaced by its Hermitian part. Otherwise (``assume_hermitian=False``, or
    Rouchon1, which assumes nothing: ``dynamiqs/integrators/apis/mesolve.py:165-171``)
    a state that is not Hermitian (to 1e-14 relative, above the rounding of a
    density matrix formed from a ket) is split as rho0 = A + i B with
    A = (rho0 + rho0^dag) / 2 and B = (rho0 - rho0^dag) / 2i, both Hermitian: the
    master equation and Tr(E rho) are linear, so rho(t) = A(t) + i B(t) and the two
    parts are solved as separate instances. Returns the list of states to solve and, per input, the index of A
    and of B (-1 when there is no B part)."""
    parts, index = [], []
    for r in rhos:
        r = np.asarray(r)
        a = 0.5 * (r + r.conj().T)
        ia = len(parts)
        parts.append(a)
        ib = -1
        asym = np.max(np.abs(r - r.conj().T), initial=0.0)
        if not assume_hermitian and asym > 1e-14 * max(1.0, np.max(np.abs(r), initial=0.0)):
            ib = len(parts)
            parts.append(-0.5j * (r - r.conj().T))
        index.append((ia, ib))
    return parts, np.asarray(index, dtype=np.int64).reshape(-1, 2)


def combine_parts(x: np.ndarray, index: np.ndarray) -> np.ndarray:
    """x[A] + i x[B] over the first axis (the inverse of ``hermitian_parts``)."""
    out = x[index[:, 0]].astype(np.result_type(x.dtype, np.complex64))
    has_b = index[:, 1] >= 0
    if np.any(has_b):
        out[has_b] = out[has_b] + 1j * x[index[has_b, 1]]
    return out


def _dia_groups(Ht, Lts, rho0: QArray, cartesian: bool):
    """The reference's batching rules over batched DIA operators: the output shape and,
    per distinct operator set (H index, jump-operator indices), its dense operators and
    the (output position, rho0 index) pairs that use it."""
    r = rho0.to_numpy()
    rb = r.shape[:-2]
    shapes = [Ht.batch, *[Lt.batch for Lt in Lts], rb]
    if cartesian:
        out_shape = tuple(d for sh in shapes for d in sh)
    else:
        out_shape = tuple(np.broadcast_shapes(*shapes))
    groups = {}
    for pos, idx in enumerate(np.ndindex(*out_shape) if out_shape else [()]):
        if cartesian:
            parts, o = [], 0
            for sh in shapes:
                parts.append(tuple(idx[o : o + len(sh)]))
                o += len(sh)
        else:
            parts = [_broadcast_index(idx, sh) for sh in shapes]
        groups.setdefault((parts[0], tuple(parts[1:-1])), []).append((pos, parts[-1]))
    out = []
    for (iH, iLs), members in groups.items():
        ops = (Ht.const_at(iH), [Lt.const_at(i) for Lt, i in zip(Lts, iLs)])
        terms = {}
        if Ht.pwc or Ht.mod:  # time-dependent Hamiltonian terms of this batch element
            terms = dict(pwc=Ht.pwc_at(iH), mod=[(g, o) for (g, o, _) in Ht.mod_at(iH)])
        out.append((ops, members, terms))
    return out_shape, out


def _mesolve_dia(route, rho0: QArray, tsave, exp_ops, method, gradient, options) -> MESolveResult:
    """Tsit5 (``dqh_mesolve_tsit5_dia``, per-instance adaptive steps) or Rouchon1
    (``dqh_mesolve_rouchon1_dia``, lockstep) on the DIA sweeps: the batch is the
    batch of ``rho0`` for shared operators; batched operators run one plan per
    distinct operator set (each launch is grid-wide over the n^2 elements)."""
    r_all = rho0.to_numpy().astype(np.complex128)
    n = r_all.shape[-2]
    if isinstance(route[0], str):  # 'batched' or 'timedep': one plan per batch element of the operators
        out_shape, groups = _dia_groups(route[1], route[2], rho0, options.cartesian_batching)
    else:
        out_shape = r_all.shape[:-2]
        rb = out_shape
        groups = [(route, [(pos, idx) for pos, idx in enumerate(np.ndindex(*rb) if rb else [()])], {})]
    Es = [E.to_numpy() for E in exp_ops] if exp_ops is not None else None
    nout = int(np.prod(out_shape)) if out_shape else 1
    states_all = None
    expects_all = np.zeros((nout, len(Es), len(tsave)), dtype=np.complex64) if Es is not None else None
    infos_all = np.zeros((3, nout), dtype=np.int64)
    plan = None
    for (Hd, Ld), members, terms in groups:
        r = np.stack([r_all[ir] for (_, ir) in members])
        st, ex, inf, plan = _dia_solve(Hd, Ld, r, tsave, Es, method, options, **terms)
        pos = np.array([p for (p, _) in members])
        if states_all is None:
            states_all = np.zeros((nout, *st.shape[1:]), dtype=st.dtype)
        states_all[pos] = st
        if Es is not None:
            expects_all[pos] = ex
        infos_all[:, pos] = inf
    states = states_all
    extra = None
    if options.save_extra is not None:
        extra = save_extra_outputs(options.save_extra, states, rho0.dims, out_shape)
    states = states.reshape(*out_shape, -1, n, n)
    if not options.save_states:  # the final state only, nsave = 1 (save_mixin.py:62-66, mesolve.py:184-187)
        states = states[..., -1:, :, :]
    ysave = QArray(states, dims=rho0.dims)
    Esave = None
    if Es is not None:
        Esave = expects_all.reshape(*out_shape, len(Es), len(tsave))

    def shaped(x):
        x = np.asarray(x).reshape(out_shape)
        return x if out_shape else x[()]

    if isinstance(method, ADAPTIVE):
        infos = AdaptiveStepInfos(shaped(infos_all[0]), shaped(infos_all[1]), shaped(infos_all[2]))
    else:
        infos = FixedStepInfos(shaped(np.full(nout, plan.nsteps, dtype=np.int64)))
    return MESolveResult(tsave, method, gradient, options, ysave, Esave, extra, infos)


def _dia_solve(Hd, Ld, r: np.ndarray, tsave, Es, method, options, pwc=None, mod=None):
    """One DIA plan over the initial states ``r`` (m, n, n|1) with shared operators:
    (states (m, nsave, n, n), expects (m, nE, ntsave) or None, infos (3, m), plan)."""
    n = r.shape[-2]
    if r.shape[-1] == 1:
        r = r @ np.conj(np.swapaxes(r, -1, -2))
    t0 = None if options.t0 is None else float(options.t0)
    if isinstance(method, ADAPTIVE) and options.assume_hermitian:
        for x in r:
            check_hermitian(x, 'y0')
    parts, pindex = hermitian_parts(r, options.assume_hermitian and isinstance(method, ADAPTIVE))
    r = np.stack(parts)
    if isinstance(method, ADAPTIVE):
        plan = dia_path.tsit5_plan(
            Hd,
            Ld,
            r,
            tsave,
            t0=t0,
            rtol=method.rtol,
            atol=method.atol,
            safety=method.safety_factor,
            factormin=method.min_factor,
            factormax=method.max_factor,
            max_steps=method.max_steps,
            exp_ops=Es,
            tableau='dopri5' if isinstance(method, Dopri5) else 'tsit5',
            pwc=pwc,
            mod=mod,
        )
    else:
        plan = dia_path.build_plan(Hd, Ld, r, tsave, method.dt, t0=t0, exp_ops=Es)
    plan.run()
    res = plan.results()
    if 'status' in res:
        if np.any(res['status'] == -3):
            raise max_steps_error()
        if np.any(res['status'] == -4):
            raise RuntimeError('The solver stopped on a non-finite error estimate or a vanishing step size.')
    states = combine_parts(res['states'], pindex)
    m = len(pindex)
    states = states.reshape(m, -1, n, n)
    ex = combine_parts(res['expects'], pindex).reshape(m, len(Es), len(tsave)) if Es is not None else None
    if isinstance(method, ADAPTIVE):
        inf = np.stack([np.asarray(res[k])[pindex[:, 0]] for k in ('nsteps', 'naccepted', 'nrejected')])
    else:
        inf = np.zeros((3, m), dtype=np.int64)
        inf[0] = plan.nsteps
    return states, ex, inf, plan
