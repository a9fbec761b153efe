"""Drop-in ``sesolve`` running on the MI355X.

Mirrors ``dynamiqs/integrators/apis/sesolve.py``: argument conversion (:214-219),
checks (:222-225, :307-318), cartesian / flat batching over ``H`` and ``psi0``
(:234-260) and the method registry (:272-283), here ``Tsit5`` and ``Dopri5``. The
integrator seam (``integrator_constructor(...).run()``, :285-301) is the HIP
kernel ``dqh_sesolve_tsit5``: every batch element is one instance, integrated by
one workgroup with its own adaptive step.
"""

from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from .. import _native
from .._checks import check_shape, check_times
from ..gradient import Gradient, reject_gradient
from ..method import Dopri5, Euler, Expm, Method, Tsit5
from ..options import Options, check_options
from ..qarrays import QArray, asqarray
from ..result import AdaptiveStepInfos, FixedStepInfos, SESolveResult
from ..time_qarray import TimeQArray, astimeqarray
from .dense import _Arena, _device, _pad, check_descriptors, pad16, rouchon_step_times
from .mesolve import assert_method_supported, max_steps_error, save_extra_outputs
from .operators import decompose, resample_pwc, resolve_callables, spline_table, union_breakpoints

SUPPORTED_METHODS = (Tsit5, Dopri5, Euler, Expm)
MAX_PWC = 8
MAX_NP = 1024  # padded dimension of the largest sesolve kernel variant
SE_FIELDS = ('tsave', 'Hnh', 'EopsT', 'Hpwc', 'pwc_times', 'pwc_vals', 'Hmod', 'mod_coef', 'step_t')


@dataclass
class SEInstance:
    """One Schrodinger solve: constant part of H, its PWC and modulated terms, psi0."""

    H: np.ndarray  # (n, n) complex128
    psi0: np.ndarray  # (n,) complex128
    pwc: list = field(default_factory=list)  # [(times, values (nv,), op (n, n))]
    mod: list = field(default_factory=list)  # [(g, op (n, n), key)]
    key_H: object = None  # equal keys share device copies of the operators


class SEPlan:
    """A packed batch of ``sesolve`` instances (inputs resident in HBM)."""

    def __init__(self, n: int, ntsave: int, nE: int, ninst: int, save_states: bool):
        self.n, self.np, self.ntsave, self.nE, self.ninst = n, pad16(n), ntsave, nE, ninst
        self.save_states = save_states
        self.tensors = []
        self.descs = self.d_descs = None
        self.states = self.final = self.expects = self.stats = self.step_log = None
        self.y = self.y0 = None

    def reset(self):
        self.y.copy_(self.y0)

    def run(self, stream=None):
        import torch

        lib = _native.lib()
        s = stream if stream is not None else torch.cuda.current_stream()
        _native.check(
            lib.dqh_sesolve_tsit5(_native.ctypes.c_void_p(self.d_descs.data_ptr()), self.ninst, self.np,
                                  _native.ctypes.c_void_p(s.cuda_stream)),
            'dqh_sesolve_tsit5',
        )

    def results(self) -> dict:
        import torch

        torch.cuda.synchronize()
        n, ns = self.n, (self.ntsave if self.save_states else 1)
        c64 = lambda t: t.cpu().numpy().view(np.complex64)  # noqa: E731
        out = {
            'states': c64(self.states).reshape(self.ninst, ns, n),
            'final_state': c64(self.final).reshape(self.ninst, n),
        }
        if self.nE:
            out['expects'] = c64(self.expects).reshape(self.ninst, self.nE, self.ntsave)
        st = self.stats.cpu().numpy()
        out['nsteps'], out['naccepted'], out['nrejected'] = st[:, 0], st[:, 1], st[:, 2]
        out['status'], out['nrhs'] = st[:, 3], st[:, 4]
        if self.step_log is not None:
            log = self.step_log.cpu().numpy()
            out['step_ends'] = [log[b, : min(int(st[b, 1]), log.shape[1])] for b in range(self.ninst)]
        return out


def _header(n: int, tsave, Es, t0, rtol, atol, safety, factormin, factormax, max_steps, save_states, tableau,
            B: int, arena, euler_dt=None):
    """Descriptor columns common to a batch (arena offsets for the shared inputs).
    With ``euler_dt`` the kernel takes fixed Euler steps of that size (the last one
    clipped to tsave[-1]) instead of the adaptive pair."""
    if tableau not in ('tsit5', 'dopri5'):
        raise ValueError(f'unknown Runge-Kutta pair {tableau!r}')
    npad = pad16(n)
    if npad > MAX_NP:
        raise ValueError(f'the MI355X sesolve supports n <= {MAX_NP} (got n={n})')
    descs = np.zeros(B, dtype=_native.ME_INST_DTYPE)
    descs['n'], descs['np'], descs['nE'], descs['ntsave'] = n, npad, len(Es), len(tsave)
    descs['rtol'], descs['atol'], descs['safety'] = rtol, atol, safety
    descs['factormin'], descs['factormax'], descs['max_steps'] = factormin, factormax, max_steps
    descs['t0'], descs['t1'] = float(tsave[0] if t0 is None else t0), float(tsave[-1])
    descs['tsave'] = arena.put(tsave)
    descs['EopsT'] = arena.put(_pad(np.stack([E.T for E in Es]), npad).astype(np.complex64)) if Es else 0
    flags = _native.FLAG_SAVE_STATES if save_states else 0
    if tableau == 'dopri5':
        flags |= _native.FLAG_DOPRI5
    if euler_dt is not None:
        step_t = rouchon_step_times(float(descs['t0'][0]), float(descs['t1'][0]), float(euler_dt))
        flags |= _native.FLAG_EULER
        descs['step_t'] = arena.put(step_t)
        descs['max_steps'] = len(step_t) - 1
    descs['flags'] = flags
    return descs


def _finish(plan: SEPlan, descs, arena, y0: np.ndarray, dev, step_log_cap: int) -> SEPlan:
    """Upload the staged inputs, allocate the state and the outputs, fill the
    pointer columns, check every pointer against the allocations."""
    import torch

    B, npad, n, nt, nE = plan.ninst, plan.np, plan.n, plan.ntsave, plan.nE
    arena.upload()
    u = np.uint64
    for f in SE_FIELDS:
        v = descs[f]
        descs[f] = np.where(v != 0, v + u(arena.base) - u(1), u(0))
    plan.y0 = torch.from_numpy(np.ascontiguousarray(y0).view(np.float64).reshape(B, npad, 2)).to(dev)
    plan.y = plan.y0.clone()
    ns = nt if plan.save_states else 1
    plan.states = torch.empty((B, ns * n * 2), dtype=torch.float32, device=dev)
    plan.final = torch.empty((B, n * 2), dtype=torch.float32, device=dev)
    plan.expects = torch.zeros((B, max(nE, 1) * nt * 2), dtype=torch.float32, device=dev)
    plan.stats = torch.zeros((B, 8), dtype=torch.int64, device=dev)
    idx = np.arange(B, dtype=u)
    descs['y'] = u(plan.y.data_ptr()) + idx * u(npad * 16)
    descs['states'] = u(plan.states.data_ptr()) + idx * u(ns * n * 8)
    descs['final_state'] = u(plan.final.data_ptr()) + idx * u(n * 8)
    descs['expects'] = u(plan.expects.data_ptr()) + idx * u(max(nE, 1) * nt * 8)
    descs['stats'] = u(plan.stats.data_ptr()) + idx * u(64)
    if step_log_cap > 0:
        plan.step_log = torch.zeros((B, int(step_log_cap)), dtype=torch.float64, device=dev)
        descs['step_log'] = u(plan.step_log.data_ptr()) + idx * u(int(step_log_cap) * 8)
        descs['step_log_cap'] = int(step_log_cap)
    npd = descs['np'].astype(u)
    extents = {
        'y': npd * u(16),
        'states': np.full(B, ns * n * 8, dtype=u),
        'final_state': np.full(B, n * 8, dtype=u),
        'Hpwc': descs['npwc'].astype(u) * npd * npd * u(8),
        'pwc_times': descs['nseg'].astype(u) * u(8),
        'pwc_vals': descs['npwc'].astype(u) * np.maximum(descs['nseg'].astype(np.int64) - 1, 0).astype(u) * u(16),
        'step_t': (np.maximum(descs['max_steps'], 0).astype(u) + u(1)) * u(8),
    }
    tensors = [*arena.tensors, plan.y, plan.states, plan.final, plan.expects, plan.stats]
    if plan.step_log is not None:
        tensors.append(plan.step_log)
    check_descriptors(descs, tensors, extents)
    plan.tensors = tensors
    plan.descs = descs
    plan.d_descs = torch.from_numpy(descs.view(np.uint8).copy()).to(dev)
    return plan


def _G_T(ops, npad: int) -> np.ndarray:
    """(-i O)^T of each operator, zero-padded to npad, complex64."""
    return _pad(np.swapaxes(-1j * np.asarray(ops), -1, -2), npad).astype(np.complex64)


def build_plan(
    instances: list,
    tsave: np.ndarray,
    *,
    exp_ops=None,
    t0: float | None = None,
    rtol: float = 1e-6,
    atol: float = 1e-6,
    safety: float = 0.9,
    factormin: float = 0.2,
    factormax: float = 5.0,
    max_steps: int = 100_000,
    save_states: bool = True,
    tableau: str = 'tsit5',
    step_log_cap: int = 0,
    euler_dt: float | None = None,
) -> SEPlan:
    """Pack ``instances`` (same Hilbert dimension) for ``dqh_sesolve_tsit5``.

    The kernel multiplies by ``G = -i H`` with rows read along the transposed
    layout, so every operator is stored as ``(-i O)^T``; the expectation operators
    as ``E^T``. PWC terms are resampled on the union of their breakpoints;
    modulated terms become piecewise-cubic tables over [t0, tsave[-1]] with one
    interval count per instance."""
    n = instances[0].H.shape[-1]
    if any(inst.H.shape[-1] != n for inst in instances):
        raise ValueError('all instances of a sesolve plan must share the Hilbert dimension')
    if any(len(inst.pwc) > MAX_PWC for inst in instances):
        raise NotImplementedError(f'at most {MAX_PWC} piecewise-constant terms per Hamiltonian are supported')
    if any(len(inst.mod) > _native.MAX_MOD for inst in instances):
        raise NotImplementedError(f'at most {_native.MAX_MOD} modulated terms per Hamiltonian are supported')
    dev = _device()
    _native.lib()
    tsave = np.asarray(tsave, dtype=np.float64)
    B = len(instances)
    Es = [np.asarray(E, dtype=np.complex128) for E in (exp_ops or [])]
    arena = _Arena(dev)
    descs = _header(n, tsave, Es, t0, rtol, atol, safety, factormin, factormax, max_steps, save_states, tableau, B,
                    arena, euler_dt)
    npad = int(descs['np'][0])
    t_start, t_end = float(descs['t0'][0]), float(descs['t1'][0])
    plan = SEPlan(n, len(tsave), len(Es), B, save_states)
    cache = {}

    def shared(key, build):
        if key is None:
            return arena.put(build())
        if key not in cache:
            cache[key] = arena.put(build())
        return cache[key]

    tables = {}
    for b, inst in enumerate(instances):
        d = descs[b]
        d['Hnh'] = shared(('H', inst.key_H), lambda inst=inst: _G_T(inst.H, npad))
        if inst.pwc:
            grid = union_breakpoints([inst.pwc])
            vals = np.stack([resample_pwc(t, np.asarray(v, dtype=np.complex128), grid) for (t, v, _) in inst.pwc])
            d['npwc'], d['nseg'] = len(inst.pwc), len(grid)
            d['Hpwc'] = shared(('Hpwc', inst.key_H),
                               lambda inst=inst: _G_T(np.stack([o for (_, _, o) in inst.pwc]), npad))
            d['pwc_times'] = shared(('ptimes', inst.key_H), lambda grid=grid: grid)
            d['pwc_vals'] = arena.put(np.ascontiguousarray(vals).view(np.float64))
        if inst.mod:
            tabs = []
            for g, _, key in inst.mod:
                k = None if key is None else (key, t_start, t_end)
                if k is None or k not in tables:
                    tab = spline_table(g, t_start, t_end)
                    if k is None:
                        tabs.append(tab)
                        continue
                    tables[k] = tab
                tabs.append(tables[k])
            nint = max(len(c) for (_, _, c) in tabs)
            tabs = [spline_table(g, t_start, t_end, n0=nint, nmax=nint) if len(c) != nint else (a, dt, c)
                    for (g, _, _), (a, dt, c) in zip(inst.mod, tabs)]
            d['nmod'] = len(inst.mod)
            d['mod_t0'], d['mod_dt'], d['mod_nint'] = tabs[0][0], tabs[0][1], nint
            d['mod_coef'] = arena.put(np.ascontiguousarray(np.stack([c for (_, _, c) in tabs])).view(np.float64))
            d['Hmod'] = shared(('Hmod', inst.key_H),
                               lambda inst=inst: _G_T(np.stack([o for (_, o, _) in inst.mod]), npad))
    y0 = np.zeros((B, npad), dtype=np.complex128)
    for b, inst in enumerate(instances):
        y0[b, :n] = inst.psi0
    return _finish(plan, descs, arena, y0, dev, step_log_cap)


def build_plan_constant(
    Hs: np.ndarray,
    h_index: np.ndarray,
    psis: np.ndarray,
    p_index: np.ndarray,
    tsave: np.ndarray,
    *,
    exp_ops=None,
    t0: float | None = None,
    rtol: float = 1e-6,
    atol: float = 1e-6,
    safety: float = 0.9,
    factormin: float = 0.2,
    factormax: float = 5.0,
    max_steps: int = 100_000,
    save_states: bool = True,
    tableau: str = 'tsit5',
    step_log_cap: int = 0,
    euler_dt: float | None = None,
) -> SEPlan:
    """The same plan for constant Hamiltonians, packed as arrays: instance b
    integrates ``Hs[h_index[b]]`` (``Hs``: (nH, n, n)) from ``psis[p_index[b]]``
    (``psis``: (nP, n)). One staged block holds every (-i H)^T; the descriptor
    columns are filled without a per-instance loop, so that a batch of thousands
    of small systems packs in milliseconds."""
    Hs = np.asarray(Hs, dtype=np.complex128)
    psis = np.asarray(psis, dtype=np.complex128)
    h_index = np.asarray(h_index, dtype=np.int64).reshape(-1)
    p_index = np.asarray(p_index, dtype=np.int64).reshape(-1)
    if h_index.shape != p_index.shape or h_index.size == 0:
        raise ValueError('h_index and p_index must be non-empty and of equal length')
    if np.any((h_index < 0) | (h_index >= len(Hs))) or np.any((p_index < 0) | (p_index >= len(psis))):
        raise ValueError('instance index out of range')
    n = Hs.shape[-1]
    if Hs.shape[-2:] != (n, n) or psis.shape[-1] != n:
        raise ValueError('Hs must be (nH, n, n) and psis (nP, n)')
    dev = _device()
    _native.lib()
    tsave = np.asarray(tsave, dtype=np.float64)
    B = len(h_index)
    Es = [np.asarray(E, dtype=np.complex128) for E in (exp_ops or [])]
    arena = _Arena(dev)
    descs = _header(n, tsave, Es, t0, rtol, atol, safety, factormin, factormax, max_steps, save_states, tableau, B,
                    arena, euler_dt)
    npad = int(descs['np'][0])
    base = arena.put(_G_T(Hs, npad))
    descs['Hnh'] = np.uint64(base) + h_index.astype(np.uint64) * np.uint64(npad * npad * 8)
    y0 = np.zeros((B, npad), dtype=np.complex128)
    y0[:, :n] = psis[p_index]
    return _finish(SEPlan(n, len(tsave), len(Es), B, save_states), descs, arena, y0, dev, step_log_cap)


def _check_sesolve_args(H: TimeQArray, psi0: QArray, exp_ops) -> None:
    check_shape(H, 'H', '(..., n, n)', subs={'...': '...H'})
    check_shape(psi0, 'psi0', '(..., n, 1)', subs={'...': '...psi0'})
    if exp_ops is not None:
        for i, E in enumerate(exp_ops):
            check_shape(E, f'exp_ops[{i}]', '(n, n)')


def enumerate_instances(H: TimeQArray, psi0: QArray, cartesian: bool):
    """(out_shape, [SEInstance]) following the reference's batching rules:
    cartesian batching over (...H, ...psi0), otherwise one broadcast shape."""
    Ht = decompose(H)
    p = psi0.to_numpy().astype(np.complex128)
    shapes = [Ht.batch, p.shape[:-2]]
    if cartesian:
        out_shape = tuple(d for s in shapes for d in s)
    else:
        out_shape = tuple(np.broadcast_shapes(*shapes))
    insts = []
    for idx in np.ndindex(*out_shape) if out_shape else [()]:
        if cartesian:
            iH, ip = tuple(idx[: len(shapes[0])]), tuple(idx[len(shapes[0]) :])
        else:
            iH, ip = [tuple(i if d > 1 else 0 for i, d in zip(idx[len(idx) - len(s) :] if s else (), s)) for s in shapes]
        insts.append(
            SEInstance(
                H=Ht.const_at(iH),
                psi0=p[ip][:, 0],
                pwc=Ht.pwc_at(iH),
                mod=[(g, o, (id(f), iH if np.ndim(f(0.0)) else None)) for (g, o, f) in Ht.mod_at(iH)],
                key_H=iH,
            )
        )
    return out_shape, insts


def constant_batch(Ht, psi0: QArray, cartesian: bool):
    """Batch of a constant Hamiltonian as arrays: (out_shape, Hs (nH, n, n),
    h_index, psis (nP, n), p_index) with instance b = (Hs[h_index[b]],
    psis[p_index[b]]), in the order of ``enumerate_instances``."""
    n = Ht.n
    Hb = tuple(Ht.batch)
    const = Ht.const if Ht.const is not None else np.zeros((*Hb, n, n), dtype=np.complex128)
    Hs = np.ascontiguousarray(const).reshape(-1, n, n)
    p = psi0.to_numpy().astype(np.complex128)
    Pb = p.shape[:-2]
    psis = p.reshape(-1, n)
    nH, nP = len(Hs), len(psis)
    if cartesian:
        out_shape = Hb + tuple(Pb)
        h_index = np.repeat(np.arange(nH), nP)
        p_index = np.tile(np.arange(nP), nH)
    else:
        out_shape = tuple(np.broadcast_shapes(Hb, Pb))
        h_index = np.broadcast_to(np.arange(nH).reshape(Hb), out_shape).reshape(-1)
        p_index = np.broadcast_to(np.arange(nP).reshape(Pb), out_shape).reshape(-1)
    return out_shape, Hs, h_index, psis, p_index


def sesolve(
    H,
    psi0,
    tsave,
    *,
    exp_ops=None,
    method: Method = Tsit5(),  # noqa: B008
    gradient: Gradient | None = None,
    options: Options = Options(),  # noqa: B008
) -> SESolveResult:
    r"""Solve the Schrodinger equation on the MI355X.

    Same signature, batching rules, errors and result as ``dq.sesolve``
    (``dynamiqs/integrators/apis/sesolve.py:36-230``). Supported methods:
    ``Tsit5``, ``Dopri5``, ``Euler`` (fixed step) and ``Expm`` (constant and PWC
    Hamiltonians); supported Hamiltonians: constant (dense or DIA,
    n <= 1024), piecewise-constant (``dq.pwc``) and modulated (``dq.modulated``)
    terms and their sums.
    """
    # === convert arguments
    H = astimeqarray(H)
    psi0 = asqarray(psi0)
    tsave = np.asarray(tsave, dtype=np.float64)
    if exp_ops is not None:
        exp_ops = [asqarray(E) for E in exp_ops] if len(exp_ops) > 0 else None

    # === check arguments
    _check_sesolve_args(H, psi0, exp_ops)
    tsave = check_times(tsave, 'tsave')
    check_options(options, 'sesolve')
    options = options.initialise()
    assert_method_supported(method, SUPPORTED_METHODS)
    method.assert_supports_gradient(gradient)
    reject_gradient(gradient, '`dq.sesolve()`')

    Es = [E.to_numpy() for E in exp_ops] if exp_ops is not None else None
    if isinstance(method, Expm):  # propagators on MFMA applied to the initial kets
        from .propagators import expm_solve

        return expm_solve(H, [], psi0, tsave, Es, method, gradient, options, me=False)
    # timecallable operators become constant + modulated terms over the solve interval
    H = resolve_callables(H, float(tsave[0] if options.t0 is None else options.t0), float(tsave[-1]))
    kw = dict(
        exp_ops=Es,
        t0=None if options.t0 is None else float(options.t0),
        save_states=options.save_states or options.save_extra is not None,
    )
    if isinstance(method, Euler):
        kw['euler_dt'] = float(method.dt)
    else:
        kw.update(
            rtol=method.rtol,
            atol=method.atol,
            safety=method.safety_factor,
            factormin=method.min_factor,
            factormax=method.max_factor,
            max_steps=method.max_steps,
            tableau='dopri5' if isinstance(method, Dopri5) else 'tsit5',
        )
    Ht = decompose(H)
    if Ht.is_constant:
        out_shape, Hs, h_index, psis, p_index = constant_batch(Ht, psi0, options.cartesian_batching)
        plan = build_plan_constant(Hs, h_index, psis, p_index, tsave, **kw)
    else:
        out_shape, insts = enumerate_instances(H, psi0, options.cartesian_batching)
        plan = build_plan(insts, tsave, **kw)
    plan.run()
    res = plan.results()
    if np.any(res['status'] == -3):
        raise max_steps_error()
    if np.any(res['status'] == -4):
        raise RuntimeError('The solver stopped on a non-finite error estimate or a vanishing step size.')
    if np.any(res['status'] == -6):
        raise RuntimeError('dqh_sesolve_tsit5 rejected an instance descriptor (unsupported size or term count).')

    n = plan.n
    states = res['states'][..., None]  # (B, ns, n, 1)
    extra = None
    if options.save_extra is not None:
        extra = save_extra_outputs(options.save_extra, states, psi0.dims, out_shape)
    states = states.reshape(*out_shape, -1, n, 1)
    if not options.save_states:  # the final state only, nsave = 1 (save_mixin.py:62-66, mesolve.py:184-187)
        states = states[..., -1:, :, :]
    ysave = QArray(states, dims=psi0.dims)
    Esave = res['expects'].reshape(*out_shape, len(Es), len(tsave)) if Es is not None else None

    def shaped(x):
        x = np.asarray(x).reshape(out_shape)
        return x if out_shape else x[()]

    if isinstance(method, Euler):
        infos = FixedStepInfos(shaped(res['nsteps']))
    else:
        infos = AdaptiveStepInfos(shaped(res['nsteps']), shaped(res['naccepted']), shaped(res['nrejected']))
    return SESolveResult(tsave, method, gradient, options, ysave, Esave, extra, infos)
