"""Adjoint gradients of master-equation losses on the MI355X.

Replaces the reverse-mode path of the reference (``jax.grad`` through
``dx.RecursiveCheckpointAdjoint``, ``diffrax_integrator.py:88-99``,
``gradient.py:32-72``) for losses that are sums of expectation values at chosen
times,

    J = sum_k Tr(O_k rho(t_k)),

of a master equation with

    H(t) = H0 + sum_j u_j(t) O_j + sum_m g_m(t) M_m,   u_j constant on [times[k], times[k+1]),

by the continuous adjoint (optimise-then-discretise). With lambda_k the adjoint
state of the loss term k (d lambda_k / dt = -L^dag[lambda_k] on [t0, t_k],
lambda_k(t_k) = O_k, zero after t_k) and lambda = sum_k lambda_k,

    dJ/du_jk  = int_{times[k]}^{times[k+1]} Re Tr(lambda (-i [O_j, rho])) dt,
    dJ/dtheta = int Re Tr(lambda (-i [g(t) P, rho])) dt   for a term theta g(t) P of H
                (constant: g = 1; modulated: g the term's coefficient),
    dJ/dkappa = int Tr(lambda D[A] rho) dt                for a jump operator sqrt(kappa) A,
    dJ/drho0  = lambda(t0),
    L^dag[l] = i[H, l] + sum_k (L_k^dag l L_k - 1/2 {L_k^dag L_k, l}).

The jump conditions lambda(t_k^-) = lambda(t_k^+) + O_k of a loss over several times
are applied by linearity: every loss term is its own backward instance, started at
its time, and the instances are summed node by node.

Four GPU launches, all on the hand-written kernels:

1. forward Tsit5 of every instance, saving rho at the checkpoints (every
   ``checkpoint_every`` segments) and at the loss times, and the expectation values
   of the loss operators;
2. recomputation from the checkpoints: one instance per (control vector,
   checkpoint interval), all in one persistent launch, saving rho at the
   Gauss-Legendre nodes of each segment;
3. backward Tsit5 of the adjoint Lindbladian (Hermitian form with
   ``Hnh' = iH - 1/2 sum L^dag L``, ``D = L^dag``, ``C = L / 2``) in reversed time
   ``s = T - t``, one instance per (control vector, loss term), saving lambda at the
   same nodes;
4. ``dqh_adjoint_pwc_quadrature`` (and its DIA and jump-rate forms): the
   Gauss-Legendre sums in FP64.

No JAX is available on this platform, so the gradient is exposed as an explicit
function rather than through ``jax.custom_vjp``.
"""

from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np

from . import _native
from .integrators import dense


def _gl(nodes: int):
    x, w = np.polynomial.legendre.leggauss(nodes)
    return x, w


def _sparse_controls(ops, n: int):
    """Control operators by diagonals for ``dqh_adjoint_pwc_quadrature_dia`` when each
    has at most 16 nonzero diagonals and fewer than n / 4 (else None: dense path).
    Returns (DiaOps, device buffers to keep alive until the launch is done)."""
    import torch

    from .integrators.dia import MAXOFF, DiaOps, _dia_of

    if n > 64:
        return None
    forms = [_dia_of(o) for o in ops]
    if any(len(offs) == 0 or len(offs) > MAXOFF or 4 * len(offs) >= n for offs, _ in forms):
        return None
    dev = torch.device('cuda', torch.cuda.current_device())
    d = DiaOps()
    d.nops = len(forms)
    keep = []
    for k, (offs, diags) in enumerate(forms):
        d.noff[k] = len(offs)
        for u, off in enumerate(offs):
            d.off[k][u] = off
        t = torch.from_numpy(np.ascontiguousarray(diags.astype(np.complex64)).view(np.float32)).to(dev)
        keep.append(t)
        d.diags[k] = t.data_ptr()
    return d, keep


def _check_status(name: str, plan) -> None:
    """Raise if a solve of ``plan`` stopped before its end time (max_steps or a
    non-finite / vanishing step), naming the first such instance."""
    st = plan.stats.cpu().numpy()
    bad = np.nonzero(st[:, 3] != 0)[0]
    if len(bad):
        i = int(bad[0])
        d = plan.descs[i]
        raise RuntimeError(
            f'adjoint: {len(bad)} of {len(st)} {name} solves stopped early; first: instance {i}, status {int(st[i, 3])}'
            f' (nsteps {int(st[i, 0])}, accepted {int(st[i, 1])}, rejected {int(st[i, 2])}),'
            f' t0 {float(d["t0"])!r}, t1 {float(d["t1"])!r}'
        )


@dataclass
class AdjointResult:
    """``J`` (B,): the loss; ``grad_u`` (B, nterms, nseg): dJ/du; ``grad_rates``
    (B, nrates): dJ/dkappa_r for jump operators sqrt(kappa_r) A_r; ``lambda0``
    (B, n, n): the adjoint state at the start time, dJ/drho0 (J is linear in rho0:
    J = Tr(lambda0 rho0)); ``grad_params`` (B, nparams): dJ/dtheta_p for the
    Hamiltonian terms theta_p g_p(t) P_p named by ``param_ops``; ``timing``:
    seconds per phase."""

    J: np.ndarray
    grad_u: np.ndarray
    grad_rates: np.ndarray | None
    lambda0: np.ndarray
    timing: dict
    grad_params: np.ndarray | None = None


def _loss_terms(O, times: np.ndarray):
    """``O`` as a list of (t_k, O_k): an operator alone is the terminal loss
    Tr(O rho(T)); a list of (t_k, O_k) pairs is sum_k Tr(O_k rho(t_k)), each t_k
    one of the segment boundaries after the start time (so that no quadrature
    segment straddles a jump of the adjoint state)."""
    T = float(times[-1])
    if isinstance(O, (list, tuple)) and len(O) and isinstance(O[0], (list, tuple)):
        terms = [(float(t), np.asarray(Ok, dtype=np.complex128)) for t, Ok in O]
    else:
        terms = [(T, np.asarray(O, dtype=np.complex128))]
    out = []
    for t, Ok in terms:
        k = int(np.argmin(np.abs(times - t)))
        if not (k > 0 and abs(times[k] - t) <= 1e-12 * max(1.0, abs(t))):
            raise ValueError(f'loss time {t} must be one of the segment boundaries `times` after the start time')
        out.append((float(times[k]), Ok))
    if len(out) > 16:
        raise ValueError('at most 16 loss terms')
    return out


def _param_terms(param_ops):
    """``param_ops`` entries: an operator P (a constant term theta P of H) or a pair
    (g, P) (a modulated term theta g(t) P); returns [(g or None, P)]."""
    out = []
    for p in param_ops:
        if isinstance(p, (list, tuple)):
            g, P = p
            if not callable(g):
                raise TypeError('a modulated parameter term is given as (g, P) with g a function of t')
            out.append((g, np.asarray(P, dtype=np.complex128)))
        else:
            out.append((None, np.asarray(p, dtype=np.complex128)))
    return out


def mesolve_pwc_grad(H0, ops, times, u, Ls, rho0, O, *, return_timing: bool = False, **kw):
    """Loss ``J_b`` (``Tr(O rho_b(T))``, or a sum over loss times as in
    :func:`mesolve_adjoint`) and gradient ``dJ_b/du_b`` for a batch of control arrays
    ``u`` of shape ``(B, nterms, nseg)`` (``T = times[-1]``).

    Returns ``(J (B,), grad (B, nterms, nseg))`` (and a timing dict if requested).
    """
    r = mesolve_adjoint(H0, ops, times, u, Ls, rho0, O, **kw)
    if return_timing:
        return r.J, r.grad_u, r.timing
    return r.J, r.grad_u


def mesolve_adjoint(
    H0: np.ndarray,
    ops,
    times: np.ndarray,
    u: np.ndarray,
    Ls,
    rho0: np.ndarray,
    O,
    *,
    rate_ops=(),
    mod_terms=(),
    param_ops=(),
    rtol: float = 1e-6,
    atol: float = 1e-6,
    checkpoint_every: int = 10,
    nodes: int = 5,
) -> AdjointResult:
    """Loss, control gradient, jump-rate gradients, Hamiltonian-parameter gradients
    and adjoint initial state for a batch of control arrays ``u`` (B, nterms, nseg).

    ``O``: an operator (``J = Tr(O rho(T))``) or a list of ``(t_k, O_k)`` pairs
    (``J = sum_k Tr(O_k rho(t_k))``, each t_k one of ``times`` after the start).
    ``mod_terms``: pairs ``(g, M)`` adding ``g(t) M`` to H. ``param_ops``: terms
    whose scalar prefactor theta is differentiated at its current value: an
    operator P for a constant term ``theta P`` of H (part of ``H0``), or ``(g, P)``
    for a modulated one ``theta g(t) P`` (part of ``mod_terms``); dJ/dtheta is in
    ``grad_params``. ``rate_ops``: operators A_r whose rates are differentiated; the
    corresponding jump operators in ``Ls`` are sqrt(kappa_r) A_r (dJ/dkappa_r is
    summed over the whole interval). ``ops`` may be empty (then ``u`` has shape
    (B, 0, nseg) and ``times`` only sets the quadrature segments). One-shot form of
    :class:`AdjointPlan` (which packs once for repeated gradients).
    """
    u = np.asarray(u)
    if u.ndim == 2:
        u = u[:, None, :]
    plan = AdjointPlan(H0, ops, times, Ls, rho0, O, u.shape[0], rate_ops=rate_ops, mod_terms=mod_terms,
                       param_ops=param_ops, rtol=rtol, atol=atol, checkpoint_every=checkpoint_every, nodes=nodes)
    res = plan.run(u)
    res.timing.update({k: v for k, v in plan.pack_timing.items()})
    return res


class AdjointPlan:
    """The three Tsit5 plans and the quadrature of the adjoint gradient, packed once
    for a batch of B control vectors; :meth:`run` takes new controls and only
    rewrites their values on the device (the operators, time grids, save times and
    scratch stay resident), so repeated gradients (optimisation loops) pay the
    GPU time and a few small copies."""

    def __init__(self, H0, ops, times, Ls, rho0, O, B: int, *, rate_ops=(), mod_terms=(), param_ops=(),
                 rtol: float = 1e-6, atol: float = 1e-6, checkpoint_every: int = 10, nodes: int = 5):
        import time

        import torch

        t_begin = time.perf_counter()
        H0 = np.asarray(H0, dtype=np.complex128)
        ops = [np.asarray(o, dtype=np.complex128) for o in ops]
        times = np.asarray(times, dtype=np.float64)
        nterms, nseg = len(ops), len(times) - 1
        if nterms > 8:
            raise ValueError('at most 8 control operators')
        if nseg < 1 or np.any(np.diff(times) <= 0):
            raise ValueError('times must be strictly increasing with at least two entries')
        if nseg % checkpoint_every != 0:
            raise ValueError('checkpoint_every must divide the number of segments')
        rate_ops = [np.asarray(A, dtype=np.complex128) for A in rate_ops]
        if len(rate_ops) > 8:
            raise ValueError('at most 8 rate operators')
        Ls = [np.asarray(L, dtype=np.complex128) for L in Ls]
        rho0 = np.asarray(rho0, dtype=np.complex128)
        if rho0.shape[-1] == 1:
            rho0 = rho0 @ rho0.conj().T
        n = H0.shape[-1]
        T, t_start = float(times[-1]), float(times[0])
        losses = _loss_terms(O, times)
        mods = [(g, np.asarray(M, dtype=np.complex128)) for g, M in mod_terms]
        if len(mods) > _native.MAX_MOD:
            raise ValueError(f'at most {_native.MAX_MOD} modulated terms')
        params = _param_terms(param_ops)
        self.B, self.n, self.nterms, self.nseg, self.nodes = int(B), n, nterms, nseg, nodes
        self.ops, self.rate_ops, self.losses, self.params = ops, rate_ops, losses, params
        zeros = np.zeros(nseg)
        dev = torch.device('cuda', torch.cuda.current_device())
        self.dev = dev
        fmods = [(g, M, ('adj-mod', m)) for m, (g, M) in enumerate(mods)]

        # forward: checkpoints every checkpoint_every segments and the loss times, the
        # loss operators through exp_ops
        ck_idx = list(range(0, nseg, checkpoint_every)) + [nseg]
        self.nck = len(ck_idx) - 1
        fwd_ts = np.unique(np.concatenate([times[ck_idx], [t for t, _ in losses]]))
        self.ck_pos = np.searchsorted(fwd_ts, times[ck_idx[:-1]])
        self.loss_pos = np.searchsorted(fwd_ts, [t for t, _ in losses])
        fwd_insts = [
            dense.Instance(H=H0, Ls=Ls, rho0=rho0, pwc=[(times, zeros, ops[j]) for j in range(nterms)], mod=fmods,
                           key_H=(), key_L=(), key_rho=())
            for _ in range(B)
        ]
        t0 = time.perf_counter()
        self.fwd = dense.build_plan('tsit5', fwd_insts, fwd_ts, exp_ops=[Ok for _, Ok in losses], rtol=rtol,
                                    atol=atol)
        t1 = time.perf_counter()

        # Gauss-Legendre nodes per segment
        x, w = _gl(nodes)
        seg_len = np.diff(times)
        node_t = (times[:-1, None] + 0.5 * seg_len[:, None] * (1.0 + x[None, :])).reshape(-1)
        self.node_w = (0.5 * seg_len[:, None] * w[None, :]).reshape(-1)
        self.seg_first = (np.arange(nseg + 1) * nodes).astype(np.int32)
        self.Q = nseg * nodes

        # recompute from the checkpoints: one instance per (control vector, interval);
        # the initial states are copied from the forward's checkpoints on the device
        rec_insts = []
        for _ in range(B):
            for c in range(self.nck):
                k0, k1 = ck_idx[c], ck_idx[c + 1]
                rec_insts.append(
                    dense.Instance(H=H0, Ls=Ls, rho0=rho0, pwc=[(times, zeros, ops[j]) for j in range(nterms)],
                                   mod=fmods, tsave=node_t[k0 * nodes : k1 * nodes], t0=float(times[k0]), key_H=(),
                                   key_L=(), key_rho=())
                )
        self.rec = dense.build_plan('tsit5', rec_insts, rec_insts[0].tsave, rtol=rtol, atol=atol)
        t2 = time.perf_counter()

        # backward adjoint in reversed time s = T - t: Hnh' = iH - 1/2 sum L^dag L,
        # D = L^dag, C = L / 2, control operators -O_j on the reversed grid, modulated
        # terms -M_m with g_m(T - s); one instance per (control vector, loss term),
        # started at s_k = T - t_k with lambda = O_k. Every instance saves at every
        # node; the nodes after its loss time (s < s_k) hold its initial value and are
        # masked out when the instances are summed.
        LdL = sum((L.conj().T @ L for L in Ls), np.zeros_like(H0))
        rtimes = T - times[::-1]
        s_nodes = np.sort(T - node_t)
        bmods = [(lambda s, g=g: g(T - s), -M, ('adj-bmod', m)) for m, (g, M) in enumerate(mods)]
        bwd_ts = np.concatenate([[0.0], s_nodes, [T - t_start]])
        bwd_insts = [
            dense.Instance(H=H0, Ls=[L.conj().T for L in Ls], rho0=Ok, Hnh=1j * H0 - 0.5 * LdL,
                           pwc=[(rtimes, zeros, -ops[j]) for j in range(nterms)], mod=bmods,
                           tsave=bwd_ts, t0=T - tk, key_H=(), key_L=(), key_rho=())
            for _ in range(B) for tk, Ok in losses
        ]
        # lambda starts at O_k: the absolute tolerance scales with |O| so that the
        # backward solve keeps the forward's relative accuracy (|rho| <= 1)
        atol_b = atol * max(1.0, max(float(np.linalg.norm(Ok, 2)) for _, Ok in losses))
        self.bwd = dense.build_plan('tsit5', bwd_insts, bwd_ts, rtol=rtol, atol=atol_b)
        # node q (in t order) takes the loss terms whose time is after it
        self.loss_mask = np.array([[node_t[q] < tk for q in range(self.Q)] for tk, _ in losses])
        self.node_t = node_t
        t3 = time.perf_counter()

        # device slots of the PWC coefficients of every instance (complex128 per
        # segment, the instance's own copy in its plan's staged arena)
        self._slots = [self._pwc_slots(p) for p in (self.fwd, self.rec, self.bwd)]
        self.bstream = torch.cuda.Stream()  # default priority
        # quadrature inputs
        self.w_t = torch.from_numpy(self.node_w).to(dev)
        self.sf_t = torch.from_numpy(self.seg_first).to(dev)
        self.dia_ops = _sparse_controls(ops, n) if nterms else None
        self.o_t = (torch.from_numpy(np.ascontiguousarray(np.stack(ops).astype(np.complex64)).view(np.float32)).to(dev)
                    if nterms else None)
        c64 = lambda xs: torch.from_numpy(np.ascontiguousarray(np.stack(xs).astype(np.complex64)).view(np.float32)).to(dev)  # noqa: E731
        if rate_ops:
            self.a_t = c64(rate_ops)
            self.n_t = c64([A.conj().T @ A for A in rate_ops])
        # parameters: one quadrature over all nodes each, weights w_q g(t_q)
        self.mask_t = torch.from_numpy(self.loss_mask.astype(np.float32)).to(dev)
        if params:
            self.p_t = [c64([P]) for _, P in params]
            self.pw_t = [torch.from_numpy(self.node_w * (1.0 if g is None else np.real(np.asarray(
                [complex(g(t)) for t in node_t])))).to(dev) for g, _ in params]
            self.sf_all = torch.from_numpy(np.array([0, self.Q], dtype=np.int32)).to(dev)
        torch.cuda.synchronize()
        self.pack_timing = {'forward_pack_s': t1 - t0, 'recompute_pack_s': t2 - t1, 'backward_pack_s': t3 - t2,
                            'plan_total_s': time.perf_counter() - t_begin}

    def _pwc_slots(self, plan):
        """(arena as float64, index of every coefficient's real part) of a plan."""
        import torch

        if not self.nterms:
            return None
        arena = plan.arena.tensors[0]
        base = arena.data_ptr()
        nv = self.nseg
        per = self.nterms * nv * 2
        starts = (plan.descs['pwc_vals'].astype(np.int64) - base) // 8
        if np.any(plan.descs['nseg'] != nv + 1) or np.any(starts < 0) or np.any(starts + per > arena.numel() // 8):
            raise RuntimeError('adjoint: unexpected PWC coefficient layout')
        idx = (starts[:, None] + np.arange(per)[None, :]).reshape(-1)
        return arena.view(torch.float64), torch.from_numpy(idx).to(self.dev)

    def _write_controls(self, which: int, vals: np.ndarray):
        """vals (instances, nterms, nseg) real controls -> the plan's complex coefficients."""
        import torch

        if self._slots[which] is None:
            return
        f64, idx = self._slots[which]
        c = np.zeros(vals.shape + (2,), dtype=np.float64)
        c[..., 0] = vals
        f64[idx] = torch.from_numpy(c.reshape(-1)).to(self.dev, non_blocking=False)

    def run(self, u) -> AdjointResult:
        import time

        import torch

        u = np.asarray(u, dtype=np.float64)
        if u.ndim == 2:
            u = u[:, None, :]
        B, n, nterms, nseg, Q, nck = self.B, self.n, self.nterms, self.nseg, self.Q, self.nck
        if u.shape != (B, nterms, nseg):
            raise ValueError(f'u must have shape {(B, nterms, nseg)}')
        timing = {}
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        K = len(self.losses)
        self._write_controls(0, u)
        self._write_controls(1, np.repeat(u, nck, axis=0))
        self._write_controls(2, np.repeat(u[:, :, ::-1], K, axis=0))
        fwd, rec, bwd = self.fwd, self.rec, self.bwd
        fwd.reset()
        bwd.reset()
        main = torch.cuda.current_stream()
        ev = {k: torch.cuda.Event(enable_timing=True) for k in ('f0', 'f1', 'r1', 'b0', 'b1')}
        # the backward solve needs only the controls: it runs on its own stream,
        # concurrently with the forward solve and the recompute (one workgroup per
        # instance leaves most CUs free for it)
        self.bstream.wait_stream(main)
        ev['b0'].record(self.bstream)
        bwd.run(self.bstream)
        ev['b1'].record(self.bstream)
        ev['f0'].record(main)
        fwd.run(main)
        ev['f1'].record(main)
        # checkpoints -> initial states of the recompute instances (FP64, padded)
        npd = rec.np
        ck = torch.from_numpy(self.ck_pos).to(self.dev)
        st = fwd.states.view(B, fwd.ntsave, n, n, 2)[:, ck].reshape(B * nck, n, n, 2).to(torch.float64)
        y0 = rec.y[:, 0].view(B * nck, npd, npd, 2)
        y0.zero_()
        y0[:, :n, :n, :] = st
        rec.run(main)
        ev['r1'].record(main)
        main.wait_stream(self.bstream)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        timing.update({
            'forward_run_s': 1e-3 * ev['f0'].elapsed_time(ev['f1']),
            'recompute_run_s': 1e-3 * ev['f1'].elapsed_time(ev['r1']),
            'backward_run_s': 1e-3 * ev['b0'].elapsed_time(ev['b1']),
            'solves_wall_s': t1 - t0,
        })
        for name, plan in (('forward', fwd), ('recompute', rec), ('backward', bwd)):
            _check_status(name, plan)
        ex = fwd.expects.view(B, K, fwd.ntsave, 2).to(torch.float64).cpu().numpy()
        J = sum(ex[:, k, int(self.loss_pos[k]), 0] for k in range(K))
        rho_nodes_dev = rec.states  # (B * nck, Qc n n 2) float32 == (B, Q, n, n) complex64
        bst = bwd.states.view(B, K, 2 + Q, n * n * 2)
        lam_k = torch.flip(bst[:, :, 1 : 1 + Q], dims=[2])  # (B, K, Q, .) in t order
        lam_dev = (lam_k * self.mask_t[None, :, :, None]).sum(dim=1).contiguous()
        lambda0 = bst[:, :, 1 + Q].sum(dim=1).cpu().numpy().view(np.complex64).reshape(B, n, n)

        stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
        lib = _native.lib()
        P = ctypes.c_void_p
        t5 = time.perf_counter()
        grad = np.zeros((B, nterms, nseg))
        if nterms:
            g_t = torch.zeros((B, nterms, nseg), dtype=torch.float64, device=self.dev)
            if self.dia_ops is not None:
                dops, _ = self.dia_ops
                _native.check(
                    lib.dqh_adjoint_pwc_quadrature_dia(
                        B, n, Q, nseg, P(rho_nodes_dev.data_ptr()), P(lam_dev.data_ptr()), ctypes.byref(dops),
                        P(self.w_t.data_ptr()), P(self.sf_t.data_ptr()), P(g_t.data_ptr()), stream,
                    ),
                    'dqh_adjoint_pwc_quadrature_dia',
                )
            else:
                _native.check(
                    lib.dqh_adjoint_pwc_quadrature(
                        B, n, Q, nterms, nseg, P(rho_nodes_dev.data_ptr()), P(lam_dev.data_ptr()),
                        P(self.o_t.data_ptr()), P(self.w_t.data_ptr()), P(self.sf_t.data_ptr()), P(g_t.data_ptr()),
                        stream,
                    ),
                    'dqh_adjoint_pwc_quadrature',
                )
            grad = g_t.cpu().numpy()
        grad_rates = None
        if self.rate_ops:
            gr_t = torch.zeros((B, len(self.rate_ops), nseg), dtype=torch.float64, device=self.dev)
            _native.check(
                lib.dqh_adjoint_rate_quadrature(
                    B, n, Q, len(self.rate_ops), nseg, P(rho_nodes_dev.data_ptr()), P(lam_dev.data_ptr()),
                    P(self.a_t.data_ptr()), P(self.n_t.data_ptr()), P(self.w_t.data_ptr()), P(self.sf_t.data_ptr()),
                    P(gr_t.data_ptr()), stream,
                ),
                'dqh_adjoint_rate_quadrature',
            )
            grad_rates = gr_t.cpu().numpy().sum(axis=-1)
        grad_params = None
        if self.params:
            gp_t = torch.zeros((len(self.params), B, 1, 1), dtype=torch.float64, device=self.dev)
            for p_i in range(len(self.params)):
                _native.check(
                    lib.dqh_adjoint_pwc_quadrature(
                        B, n, Q, 1, 1, P(rho_nodes_dev.data_ptr()), P(lam_dev.data_ptr()), P(self.p_t[p_i].data_ptr()),
                        P(self.pw_t[p_i].data_ptr()), P(self.sf_all.data_ptr()), P(gp_t[p_i].data_ptr()), stream,
                    ),
                    'dqh_adjoint_pwc_quadrature',
                )
            grad_params = gp_t.cpu().numpy()[:, :, 0, 0].T.copy()
        torch.cuda.synchronize()
        timing['quadrature_s'] = time.perf_counter() - t5
        timing['forward_nrhs'] = float(fwd.stats[:, 4].sum().item())
        return AdjointResult(J=J, grad_u=grad, grad_rates=grad_rates, lambda0=lambda0, timing=timing,
                             grad_params=grad_params)
