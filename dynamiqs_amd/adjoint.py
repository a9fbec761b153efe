"""Adjoint gradients of a piecewise-constant control on the MI355X.

Replaces the reverse-mode path of the reference (``jax.grad`` through
``dx.RecursiveCheckpointAdjoint``, ``diffrax_integrator.py:88-99``,
``gradient.py:32-72``) for the loss ``J = Tr(O rho(T))`` of a master equation with

    H(t) = H0 + sum_j u_j(t) O_j,   u_j constant on [times[k], times[k+1]),

by the continuous adjoint (optimise-then-discretise):

    dJ/du_jk = int_{times[k]}^{times[k+1]} Re Tr(lambda(t) (-i [O_j, rho(t)])) dt,
    d lambda / dt = -L^dag[lambda],  lambda(T) = O,
    L^dag[l] = i[H, l] + sum_k (L_k^dag l L_k - 1/2 {L_k^dag L_k, l}).

Four GPU launches, all on the hand-written kernels:

1. forward Tsit5 of every instance, saving rho at the checkpoints (every
   ``checkpoint_every`` segments) and ``J`` through ``exp_ops``;
2. recomputation from the checkpoints: one instance per (control vector,
   checkpoint interval), all in one persistent launch, saving rho at the
   Gauss-Legendre nodes of each segment;
3. backward Tsit5 of the adjoint Lindbladian (Hermitian form with
   ``Hnh' = iH - 1/2 sum L^dag L``, ``D = L^dag``, ``C = L / 2``) in reversed time
   ``s = T - t``, saving lambda at the same nodes;
4. ``dqh_adjoint_pwc_quadrature``: the Gauss-Legendre sum per segment in FP64.

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
    """``J`` (B,): loss ``Tr(O rho(T))``; ``grad_u`` (B, nterms, nseg): dJ/du;
    ``grad_rates`` (B, nrates): dJ/dkappa_r for jump operators sqrt(kappa_r) A_r;
    ``lambda0`` (B, n, n): the adjoint state at the start time, dJ/drho0 (J is linear
    in rho0: J = Tr(lambda0 rho0)); ``timing``: seconds per phase."""

    J: np.ndarray
    grad_u: np.ndarray
    grad_rates: np.ndarray | None
    lambda0: np.ndarray
    timing: dict


def mesolve_pwc_grad(H0, ops, times, u, Ls, rho0, O, *, return_timing: bool = False, **kw):
    """Loss ``J_b = Tr(O rho_b(T))`` and gradient ``dJ_b/du_b`` for a batch of control
    arrays ``u`` of shape ``(B, nterms, nseg)`` (``T = times[-1]``).

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
    O: np.ndarray,
    *,
    rate_ops=(),
    rtol: float = 1e-6,
    atol: float = 1e-6,
    checkpoint_every: int = 10,
    nodes: int = 5,
) -> AdjointResult:
    """Loss, control gradient, jump-rate gradients and adjoint initial state of
    ``J = Tr(O rho(T))`` for a batch of control arrays ``u`` (B, nterms, nseg).

    ``rate_ops``: operators A_r whose rates are differentiated; the corresponding
    jump operators in ``Ls`` are sqrt(kappa_r) A_r (dJ/dkappa_r is summed over the
    whole interval). ``ops`` may be empty (then ``u`` has shape (B, 0, nseg) and
    ``times`` only sets the quadrature segments).
    """
    import time

    import torch

    H0 = np.asarray(H0, dtype=np.complex128)
    ops = [np.asarray(o, dtype=np.complex128) for o in ops]
    times = np.asarray(times, dtype=np.float64)
    u = np.asarray(u)
    if u.ndim == 2:
        u = u[:, None, :]
    B, nterms, nseg = u.shape
    if len(ops) != nterms or len(times) != nseg + 1:
        raise ValueError('u must have shape (B, len(ops), len(times) - 1)')
    if nterms > 8:
        raise ValueError('at most 8 control operators')
    rate_ops = [np.asarray(A, dtype=np.complex128) for A in rate_ops]
    if len(rate_ops) > 8:
        raise ValueError('at most 8 rate operators')
    Ls = [np.asarray(L, dtype=np.complex128) for L in Ls]
    rho0 = np.asarray(rho0, dtype=np.complex128)
    if rho0.shape[-1] == 1:
        rho0 = rho0 @ rho0.conj().T
    O = np.asarray(O, dtype=np.complex128)
    n = H0.shape[-1]
    T = float(times[-1])
    t_start = float(times[0])
    timing = {}

    # ---------------------------------------------------------------- forward
    ck_idx = list(range(0, nseg, checkpoint_every)) + [nseg]
    ck_t = times[ck_idx]
    fwd_insts = [
        dense.Instance(
            H=H0, Ls=Ls, rho0=rho0, pwc=[(times, u[b, j], ops[j]) for j in range(nterms)], key_H=(), key_L=(), key_rho=()
        )
        for b in range(B)
    ]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    fwd = dense.build_plan('tsit5', fwd_insts, ck_t, exp_ops=[O], rtol=rtol, atol=atol)
    t1 = time.perf_counter()
    fwd.run()
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    timing['forward_pack_s'] = t1 - t0
    timing['forward_run_s'] = t2 - t1
    fres = fwd.results()
    J = fres['expects'][:, 0, -1].real.astype(np.float64)
    rho_ck = fres['states']  # (B, nck, n, n) complex64

    # ------------------------------------------ Gauss-Legendre nodes per segment
    x, w = _gl(nodes)
    seg_len = np.diff(times)
    node_t = (times[:-1, None] + 0.5 * seg_len[:, None] * (1.0 + x[None, :])).reshape(-1)  # (nseg*nodes,)
    node_w = (0.5 * seg_len[:, None] * w[None, :]).reshape(-1)
    seg_first = (np.arange(nseg + 1) * nodes).astype(np.int32)

    # ------------------------------------------- recompute rho from checkpoints
    t3 = time.perf_counter()
    rec_insts = []
    nck = len(ck_idx) - 1
    for b in range(B):
        for c in range(nck):
            k0, k1 = ck_idx[c], ck_idx[c + 1]
            ts = node_t[k0 * nodes : k1 * nodes]
            rec_insts.append(
                dense.Instance(
                    H=H0,
                    Ls=Ls,
                    rho0=rho_ck[b, c].astype(np.complex128),
                    pwc=[(times, u[b, j], ops[j]) for j in range(nterms)],
                    tsave=ts,
                    t0=float(times[k0]),
                    key_H=(),
                    key_L=(),
                )
            )
    if len({len(i.tsave) for i in rec_insts}) != 1:
        raise ValueError('checkpoint_every must divide the number of segments')
    rec = dense.build_plan('tsit5', rec_insts, rec_insts[0].tsave, rtol=rtol, atol=atol)
    t4 = time.perf_counter()
    rec.run()
    torch.cuda.synchronize()
    t5 = time.perf_counter()
    timing['recompute_pack_s'] = t4 - t3
    timing['recompute_run_s'] = t5 - t4
    Q = nseg * nodes
    rho_nodes_dev = rec.states  # (B * nck, Qc * n * n * 2) float32 == (B, Q, n, n) complex64

    # -------------------------------------------- backward adjoint (s = T - t)
    LdL = sum((L.conj().T @ L for L in Ls), np.zeros_like(H0))
    Hnh_adj = 1j * H0 - 0.5 * LdL
    rtimes = T - times[::-1]
    s_nodes = np.sort(T - node_t)
    bwd_insts = [
        dense.Instance(
            H=H0,
            Ls=[L.conj().T for L in Ls],
            rho0=O,
            Hnh=Hnh_adj,
            pwc=[(rtimes, u[b, j][::-1].copy(), -ops[j]) for j in range(nterms)],
            key_H=(),
            key_L=(),
            key_rho=(),
        )
        for b in range(B)
    ]
    # saves: s = 0 (lambda = O), the nodes, and s = T - t0 (lambda at the start time)
    bwd_ts = np.concatenate([[0.0], s_nodes, [T - t_start]])
    # lambda starts at O: scale the absolute tolerance by |O| so that the backward
    # solve is held to the same relative accuracy as the forward one (|rho| <= 1)
    atol_b = atol * max(1.0, float(np.linalg.norm(O, 2)))
    t6 = time.perf_counter()
    bwd = dense.build_plan('tsit5', bwd_insts, bwd_ts, rtol=rtol, atol=atol_b)
    t7 = time.perf_counter()
    bwd.run()
    torch.cuda.synchronize()
    t8 = time.perf_counter()
    timing['backward_pack_s'] = t7 - t6
    timing['backward_run_s'] = t8 - t7
    for name, plan in (('forward', fwd), ('recompute', rec), ('backward', bwd)):
        _check_status(name, plan)
    # lambda at the nodes in increasing t, kept on the device
    bst = bwd.states.view(B, 2 + Q, n * n * 2)
    lam_dev = torch.flip(bst[:, 1 : 1 + Q], dims=[1]).contiguous()
    lambda0 = bst[:, 1 + Q].cpu().numpy().view(np.complex64).reshape(B, n, n)

    # ------------------------------------------------------------- quadrature
    dev = torch.device('cuda', torch.cuda.current_device())
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    w_t = torch.from_numpy(node_w).to(dev)
    sf_t = torch.from_numpy(seg_first).to(dev)
    lib = _native.lib()
    P = ctypes.c_void_p
    t9 = time.perf_counter()
    grad = np.zeros((B, nterms, nseg))
    if nterms:
        g_t = torch.zeros((B, nterms, nseg), dtype=torch.float64, device=dev)
        dia_ops = _sparse_controls(ops, n)
        if dia_ops is not None:
            # control operators with few diagonals: 2 d n^2 products per node
            dops, keep = dia_ops
            _native.check(
                lib.dqh_adjoint_pwc_quadrature_dia(
                    B, n, Q, nseg, P(rho_nodes_dev.data_ptr()), P(lam_dev.data_ptr()), ctypes.byref(dops),
                    P(w_t.data_ptr()), P(sf_t.data_ptr()), P(g_t.data_ptr()), stream,
                ),
                'dqh_adjoint_pwc_quadrature_dia',
            )
            torch.cuda.synchronize()
            del keep
        else:
            o_t = torch.from_numpy(np.ascontiguousarray(np.stack(ops).astype(np.complex64)).view(np.float32)).to(dev)
            _native.check(
                lib.dqh_adjoint_pwc_quadrature(
                    B, n, Q, nterms, nseg, P(rho_nodes_dev.data_ptr()), P(lam_dev.data_ptr()), P(o_t.data_ptr()),
                    P(w_t.data_ptr()), P(sf_t.data_ptr()), P(g_t.data_ptr()), stream,
                ),
                'dqh_adjoint_pwc_quadrature',
            )
        grad = g_t.cpu().numpy()
    grad_rates = None
    if rate_ops:
        c64 = lambda xs: torch.from_numpy(np.ascontiguousarray(np.stack(xs).astype(np.complex64)).view(np.float32)).to(dev)  # noqa: E731
        a_t = c64(rate_ops)
        n_t = c64([A.conj().T @ A for A in rate_ops])
        gr_t = torch.zeros((B, len(rate_ops), nseg), dtype=torch.float64, device=dev)
        _native.check(
            lib.dqh_adjoint_rate_quadrature(
                B, n, Q, len(rate_ops), nseg, P(rho_nodes_dev.data_ptr()), P(lam_dev.data_ptr()), P(a_t.data_ptr()),
                P(n_t.data_ptr()), P(w_t.data_ptr()), P(sf_t.data_ptr()), P(gr_t.data_ptr()), stream,
            ),
            'dqh_adjoint_rate_quadrature',
        )
        grad_rates = gr_t.cpu().numpy().sum(axis=-1)
    torch.cuda.synchronize()
    timing['quadrature_s'] = time.perf_counter() - t9
    timing['forward_nrhs'] = float(np.sum(fres['nrhs']))
    return AdjointResult(J=J, grad_u=grad, grad_rates=grad_rates, lambda0=lambda0, timing=timing)
