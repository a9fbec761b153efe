"""Rouchon1 for sparse-DIA operators on the MI355X (``dqh_mesolve_rouchon1_dia``).

Host part of the large-``n`` path: the normalisation ``S = M0^dag M0 + dt sum
L^dag L = T T^dag`` (``rouchon_integrator.py:259-263, 573``) is factorised once in
FP64 with a banded Cholesky, and the blocks of the banded solve ``T^dag x = s``
(``:578-582``) are handed to the kernels as dense ``64 x 64`` matrices:
``V_i - I`` with ``V_i = (T_ii^dag)^-1`` and ``Wn_i = -V_i T_{i+1,i}^dag``. The
Kraus operators go to the device by diagonals: ``dM = M0 - I = -i dt Hnh`` and
``A_k = sqrt(dt) L_k`` (``:250-257``).
"""

from __future__ import annotations

import ctypes

import numpy as np
import scipy.linalg as sla

from .. import _native
from .dense import _Arena, _device, rouchon_step_times

NB = 64
MAXOPS = 8
MAXOFF = 16


class DiaOps(ctypes.Structure):
    _fields_ = [
        ('nops', ctypes.c_int32),
        ('noff', ctypes.c_int32 * MAXOPS),
        ('off', (ctypes.c_int32 * MAXOFF) * MAXOPS),
        ('pad_', ctypes.c_int32),
        ('diags', ctypes.c_uint64 * MAXOPS),
    ]


class DiaRouchon1(ctypes.Structure):
    _fields_ = [
        ('n', ctypes.c_int64),
        ('n_true', ctypes.c_int64),
        ('nb', ctypes.c_int64),
        ('ninst', ctypes.c_int64),
        ('nsteps', ctypes.c_int64),
        ('nsave', ctypes.c_int64),
        ('nsets', ctypes.c_int64),
        ('pad_', ctypes.c_int64),
        ('Vp', ctypes.c_uint64 * 2),
        ('Wn', ctypes.c_uint64 * 2),
        ('ops', DiaOps * 2),
        ('rho64', ctypes.c_uint64),
        ('rho32', ctypes.c_uint64),
        ('work', ctypes.c_uint64),
        ('states', ctypes.c_uint64),
        ('save_step', ctypes.c_uint64),
        ('save_theta', ctypes.c_uint64),
        ('nE', ctypes.c_int64),
        ('EopsT', ctypes.c_uint64),
        ('expects', ctypes.c_uint64),
        ('variant', ctypes.c_int64),
        ('solve_panel', ctypes.c_int64),
        ('timing', ctypes.c_uint64),
        ('nblk', ctypes.c_int64),
        ('blk_maxd', ctypes.c_int64),
        ('blk_of', ctypes.c_uint64),
        ('blk_loc', ctypes.c_uint64),
        ('blk_start', ctypes.c_uint64),
        ('blk_idx', ctypes.c_uint64),
        ('blk_mat', ctypes.c_uint64),
        ('Lm', ctypes.c_uint64 * 2),
        ('Rm', ctypes.c_uint64 * 2),
    ]


SWEEP_AUTO, SWEEP_GLOBAL, SWEEP_TILED, SWEEP_TILED_H8, SWEEP_TILED_H16, SWEEP_TILED_H32 = 0, 1, 2, 3, 4, 5
SWEEP_HALO = {SWEEP_TILED_H8: 8, SWEEP_TILED_H16: 16, SWEEP_TILED_H32: 32}
SWEEP_SPLIT = 6  # accepted by the Rouchon1 Kraus sweep, runs its global form
MAXBLK = 64  # largest index set of the block form of the normalisation (DQH_DIA_MAXBLK)
MAXN_BLOCKS = 4096  # the block form's strips of n elements fit 64 KiB of LDS


class DiaRhs(ctypes.Structure):
    _fields_ = [
        ('n', ctypes.c_int64),
        ('ninst', ctypes.c_int64),
        ('variant', ctypes.c_int64),
        ('pad_', ctypes.c_int64),
        ('ops', DiaOps),
        ('rho', ctypes.c_uint64),
        ('drho', ctypes.c_uint64),
    ]


def _dia_of(x: np.ndarray, tol: float = 0.0):
    n = x.shape[-1]
    offs = [k for k in range(-(n - 1), n) if np.max(np.abs(np.diagonal(x, offset=k)), initial=0.0) > tol]
    d = np.zeros((len(offs), n), dtype=np.complex128)
    for i, k in enumerate(offs):
        if k >= 0:
            d[i, k:] = np.diagonal(x, offset=k)
        else:
            d[i, : n + k] = np.diagonal(x, offset=k)
    return offs, d


def half_bandwidth(S: np.ndarray) -> int:
    offs, _ = _dia_of(S)
    return max(abs(o) for o in offs) if offs else 0


def normalisation_S(H: np.ndarray, Ls, dt: float, n_pad: int) -> np.ndarray:
    """S = M0^dag M0 + dt sum L^dag L of one Rouchon1 step (rouchon_integrator.py:259-263),
    padded with the identity to n_pad."""
    n = H.shape[-1]
    LdL = sum((L.conj().T @ L for L in Ls), np.zeros_like(H))
    M0 = np.eye(n) - 1j * dt * (H - 0.5j * LdL)
    Sp = np.eye(n_pad, dtype=np.complex128)
    Sp[:n, :n] = M0.conj().T @ M0 + dt * LdL
    return Sp


def coupled_sets(S: np.ndarray):
    """Index sets of the connected components of the graph with an edge i-j where
    S[i, j] != 0 (each set ascending), or None when they do not make the block form
    of the normalisation worthwhile: one set only, or a set above MAXBLK indices."""
    from scipy.sparse import csr_matrix
    from scipy.sparse.csgraph import connected_components

    ncomp, labels = connected_components(csr_matrix(S != 0), directed=False)
    if ncomp < 2:
        return None
    sets = [np.flatnonzero(labels == c) for c in range(ncomp)]
    if max(len(x) for x in sets) > MAXBLK:
        return None
    return sets


def block_normalisation(S: np.ndarray, sets):
    """Device arrays of the block form: (of, loc, start, idx, mat, L, R) with, per set b
    of d indices, L_b = T_b^-dag - I (row-major) and R_b = (T_b^-1 - I)^T (row-major),
    T_b the lower Cholesky factor of S on the set (FP64, rounded to complex64). The
    factor assembled from the sets' factors is lower triangular in the original order
    and reproduces S, so it is the Cholesky factor of S."""
    n = S.shape[0]
    of = np.zeros(n, dtype=np.int32)
    loc = np.zeros(n, dtype=np.int32)
    start = np.zeros(len(sets) + 1, dtype=np.int32)
    mat = np.zeros(len(sets) + 1, dtype=np.int64)
    Ls, Rs = [], []
    for b, idx in enumerate(sets):
        d = len(idx)
        of[idx] = b
        loc[idx] = np.arange(d, dtype=np.int32)
        start[b + 1] = start[b] + d
        mat[b + 1] = mat[b] + d * d
        T = np.linalg.cholesky(S[np.ix_(idx, idx)])
        Ti = sla.solve_triangular(T, np.eye(d), lower=True)
        Ls.append((Ti.conj().T - np.eye(d)).reshape(-1))
        Rs.append((Ti - np.eye(d)).T.reshape(-1))
    idx_all = np.concatenate(sets).astype(np.int32)
    return of, loc, start, idx_all, mat, np.concatenate(Ls).astype(np.complex64), np.concatenate(Rs).astype(np.complex64)


def block_solve_operators(H: np.ndarray, Ls, dt: float, n_pad: int, banded: bool = True):
    """FP64 host precomputation of one step length: (Vp, Wn, dM diagonals, A_k
    diagonals); Vp and Wn are None without ``banded``."""
    n = H.shape[-1]
    LdL = sum((L.conj().T @ L for L in Ls), np.zeros_like(H))
    Hnh = H - 0.5j * LdL
    if not banded:
        ops = [_dia_of(-1j * dt * Hnh)] + [_dia_of(np.sqrt(dt) * L) for L in Ls]
        return None, None, ops
    M0 = np.eye(n) - 1j * dt * Hnh
    S = M0.conj().T @ M0 + dt * LdL
    w = half_bandwidth(S)
    if w > NB:
        raise NotImplementedError(f'DIA Rouchon1: half-bandwidth {w} of the normalisation exceeds {NB}')
    Sp = np.eye(n_pad, dtype=np.complex128)
    Sp[:n, :n] = S
    ab = np.zeros((w + 1, n_pad), dtype=np.complex128)
    for k in range(w + 1):
        ab[k, : n_pad - k] = np.diagonal(Sp, offset=-k)
    Tab = sla.cholesky_banded(ab, lower=True)  # T[j + k, j] = Tab[k, j]
    T = np.zeros((n_pad, n_pad), dtype=np.complex128)
    for k in range(w + 1):
        idx = np.arange(n_pad - k)
        T[idx + k, idx] = Tab[k, : n_pad - k]
    nblk = n_pad // NB
    Vp = np.zeros((nblk, NB, NB), dtype=np.complex128)
    Wn = np.zeros((nblk, NB, NB), dtype=np.complex128)
    for i in range(nblk):
        s = slice(i * NB, (i + 1) * NB)
        V = sla.solve_triangular(T[s, s].conj().T, np.eye(NB), lower=False)
        # inverse of an upper-triangular block: upper triangular (the solve kernel
        # skips the zero part below the diagonal)
        Vp[i] = np.triu(V - np.eye(NB))
        if i + 1 < nblk:
            s1 = slice((i + 1) * NB, (i + 2) * NB)
            Wn[i] = -V @ T[s1, s].conj().T
    dM_off, dM_d = _dia_of(-1j * dt * Hnh)
    ops = [(dM_off, dM_d)]
    for L in Ls:
        ops.append(_dia_of(np.sqrt(dt) * L))
    return Vp, Wn, ops


def save_schedule(tsave: np.ndarray, step_t: np.ndarray):
    steps, thetas = [], []
    for ts in tsave:
        if ts <= step_t[0]:
            steps.append(-1)
            thetas.append(0.0)
            continue
        m = int(np.searchsorted(step_t, ts, side='left')) - 1
        m = min(max(m, 0), len(step_t) - 2)
        h = step_t[m + 1] - step_t[m]
        steps.append(m)
        thetas.append(float(min(max((ts - step_t[m]) / h, 0.0), 1.0)))
    return np.asarray(steps, dtype=np.int64), np.asarray(thetas, dtype=np.float64)


class DiaPlan:
    def __init__(self):
        self.problem = None
        self.arena = None
        self.keep = []

    def run(self, stream=None):
        import torch

        s = stream if stream is not None else torch.cuda.current_stream()
        lib = _native.lib()
        _native.check(lib.dqh_mesolve_rouchon1_dia(ctypes.byref(self.problem), ctypes.c_void_p(s.cuda_stream)), 'dqh_mesolve_rouchon1_dia')

    def reset(self):
        self.rho64.copy_(self.rho64_0)

    def phase_us(self, stream=None) -> dict:
        """Reset and run once with HIP events around the block solves and the Kraus
        sweep of steps 1..8: mean microseconds per step of each phase."""
        import torch

        t = np.zeros(2, dtype=np.float64)
        self.reset()
        self.problem.timing = t.ctypes.data
        try:
            self.run(stream)
        finally:
            self.problem.timing = 0
        torch.cuda.synchronize()
        return {'solve': 1e3 * float(t[0]), 'sweep': 1e3 * float(t[1])}

    def results(self) -> dict:
        import torch

        torch.cuda.synchronize()
        st = self.states.cpu().numpy().view(np.complex64).reshape(self.ninst, self.nsave, self.n_true, self.n_true)
        fin = self.rho64.cpu().numpy().view(np.complex128).reshape(self.ninst, self.n, self.n)[:, : self.n_true, : self.n_true]
        out = {'states': st, 'final_state': fin, 'nsteps': np.full(self.ninst, self.nsteps)}
        if self.nE:
            out['expects'] = self.expects.cpu().numpy().view(np.complex64).reshape(self.ninst, self.nE, self.nsave)
        return out


def build_plan(
    H: np.ndarray,
    Ls,
    rho0: np.ndarray,
    tsave: np.ndarray,
    dt: float,
    t0: float | None = None,
    exp_ops=None,
    variant: int = SWEEP_AUTO,
    solve_panel: int = 0,
    normalisation: str = 'auto',
) -> DiaPlan:
    """Pack a batch of Rouchon1 solves that share H and Ls (dense FP64 host arrays
    whose nonzeros lie on few diagonals) and differ in rho0 (B, n, n). ``exp_ops``
    (dense (n, n) arrays) are evaluated at every save time on the device."""
    import torch

    dev = _device()
    _native.lib()
    tsave = np.asarray(tsave, dtype=np.float64)
    rho0 = np.asarray(rho0, dtype=np.complex128)
    if rho0.ndim == 2:
        rho0 = rho0[None]
    B, n = rho0.shape[0], H.shape[-1]
    n_pad = NB * ((n + NB - 1) // NB)
    t_start = float(tsave[0] if t0 is None else t0)
    step_t = rouchon_step_times(t_start, float(tsave[-1]), float(dt))
    hs = np.diff(step_t)
    nsteps = len(hs)
    sets = [float(hs[0])]
    if abs(hs[-1] - hs[0]) > 1e-12 * max(1.0, abs(hs[0])):
        sets.append(float(hs[-1]))
    arena = _Arena(dev)
    prob = DiaRouchon1()
    prob.n, prob.n_true, prob.nb, prob.ninst = n_pad, n, NB, B
    if variant == SWEEP_AUTO:
        from .. import autotune

        variant = autotune.dia_sweep(H, Ls, B, kind='kraus', dt=float(dt))
    prob.nsteps, prob.nsets, prob.variant, prob.solve_panel = nsteps, len(sets), variant, int(solve_panel)
    # normalisation: block form when S decouples into small index sets (the same
    # sets for every step length: they follow the sparsity of Hnh and L^dag L)
    if normalisation not in ('auto', 'banded', 'blocks'):
        raise ValueError("normalisation must be 'auto', 'banded' or 'blocks'")
    isets = None
    if normalisation != 'banded' and n_pad <= MAXN_BLOCKS:
        isets = coupled_sets(normalisation_S(H, Ls, sets[0], n_pad))
    if normalisation == 'blocks' and isets is None:
        raise ValueError('the normalisation matrix S does not decouple into sets of at most '
                         f'{MAXBLK} indices (or n > {MAXN_BLOCKS})')
    for si, h in enumerate(sets):
        Vp, Wn, ops = block_solve_operators(H, Ls, h, n_pad, banded=isets is None)
        if isets is None:
            prob.Vp[si] = arena.put_now(Vp.astype(np.complex64))
            prob.Wn[si] = arena.put_now(Wn.astype(np.complex64))
        else:
            of, loc, start, idx, mat, Lm, Rm = block_normalisation(normalisation_S(H, Ls, h, n_pad), isets)
            if si == 0:
                prob.nblk, prob.blk_maxd = len(isets), max(len(x) for x in isets)
                prob.blk_of, prob.blk_loc = arena.put_now(of), arena.put_now(loc)
                prob.blk_start, prob.blk_idx = arena.put_now(start), arena.put_now(idx)
                prob.blk_mat = arena.put_now(mat)
            prob.Lm[si], prob.Rm[si] = arena.put_now(Lm), arena.put_now(Rm)
        _fill_ops(prob.ops[si], ops, n_pad, arena)
    r64 = np.zeros((B, n_pad, n_pad), dtype=np.complex128)
    r64[:, :n, :n] = rho0
    rho64_0 = torch.from_numpy(r64.view(np.float64)).to(dev)
    rho64 = rho64_0.clone()
    rho32 = torch.empty((2 * B * n_pad * n_pad * 2,), dtype=torch.float32, device=dev)
    work = torch.empty((2 * B * n_pad * n_pad * 2,), dtype=torch.float32, device=dev)
    steps_h, thetas_h = save_schedule(tsave, step_t)
    states = torch.zeros((B * len(tsave) * n * n * 2,), dtype=torch.float32, device=dev)
    prob.rho64, prob.rho32, prob.work, prob.states = rho64.data_ptr(), rho32.data_ptr(), work.data_ptr(), states.data_ptr()
    prob.nsave = len(tsave)
    prob.save_step = steps_h.ctypes.data
    prob.save_theta = thetas_h.ctypes.data
    nE = 0 if exp_ops is None else len(exp_ops)
    expects = torch.zeros((B * max(nE, 1) * len(tsave) * 2,), dtype=torch.float32, device=dev)
    prob.nE = nE
    if nE:
        ET = np.stack([np.asarray(E, dtype=np.complex128).T for E in exp_ops]).astype(np.complex64)
        prob.EopsT = arena.put_now(ET)
        prob.expects = expects.data_ptr()
    plan = DiaPlan()
    plan.problem, plan.arena = prob, arena
    plan.keep = [steps_h, thetas_h, rho32, work]
    plan.rho64, plan.rho64_0, plan.states, plan.expects = rho64, rho64_0, states, expects
    plan.ninst, plan.n, plan.n_true, plan.nsave, plan.nsteps, plan.nE = B, n_pad, n, len(tsave), nsteps, nE
    # roofline units: the Kraus sweep reads rho and writes rho' (2 n^2 complex64) per step and instance
    plan.sweep_bytes_per_step = 2.0 * n * n * 8 * B
    plan.solve_flops_per_step = 2 * 2 * 8.0 * NB * NB * n_pad * (n_pad // NB) * B
    plan.normalisation = 'banded' if isets is None else 'blocks'
    if isets is not None:  # two one-sided products of d x d triangular blocks per row / column
        d2 = float(sum(len(x) * (len(x) + 1) / 2 for x in isets))
        plan.solve_flops_per_step = 2 * 8.0 * d2 * n_pad * B
    return plan


def _fill_ops(o: DiaOps, ops, n_pad: int, arena: _Arena) -> None:
    if len(ops) > MAXOPS:
        raise NotImplementedError(f'DIA sweeps support at most {MAXOPS} operators (Hnh and {MAXOPS - 1} jump operators)')
    o.nops = len(ops)
    for k, (offs, d) in enumerate(ops):
        if len(offs) > MAXOFF:
            raise NotImplementedError(f'DIA sweeps support at most {MAXOFF} diagonals per operator')
        if len(offs) == 0:  # an all-zero operator: one zero diagonal keeps the kernel loops uniform
            offs, d = [0], np.zeros((1, d.shape[-1]), dtype=np.complex128)
        o.noff[k] = len(offs)
        for u, off in enumerate(offs):
            o.off[k][u] = off
        dp = np.zeros((len(offs), n_pad), dtype=np.complex64)
        dp[:, : d.shape[-1]] = d
        o.diags[k] = arena.put_now(dp)


class DiaRhsPlan:
    """A batch of DIA Lindblad vector fields (``dqh_lindblad_rhs_dia``) on
    resident buffers: ``rho`` in, ``drho`` out, both (B, n_pad, n_pad) complex64."""

    def run(self, stream=None):
        import torch

        s = stream if stream is not None else torch.cuda.current_stream()
        _native.check(_native.lib().dqh_lindblad_rhs_dia(ctypes.byref(self.problem), ctypes.c_void_p(s.cuda_stream)), 'dqh_lindblad_rhs_dia')

    def result(self) -> np.ndarray:
        import torch

        torch.cuda.synchronize()
        out = self.drho.cpu().numpy().view(np.complex64).reshape(self.ninst, self.n, self.n)
        return out[:, : self.n_true, : self.n_true]


def rhs_plan(H: np.ndarray, Ls, rho: np.ndarray, variant: int = SWEEP_AUTO) -> DiaRhsPlan:
    """Pack the DIA vector field of (H, Ls) (dense FP64 host arrays whose nonzeros lie
    on few diagonals) for a batch of states ``rho`` (B, n, n)."""
    import torch

    dev = _device()
    _native.lib()
    rho = np.asarray(rho, dtype=np.complex128)
    if rho.ndim == 2:
        rho = rho[None]
    B, n = rho.shape[0], H.shape[-1]
    n_pad = NB * ((n + NB - 1) // NB)
    LdL = sum((L.conj().T @ L for L in Ls), np.zeros_like(H))
    Hnh = -1j * H - 0.5 * LdL
    arena = _Arena(dev)
    prob = DiaRhs()
    prob.n, prob.ninst, prob.variant = n_pad, B, variant
    _fill_ops(prob.ops, [_dia_of(Hnh)] + [_dia_of(L) for L in Ls], n_pad, arena)
    r = np.zeros((B, n_pad, n_pad), dtype=np.complex64)
    r[:, :n, :n] = rho
    rho_d = torch.from_numpy(r.view(np.float32)).to(dev)
    drho = torch.empty_like(rho_d)
    prob.rho, prob.drho = rho_d.data_ptr(), drho.data_ptr()
    plan = DiaRhsPlan()
    plan.problem, plan.arena, plan.rho, plan.drho = prob, arena, rho_d, drho
    plan.ninst, plan.n, plan.n_true = B, n_pad, n
    # roofline unit: one vector field reads rho and writes drho (2 n^2 complex64) per instance
    plan.bytes_per_call = 2.0 * n * n * 8 * B
    return plan


class DiaTsit5(ctypes.Structure):
    _fields_ = [
        ('n', ctypes.c_int64),
        ('n_true', ctypes.c_int64),
        ('ninst', ctypes.c_int64),
        ('nsave', ctypes.c_int64),
        ('max_steps', ctypes.c_int64),
        ('nE', ctypes.c_int64),
        ('variant', ctypes.c_int64),
        ('graph_steps', ctypes.c_int64),
        ('rtol', ctypes.c_double),
        ('atol', ctypes.c_double),
        ('safety', ctypes.c_double),
        ('factormin', ctypes.c_double),
        ('factormax', ctypes.c_double),
        ('t0', ctypes.c_double),
        ('t1', ctypes.c_double),
        ('flags', ctypes.c_int64),
        ('ops', DiaOps),
        ('y', ctypes.c_uint64),
        ('work', ctypes.c_uint64),
        ('tsave', ctypes.c_uint64),
        ('states', ctypes.c_uint64),
        ('EopsT', ctypes.c_uint64),
        ('expects', ctypes.c_uint64),
        ('stats', ctypes.c_uint64),
        ('host_flags', ctypes.c_uint64),
        ('npwc', ctypes.c_int64),
        ('nseg', ctypes.c_int64),
        ('nmod', ctypes.c_int64),
        ('Hterms', ctypes.c_uint64),
        ('pwc_times', ctypes.c_uint64),
        ('pwc_vals', ctypes.c_uint64),
        ('mod_coef', ctypes.c_uint64),
        ('mod_t0', ctypes.c_double),
        ('mod_dt', ctypes.c_double),
        ('mod_nint', ctypes.c_int64),
        ('hcur', ctypes.c_uint64),
    ]


def _diags_on(x: np.ndarray, offs) -> np.ndarray:
    """The diagonals of ``x`` at the offsets ``offs`` in the column-indexed DIA layout
    (``diags[k, j] = x[j - offs[k], j]``, zero outside the matrix)."""
    n = x.shape[-1]
    d = np.zeros((len(offs), n), dtype=np.complex128)
    for i, k in enumerate(offs):
        if k >= 0:
            d[i, k:] = np.diagonal(x, offset=k)
        else:
            d[i, : n + k] = np.diagonal(x, offset=k)
    return d


class DiaTsit5Plan:
    """A batch of adaptive Tsit5 solves on the DIA sweeps (``dqh_mesolve_tsit5_dia``)."""

    def run(self, stream=None):
        import torch

        s = stream if stream is not None else torch.cuda.current_stream()
        _native.check(_native.lib().dqh_mesolve_tsit5_dia(ctypes.byref(self.problem), ctypes.c_void_p(s.cuda_stream)), 'dqh_mesolve_tsit5_dia')

    def reset(self):
        self.y.copy_(self.y0)

    def results(self) -> dict:
        import torch

        torch.cuda.synchronize()
        st = self.states.cpu().numpy().view(np.complex64).reshape(self.ninst, self.nsave, self.n_true, self.n_true)
        fin = self.y.cpu().numpy().view(np.complex128).reshape(self.ninst, self.n, self.n)[:, : self.n_true, : self.n_true]
        stats = self.stats.cpu().numpy()
        out = {
            'states': st,
            'final_state': fin,
            'nsteps': stats[:, 0],
            'naccepted': stats[:, 1],
            'nrejected': stats[:, 2],
            'status': stats[:, 3],
            'nrhs': stats[:, 4],
        }
        if self.nE:
            out['expects'] = self.expects.cpu().numpy().view(np.complex64).reshape(self.ninst, self.nE, self.nsave)
        return out


def _tableau_flag(tableau: str) -> int:
    if tableau not in ('tsit5', 'dopri5'):
        raise ValueError(f'unknown Runge-Kutta pair {tableau!r}')
    return _native.FLAG_DOPRI5 if tableau == 'dopri5' else 0


def tsit5_plan(
    H: np.ndarray,
    Ls,
    rho0: np.ndarray,
    tsave: np.ndarray,
    *,
    t0: float | None = None,
    rtol: float = 1e-6,
    atol: float = 1e-6,
    safety: float = 0.9,
    factormin: float = 0.2,
    factormax: float = 5.0,
    max_steps: int = 100_000,
    tableau: str = 'tsit5',
    exp_ops=None,
    variant: int = SWEEP_AUTO,
    graph_steps: int = 16,
    pwc=None,
    mod=None,
) -> DiaTsit5Plan:
    """Pack a batch of Tsit5 solves that share H and Ls (dense FP64 host arrays whose
    nonzeros lie on few diagonals) and differ in rho0 (B, n, n). ``tableau='dopri5'``
    integrates with the Dopri5 pair instead. Time-dependent Hamiltonian terms, shared
    by the batch: ``pwc`` [(times, values (nv,), O)] (``dq.pwc``) and ``mod`` [(f, O)]
    (``dq.modulated``, f(t) scalar), with O on few diagonals: H(t) = H + sum c_j(t) O_j."""
    import torch

    dev = _device()
    lib = _native.lib()
    tsave = np.asarray(tsave, dtype=np.float64)
    rho0 = np.asarray(rho0, dtype=np.complex128)
    if rho0.ndim == 2:
        rho0 = rho0[None]
    B, n = rho0.shape[0], H.shape[-1]
    n_pad = NB * ((n + NB - 1) // NB)
    LdL = sum((L.conj().T @ L for L in Ls), np.zeros_like(H))
    Hnh = -1j * H - 0.5 * LdL
    if variant == SWEEP_AUTO:
        from .. import autotune

        # the sweep form must admit every offset of the time-dependent terms too
        Hpat = H + sum((np.abs(np.asarray(O)) for O in [o for (_, _, o) in (pwc or [])] + [o for (_, o) in (mod or [])]),
                       np.zeros(H.shape))
        variant = autotune.dia_sweep(Hpat, Ls, B)
    arena = _Arena(dev)
    prob = DiaTsit5()
    prob.n, prob.n_true, prob.ninst, prob.nsave = n_pad, n, B, len(tsave)
    prob.max_steps, prob.variant, prob.graph_steps = int(max_steps), variant, int(graph_steps)
    prob.rtol, prob.atol, prob.safety = float(rtol), float(atol), float(safety)
    prob.factormin, prob.factormax = float(factormin), float(factormax)
    prob.t0 = float(tsave[0] if t0 is None else t0)
    prob.t1 = float(tsave[-1])
    prob.flags = _tableau_flag(tableau)
    pwc, mod = list(pwc or []), list(mod or [])
    terms = [O for (_, _, O) in pwc] + [O for (_, O) in mod]
    if terms:
        # operator 0 on the union of the offsets of Hnh and of every term
        offs0 = sorted(set(_dia_of(Hnh)[0]).union(*[set(_dia_of(O)[0]) for O in terms])) or [0]
        if len(offs0) > MAXOFF:
            raise NotImplementedError(f'DIA sweeps support at most {MAXOFF} diagonals per operator')
        _fill_ops(prob.ops, [(offs0, _diags_on(Hnh, offs0))] + [_dia_of(L) for L in Ls], n_pad, arena)
        Ht = np.zeros((len(terms), len(offs0), n_pad), dtype=np.complex64)
        for j, O in enumerate(terms):
            Ht[j, :, :n] = _diags_on(-1j * np.asarray(O, dtype=np.complex128), offs0)
        prob.Hterms = arena.put_now(Ht)
        prob.hcur = arena.put_now(np.zeros((B, len(offs0), n_pad), dtype=np.complex64))
        prob.npwc, prob.nmod = len(pwc), len(mod)
        if pwc:
            from .dense import pwc_rows

            grid, rows = pwc_rows([(t, v, None) for (t, v, _) in pwc], [])
            prob.nseg = len(grid)
            prob.pwc_times = arena.put_now(grid)
            prob.pwc_vals = arena.put_now(np.ascontiguousarray(np.stack(rows)).view(np.float64))
        if mod:
            from .operators import spline_table

            tabs = [spline_table(f, prob.t0, prob.t1) for (f, _) in mod]
            nint = max(len(c) for (_, _, c) in tabs)
            tabs = [spline_table(f, prob.t0, prob.t1, n0=nint, nmax=nint) if len(tab[2]) != nint else tab
                    for (f, _), tab in zip(mod, tabs)]
            prob.mod_t0, prob.mod_dt, prob.mod_nint = tabs[0][0], tabs[0][1], nint
            prob.mod_coef = arena.put_now(np.ascontiguousarray(np.stack([c for (_, _, c) in tabs])).view(np.float64))
    else:
        _fill_ops(prob.ops, [_dia_of(Hnh)] + [_dia_of(L) for L in Ls], n_pad, arena)
    r64 = np.zeros((B, n_pad, n_pad), dtype=np.complex128)
    r64[:, :n, :n] = rho0
    y0 = torch.from_numpy(r64.view(np.float64)).to(dev)
    y = y0.clone()
    work = torch.empty((int(lib.dqh_dia_tsit5_work_bytes(n_pad, B)),), dtype=torch.uint8, device=dev)
    ts_d = torch.from_numpy(tsave.copy()).to(dev)
    states = torch.zeros((B * len(tsave) * n * n * 2,), dtype=torch.float32, device=dev)
    stats = torch.zeros((B, 8), dtype=torch.int64, device=dev)
    flags = torch.zeros((2,), dtype=torch.int64, pin_memory=True)
    nE = 0 if exp_ops is None else len(exp_ops)
    expects = torch.zeros((B * max(nE, 1) * len(tsave) * 2,), dtype=torch.float32, device=dev)
    prob.nE = nE
    if nE:
        ET = np.stack([np.asarray(E, dtype=np.complex128).T for E in exp_ops]).astype(np.complex64)
        prob.EopsT = arena.put_now(ET)
        prob.expects = expects.data_ptr()
    prob.y, prob.work, prob.tsave = y.data_ptr(), work.data_ptr(), ts_d.data_ptr()
    prob.states, prob.stats, prob.host_flags = states.data_ptr(), stats.data_ptr(), flags.data_ptr()
    plan = DiaTsit5Plan()
    plan.problem, plan.arena = prob, arena
    plan.y, plan.y0, plan.work, plan.tsave_d, plan.states, plan.stats, plan.flags = y, y0, work, ts_d, states, stats, flags
    plan.expects, plan.nE = expects, nE
    plan.ninst, plan.n, plan.n_true, plan.nsave = B, n_pad, n, len(tsave)
    # roofline unit: one vector field reads rho and writes drho (2 n^2 complex64) per instance
    plan.bytes_per_rhs = 2.0 * n * n * 8
    return plan
