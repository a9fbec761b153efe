"""Packing of dense master-equation instances into ``dqh_me_inst`` descriptors and
launch of the HIP kernels (``dqh_mesolve_tsit5_dense``,
``dqh_mesolve_rouchon1_dense``, ``dqh_lindblad_rhs_dense``).

Every instance of a batch gets its own descriptor; operators shared between
instances (e.g. un-batched jump operators) are stored once on the device and
referenced by every descriptor. Matrices are zero-padded to ``np = 16 *
ceil(n / 16)`` so that MFMA tiles never need bounds checks.
"""

from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np
import scipy.linalg as sla

from .. import _native
from .operators import spline_table

MAX_NP = 1024


def pad16(n: int) -> int:
    return 16 * ((n + 15) // 16)


def _pad(x: np.ndarray, np_: int) -> np.ndarray:
    n = x.shape[-1]
    if n == np_:
        return x
    out = np.zeros((*x.shape[:-2], np_, np_), dtype=x.dtype)
    out[..., :n, :n] = x
    return out


def _dag(x: np.ndarray) -> np.ndarray:
    return np.swapaxes(x.conj(), -1, -2)


@dataclass
class Instance:
    """One solve: constant part of H, PWC terms of H, jump operators, initial state."""

    H: np.ndarray  # (n, n) complex128
    Ls: list  # [(n, n) complex128]
    rho0: np.ndarray  # (n, n) complex128
    pwc: list = field(default_factory=list)  # [(times, values (nv,), op (n, n))]
    key_H: object = None  # sharing keys: equal keys share device copies
    key_L: object = None
    key_rho: object = None
    Hnh: np.ndarray | None = None  # explicit constant generator part (overrides -iH - 1/2 sum L^dag L)
    tsave: np.ndarray | None = None  # per-instance save times (all instances of a plan: same length)
    t0: float | None = None  # per-instance start time
    mod: list = field(default_factory=list)  # [(g, op (n, n), key)]: + g(t) op, g scalar; equal keys share tables
    Lmod: list = field(default_factory=list)  # per jump operator: None (constant) or (g, key): L_k(t) = g(t) Ls[k]
    Lpwc: list = field(default_factory=list)  # per jump operator: None or (times, values): L_k(t) = c(t) Ls[k], c PWC
    # per jump operator: None or a sum of terms [(coef, op)], coef ('const', v) | ('mod', g, key) |
    # ('pwc', times, values); expand_jump_sums turns such instances into D/C pairs
    Lsum: list = field(default_factory=list)
    Cs: list | None = None  # explicit right factors C_k (default 1/2 L_k^dag)
    Lbmod: list = field(default_factory=list)  # per D/C pair: the coefficient of C when it differs from D's
    Lbpwc: list = field(default_factory=list)
    expanded: bool = False  # the generator terms of the jump operators are in H / mod / pwc already


class _Arena:
    """Device buffers of a plan. ``put`` stages host arrays into one pinned
    buffer that ``upload`` copies to the device in a single transfer; it returns
    ``offset + 1`` (0 stays the null pointer) and ``patch`` turns such values
    into device addresses. ``put_now`` uploads immediately (one tensor each)."""

    def __init__(self, device):
        self.device = device
        self.tensors = []
        self._chunks = []
        self._size = 0
        self.base = None

    def put(self, arr: np.ndarray) -> int:
        a = np.ascontiguousarray(arr)
        off = (self._size + 255) // 256 * 256
        self._chunks.append((off, a.view(np.uint8).reshape(-1)))
        self._size = off + a.nbytes
        return off + 1

    def upload(self) -> int:
        import torch

        host = torch.empty(max(self._size, 16), dtype=torch.uint8, pin_memory=True)
        hn = host.numpy()
        for off, b in self._chunks:
            hn[off : off + b.size] = b
        dev = host.to(self.device, non_blocking=True)
        self.tensors.append(dev)
        self.tensors.append(host)
        self._chunks = []
        self.base = dev.data_ptr()
        return self.base

    def patch(self, v: int) -> int:
        return 0 if v == 0 else self.base + int(v) - 1

    def put_now(self, arr: np.ndarray) -> int:
        import torch

        a = np.ascontiguousarray(arr)
        if a.dtype == np.complex64:
            t = torch.from_numpy(a.view(np.float32)).to(self.device)
        elif a.dtype == np.complex128:
            t = torch.from_numpy(a.view(np.float64)).to(self.device)
        else:
            t = torch.from_numpy(a).to(self.device)
        self.tensors.append(t)
        return t.data_ptr()

    def empty(self, nbytes: int, zero: bool = False):
        import torch

        t = (torch.zeros if zero else torch.empty)(max(int(nbytes), 16), dtype=torch.uint8, device=self.device)
        self.tensors.append(t)
        return t


ARENA_FIELDS = ('tsave', 'Hnh', 'Dops', 'Cops', 'EopsT', 'Hpwc', 'pwc_times', 'pwc_vals', 'step_t', 'step_set',
                'Hmod', 'mod_coef', 'hdiag')
POINTER_FIELDS = ARENA_FIELDS + ('y', 'work', 'states', 'expects', 'final_state', 'stats', 'step_log')


def check_descriptors(descs: np.ndarray, tensors, extents: dict | None = None) -> None:
    """Host-side guard before a launch: every non-null pointer of every descriptor
    must lie inside one of the device allocations of the plan, and the buffers
    whose extent the descriptor fixes must end inside the same allocation.
    ``extents`` replaces the byte extents of the named fields (per descriptor) for
    kernels that read the descriptor with another layout."""
    spans = sorted((t.data_ptr(), t.data_ptr() + t.numel() * t.element_size()) for t in tensors)
    lo = np.array([s[0] for s in spans], dtype=np.uint64)
    hi = np.array([s[1] for s in spans], dtype=np.uint64)
    npd = descs['np'].astype(np.uint64)
    n = descs['n'].astype(np.uint64)
    nt = descs['ntsave'].astype(np.uint64)
    nE = descs['nE'].astype(np.uint64)
    one = np.uint64(1)
    extent = {
        'Hnh': npd * npd * np.uint64(8),
        'EopsT': nE * npd * npd * np.uint64(8),
        'y': np.uint64(2) * npd * npd * np.uint64(16),
        'tsave': nt * np.uint64(8),
        'final_state': n * n * np.uint64(8),
        'expects': np.maximum(nE, one) * nt * np.uint64(8),
        'stats': np.full(len(descs), 64, dtype=np.uint64),
        'step_log': descs['step_log_cap'].astype(np.uint64) * np.uint64(8),
        'Hmod': descs['nmod'].astype(np.uint64) * npd * npd * np.uint64(8),
        'mod_coef': descs['nmod'].astype(np.uint64) * descs['mod_nint'].astype(np.uint64) * np.uint64(64),
        'hdiag': npd * np.uint64(16),
    }
    extent.update(extents or {})
    # every (field, descriptor) pair in one vectorised pass
    B = len(descs)
    ones = np.ones(B, dtype=np.uint64)
    v = np.concatenate([descs[f].astype(np.uint64) for f in POINTER_FIELDS])
    end = v + np.concatenate([np.broadcast_to(np.asarray(extent.get(f, ones), dtype=np.uint64), (B,))
                              for f in POINTER_FIELDS])
    idx = np.clip(np.searchsorted(lo, v, side='right') - 1, 0, len(hi) - 1)
    bad = (v != 0) & ((v < lo[idx]) | (end > hi[idx]))
    if bad.any():
        j = int(np.flatnonzero(bad)[0])
        raise RuntimeError(f'descriptor field {POINTER_FIELDS[j // B]} = {int(v[j]):#x} is outside the plan allocations')


def _device():
    import torch

    if not torch.cuda.is_available():
        raise RuntimeError('dynamiqs_amd needs an MI355X GPU (torch.cuda.is_available() is False)')
    return torch.device('cuda', torch.cuda.current_device())


class DensePlan:
    """A packed batch of instances of one kind, ready to launch (inputs resident in HBM)."""

    def __init__(self, kind: str, n: int, ntsave: int, nE: int, ninst: int, save_states: bool):
        self.kind = kind
        self.n = n
        self.np = pad16(n)
        self.ntsave = ntsave
        self.nE = nE
        self.ninst = ninst
        self.save_states = save_states
        self.arena = None
        self.descs = None
        self.d_descs = None
        self.states = None
        self.expects = None
        self.final = None
        self.stats = None
        self.y = None
        self.y0 = None
        self.nrhs_flops = 0.0
        self.grid = None  # (scratch, pinned flags, graph steps) for the grid-wide Tsit5
        self.uniform_np = 0  # every instance takes the resident Tsit5 body of this np
        self.fp64_diag = False  # ... with the generator's diagonal in FP64 (hdiag)
        self.step_log = None

    # ------------------------------------------------------------------ launch
    def reset(self):
        """Restore the initial states so that the plan can be run again."""
        self.y[:, 0].copy_(self.y0)

    def run(self, stream=None):
        import torch

        lib = _native.lib()
        s = stream if stream is not None else torch.cuda.current_stream()
        sp = _native.ctypes.c_void_p(s.cuda_stream)
        if self.grid is not None:
            c_p = _native.ctypes.c_void_p
            scratch, flags, steps = self.grid
            _native.check(
                lib.dqh_mesolve_tsit5_dense_grid(
                    c_p(self.d_descs.data_ptr()), self.descs.ctypes.data_as(c_p), self.ninst, c_p(scratch.data_ptr()),
                    scratch.numel(), steps, c_p(flags.data_ptr()), sp,
                ),
                'dqh_mesolve_tsit5_dense_grid',
            )
            return
        if self.uniform_np:
            fn = lib.dqh_mesolve_tsit5_dense_uniform_fp64diag if self.fp64_diag else lib.dqh_mesolve_tsit5_dense_uniform
            _native.check(fn(_native.ctypes.c_void_p(self.d_descs.data_ptr()), self.ninst, self.uniform_np, sp),
                          'dqh_mesolve_tsit5_dense_uniform')
            return
        fn = {
            'tsit5': lib.dqh_mesolve_tsit5_dense,
            'rouchon1': lib.dqh_mesolve_rouchon1_dense,
            'rhs': lib.dqh_lindblad_rhs_dense,
        }[self.kind]
        _native.check(fn(_native.ctypes.c_void_p(self.d_descs.data_ptr()), self.ninst, sp), f'dqh {self.kind}')

    def results(self) -> dict:
        import torch

        torch.cuda.synchronize()
        n = self.n
        out = {}
        c64 = lambda t: t.cpu().numpy().view(np.complex64)  # noqa: E731
        if self.states is not None:
            shp = (self.ninst, self.ntsave, n, n) if self.save_states else (self.ninst, 1, n, n)
            out['states'] = c64(self.states).reshape(shp)
        out['final_state'] = c64(self.final).reshape(self.ninst, n, n)
        if self.nE:
            out['expects'] = c64(self.expects).reshape(self.ninst, self.nE, self.ntsave)
        st = self.stats.cpu().numpy()
        out['nsteps'], out['naccepted'], out['nrejected'] = st[:, 0], st[:, 1], st[:, 2]
        if getattr(self, 'step_log', None) is not None:
            log = self.step_log.cpu().numpy()
            out['step_ends'] = [log[b, : min(int(st[b, 1]), log.shape[1])] for b in range(self.ninst)]
        out['status'], out['nrhs'] = st[:, 3], st[:, 4]
        return out


def kraus_operators(H: np.ndarray, Ls, h: float, order: int = 1) -> list:
    """The Kraus operators M_k of one Rouchon step, the first one the no-jump map.

    order 1 (``KrausEuler``, rouchon_integrator.py:250-263): M0 = I - i h Hnh and
    sqrt(h) L_k. order 2 (``KrausHeun2`` :272-284 with the midpoint no-jump
    propagator of :441-443): U = I + h G (I + h/2 G), G = -i H - 1/2 sum L^dag L,
    and the map U rho U^dag + h/2 U D(rho) U^dag + h/2 D(U rho U^dag)
    + h^2/2 D(U D(rho) U^dag) of the Heun tableau, whose Kraus operators are U,
    sqrt(h/2) U L_k, sqrt(h/2) L_k U and h/sqrt(2) L_k U L_j."""
    n = H.shape[-1]
    LdL = sum((_dag(L) @ L for L in Ls), np.zeros_like(H))
    if order == 1:
        M0 = np.eye(n) - 1j * h * (H - 0.5j * LdL)
        return [M0, *[np.sqrt(h) * L for L in Ls]]
    if order == 2:
        G = -1j * H - 0.5 * LdL
        U = np.eye(n) + h * G @ (np.eye(n) + 0.5 * h * G)
        out = [U]
        out += [np.sqrt(0.5 * h) * U @ L for L in Ls]
        out += [np.sqrt(0.5 * h) * L @ U for L in Ls]
        out += [h / np.sqrt(2.0) * Lk @ U @ Lj for Lk in Ls for Lj in Ls]
        return out
    if order == 3:
        # KrausHeun3 (:288-302, c = (0, 1/3, 2/3), a10 = 1/3, a21 = 2/3,
        # b = (1/4, 0, 3/4)) with the Bosh3 no-jump propagator (:445-453): U1 after
        # one Bogacki-Shampine step, U(1/3), U(2/3) on its cubic Hermite dense output.
        # The stage recursion expands to the Kraus operators U1, sqrt(h/4) U1 L_k,
        # sqrt(3h/4) Q2 L_k Ub, h/sqrt(2) Q2 L_k P L_j Ua and
        # sqrt(h^3/6) Q2 L_k P L_j Ua L_i, with P = Ub Ua^-1 and Q2 = U1 Ub^-1.
        eye = np.eye(n)
        G = -1j * H - 0.5 * LdL
        k1 = G
        k2 = G @ (eye + 0.5 * h * k1)
        k3 = G @ (eye + 0.75 * h * k2)
        U1 = eye + h * (2.0 / 9.0 * k1 + 1.0 / 3.0 * k2 + 4.0 / 9.0 * k3)
        d0, d1 = h * G, h * G @ U1

        def herm(th):
            t2, t3 = th * th, th * th * th
            return (2 * t3 - 3 * t2 + 1) * eye + (t3 - 2 * t2 + th) * d0 + (-2 * t3 + 3 * t2) * U1 + (t3 - t2) * d1

        Ua, Ub = herm(1.0 / 3.0), herm(2.0 / 3.0)
        P = np.linalg.solve(Ua.T, Ub.T).T
        Q2 = np.linalg.solve(Ub.T, U1.T).T
        out = [U1]
        out += [np.sqrt(0.25 * h) * U1 @ L for L in Ls]
        out += [np.sqrt(0.75 * h) * Q2 @ L @ Ub for L in Ls]
        out += [h / np.sqrt(2.0) * Q2 @ Lk @ P @ Lj @ Ua for Lk in Ls for Lj in Ls]
        out += [np.sqrt(h**3 / 6.0) * Q2 @ Lk @ P @ Lj @ Ua @ Li for Lk in Ls for Lj in Ls for Li in Ls]
        return out
    raise ValueError(f'Rouchon order {order} is not supported')


def _kraus_sets(H: np.ndarray, Ls, h: float, normalize: bool, order: int = 1):
    """Rouchon operators with the Cholesky normalisation folded in
    (rouchon_integrator.py:394-407, 552-582): S = sum_k M_k^dag M_k = T T^dag and
    A_k = M_k T^-dag. One step adds tmp + tmp^dag with
    tmp = dA rho (I + dA^dag / 2) + sum_{k>0} A_k rho A_k^dag / 2, dA = A_0 - I,
    i.e. D = [dA, A_k], C = [I + dA^dag / 2, A_k^dag / 2]."""
    n = H.shape[-1]
    if order == 0:
        # Euler on the Lindblad equation (diffrax Euler in the reference's Euler
        # integrator): rho' = rho + h (G rho + rho G^dag + sum L rho L^dag), G = -i H -
        # 1/2 sum L^dag L, i.e. tmp = h G rho + sum (sqrt(h) L) rho (sqrt(h) L)^dag / 2
        LdL = sum((_dag(L) @ L for L in Ls), np.zeros_like(H))
        G = -1j * H - 0.5 * LdL
        return [h * G, *[np.sqrt(h) * L for L in Ls]], [np.eye(n), *[0.5 * np.sqrt(h) * _dag(L) for L in Ls]]
    Mall = kraus_operators(H, Ls, h, order)
    M0, Ms = Mall[0], Mall[1:]
    if normalize:
        S = sum((_dag(M) @ M for M in Mall), np.zeros_like(H))
        T = np.linalg.cholesky(S)
        Tinv_dag = _dag(sla.solve_triangular(T, np.eye(n), lower=True))
        M0 = M0 @ Tinv_dag
        Ms = [M @ Tinv_dag for M in Ms]
    dA = M0 - np.eye(n)
    D = [dA, *Ms]
    C = [np.eye(n) + 0.5 * _dag(dA), *[0.5 * _dag(M) for M in Ms]]
    return D, C


# device bytes the per-step Kraus sets (D and C) of one batch may take
KRAUS_SET_BUDGET = 2 << 30


def rouchon_step_times(t0: float, t1: float, dt: float) -> np.ndarray:
    """Boundaries ``t0 + k dt`` with the last step clipped to ``t1`` (diffrax
    ``ConstantStepSize`` with ``dt0 = method.dt``)."""
    nf = (t1 - t0) / dt
    nr = int(round(nf))
    if abs(nr - nf) <= 1e-9 * max(1.0, nf):
        nsteps = max(nr, 1)
    else:
        nsteps = int(np.ceil(nf))
    ts = t0 + dt * np.arange(nsteps + 1, dtype=np.float64)
    ts[-1] = t1
    return ts


def _abs2(g):
    """|g(t)|^2 as a function (the rate of a modulated jump operator)."""

    def h(t, g=g):
        return np.abs(np.asarray(g(t))) ** 2

    h.table_tol = getattr(g, 'table_tol', 0.0)
    return h


def _one(t):
    return np.ones(np.shape(t)) + 0j


def _coef_parts(c):
    """(constant factor, modulated (g, key) or None, PWC (times, values) or None) of a term coefficient."""
    if c[0] == 'const':
        return complex(c[1]), None, None
    if c[0] == 'mod':
        return 1.0, (c[1], c[2]), None
    return 1.0, None, (c[1], c[2])


def coef_value(c, t: float) -> complex:
    """A term coefficient at time t (PWC: the reference's prefactor rule)."""
    if c[0] == 'const':
        return complex(c[1])
    if c[0] == 'mod':
        return complex(c[1](t))
    return pwc_value(c[1], c[2], t)


def expand_jump_sums(inst: Instance) -> Instance:
    """Rewrite an instance whose jump operators are sums L_k(t) = sum_j c_j(t) L_j
    (constant, modulated or PWC terms) as D/C pairs for the kernels' per-stage rebuild
    (DQH_FLAG_LPAIR): (c_j L_j, conj(c_j) L_j^dag / 2) and, for j < l,
    (c_j L_j, conj(c_l) L_l^dag), which the vector field's tmp + tmp^dag completes to
    L rho L^dag. The products -1/2 conj(c_j) c_l L_j^dag L_l go into the constant
    generator, the modulated terms (product functions) or the PWC terms (products on
    the merged breakpoints). A sum that mixes modulated and PWC terms raises."""
    from dataclasses import replace

    if inst.Hnh is not None:
        raise NotImplementedError('sums of jump-operator terms with an explicit generator are not supported')
    K = len(inst.Ls)
    per = []
    for k in range(K):
        if inst.Lsum and inst.Lsum[k] is not None:
            per.append(list(inst.Lsum[k]))
        elif inst.Lmod and inst.Lmod[k] is not None:
            per.append([(('mod',) + tuple(inst.Lmod[k]), inst.Ls[k])])
        elif inst.Lpwc and inst.Lpwc[k] is not None:
            per.append([(('pwc',) + tuple(inst.Lpwc[k]), inst.Ls[k])])
        else:
            per.append([(('const', 1.0), inst.Ls[k])])
    Hc = -1j * np.asarray(inst.H, dtype=np.complex128)
    mod, pwc = list(inst.mod), list(inst.pwc)
    D, C, amod, apwc, bmod, bpwc = [], [], [], [], [], []
    for terms in per:
        parts = [(_coef_parts(c), np.asarray(L, dtype=np.complex128)) for c, L in terms]
        if any(m is not None for (_, m, _), _ in parts) and any(q is not None for (_, _, q), _ in parts):
            raise NotImplementedError('a jump operator that sums modulated and piecewise-constant terms is not supported')
        J = len(parts)
        for j in range(J):
            for l in range(j, J):
                (vj, mj, pj), Lj = parts[j]
                (vl, ml, pl), Ll = parts[l]
                D.append(vj * Lj)
                C.append(np.conj(vl) * (0.5 if j == l else 1.0) * _dag(Ll))
                amod.append(mj)
                apwc.append(pj)
                bmod.append(ml)
                bpwc.append(pl)
        for j in range(J):
            for l in range(J):
                (vj, mj, pj), Lj = parts[j]
                (vl, ml, pl), Ll = parts[l]
                O = np.conj(vj) * vl * (_dag(Lj) @ Ll)
                if mj is None and pj is None and ml is None and pl is None:
                    Hc = Hc - 0.5 * O
                elif pj is None and pl is None:
                    # -1/2 conj(g_j) g_l O as a modulated term: -i op with op = -i/2 O
                    def g(t, mj=mj, ml=ml):
                        a = np.conj(mj[0](t)) if mj is not None else 1.0
                        return a * (ml[0](t) if ml is not None else 1.0)

                    mod.append((g, -0.5j * O, None))
                else:
                    rows_in = [(q[0], np.conj(q[1]) if side == 0 else q[1], None)
                               for side, q in ((0, pj), (1, pl)) if q is not None]
                    grid, rows = pwc_rows(rows_in, [])
                    vals = rows[0] * (rows[1] if len(rows) > 1 else 1.0)
                    pwc.append((grid, vals, -0.5j * O))
    return replace(inst, Ls=D, Cs=C, Hnh=Hc, mod=mod, pwc=pwc, Lmod=amod, Lpwc=apwc, Lbmod=bmod, Lbpwc=bpwc,
                   Lsum=[], expanded=True)


def pwc_value(times, values, t: float) -> complex:
    """values[searchsorted(times, t, 'right') - 1], zero outside [times[0], times[-1])
    (the reference's PWC prefactor, time_qarray.py:649-659)."""
    times = np.asarray(times, dtype=np.float64)
    if not (times[0] <= t < times[-1]):
        return 0j
    return complex(np.asarray(values)[int(np.searchsorted(times, t, side='right')) - 1])


def pwc_rows(terms, extra):
    """Union breakpoint grid of the PWC terms [(times, values, _)] and the times in
    ``extra``, and each term's value on every grid interval (zero outside its own
    [times[0], times[-1])), as complex rows."""
    grid = np.unique(np.concatenate([np.asarray(t, dtype=np.float64) for (t, _, _) in terms]
                                    + [np.asarray(extra, dtype=np.float64)]))
    left = grid[:-1]
    rows = []
    for t, v, _ in terms:
        t = np.asarray(t, dtype=np.float64)
        v = np.asarray(v, dtype=np.complex128)
        idx = np.clip(np.searchsorted(t, left, side='right') - 1, 0, len(v) - 1)
        rows.append(np.where((left >= t[0]) & (left < t[-1]), v[idx], 0.0))
    return grid, rows


def _hmods(inst: Instance) -> list:
    """The modulated Hamiltonian-side terms of an instance: its own, plus for every
    modulated jump operator g_k(t) L_k the term -1/2 |g_k(t)|^2 L_k^dag L_k of
    -1/2 sum L^dag L, written as |g_k|^2 times O = -i/2 L_k^dag L_k (the kernels add
    -i O times the coefficient)."""
    out = list(inst.mod)
    if inst.expanded:
        return out
    for L, lm in zip(inst.Ls, inst.Lmod or []):
        if lm is not None:
            g, key = lm
            out.append((_abs2(g), -0.5j * (_dag(L) @ L), None if key is None else ('abs2', key)))
    return out


def build_plan(
    kind: str,
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
    dt: float | None = None,
    normalize: bool = True,
    save_states: bool = True,
    form: str = 'auto',
    grid_tsit5: bool | str = 'auto',
    graph_steps: int = 16,
    step_log_cap: int = 0,
    tableau: str = 'tsit5',
    rouchon_order: int = 1,
    fp64_diag: bool = True,
) -> DensePlan:
    """Pack ``instances`` (same Hilbert dimension) for ``kind`` in {'tsit5', 'rouchon1', 'rhs'}.

    ``form`` picks the kernel form of the Tsit5 / RHS kernels: 'resident' (stages
    and rho in registers / LDS, n <= 64 and at most 3 jump operators), 'global'
    (stages in global memory) or 'auto' (the autotuner's choice for (n, K)).

    ``grid_tsit5`` runs Tsit5 as grid-wide launches per stage with the controller on the
    device (``dqh_mesolve_tsit5_dense_grid``) instead of one persistent workgroup
    per instance: True, False, or 'auto' (when the batch is too small to fill the
    GPU with one workgroup per instance and n >= GRID_MIN_N).

    ``tableau`` is the embedded pair of the adaptive kinds: 'tsit5' (``Tsit5``),
    'dopri5' (``Dopri5``, ``DQH_FLAG_DOPRI5``), both with the same controller, error
    norm and dense-output saving, or 'dopri8' (``Dopri8``, ``DQH_FLAG_DOPRI8``: the
    Dormand-Prince 8(7) pair on the global-memory form, exponent 1/8, steps landing
    on the save times).

    ``rouchon_order`` selects the fixed-step map of kind 'rouchon1': 1, 2, 3 for
    Rouchon1/2/3 (Kraus operators at the step midpoint, optionally normalised) and
    0 for Euler (the Lindblad vector field at the step start, no normalisation).

    ``fp64_diag`` (Tsit5 instances on the resident body): the diagonal of the
    constant generator part goes to the kernel in FP64 (``hdiag``) and is applied in
    FP64 at every vector-field evaluation, the off-diagonal part stays on FP32 MFMA.
    The diagonal holds the largest single terms (energies and decay rates), so this
    removes their FP32 rounding from every stage (about 3x lower error against the
    FP64 oracle on stiff problems such as the driven Kerr batch)."""
    import torch

    if tableau not in ('tsit5', 'dopri5', 'dopri8'):
        raise ValueError(f'unknown Runge-Kutta pair {tableau!r}')
    dopri8 = tableau == 'dopri8' and kind == 'tsit5'
    if kind in ('tsit5', 'rhs'):
        instances = [expand_jump_sums(i) if any(x is not None for x in (i.Lsum or [])) else i for i in instances]
    lpwc_any = any(lp is not None for inst in instances for lp in list(inst.Lpwc or []) + list(inst.Lbpwc or []))
    lmod_any = (any(lm is not None for inst in instances for lm in (inst.Lmod or [])) or lpwc_any
                or any(inst.expanded for inst in instances))
    if dopri8 and lmod_any:
        raise NotImplementedError('Dopri8 with time-dependent jump operators is not supported.')
    if (any(inst.mod for inst in instances) or lmod_any) and kind != 'rouchon1':
        if max(len(_hmods(inst)) for inst in instances) > _native.MAX_MOD:
            raise NotImplementedError(
                f'at most {_native.MAX_MOD} modulated terms per operator are supported (a modulated jump operator'
                ' counts one for its rate in the effective Hamiltonian)'
            )

    dev = _device()
    _native.lib()
    n = instances[0].H.shape[-1]
    if any(inst.H.shape[-1] != n for inst in instances):
        raise ValueError('all instances of a dense plan must share the Hilbert dimension')
    npad = pad16(n)
    if npad > MAX_NP:
        raise ValueError(f'dense MI355X path supports n <= {MAX_NP} (got n={n}); use DIA operators')
    tsave = np.asarray(tsave, dtype=np.float64)
    B = len(instances)
    Es = [np.asarray(E, dtype=np.complex128) for E in (exp_ops or [])]
    nE = len(Es)
    t_start = float(tsave[0] if t0 is None else t0)
    t_end = float(tsave[-1])
    per_inst_ts = instances[0].tsave is not None
    if per_inst_ts:
        if kind != 'tsit5' or any(inst.tsave is None or len(inst.tsave) != len(tsave) for inst in instances):
            raise ValueError('per-instance save times need kind="tsit5" and equal lengths')
    plan = DensePlan(kind, n, len(tsave), nE, B, save_states)
    arena = _Arena(dev)
    plan.arena = arena
    nn = npad * npad

    K = len(instances[0].Ls)
    if any(len(inst.Ls) != K for inst in instances):
        raise ValueError('all instances of a dense plan must have the same number of jump operators')
    if form not in ('auto', 'resident', 'global'):
        raise ValueError("form must be 'auto', 'resident' or 'global'")
    use_global = form == 'global'
    if form == 'auto' and kind in ('tsit5', 'rhs') and not dopri8:
        from .. import autotune

        use_global = autotune.dense_global(npad, K, B)
    if dopri8:  # the 13-stage pair runs on the global-memory form only
        use_global = True
    descs = np.zeros(B, dtype=_native.ME_INST_DTYPE)

    p_tsave = arena.put(tsave)
    p_E = arena.put(_pad(np.stack([E.T for E in Es]), npad).astype(np.complex64)) if nE else 0

    # ---- operators, shared by key
    cache = {}

    sampling_cache = {}

    def pwc_sampling(pwc):
        """Union breakpoint grid of the PWC terms and, per term, the value index on
        each grid interval (resample_pwc's rule), cached for terms that share their
        times arrays (e.g. a batch of control vectors on one grid)."""
        key = tuple(id(t) for (t, _, _) in pwc) + tuple(len(v) for (_, v, _) in pwc)
        if key not in sampling_cache:
            grid = np.unique(np.concatenate([np.asarray(t, dtype=np.float64) for (t, _, _) in pwc]))
            left = grid[:-1]
            samplers = []
            for t, v, _ in pwc:
                t = np.asarray(t, dtype=np.float64)
                idx = np.clip(np.searchsorted(t, left, side='right') - 1, 0, len(v) - 1)
                samplers.append((idx, (left >= t[0]) & (left < t[-1])))
            sampling_cache[key] = (grid, samplers, [t for (t, _, _) in pwc])  # keep the arrays alive
        grid, samplers, _ = sampling_cache[key]
        return grid, samplers

    def shared(key, build):
        # instances share a device copy only when every component of the key is known
        if key is None or any(k is None for k in key[1:]):
            return arena.put(build())
        if key not in cache:
            cache[key] = arena.put(build())
        return cache[key]

    npwc_max = 0
    kraus_index = {}  # ('kraus', key_H, key_L) -> step key -> set index, for sets already built
    kraus_bytes = [0]  # device bytes of the Kraus sets built for this batch
    nmod_max = max((len(_hmods(inst)) for inst in instances), default=0)
    # the FP64 diagonal goes with the single-size resident kernel (every instance of
    # the plan on it), never with the grid-wide kernels
    npwc_all = max((len(inst.pwc) + sum(lp is not None for lp in (inst.Lpwc or [])) for inst in instances), default=0)
    grid_plan = kind == 'tsit5' and not dopri8 and use_grid(grid_tsit5, n, B, npwc_all + nmod_max, nE, save_states)
    hdiag_plan = (fp64_diag and kind == 'tsit5' and npad in (16, 32, 48, 64) and K <= 3 and not use_global
                  and not lmod_any and not grid_plan)
    t_stop = float(tsave[-1])
    if instances[0].tsave is not None:
        t_stop = max(float(inst.tsave[-1]) for inst in instances)
    t_tab = min([t_start] + [float(inst.t0) for inst in instances if inst.t0 is not None])
    mod_tables = {}

    def mod_table(g, key):
        # one table per (function, batch element), shared by the instances that use it
        k = (key, t_tab, t_stop) if key is not None else None
        if k is None or k not in mod_tables:
            tab = spline_table(g, t_tab, t_stop)
            if k is None:
                return tab
            mod_tables[k] = tab
        return mod_tables[k]

    if kind == 'rouchon1':
        if dt is None:
            raise ValueError('Rouchon1 needs a fixed step dt')
        step_t = rouchon_step_times(t_start, t_end, float(dt))
        nsteps = len(step_t) - 1
        p_step_t = arena.put(step_t)
        hs = np.diff(step_t)
        mids = step_t[:-1] + 0.5 * hs
        nLdev = {0: K + 1, 1: K + 1, 2: 1 + 2 * K + K * K, 3: 1 + 2 * K + K * K + K**3}[rouchon_order]
    else:
        nLdev = K
    # fields common to the batch, as whole columns
    descs['n'], descs['np'], descs['nE'], descs['ntsave'] = n, npad, nE, len(tsave)
    descs['rtol'], descs['atol'], descs['safety'] = rtol, atol, safety
    descs['factormin'], descs['factormax'] = factormin, factormax
    descs['max_steps'] = max_steps
    descs['t0'], descs['t1'] = t_start, t_end
    descs['tsave'] = p_tsave
    for b, inst in enumerate(instances):
        d = descs[b]
        if per_inst_ts:
            ts_i = np.asarray(inst.tsave, dtype=np.float64)
            d['tsave'] = arena.put(ts_i)
            d['t0'] = float(ts_i[0] if inst.t0 is None else inst.t0)
            d['t1'] = float(ts_i[-1])
        d['EopsT'] = p_E
        d['flags'] = _native.FLAG_SAVE_STATES if save_states else 0
        if kind in ('tsit5', 'rhs'):
            Ls = inst.Ls
            d['nL'] = K
            d['flags'] |= _native.FLAG_HAS_HNH
            if tableau == 'dopri5':
                d['flags'] |= _native.FLAG_DOPRI5
            elif tableau == 'dopri8' and kind == 'tsit5':
                d['flags'] |= _native.FLAG_DOPRI8
            lpwc = [k < len(inst.Lpwc or []) and inst.Lpwc[k] is not None for k in range(K)]
            lbpwc = [k < len(inst.Lbpwc or []) and inst.Lbpwc[k] is not None for k in range(K)]
            lmod = [(k < len(inst.Lmod or []) and inst.Lmod[k] is not None) or lpwc[k] or inst.expanded
                    for k in range(K)]
            if use_global or any(lmod):
                d['flags'] |= _native.FLAG_GLOBAL
            if any(lmod):
                d['flags'] |= _native.FLAG_LMOD
            if any(lpwc) or any(lbpwc):
                d['flags'] |= _native.FLAG_LPWC
            if inst.expanded:
                d['flags'] |= _native.FLAG_LPAIR

            def bH128(inst=inst, lmod=lmod):
                if inst.Hnh is not None:
                    return _pad(np.asarray(inst.Hnh, dtype=np.complex128), npad)
                # the time-dependent jump operators' L^dag L terms are in Hmod / Hpwc
                LdL = sum((_dag(L) @ L for k, L in enumerate(inst.Ls) if not lmod[k]), np.zeros_like(inst.H))
                return _pad(-1j * inst.H - 0.5 * LdL, npad)

            hkey = ('Hnh', inst.key_H, inst.key_L, tuple(lmod))
            if hdiag_plan:
                # the resident body adds the diagonal in FP64 (hdiag); Hnh keeps the rest
                def bHoff(bH128=bH128):
                    a = np.array(bH128(), dtype=np.complex128)  # a copy: bH128 may return the caller's Hnh
                    np.fill_diagonal(a, 0.0)
                    return a.astype(np.complex64)

                d['Hnh'] = shared(hkey + ('off',), bHoff)
                d['hdiag'] = shared(hkey + ('diag',), lambda bH128=bH128: np.ascontiguousarray(np.diag(bH128())))
            else:
                d['Hnh'] = shared(hkey, lambda bH128=bH128: bH128().astype(np.complex64))
            if K and inst.Cs is not None:  # D/C pairs of expanded jump-operator sums
                d['Dops'] = shared(('Dpairs', inst.key_L), lambda Ls=Ls: _pad(np.stack(Ls), npad).astype(np.complex64))
                d['Cops'] = shared(('Cpairs', inst.key_L),
                                   lambda Cs=inst.Cs: _pad(np.stack(Cs), npad).astype(np.complex64))
            elif K:
                d['Dops'] = shared(('D', inst.key_L), lambda Ls=Ls: _pad(np.stack(Ls), npad).astype(np.complex64))
                d['Cops'] = shared(
                    ('C', inst.key_L), lambda Ls=Ls: _pad(0.5 * np.stack([_dag(L) for L in Ls]), npad).astype(np.complex64)
                )
            if any(lpwc) or any(lbpwc):
                # PWC jump operators c_k(t) L_k: the terms -1/2 |c_k|^2 L_k^dag L_k join the
                # PWC Hamiltonian terms (O = -i/2 L^dag L, so that -i O = -1/2 L^dag L), and
                # rows npwc + k of pwc_vals hold every jump operator's coefficient (1 for
                # the others) on a grid that covers the whole solve interval; expanded sums
                # carry their generator terms already and add the rows of the C factors
                jterms = [] if inst.expanded else [
                    (inst.Lpwc[k][0], np.abs(np.asarray(inst.Lpwc[k][1], dtype=np.complex128)) ** 2,
                     -0.5j * (_dag(L) @ L)) for k, L in enumerate(inst.Ls) if lpwc[k]]
                pwc_eff = list(inst.pwc) + jterms
                span = np.array([t_tab, t_stop])
                coef = [(inst.Lpwc[k][0], inst.Lpwc[k][1], None) if lpwc[k] else (span, np.ones(1), None)
                        for k in range(K)]
                if inst.expanded:
                    coef += [(inst.Lbpwc[k][0], inst.Lbpwc[k][1], None) if lbpwc[k] else (span, np.ones(1), None)
                             for k in range(K)]
                grid, rows = pwc_rows(pwc_eff + coef, span)
                d['npwc'] = len(pwc_eff)
                d['nseg'] = len(grid)
                npwc_max = max(npwc_max, len(pwc_eff))
                d['Hpwc'] = shared(
                    ('Hpwc', inst.key_H, inst.key_L, tuple(lpwc)),
                    lambda pwc_eff=pwc_eff: _pad(np.stack([-1j * o for (_, _, o) in pwc_eff]), npad).astype(np.complex64),
                )
                d['pwc_times'] = shared(('ptimes', inst.key_H, inst.key_L, tuple(lpwc)), lambda grid=grid: grid)
                d['pwc_vals'] = arena.put(np.ascontiguousarray(np.stack(rows)).view(np.float64))
            elif inst.pwc:
                grid, samplers = pwc_sampling(inst.pwc)
                vals = np.stack([np.where(inside, np.asarray(v, dtype=np.complex128)[idx], 0.0)
                                 for (_, v, _), (idx, inside) in zip(inst.pwc, samplers)])
                d['npwc'] = len(inst.pwc)
                d['nseg'] = len(grid)
                npwc_max = max(npwc_max, len(inst.pwc))
                d['Hpwc'] = shared(
                    ('Hpwc', inst.key_H), lambda inst=inst: _pad(np.stack([-1j * o for (_, _, o) in inst.pwc]), npad).astype(np.complex64)
                )
                d['pwc_times'] = shared(('ptimes', inst.key_H), lambda grid=grid: grid)
                d['pwc_vals'] = arena.put(np.ascontiguousarray(vals).view(np.float64))
            hm = _hmods(inst)
            if hm or any(lmod):
                # tables: the nmod Hamiltonian-side terms, then (DQH_FLAG_LMOD) one per jump
                # operator, g_k or a constant 1, all on one grid
                fns = [(g, key) for (g, _, key) in hm]
                if any(lmod):
                    fns += [inst.Lmod[k] if (k < len(inst.Lmod or []) and inst.Lmod[k] is not None)
                            else (_one, ('one',)) for k in range(K)]
                if inst.expanded:
                    fns += [inst.Lbmod[k] if inst.Lbmod[k] is not None else (_one, ('one',)) for k in range(K)]
                tabs = [mod_table(g, key) for (g, key) in fns]
                nint = max(len(c) for (_, _, c) in tabs)
                tabs = [spline_table(g, t_tab, t_stop, n0=nint, nmax=nint) if len(c) != nint else (a, dtt, c)
                        for (g, _), (a, dtt, c) in zip(fns, tabs)]
                d['nmod'] = len(hm)
                d['mod_t0'], d['mod_dt'], d['mod_nint'] = tabs[0][0], tabs[0][1], nint
                d['mod_coef'] = arena.put(np.ascontiguousarray(np.stack([c for (_, _, c) in tabs])).view(np.float64))
                if hm:
                    d['Hmod'] = shared(
                        ('Hmod', inst.key_H, inst.key_L if any(lmod) else None),
                        lambda hm=hm: _pad(np.stack([-1j * o for (_, o, _) in hm]), npad).astype(np.complex64),
                    )
        else:  # rouchon1: Kraus sets per (PWC segment, step length)
            d['nL'] = nLdev
            grid = None
            if inst.pwc:
                grid = np.unique(np.concatenate([np.asarray(t, dtype=np.float64) for (t, _, _) in inst.pwc]))
            # smooth (modulated) terms make every step's operators distinct: one set per step
            lmods = list(inst.Lmod or [None] * K)
            lpwcs = list(inst.Lpwc or [None] * K)
            lsums = list(inst.Lsum or [None] * K)
            smooth = (bool(inst.mod) or any(lm is not None for lm in lmods) or any(lp is not None for lp in lpwcs)
                      or any(ls is not None for ls in lsums))
            key = ('kraus', inst.key_H, inst.key_L)
            cacheable = not any(k is None for k in key[1:])
            keys = []
            for m in range(nsteps):
                # Rouchon takes the operators at the step midpoint, Euler (order 0) at its start
                tsel = step_t[m] if rouchon_order == 0 else mids[m]
                seg = -1 if grid is None else int(np.searchsorted(grid, tsel, side='right')) - 1
                keys.append((seg, float(hs[m]), float(tsel) if smooth else 0.0))
            if cacheable and key in kraus_index:
                # instances sharing key_H / key_L (a batched rho0, the Hermitian split,
                # propagator basis solves) share the sets already on the device
                index = kraus_index[key]
                d['Dops'], d['Cops'] = cache[key + ('D',)], cache[key + ('C',)]
                d['step_t'] = p_step_t
                d['step_set'] = arena.put(np.asarray([index[k] for k in keys], dtype=np.int32))
                d['max_steps'] = nsteps
                continue
            uniq = sorted(set(keys))
            index = {k: i for i, k in enumerate(uniq)}
            # device bytes of D and C sets together, over every distinct set built so far
            set_bytes = 2 * len(uniq) * nLdev * nn * 8
            if kraus_bytes[0] + set_bytes > KRAUS_SET_BUDGET:
                raise NotImplementedError(
                    f'{len(uniq)} Kraus sets of {nLdev} operators ({(kraus_bytes[0] + set_bytes) / 2**30:.1f} GiB with '
                    f'the sets of the batch built so far) exceed {KRAUS_SET_BUDGET / 2**30:.0f} GiB: modulated '
                    'operators need one set per step; use a larger dt, fewer instances per call or an adaptive method'
                )
            kraus_bytes[0] += set_bytes
            if cacheable:
                kraus_index[key] = index
            sets_D, sets_C = [], []
            for seg, h, tk in uniq:
                tm = None
                Hm = inst.H.copy()
                if inst.pwc:
                    tm = grid[seg] if seg >= 0 else -np.inf
                    for (times, vals, op) in inst.pwc:
                        if tm >= times[0] and tm < times[-1]:
                            idx = int(np.searchsorted(times, tm, side='right')) - 1
                            Hm = Hm + vals[idx] * op
                Lm = inst.Ls
                if smooth:
                    for g, op, _ in inst.mod:
                        Hm = Hm + complex(g(tk)) * op
                    Lm = [L if lm is None else complex(lm[0](tk)) * L for L, lm in zip(inst.Ls, lmods)]
                    Lm = [L if lp is None else pwc_value(lp[0], lp[1], tk) * L for L, lp in zip(Lm, lpwcs)]
                    Lm = [L if ls is None else sum(coef_value(c, tk) * np.asarray(o, dtype=np.complex128) for c, o in ls)
                          for L, ls in zip(Lm, lsums)]
                Dk, Ck = _kraus_sets(Hm, Lm, h, normalize, rouchon_order)
                sets_D.append(np.stack(Dk))
                sets_C.append(np.stack(Ck))
            d['Dops'] = shared(key + ('D',), lambda: _pad(np.stack(sets_D), npad).astype(np.complex64))
            d['Cops'] = shared(key + ('C',), lambda: _pad(np.stack(sets_C), npad).astype(np.complex64))
            d['step_t'] = p_step_t
            d['step_set'] = arena.put(np.asarray([index[k] for k in keys], dtype=np.int32))
            d['max_steps'] = nsteps

    # ---- one host-to-device transfer for every staged operator, then real addresses
    arena.upload()
    for f in ARENA_FIELDS:
        descs[f] = [arena.patch(int(v)) for v in descs[f]]

    # ---- state, scratch, outputs
    welems = _native.work_elems(npad, nLdev, npwc_max + nmod_max)
    if lmod_any:  # per-stage copies of D and C (DQH_FLAG_LMOD)
        welems += 2 * nLdev * nn
    if dopri8:  # stages 7..12 after the generator slot: (16 + nL) np^2
        welems = max(welems, (16 + nLdev) * nn)
    work = arena.empty(B * welems * 8)
    y = torch.zeros((B, 2, nn, 2), dtype=torch.float64, device=dev)
    y0 = np.zeros((B, npad, npad), dtype=np.complex128)
    rho_cache = {}
    for b, inst in enumerate(instances):
        key = inst.key_rho
        y0[b, :n, :n] = inst.rho0
        if key is not None:
            rho_cache[key] = b
    y0_t = torch.from_numpy(y0.reshape(B, nn).view(np.float64).reshape(B, nn, 2)).to(dev)
    y[:, 0].copy_(y0_t)
    plan.y, plan.y0 = y, y0_t
    arena.tensors.append(y)
    ns = len(tsave) if save_states else 1
    states = torch.empty((B, ns * n * n * 2), dtype=torch.float32, device=dev)
    final = torch.empty((B, n * n * 2), dtype=torch.float32, device=dev)
    expects = torch.zeros((B, max(nE, 1) * len(tsave) * 2), dtype=torch.float32, device=dev)
    stats = torch.zeros((B, 8), dtype=torch.int64, device=dev)
    plan.states, plan.final, plan.expects, plan.stats = states, final, expects, stats
    for b in range(B):
        d = descs[b]
        d['y'] = y.data_ptr() + b * 2 * nn * 16
        d['work'] = work.data_ptr() + b * welems * 8
        d['states'] = states.data_ptr() + b * ns * n * n * 8
        d['final_state'] = final.data_ptr() + b * n * n * 8
        d['expects'] = expects.data_ptr() + b * max(nE, 1) * len(tsave) * 8
        d['stats'] = stats.data_ptr() + b * 64
    step_log = None
    if step_log_cap > 0 and kind == 'tsit5':
        step_log = torch.zeros((B, int(step_log_cap)), dtype=torch.float64, device=dev)
        for b in range(B):
            descs[b]['step_log'] = step_log.data_ptr() + b * int(step_log_cap) * 8
            descs[b]['step_log_cap'] = int(step_log_cap)
    plan.step_log = step_log
    check_descriptors(descs, [*arena.tensors, work, states, final, expects, stats] + ([step_log] if step_log is not None else []))
    plan.descs = descs
    plan.d_descs = torch.from_numpy(descs.view(np.uint8).copy()).to(dev)
    plan.K = K
    if (
        kind == 'tsit5'
        and npad in (16, 32, 48, 64)
        and np.all(descs['nL'] <= 3)
        and not np.any(descs['flags'] & _native.FLAG_GLOBAL)
    ):
        plan.uniform_np = npad
        plan.fp64_diag = bool(hdiag_plan)
    elif hdiag_plan:
        raise AssertionError('FP64 diagonal set on a plan that does not run the single-size resident kernel')
    if grid_plan:
        lib = _native.lib()
        nbytes = int(lib.dqh_tsit5_grid_scratch_bytes(npad, B))
        scratch = torch.empty((nbytes,), dtype=torch.uint8, device=dev)
        flags = torch.zeros((2,), dtype=torch.int64, pin_memory=True)
        plan.grid = (scratch, flags, int(graph_steps))
    return plan


# The grid-wide Tsit5 pays off once one instance has enough tiles to spread and the
# batch leaves most CUs idle with one workgroup per instance.
GRID_MIN_N = 128
GRID_MAX_BATCH = 64
# above this dimension one workgroup per instance is far too slow for any batch size
GRID_ALWAYS_N = 256


def use_grid(grid, n: int, B: int, npwc: int, nE: int, save_states: bool) -> bool:
    """Whether a Tsit5 plan runs on the grid-wide kernels (constant generators only;
    expectation values are taken from the saved states)."""
    ok = npwc == 0 and (nE == 0 or save_states)
    if grid == 'auto':
        return ok and n >= GRID_MIN_N and (B <= GRID_MAX_BATCH or n > GRID_ALWAYS_N)
    if grid and not ok:
        raise ValueError('the grid-wide Tsit5 needs constant generators and saved states for expectation values')
    return bool(grid)
