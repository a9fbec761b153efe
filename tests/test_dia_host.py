"""CPU checks of the host precomputation of the DIA Rouchon1 path: a NumPy
emulation of the three kernels (two banded block solves and the Kraus stencil
sweep, increment form) must reproduce the oracle's dense Rouchon1 step
(rouchon_integrator.py:226-269, 552-582)."""

import numpy as np
import pytest

from dynamiqs_amd.integrators import dia
from oracle import operators as op
from oracle import rouchon


def _two_mode(na, nb_):
    a, b = op.destroy(na, nb_)
    ad, bd = a.conj().T, b.conj().T
    H = 5 * ad @ a + 6 * bd @ b + 0.1 * (ad @ b + a @ bd) - 0.01 * ad @ a @ bd @ b
    Ls = [np.sqrt(0.02) * a, np.sqrt(0.05) * b]
    return H, Ls


def _emulate_step(rho, Vp, Wn, ops, n_pad):
    """Increment-form step exactly as k_lsolve (x2) and k_kraus compute it: mode 0
    gives D1 = T^-dag rho - rho; mode 1 solves on (rho + D1)^dag and writes
    E = D1 + d^dag (its output transposed)."""
    NB = dia.NB
    nblk = n_pad // NB

    def lsolve(s):
        d = np.zeros_like(s)
        prev = np.zeros((NB, n_pad), dtype=complex)
        for i in reversed(range(nblk)):
            r = slice(i * NB, (i + 1) * NB)
            di = Vp[i] @ s[r]
            if i + 1 < nblk:
                di = di + Wn[i] @ prev
            d[r] = di
            prev = s[r] + di
        return d

    D1 = lsolve(rho)
    # second solve on Z = (rho + D1)^dag: X = rho + E = (T^-dag Z)^dag = T^-dag rho T^-1
    # for any rho (no Hermitian assumption)
    Z = (rho + D1).conj().T
    E = D1 + lsolve(Z).conj().T
    X = rho + E
    n = n_pad

    def dense(offs, d):
        out = np.zeros((n, n), dtype=complex)
        for u, o in enumerate(offs):
            for c in range(n):
                r = c - o
                if 0 <= r < n:
                    out[r, c] = d[u, c]
        return out

    dM = dense(*ops[0])
    inc = E + dM @ X + X @ dM.conj().T
    for offs, d in ops:
        A = dense(offs, d)
        inc = inc + A @ X @ A.conj().T
    return rho + inc


@pytest.mark.parametrize('na,nb_', [(4, 5), (8, 9), (12, 12)])
def test_block_solve_emulation_matches_oracle_step(na, nb_):
    H, Ls = _two_mode(na, nb_)
    n = na * nb_
    n_pad = dia.NB * ((n + dia.NB - 1) // dia.NB)
    dt = 1e-3
    rng = np.random.default_rng(0)
    psi = rng.normal(size=(n, 1)) + 1j * rng.normal(size=(n, 1))
    rho = psi @ psi.conj().T / np.vdot(psi, psi)
    Vp, Wn, ops = dia.block_solve_operators(H, Ls, dt, n_pad)
    ops_p = []
    for offs, d in ops:
        dp = np.zeros((len(offs), n_pad), dtype=complex)
        dp[:, :n] = d
        ops_p.append((offs, dp))
    rp = np.zeros((n_pad, n_pad), dtype=complex)
    rp[:n, :n] = rho
    got = _emulate_step(rp, Vp, Wn, ops_p, n_pad)
    want = rouchon.kraus_step_dense(H, Ls, rho, dt)
    np.testing.assert_allclose(got[:n, :n], want, atol=1e-12)
    assert np.abs(got[n:, :]).max() < 1e-14 and np.abs(got[:, n:]).max() < 1e-14


def test_c3_operator_structure():
    # the 32 x 32 two-mode config: H has offsets {-31, 0, 31}, L_a {+32}, L_b {+1}, w = 62
    a, b = op.destroy(32, 32)
    ad, bd = a.conj().T, b.conj().T
    H = 5 * ad @ a + 6 * bd @ b + 0.1 * (ad @ b + a @ bd) - 0.01 * ad @ a @ bd @ b
    assert dia._dia_of(H)[0] == [-31, 0, 31]
    assert dia._dia_of(a)[0] == [32] and dia._dia_of(b)[0] == [1]
    Ls = [np.sqrt(0.02) * a, np.sqrt(0.05) * b]
    LdL = sum(L.conj().T @ L for L in Ls)
    M0 = np.eye(1024) - 1e-3j * (H - 0.5j * LdL)
    S = M0.conj().T @ M0 + 1e-3 * LdL
    assert dia.half_bandwidth(S) == 62


def test_save_schedule():
    step_t = np.linspace(0, 1, 11)
    steps, th = dia.save_schedule(np.array([0.0, 0.1, 0.25, 1.0]), step_t)
    assert list(steps) == [-1, 0, 2, 9]
    np.testing.assert_allclose(th, [0.0, 1.0, 0.5, 1.0])


def _emulate_tiled_sweep(X, ops, sand0, n, H):
    """The LDS-tiled sweep as k_rhs_tiled / k_kraus_tiled index it: workgroup id ->
    (XCD remap) -> 64 x 64 tile, (64 + 2H)^2 halo tile zero outside the matrix,
    X(p, q) read at halo position (ii + H + p, jj + H + q), diagonal indices
    clamped to [0, n). Every output element must be written exactly once."""
    TI = 64
    tpr = n // TI
    ntiles = tpr * tpr
    out = np.zeros((n, n), dtype=complex)
    hits = np.zeros((n, n), dtype=int)
    for bid in range(ntiles):
        b = (bid & 7) * (ntiles >> 3) + (bid >> 3) if ntiles % 8 == 0 else bid
        assert 0 <= b < ntiles
        i0, j0 = (b // tpr) * TI, (b % tpr) * TI
        sx = np.zeros((TI + 2 * H, TI + 2 * H), dtype=complex)
        gi = np.arange(i0 - H, i0 + TI + H)
        gj = np.arange(j0 - H, j0 + TI + H)
        mi, mj = (gi >= 0) & (gi < n), (gj >= 0) & (gj < n)
        sx[np.ix_(mi, mj)] = X[np.ix_(gi[mi], gj[mj])]
        ii, jj = np.meshgrid(np.arange(TI), np.arange(TI), indexing='ij')
        i, j = i0 + ii, j0 + jj
        acc = np.zeros((TI, TI), dtype=complex)

        def Xat(p, q):
            return sx[ii + H + p, jj + H + q]

        offs0, D0 = ops[0]
        for u, p in enumerate(offs0):
            acc += D0[u][np.clip(i + p, 0, n - 1)] * Xat(p, 0)
            acc += Xat(0, p) * np.conj(D0[u][np.clip(j + p, 0, n - 1)])
        for s in range(0 if sand0 else 1, len(ops)):
            offs, A = ops[s]
            for u, p in enumerate(offs):
                z = sum(Xat(p, q) * np.conj(A[v][np.clip(j + q, 0, n - 1)]) for v, q in enumerate(offs))
                acc += A[u][np.clip(i + p, 0, n - 1)] * z
        out[i0 : i0 + TI, j0 : j0 + TI] = acc
        hits[i0 : i0 + TI, j0 : j0 + TI] += 1
    assert np.all(hits == 1)
    return out


@pytest.mark.parametrize('na,nb_,halo', [(12, 12, 16), (12, 12, 32), (32, 32, 32)])
def test_tiled_sweep_indexing_reproduces_dense_rhs(na, nb_, halo):
    """The tiled sweep's index arithmetic (n = 192 without and n = 1024 with the XCD
    remap; halos 16 and 32) gives the dense Hermitian-form Lindblad RHS."""
    from oracle import lindblad

    H, Ls = _two_mode(na, nb_)
    n = H.shape[0]
    n_pad = dia.NB * ((n + dia.NB - 1) // dia.NB)
    rng = np.random.default_rng(3)
    x = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
    rho = x @ x.conj().T
    rho /= np.trace(rho).real
    LdL = sum(L.conj().T @ L for L in Ls)
    Hnh = -1j * H - 0.5 * LdL

    def padded(opr):
        offs, d = dia._dia_of(opr)
        dp = np.zeros((len(offs), n_pad), dtype=complex)
        dp[:, :n] = d
        return offs, dp

    assert max(abs(k) for op_ in [Hnh, *Ls] for k in dia._dia_of(op_)[0]) <= halo
    X = np.zeros((n_pad, n_pad), dtype=complex)
    X[:n, :n] = rho
    out = _emulate_tiled_sweep(X, [padded(Hnh)] + [padded(L) for L in Ls], False, n_pad, halo)
    ref = lindblad.rhs_hermitian(H, Ls, rho)
    np.testing.assert_allclose(out[:n, :n], ref, atol=1e-12)
    assert np.abs(out[n:, :]).max(initial=0.0) == 0 and np.abs(out[:, n:]).max(initial=0.0) == 0


def test_hi_lo_state_and_fused_sweep_emulation_matches_oracle():
    """The kernels' arithmetic of a Rouchon1 step, emulated in the precisions they
    use: the state as FP32 planes hi + lo, the block solves on hi (complex64
    outputs), the second solve writing E = D1 + d^dag and X = (hi + D1) + d^dag in
    complex64, the fused sweep's increment E + dM X + X dM^dag + dM X dM^dag +
    sum A X A^dag added to hi + lo in FP64 and split back. Twenty steps stay within
    2e-6 of the oracle's dense FP64 Rouchon1; hi + lo carries the state to ~1e-14."""
    H, Ls = _two_mode(12, 12)
    n = 144
    n_pad = 192
    dt = 1e-3
    NB = dia.NB
    nblk = n_pad // NB
    Vp, Wn, ops = dia.block_solve_operators(H, Ls, dt, n_pad)

    def dense(offs, d):
        out = np.zeros((n_pad, n_pad), dtype=complex)
        for u, o in enumerate(offs):
            c = np.arange(max(o, 0), min(n_pad, n_pad + o))
            out[c - o, c] = d[u, c]
        return out

    mats = []
    for offs, d in ops:
        dp = np.zeros((len(offs), n_pad), dtype=complex)
        dp[:, :n] = d
        mats.append(dense(offs, dp))

    def lsolve(s):
        d = np.zeros_like(s)
        prev = np.zeros((NB, n_pad), dtype=complex)
        for i in reversed(range(nblk)):
            r = slice(i * NB, (i + 1) * NB)
            di = Vp[i].astype(np.complex64) @ s[r] + (Wn[i].astype(np.complex64) @ prev if i + 1 < nblk else 0)
            d[r] = di
            prev = s[r] + di
        return d.astype(np.complex64)

    rng = np.random.default_rng(2)
    psi = rng.normal(size=(n, 1)) + 1j * rng.normal(size=(n, 1))
    rho = np.zeros((n_pad, n_pad), dtype=complex)
    rho[:n, :n] = psi @ psi.conj().T / np.vdot(psi, psi).real
    hi = rho.astype(np.complex64)
    lo = (rho - hi).astype(np.complex64)
    assert np.abs(hi.astype(complex) + lo - rho).max() < 1e-14
    want = rho[:n, :n]
    for _ in range(20):
        D1 = lsolve(hi)
        T1 = (hi + D1).astype(np.complex64)
        d = lsolve(T1.conj().T)
        E = (D1 + d.conj().T).astype(np.complex64)
        X = (T1 + d.conj().T).astype(np.complex64)
        Xc = X.astype(complex)
        dM = mats[0]
        acc = dM @ Xc + Xc @ dM.conj().T + dM @ Xc @ dM.conj().T
        for A in mats[1:]:
            acc = acc + A @ Xc @ A.conj().T
        new = hi.astype(complex) + lo + (E.astype(complex) + acc.astype(np.complex64))
        hi = new.astype(np.complex64)
        lo = (new - hi).astype(np.complex64)
        want = rouchon.kraus_step_dense(H, Ls, want, dt)
    got = hi.astype(complex) + lo
    err = np.linalg.norm(got[:n, :n] - want)
    assert err < 2e-6, err
    assert abs(np.trace(got) - 1) < 1e-6
    assert np.abs(got[n:, :]).max() < 1e-12 and np.abs(got[:, n:]).max() < 1e-12
    # the normalisation is what X carries: forming X without the second solve's
    # correction (X = hi + D1) is off by far more than rounding
    D1 = lsolve(hi)
    bad = (hi + D1).astype(complex)
    good = (hi + D1).astype(np.complex64) + lsolve((hi + D1).astype(np.complex64).conj().T).conj().T
    assert np.abs(bad - good).max() > 1e-4


@pytest.mark.parametrize('n', [64, 100, 192, 256, 1024])
def test_row_blocked_grid_covers_each_element_once(n):
    """diat::row_grid / row_elem: grid.x = n * ceil(n / 256), workgroup b covers
    columns 256 (b mod cpr) + t of row b / cpr; every (i, j) once, none outside."""
    cpr = (n + 255) // 256
    hits = np.zeros((n, n), dtype=int)
    b = np.arange(n * cpr)[:, None]
    t = np.arange(256)[None, :]
    i = b // cpr + 0 * t
    j = (b - (b // cpr) * cpr) * 256 + t
    ok = j < n
    np.add.at(hits, (i[ok], j[ok]), 1)
    assert np.all(hits == 1)
    assert np.all(i < n)


@pytest.mark.parametrize('na', [12, 16])
def test_block_solve_step_is_exact_off_the_hermitian_matrices(na):
    """The kernels' step equals the exact Rouchon1 map on non-Hermitian matrices
    too, so rounding that leaves an anti-Hermitian part in the state is contracted
    by the map, not amplified (a Hermitian-only shortcut, T^-dag rho T^-1 =
    T^-dag (T^-dag rho)^dag, amplifies it by ~1 + 20 dt per step at n = 256)."""
    H, Ls = _two_mode(na, na)
    n = na * na
    n_pad = dia.NB * ((n + dia.NB - 1) // dia.NB)
    dt = 1e-3
    Vp, Wn, ops = dia.block_solve_operators(H, Ls, dt, n_pad)
    ops_p = []
    for offs, d in ops:
        dp = np.zeros((len(offs), n_pad), dtype=complex)
        dp[:, :n] = d
        ops_p.append((offs, dp))
    rng = np.random.default_rng(1)
    x = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
    for rho in (x, x - x.conj().T):
        rp = np.zeros((n_pad, n_pad), dtype=complex)
        rp[:n, :n] = rho
        got = _emulate_step(rp, Vp, Wn, ops_p, n_pad)
        np.testing.assert_allclose(got[:n, :n], rouchon.kraus_step_dense(H, Ls, rho, dt), atol=1e-11)
    # power iteration on anti-Hermitian inputs: no growth
    A = np.zeros((n_pad, n_pad), dtype=complex)
    A[:n, :n] = x - x.conj().T
    A /= np.linalg.norm(A)
    for _ in range(200):
        A = _emulate_step(A, Vp, Wn, ops_p, n_pad)
        assert np.linalg.norm(A) <= 1.0 + 1e-9
        A /= np.linalg.norm(A)


def test_block_normalisation_emulation_matches_cholesky():
    """The block form of the normalisation as the strip kernels compute it
    (k_blk_left: D1[r] = sum_k (T_b^-dag - I)[loc r][k] hi[idx_k]; k_blk_right:
    E = D1 + Y (T_b^-1 - I) on the columns of each set, X = Y + the same sum, Y = hi + D1)
    gives hi + E = T^-dag hi T^-1 with T the dense Cholesky factor of S, on the
    two-mode problem padded to a multiple of 64; its sets are the excitation-number
    sectors (at most 32 indices at 32 x 32)."""
    H, Ls = _two_mode(12, 12)
    n_pad = 192
    S = dia.normalisation_S(H, Ls, 1e-3, n_pad)
    sets = dia.coupled_sets(S)
    assert sets is not None and max(len(x) for x in sets) == 12
    of, loc, start, idx, mat, Lm, Rm = dia.block_normalisation(S, sets)
    rng = np.random.default_rng(7)
    hi = (rng.normal(size=(n_pad, n_pad)) + 1j * rng.normal(size=(n_pad, n_pad))).astype(np.complex64)
    D1 = np.zeros_like(hi)
    for r in range(n_pad):
        b, l = of[r], loc[r]
        s0, d = start[b], start[b + 1] - start[b]
        Lrow = Lm[mat[b] + l * d : mat[b] + (l + 1) * d]
        D1[r] = Lrow[l:] @ hi[idx[s0 + l : s0 + d]]
    Y = hi + D1
    E = np.zeros_like(hi)
    X = np.zeros_like(hi)
    for c in range(n_pad):
        b, l = of[c], loc[c]
        s0, d = start[b], start[b + 1] - start[b]
        Rcol = Rm[mat[b] + l * d : mat[b] + (l + 1) * d]
        acc = Y[:, idx[s0 + l : s0 + d]] @ Rcol[l:]
        E[:, c] = D1[:, c] + acc
        X[:, c] = Y[:, c] + acc
    T = np.linalg.cholesky(S)
    Ti = np.linalg.inv(T)
    want = Ti.conj().T @ hi.astype(complex) @ Ti
    assert np.abs(X - want).max() < 2e-5 * np.abs(want).max()
    assert np.abs((hi.astype(complex) + E) - want).max() < 2e-5 * np.abs(want).max()
    # the C3 operators: 63 sectors of at most 32 indices (plus nothing to pad at 1024)
    a, bb = op.destroy(32, 32)
    ad, bd = a.conj().T, bb.conj().T
    H3 = 5 * ad @ a + 6 * bd @ bb + 0.1 * (ad @ bb + a @ bd) - 0.01 * ad @ a @ bd @ bb
    sets3 = dia.coupled_sets(dia.normalisation_S(H3, [np.sqrt(0.02) * a, np.sqrt(0.05) * bb], 1e-3, 1024))
    assert len(sets3) == 63 and max(len(x) for x in sets3) == 32
