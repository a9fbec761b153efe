"""Rouchon1 (Kraus-Euler), Rouchon2 (Kraus-Heun) and Rouchon3 (Kraus-Heun3) with
Cholesky normalisation, in complex128 (oracle).

Restates ``dynamiqs/integrators/core/rouchon_integrator.py``:

* ``KrausEuler.U`` :250-253: ``Hnh = H(t_mid) - 1/2 i sum L^dag L(t_mid)``,
  ``M0 = I - i dt Hnh`` with ``t_mid = t + dt/2``;
* ``KrausEuler.__call__`` :255-257: ``rho' = M0 rho M0^dag + dt sum L rho L^dag``
  (jump operators at ``t_mid``);
* ``KrausEuler.S`` :259-263: ``S = M0^dag M0 + dt sum L^dag L``;
* ``cholesky_normalize`` :552-582: ``S = T T^dag`` (T lower), then
  ``rho <- T^{-dag} rho T^{-1}`` by two triangular solves;
* ``rouchon_step`` :394-407 (normalise, then apply the Kraus map), fixed steps
  of ``method.dt`` (``ConstantStepSize``, ``diffrax_integrator.py:61-62, 74-76``)
  with linear interpolation to ``tsave`` (``RouchonDXSolver.interpolation_cls =
  LocalLinearInterpolation`` :44).

Two implementations of the same step: :func:`kraus_step_dense` (dense LAPACK
Cholesky / triangular solves, exactly the reference's calls) and
:func:`kraus_step_banded` (banded Cholesky and banded solves for DIA operators).
Pinned in ``tests/test_oracle.py``: the two agree to rounding; trace
preservation to machine precision; complete positivity (normalised Kraus
operators satisfy ``sum A_k^dag A_k = I``); first-order convergence to the exact
Lindblad solution (``expm`` of the superoperator) on the lossy oscillator.
"""

from __future__ import annotations

import numpy as np
import scipy.linalg as sla

from .lindblad import dag


def kraus_ops(H: np.ndarray, Ls, dt: float):
    """Return ``(M0, [sqrt(dt) L_k], S)`` of one Kraus-Euler step."""
    n = H.shape[-1]
    LdL = sum((dag(L) @ L for L in Ls), np.zeros_like(H))
    Hnh = H - 0.5j * LdL
    M0 = np.eye(n) - 1j * dt * Hnh
    S = dag(M0) @ M0 + dt * LdL
    return M0, [np.sqrt(dt) * L for L in Ls], S


def cholesky_normalize(S: np.ndarray, rho: np.ndarray) -> np.ndarray:
    T = np.linalg.cholesky(S)  # lower triangular, S = T T^dag
    x = sla.solve_triangular(T, rho, lower=True, trans='C')  # T^dag x = rho
    # x T = rho'  <=>  T^T x^T = rho'^T
    return sla.solve_triangular(T, x.T, lower=True, trans='T').T


def kraus_step_dense(H: np.ndarray, Ls, rho: np.ndarray, dt: float, normalize: bool = True) -> np.ndarray:
    M0, Ms, S = kraus_ops(H, Ls, dt)
    if normalize:
        rho = cholesky_normalize(S, rho)
    out = M0 @ rho @ dag(M0)
    for M in Ms:
        out = out + M @ rho @ dag(M)
    return out


def normalized_kraus(H: np.ndarray, Ls, dt: float, normalize: bool = True):
    """Kraus operators with the normalisation folded in: ``A_k = M_k T^{-dag}``,
    so that one step is ``rho' = sum_k A_k rho A_k^dag`` and ``sum A_k^dag A_k = I``."""
    M0, Ms, S = kraus_ops(H, Ls, dt)
    if not normalize:
        return [M0, *Ms]
    T = np.linalg.cholesky(S)
    Tinv_dag = dag(sla.solve_triangular(T, np.eye(T.shape[0]), lower=True))
    return [M @ Tinv_dag for M in (M0, *Ms)]


# --------------------------------------------------------------------- Rouchon2
# KrausMapRK (rouchon_integrator.py:85-224) with the Heun tableau of KrausHeun2
# (:272-284: c = (0, 1), A = ((0, 0), (1, 0)), b = (1/2, 1/2)) and the no-jump
# propagator of MESolveFixedRouchon2Integrator (:435-443): one step of the explicit
# midpoint method on dU/dt = G U, U(t) = I, G = -i H - 1/2 sum L^dag L, i.e.
# U(1) = I + dt G (I + dt/2 G) for an operator constant over the step.
HEUN2_C = (0.0, 1.0)
HEUN2_A = ((0.0, 0.0), (1.0, 0.0))
HEUN2_B = (0.5, 0.5)


def rouchon2_propagator(H: np.ndarray, Ls, dt: float) -> np.ndarray:
    n = H.shape[-1]
    G = -1j * H - 0.5 * sum((dag(L) @ L for L in Ls), np.zeros_like(H))
    return np.eye(n) + dt * G @ (np.eye(n) + 0.5 * dt * G)


def kraus_rk_step(U: dict, tab_c, tab_a, tab_b, Ls, rho: np.ndarray, dt: float, normalize: bool = True) -> np.ndarray:
    """One step of ``KrausMapRK`` (rouchon_integrator.py:83-224) for the tableau
    (c, A, b) and the no-jump propagators ``U[c]`` at the nodes and at 1: the stages
    rho^(i) = U(c_i) rho U(c_i)^dag + dt sum_j a_ij P_ij D(rho^(j)) P_ij^dag
    (``compute_stages`` :160-175), the update
    rho' = U(1) rho U(1)^dag + dt sum_i b_i Q_i D(rho^(i)) Q_i^dag (``__call__``
    :177-190) with P_ij = U(c_i) U(c_j)^-1, Q_i = U(1) U(c_i)^-1 (linear solves, as
    ``propagator_ratio`` :140-154), and, with ``normalize``, rho first mapped to
    T^-dag rho T^-1 where T T^dag = S from the adjoint pass (``S_stage`` / ``S``
    :196-224)."""
    n = rho.shape[-1]
    I = np.eye(n)

    def ratio(ci, cj):
        if ci == cj:
            return I
        if cj == 0.0:
            return U[ci]
        return np.linalg.solve(U[cj].T, U[ci].T).T

    def D(r):
        return sum((L @ r @ dag(L) for L in Ls), np.zeros_like(r))

    def Dadj(o):
        return sum((dag(L) @ o @ L for L in Ls), np.zeros_like(o))

    def S_stage(i, O):
        ci = tab_c[i]
        res = dag(U[ci]) @ O @ U[ci]
        for j in range(i):
            if tab_a[i][j] == 0.0:
                continue
            P = ratio(ci, tab_c[j])
            res = res + S_stage(j, dt * tab_a[i][j] * Dadj(dag(P) @ O @ P))
        return res

    if normalize:
        S = dag(U[1.0]) @ U[1.0]
        for i, bi in enumerate(tab_b):
            if bi == 0.0:
                continue
            Q = ratio(1.0, tab_c[i])
            S = S + dt * bi * S_stage(i, Dadj(dag(Q) @ Q))
        rho = cholesky_normalize(S, rho)
    stages = []
    for i, ci in enumerate(tab_c):
        r = U[ci] @ rho @ dag(U[ci])
        for j in range(i):
            if tab_a[i][j] == 0.0:
                continue
            P = ratio(ci, tab_c[j])
            r = r + dt * tab_a[i][j] * P @ D(stages[j]) @ dag(P)
        stages.append(r)
    out = U[1.0] @ rho @ dag(U[1.0])
    for i, bi in enumerate(tab_b):
        if bi == 0.0:
            continue
        Q = ratio(1.0, tab_c[i])
        out = out + dt * bi * Q @ D(stages[i]) @ dag(Q)
    return out


def rouchon2_step(H: np.ndarray, Ls, rho: np.ndarray, dt: float, normalize: bool = True) -> np.ndarray:
    """One Rouchon2 step: ``kraus_rk_step`` with the Heun tableau and U(1) of the
    midpoint no-jump propagator."""
    n = H.shape[-1]
    U = {0.0: np.eye(n), 1.0: rouchon2_propagator(H, Ls, dt)}
    return kraus_rk_step(U, HEUN2_C, HEUN2_A, HEUN2_B, Ls, rho, dt, normalize)


# --------------------------------------------------------------------- Rouchon3
# KrausHeun3 (rouchon_integrator.py:288-302: c = (0, 1/3, 2/3),
# A = ((0, 0, 0), (1/3, 0, 0), (0, 2/3, 0)), b = (1/4, 0, 3/4)) with the no-jump
# propagator of MESolveFixedRouchon3Integrator (:445-453): one Bogacki-Shampine
# (Bosh3) step on dU/dt = G U from U = I, whose dense output (diffrax's
# ThirdOrderHermitePolynomialInterpolation) is the cubic Hermite polynomial through
# (I, dt G) at theta = 0 and (U(1), dt G U(1)) at theta = 1 (make_nojump_propagator
# :326-354). Bosh3 tableau (Bogacki & Shampine 1989): c = (0, 1/2, 3/4, 1),
# a21 = 1/2, a32 = 3/4, b = (2/9, 1/3, 4/9), the fourth stage at y1 (FSAL).
HEUN3_C = (0.0, 1.0 / 3.0, 2.0 / 3.0)
HEUN3_A = ((0.0, 0.0, 0.0), (1.0 / 3.0, 0.0, 0.0), (0.0, 2.0 / 3.0, 0.0))
HEUN3_B = (0.25, 0.0, 0.75)


def hermite3(y0, y1, k0, k1, theta: float):
    """Cubic through y0, y1 with theta-derivatives k0, k1 (dt-scaled slopes)."""
    t2, t3 = theta * theta, theta * theta * theta
    return (2 * t3 - 3 * t2 + 1) * y0 + (t3 - 2 * t2 + theta) * k0 + (-2 * t3 + 3 * t2) * y1 + (t3 - t2) * k1


def rouchon3_propagators(H: np.ndarray, Ls, dt: float) -> dict:
    """No-jump propagators U(c) at c = 0, 1/3, 2/3, 1 for Rouchon3."""
    n = H.shape[-1]
    I = np.eye(n)
    G = -1j * H - 0.5 * sum((dag(L) @ L for L in Ls), np.zeros_like(H))
    k1 = G
    k2 = G @ (I + 0.5 * dt * k1)
    k3 = G @ (I + 0.75 * dt * k2)
    U1 = I + dt * (2.0 / 9.0 * k1 + 1.0 / 3.0 * k2 + 4.0 / 9.0 * k3)
    d0, d1 = dt * G, dt * G @ U1
    return {0.0: I, HEUN3_C[1]: hermite3(I, U1, d0, d1, HEUN3_C[1]), HEUN3_C[2]: hermite3(I, U1, d0, d1, HEUN3_C[2]),
            1.0: U1}


def rouchon3_step(H: np.ndarray, Ls, rho: np.ndarray, dt: float, normalize: bool = True) -> np.ndarray:
    """One Rouchon3 step: ``kraus_rk_step`` with the Heun3 tableau."""
    return kraus_rk_step(rouchon3_propagators(H, Ls, dt), HEUN3_C, HEUN3_A, HEUN3_B, Ls, rho, dt, normalize)


def _to_lower_band(S: np.ndarray, w: int) -> np.ndarray:
    n = S.shape[0]
    ab = np.zeros((w + 1, n), dtype=S.dtype)
    for k in range(w + 1):
        ab[k, : n - k] = np.diagonal(S, offset=-k)
    return ab


def banded_factor(S: np.ndarray, w: int):
    """Lower banded Cholesky factor of a Hermitian matrix with half-bandwidth ``w``
    in LAPACK lower-band storage (``ab[k, j] = T[j + k, j]``)."""
    return sla.cholesky_banded(_to_lower_band(S, w), lower=True)


def banded_normalize(Tab: np.ndarray, rho: np.ndarray) -> np.ndarray:
    """``T^{-dag} rho T^{-1}`` with ``T`` given in lower-band storage."""
    w = Tab.shape[0] - 1
    n = Tab.shape[1]
    # upper-band storage of T^dag: ab_u[w + i - j, j] = T^dag[i, j] = conj(T[j, i])
    ab_u = np.zeros_like(Tab)
    for k in range(w + 1):
        ab_u[w - k, k:] = Tab[k, : n - k].conj()
    x = sla.solve_banded((0, w), ab_u, rho)  # T^dag x = rho
    # x T = y  <=>  T^dag x^dag = y^dag, and x^dag = x here up to rounding; solve exactly
    y = sla.solve_banded((0, w), ab_u, dag(x))
    return dag(y)


def kraus_step_banded(H: np.ndarray, Ls, rho: np.ndarray, dt: float, w: int, Tab=None) -> np.ndarray:
    M0, Ms, S = kraus_ops(H, Ls, dt)
    if Tab is None:
        Tab = banded_factor(S, w)
    rho = banded_normalize(Tab, rho)
    out = M0 @ rho @ dag(M0)
    for M in Ms:
        out = out + M @ rho @ dag(M)
    return out


def step_times(t0: float, t1: float, dt: float) -> np.ndarray:
    """Step boundaries ``t0, t0+dt, ...`` with the last step clipped to ``t1``."""
    n = (t1 - t0) / dt
    nr = int(round(n))
    if abs(nr - n) <= 1e-9 * max(1.0, n):
        nsteps = max(nr, 1)
    else:
        nsteps = int(np.ceil(n))
    ts = t0 + dt * np.arange(nsteps + 1, dtype=np.float64)
    ts[-1] = t1
    return ts


def integrate(step, rho0: np.ndarray, tsave: np.ndarray, dt: float, *, t0=None, save=None):
    """Fixed-step loop; ``step(t, h, rho) -> rho'``. Linear interpolation at tsave."""
    tsave = np.asarray(tsave, dtype=np.float64)
    t_start = float(tsave[0] if t0 is None else t0)
    ts = step_times(t_start, float(tsave[-1]), dt)
    save = save or (lambda y: y.copy())
    rho = np.array(rho0, dtype=np.complex128)
    saved = []
    isave = 0
    while isave < len(tsave) and tsave[isave] <= t_start:
        saved.append(save(rho))
        isave += 1
    for i in range(len(ts) - 1):
        t, tn = ts[i], ts[i + 1]
        new = step(t, tn - t, rho)
        while isave < len(tsave) and tsave[isave] <= tn + 1e-12 * max(1.0, abs(tn)):
            theta = (tsave[isave] - t) / (tn - t)
            theta = min(max(theta, 0.0), 1.0)
            saved.append(save(rho + theta * (new - rho)))
            isave += 1
        rho = new
    return saved, rho, len(ts) - 1


# ------------------------------------------------- block-diagonal normalisation
# When S splits into independent blocks (index sets that S never couples: for the
# two-mode problem with an excitation-number-conserving H and jump operators whose
# L^dag L is diagonal, the sets of fixed a^dag a + b^dag b), the Cholesky factor of
# S is block-diagonal on the same sets, provided each set keeps the original
# relative order of its indices: the factor assembled from the blocks' factors is
# lower triangular in the original order and satisfies T T^dag = S, so by
# uniqueness of the Cholesky factorisation it IS chol(S). The normalisation
# T^-dag rho T^-1 is then a set of small products on row and column slices. Same
# step as kraus_step_dense, at a cost that allows 5000 FP64 steps at n = 1024.


def coupled_blocks(S: np.ndarray, tol: float = 0.0):
    """Index sets (each sorted ascending) of the connected components of the graph
    with an edge i-j wherever |S[i, j]| > tol."""
    from scipy.sparse import csr_matrix
    from scipy.sparse.csgraph import connected_components

    ncomp, labels = connected_components(csr_matrix(np.abs(S) > tol), directed=False)
    return [np.flatnonzero(labels == c) for c in range(ncomp)]


def block_factor(S: np.ndarray, blocks):
    """``[(idx, T_b^-1)]``: the inverse Cholesky factor of every diagonal block."""
    out = []
    for idx in blocks:
        Tb = np.linalg.cholesky(S[np.ix_(idx, idx)])
        out.append((idx, sla.solve_triangular(Tb, np.eye(len(idx)), lower=True)))
    return out


def block_normalize(fac, rho: np.ndarray) -> np.ndarray:
    """``T^-dag rho T^-1`` with ``T`` block-diagonal on the index sets of ``fac``."""
    x = np.empty_like(rho)
    for idx, Ti in fac:
        x[idx, :] = dag(Ti) @ rho[idx, :]
    y = np.empty_like(rho)
    for idx, Ti in fac:
        y[:, idx] = x[:, idx] @ Ti
    return y


def kraus_step_dia_blocks(dM, Ls_dia, rho: np.ndarray, fac) -> np.ndarray:
    """One Rouchon1 step with DIA operators: ``X = T^-dag rho T^-1``,
    ``rho' = M0 X M0^dag + sum_k M_k X M_k^dag`` with ``M0 = I + dM`` and ``dM``,
    ``M_k = sqrt(dt) L_k`` given as ``(offsets, diags)`` (products as in
    ``oracle.dia``)."""
    from .dia import dense_matmul_dia, dia_dag, dia_matmul_dense

    X = block_normalize(fac, rho)
    Y = X + dia_matmul_dense(*dM, X)  # M0 X
    out = Y + dense_matmul_dia(Y, *dia_dag(*dM))  # M0 X M0^dag
    for o, d in Ls_dia:
        out = out + dense_matmul_dia(dia_matmul_dense(o, d, X), *dia_dag(o, d))
    return out
