"""Write ``c5_adjoint_n32_fd.npz``: central finite-difference gradients (h = 1e-4,
complex128) of the driven-oscillator loss at full size, the fixture of the
full-size GPU adjoint test.

Problem: n = 32, H(t) = a^dag a + u_k (a + a^dag) with u piecewise constant on 100
segments of [0, 10], jump operator sqrt(0.05) a, rho0 = |0><0|, loss
J = Tr(a^dag a rho(10)); the first two control vectors of u ~ N(0, 0.5) drawn with
numpy's default_rng(4) as an array of shape (64, 1, 100).

The solution is exact for piecewise-constant generators: on segment k,
vec(rho) <- expm(dt (L0 + u_k L1)) vec(rho) with the column-stacked Liouvillians
L0 = slindbladian(a^dag a, [sqrt(0.05) a]) and L1 = slindbladian(a + a^dag, []),
applied with scipy.sparse.linalg.expm_multiply. With the forward states v_k and the
adjoint row vectors r_{k+1} = vec(O^T)^T P_N ... P_{k+2}, the central difference
of segment k is r_{k+1} . (expm(dt A(u_k + h)) - expm(dt A(u_k - h))) v_k / 2h. The
exact derivative (the Frechet derivative from the block-triangular generator
[[A, L1], [0, A]]) is stored beside it; the two agree to O(h^2).

Run once on a CPU (about a minute): ``python tests/golden/make_c5_fixture.py``.
"""

from __future__ import annotations

import os
import sys

import numpy as np
import scipy.sparse as sp
from scipy.sparse.linalg import expm_multiply

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import lindblad, operators as op  # noqa: E402

N, NSEG, T_END, H_FD = 32, 100, 10.0, 1e-4
OUT = os.path.join(HERE, 'c5_adjoint_n32_fd.npz')


def problem():
    a = op.destroy(N)
    H0 = a.conj().T @ a
    X = a + a.conj().T
    Ls = [np.sqrt(0.05) * a]
    rho0 = op.fock_dm(N, 0)
    O = a.conj().T @ a
    u = np.random.default_rng(4).normal(0.0, 0.5, size=(64, 1, NSEG))[:2]
    return H0, X, Ls, rho0, O, u


def gradients(L0, L1, v0, w, u, dt):
    """(J, central FD, exact derivative) for one control vector u (nseg,)."""
    nseg = len(u)
    gen = [(dt * (L0 + u[k] * L1)).tocsr() for k in range(nseg)]
    vs = [v0]
    for k in range(nseg):
        vs.append(expm_multiply(gen[k], vs[-1]))
    J = w @ vs[-1]
    rs = [None] * (nseg + 1)
    rs[nseg] = w
    for k in range(nseg - 1, -1, -1):
        rs[k] = expm_multiply(gen[k].T.tocsr(), rs[k + 1])
    fd = np.zeros(nseg)
    ex = np.zeros(nseg)
    M = L0.shape[0]
    Z = sp.csr_matrix((M, M), dtype=np.complex128)
    for k in range(nseg):
        up = expm_multiply((dt * (L0 + (u[k] + H_FD) * L1)).tocsr(), vs[k])
        um = expm_multiply((dt * (L0 + (u[k] - H_FD) * L1)).tocsr(), vs[k])
        fd[k] = np.real(rs[k + 1] @ (up - um)) / (2 * H_FD)
        A = L0 + u[k] * L1
        aug = (dt * sp.bmat([[A, L1], [Z, A]])).tocsr()
        top = expm_multiply(aug, np.concatenate([np.zeros(M, dtype=np.complex128), vs[k]]))[:M]
        ex[k] = np.real(rs[k + 1] @ top)
    return float(np.real(J)), fd, ex


def main():
    H0, X, Ls, rho0, O, u = problem()
    L0 = sp.csr_matrix(lindblad.slindbladian(H0, Ls))
    L1 = sp.csr_matrix(lindblad.slindbladian(X, []))
    v0 = lindblad.vectorize(rho0)[:, 0]
    w = lindblad.vectorize(O.T)[:, 0]  # Tr(O rho) = vec(O^T) . vec(rho)
    dt = T_END / NSEG
    Js, fds, exs = [], [], []
    for b in range(u.shape[0]):
        J, fd, ex = gradients(L0, L1, v0, w, u[b, 0], dt)
        print(f'vector {b}: J = {J:.10f}, max |fd| {np.abs(fd).max():.4e}, max |fd - exact| {np.abs(fd - ex).max():.2e}',
              flush=True)
        Js.append(J)
        fds.append(fd)
        exs.append(ex)
    np.savez_compressed(OUT, u=u, times=np.linspace(0.0, T_END, NSEG + 1), h=H_FD, J=np.array(Js),
                        grad_fd=np.stack(fds)[:, None, :], grad_exact=np.stack(exs)[:, None, :])
    print(f'wrote {OUT}', flush=True)


if __name__ == '__main__':
    main()
