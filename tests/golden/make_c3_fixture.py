"""Write ``c3_dia_rouchon1_alpha2.npz``: the complex128 oracle's Rouchon1 solve of
the 1024-dimensional two-mode problem over its full length, as the fixture of the
full-size GPU parity test.

Problem: two coupled 32-level modes, H = 5 a^dag a + 6 b^dag b + 0.1 (a^dag b + a b^dag)
- 0.01 a^dag a b^dag b, jump operators sqrt(0.02) a and sqrt(0.05) b,
rho0 = |alpha, 0><alpha, 0| with alpha = 2.0, Rouchon1 with dt = 1e-3 on [0, 5]
(5000 steps), 51 save times. The oracle is ``oracle.rouchon.kraus_step_dia_blocks``
(the reference's Rouchon1 step, rouchon_integrator.py:250-263, 552-582, with the
Cholesky factor of S taken block by block on the sets of fixed excitation number
that S never couples; equal to the dense step, checked in tests/test_oracle.py).

Saved: tsave; <a^dag a>, <b^dag b>, Tr(rho), Tr(rho^2) at every save time; the final
state on the index set where its diagonal exceeds 1e-15 (rounded to complex64, 6e-8
relative) and the Frobenius norm of the final state outside that block. Run once on a CPU
(about 15 minutes): ``python tests/golden/make_c3_fixture.py``.
"""

from __future__ import annotations

import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import dia, operators as op, rouchon  # noqa: E402

NA = NB_ = 32
ALPHA = 2.0
DT = 1e-3
T_END = 5.0
NSAVE = 51
DIAG_CUT = 1e-15
OUT = os.path.join(HERE, 'c3_dia_rouchon1_alpha2.npz')


def two_mode():
    a, b = op.destroy(NA, NB_)
    ad, bd = a.conj().T, b.conj().T
    H = 5 * ad @ a + 6 * bd @ b + 0.1 * (ad @ b + a @ bd) - 0.01 * ad @ a @ bd @ b
    return H, [np.sqrt(0.02) * a, np.sqrt(0.05) * b], a, b


def main():
    H, Ls, a, b = two_mode()
    n = H.shape[0]
    na_diag = np.real(np.diag(a.conj().T @ a))
    nb_diag = np.real(np.diag(b.conj().T @ b))
    cache = {}

    def ops_for(h):
        key = round(h / DT, 9)
        if key not in cache:
            M0, Ms, S = rouchon.kraus_ops(H, Ls, h)
            fac = rouchon.block_factor(S, rouchon.coupled_blocks(S))
            cache[key] = (dia.dia_from_dense(M0 - np.eye(n)), [dia.dia_from_dense(M) for M in Ms], fac)
        return cache[key]

    def step(t, h, rho):
        dM, Ld, fac = ops_for(h)
        return rouchon.kraus_step_dia_blocks(dM, Ld, rho, fac)

    def save(r):
        d = np.real(np.diag(r))
        return np.array([d @ na_diag, d @ nb_diag, d.sum(), np.sum(np.abs(r) ** 2)])

    tsave = np.linspace(0.0, T_END, NSAVE)
    rho0 = op.coherent_dm((NA, NB_), (ALPHA, 0.0))
    t0 = time.time()
    saved, rho, nsteps = rouchon.integrate(step, rho0, tsave, DT, save=save)
    print(f'{nsteps} steps in {time.time() - t0:.0f} s', flush=True)
    saved = np.stack(saved)
    idx = np.flatnonzero(np.real(np.diag(rho)) > DIAG_CUT)
    block = rho[np.ix_(idx, idx)]
    tail = rho.copy()
    tail[np.ix_(idx, idx)] = 0.0
    np.savez_compressed(
        OUT,
        tsave=tsave,
        alpha=ALPHA,
        dt=DT,
        nsteps=nsteps,
        expect_a=saved[:, 0],
        expect_b=saved[:, 1],
        trace=saved[:, 2],
        purity=saved[:, 3],
        final_index=idx.astype(np.int32),
        final_block=block.astype(np.complex64),
        final_tail_frobenius=np.linalg.norm(tail),
    )
    print(f'wrote {OUT}: block {len(idx)} of {n}, tail {np.linalg.norm(tail):.2e}', flush=True)


if __name__ == '__main__':
    main()
