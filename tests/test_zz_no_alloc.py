"""dqh_expm_propagators_noalloc (the Expm exponentials and chained products with the
library's fixed device work array instead of a caller-allocated one) against
dqh_expm_propagators on the same inputs, through the C-ABI: two calls in a row,
the second with a smaller problem, each equal to the allocating entry point's
result; and the size limit rejected with an error before any launch."""

import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TAYLOR_DEGREE = 8


def _generators(rng, nchain, nint, N, scale):
    """Random anti-Hermitian generators -i H dt / 2^s of 1-norm below 1/2 (the scaled
    input; their exponentials are unitary, so products of them stay of norm 1)."""
    A = rng.normal(size=(nchain, nint, N, N)) + 1j * rng.normal(size=(nchain, nint, N, N))
    X = -0.5j * (A + np.conj(np.swapaxes(A, -1, -2)))
    X *= scale / np.max(np.sum(np.abs(X), axis=-2))
    return X.astype(np.complex64)


def _run_both(lib, torch, X, s):
    nchain, nint, N = X.shape[0], X.shape[1], X.shape[-1]
    c = ctypes.c_void_p
    x = torch.from_numpy(np.ascontiguousarray(X).view(np.float32)).cuda()
    work = torch.empty(int(lib.dqh_expm_work_bytes(nchain, nint, N)), dtype=torch.uint8, device='cuda')
    U_ref = torch.empty_like(x)
    U_na = torch.full_like(x, float('nan'))
    st = torch.cuda.current_stream().cuda_stream
    rc = lib.dqh_expm_propagators(c(x.data_ptr()), c(work.data_ptr()), c(U_ref.data_ptr()), nchain, nint, N,
                                  TAYLOR_DEGREE, s, c(st))
    assert rc == 0, lib.dqh_last_error()
    rc = lib.dqh_expm_propagators_noalloc(c(x.data_ptr()), c(U_na.data_ptr()), nchain, nint, N, TAYLOR_DEGREE, s,
                                          c(st))
    assert rc == 0, lib.dqh_last_error()
    torch.cuda.synchronize()
    f = lambda t: t.cpu().numpy().view(np.complex64).reshape(nchain, nint, N, N)
    return f(U_ref), f(U_na)


def test_noalloc_expm_equals_allocating_entry_point_twice():
    import torch

    from dynamiqs_amd import _native

    assert torch.cuda.is_available()
    lib = _native.lib()
    rng = np.random.default_rng(11)
    # first call: 3 chains of 5 intervals at N = 64 with two squarings;
    # second call: a smaller problem (1 chain, 3 intervals, N = 32, no squaring)
    for (nchain, nint, N, s) in ((3, 5, 64, 2), (1, 3, 32, 0)):
        print(f'noalloc Expm: nchain={nchain} nint={nint} N={N} s={s}', flush=True)
        X = _generators(rng, nchain, nint, N, 0.45)
        U_ref, U_na = _run_both(lib, torch, X, s)
        assert np.all(np.isfinite(U_na))
        np.testing.assert_array_equal(U_na, U_ref)
        # and the products are the exponentials: U[b, k] = expm(2^s X_k) ... expm(2^s X_0)
        import scipy.linalg as sla

        for b in range(nchain):
            P = np.eye(N, dtype=np.complex128)
            for k in range(nint):
                P = sla.expm((2.0**s) * X[b, k].astype(np.complex128)) @ P
                assert np.abs(U_na[b, k] - P).max() < 2e-5


def test_noalloc_expm_rejects_problems_above_its_fixed_work_size():
    from dynamiqs_amd import _native

    lib = _native.lib()
    cap = 64 << 20
    N, nchain = 1024, 1
    nint = cap // (2 * N * N * 8) + 1  # one interval more than fits
    assert lib.dqh_expm_work_bytes(nchain, nint, N) > cap
    dummy = ctypes.c_void_p(16)  # never dereferenced: the size check comes first
    rc = lib.dqh_expm_propagators_noalloc(dummy, dummy, nchain, nint, N, TAYLOR_DEGREE, 0, None)
    assert rc == -2
    assert b'above its fixed' in lib.dqh_last_error()
