"""Argument checks with the reference's messages (``dynamiqs/_checks.py:44-117``)."""

from __future__ import annotations

import numpy as np

_cases = {
    '(..., n, 1)': lambda x: len(x.shape) >= 2 and x.shape[-1] == 1,
    '(..., n, n)': lambda x: len(x.shape) >= 2 and x.shape[-2] == x.shape[-1],
    '(n, n)': lambda x: len(x.shape) == 2 and x.shape[-2] == x.shape[-1],
    '(?, n, n)': lambda x: 2 <= len(x.shape) <= 3 and x.shape[-2] == x.shape[-1],
}


def check_shape(x, argname: str, *shapes: str, subs: dict | None = None) -> None:
    for shape in shapes:
        if _cases[shape](x):
            return
    shapes_str = shapes[0] if len(shapes) == 1 else ', '.join(shapes[:-1]) + ' or ' + shapes[-1]
    if subs is not None:
        for k, v in subs.items():
            shapes_str = shapes_str.replace(k, v)
    raise ValueError(
        f'Argument `{argname}` must have shape {shapes_str}, but has shape {argname}.shape={tuple(x.shape)}.'
    )


def check_times(x: np.ndarray, argname: str, allow_empty: bool = False) -> np.ndarray:
    if x.ndim != 1:
        raise ValueError(f'Argument {argname} must be a 1D array, but is a {x.ndim}D array.')
    if not allow_empty and len(x) == 0:
        raise ValueError(f'Argument {argname} must contain at least one element.')
    if np.any(x[1:] < x[:-1]):
        raise ValueError(f'Argument {argname} must be sorted in strictly ascending order.')
    return x


def check_hermitian(x: np.ndarray, argname: str) -> None:
    rtol, atol = 1e-5, 1e-5
    if not np.allclose(x, np.swapaxes(x.conj(), -1, -2), rtol=rtol, atol=atol):
        raise ValueError(f'Argument {argname} is not hermitian.')


def check_qarray_is_dense(x, argname: str) -> None:
    from .qarrays import dense

    if x.layout is not dense:
        raise ValueError(f'Argument {argname} must have layout `dense` but has layout {argname}.layout={x.layout}.')
