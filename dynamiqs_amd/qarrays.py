"""Quantum arrays with a dense or sparse-DIA layout (host side).

Mirrors the data layer of the reference (``dynamiqs/qarrays/qarray.py:134-714``,
``dense_dataarray.py``, ``sparsedia_dataarray.py:42-345``,
``sparsedia_primitives.py``): a :class:`QArray` is a batched ``(..., m, n)``
operator or state with Hilbert-space ``dims`` and a ``layout`` (``dense`` or
``dia``). DIA storage is ``offsets`` (tuple) and ``diags[..., D, n]`` with
``A[i, j] = diags[k, j]`` for ``j - i == offsets[k]`` (out-of-bounds entries are
zero), the ``scipy.sparse.dia_matrix`` convention. The default layout is DIA as in
the reference (``qarrays/layout.py:20``).

Arithmetic here builds operators on the host in NumPy; the solve itself runs on
the GPU through ``dynamiqs_amd.integrators``.
"""

from __future__ import annotations

from typing import Any

import numpy as np
import scipy.linalg as sla


class Layout:
    def __init__(self, name: str):
        self.name = name

    def __repr__(self) -> str:
        return self.name

    __str__ = __repr__


dense = Layout('dense')
dia = Layout('dia')

_settings = {'layout': dia, 'cdtype': np.complex64}


def set_layout(layout) -> None:
    if isinstance(layout, str):
        layout = {'dense': dense, 'dia': dia}[layout]
    if layout not in (dense, dia):
        raise ValueError(f'Argument `layout` must be `dq.dense` or `dq.dia`, got {layout}.')
    _settings['layout'] = layout


def get_layout(layout=None) -> Layout:
    return _settings['layout'] if layout is None else layout


def set_precision(precision: str) -> None:
    if precision not in ('simple', 'double'):
        raise ValueError(f"Argument `precision` should be a string 'simple' or 'double', but is '{precision}'.")
    _settings['cdtype'] = np.complex64 if precision == 'simple' else np.complex128


def cdtype(dtype=None):
    if dtype is not None:
        return np.dtype(dtype)
    return np.dtype(_settings['cdtype'])


def _valid_cols(offset: int, n: int) -> slice:
    return slice(offset, None) if offset >= 0 else slice(None, n + offset)


def _zero_oob(offsets, diags):
    n = diags.shape[-1]
    for i, o in enumerate(offsets):
        if o > 0:
            diags[..., i, :o] = 0
        elif o < 0:
            diags[..., i, n + o :] = 0
    return diags


class QArray:
    """Batched quantum operator or state, dense or sparse-DIA."""

    __array_priority__ = 1000

    def __init__(self, data=None, *, dims=None, offsets=None, diags=None):
        if data is not None:
            self._data = np.asarray(data)
            self._offsets = None
            self._diags = None
            if dims is not None:
                n = None
            elif self._data.ndim >= 2:
                n = self._data.shape[-1] if self._data.shape[-1] != 1 else self._data.shape[-2]
            else:  # an element or a row taken out of a qarray with explicit dims
                n = self._data.shape[-1] if self._data.ndim == 1 else 1
        else:
            self._data = None
            self._offsets = tuple(int(o) for o in offsets)
            self._diags = np.asarray(diags)
            n = self._diags.shape[-1]
        self.dims = tuple(dims) if dims is not None else (int(n),)

    # ------------------------------------------------------------ properties
    @property
    def layout(self) -> Layout:
        return dense if self._data is not None else dia

    @property
    def shape(self) -> tuple:
        if self._data is not None:
            return self._data.shape
        n = self._diags.shape[-1]
        return (*self._diags.shape[:-2], n, n)

    @property
    def ndim(self) -> int:
        return len(self.shape)

    @property
    def dtype(self):
        return (self._data if self._data is not None else self._diags).dtype

    @property
    def offsets(self):
        return self._offsets

    @property
    def diags(self):
        return self._diags

    @property
    def ndiags(self) -> int:
        return len(self._offsets) if self._offsets is not None else 0

    def __len__(self) -> int:
        return self.shape[0]

    # ------------------------------------------------------------ conversion
    def to_numpy(self) -> np.ndarray:
        if self._data is not None:
            return self._data
        n = self._diags.shape[-1]
        out = np.zeros((*self._diags.shape[:-2], n, n), dtype=self._diags.dtype)
        rows = np.arange(n)
        for i, k in enumerate(self._offsets):
            cols = rows + k
            m = (cols >= 0) & (cols < n)
            out[..., rows[m], cols[m]] = self._diags[..., i, cols[m]]
        return out

    to_jax = to_numpy

    def __array__(self, dtype=None, copy=None):
        x = self.to_numpy()
        return x.astype(dtype) if dtype is not None else x

    def asdense(self) -> QArray:
        return self if self._data is not None else QArray(self.to_numpy(), dims=self.dims)

    def asdia(self) -> QArray:
        if self._data is None:
            return self
        x = self._data
        n = x.shape[-1]
        offsets = [
            k for k in range(-(n - 1), n) if np.any(np.diagonal(x, offset=k, axis1=-2, axis2=-1) != 0)
        ]
        diags = np.zeros((*x.shape[:-2], len(offsets), n), dtype=x.dtype)
        for i, k in enumerate(offsets):
            diags[..., i, _valid_cols(k, n)] = np.diagonal(x, offset=k, axis1=-2, axis2=-1)
        return QArray(offsets=offsets, diags=diags, dims=self.dims)

    def to_layout(self, layout) -> QArray:
        return self.asdense() if layout is dense else self.asdia()

    def astype(self, dtype) -> QArray:
        if self._data is not None:
            return QArray(self._data.astype(dtype), dims=self.dims)
        return QArray(offsets=self._offsets, diags=self._diags.astype(dtype), dims=self.dims)

    # ------------------------------------------------------------ structure
    def isket(self) -> bool:
        return self.shape[-1] == 1

    def isbra(self) -> bool:
        return self.shape[-2] == 1

    def isop(self) -> bool:
        return self.shape[-1] == self.shape[-2]

    isdm = isop

    def isherm(self, rtol: float = 1e-5, atol: float = 1e-8) -> bool:
        x = self.to_numpy()
        return bool(np.allclose(x, np.swapaxes(x.conj(), -1, -2), rtol=rtol, atol=atol))

    def todm(self) -> QArray:
        if self.isket():
            x = self.to_numpy()
            return QArray(x @ np.swapaxes(x.conj(), -1, -2), dims=self.dims)
        if self.isbra():
            x = self.to_numpy()
            return QArray(np.swapaxes(x.conj(), -1, -2) @ x, dims=self.dims)
        return self

    def dag(self) -> QArray:
        if self._data is not None:
            return QArray(np.swapaxes(self._data.conj(), -1, -2), dims=self.dims)
        n = self._diags.shape[-1]
        out = np.zeros_like(self._diags)
        for i, o in enumerate(self._offsets):
            if o >= 0:
                out[..., i, : n - o] = self._diags[..., i, o:].conj()
            else:
                out[..., i, -o:] = self._diags[..., i, : n + o].conj()
        return QArray(offsets=tuple(-o for o in self._offsets), diags=out, dims=self.dims)

    def conj(self) -> QArray:
        if self._data is not None:
            return QArray(self._data.conj(), dims=self.dims)
        return QArray(offsets=self._offsets, diags=self._diags.conj(), dims=self.dims)

    @property
    def mT(self) -> QArray:
        return self.dag().conj()

    def trace(self) -> np.ndarray:
        return np.trace(self.to_numpy(), axis1=-2, axis2=-1)

    def norm(self) -> np.ndarray:
        x = self.to_numpy()
        if self.isket() or self.isbra():
            return np.sqrt(np.sum(np.abs(x) ** 2, axis=(-2, -1)))
        return np.real(np.trace(x, axis1=-2, axis2=-1))

    def expm(self) -> QArray:
        x = self.to_numpy()
        flat = x.reshape(-1, *x.shape[-2:])
        out = np.stack([sla.expm(m) for m in flat]).reshape(x.shape).astype(x.dtype)
        return QArray(out, dims=self.dims)

    def reshape(self, *shape) -> QArray:
        if self._data is not None:
            return QArray(self._data.reshape(*shape), dims=self.dims)
        return QArray(offsets=self._offsets, diags=self._diags.reshape(*shape[:-2], *self._diags.shape[-2:]), dims=self.dims)

    def broadcast_to(self, *shape) -> QArray:
        if self._data is not None:
            return QArray(np.broadcast_to(self._data, shape), dims=self.dims)
        return QArray(
            offsets=self._offsets, diags=np.broadcast_to(self._diags, (*shape[:-2], *self._diags.shape[-2:])), dims=self.dims
        )

    def __getitem__(self, key) -> QArray:
        if self._data is not None:
            return QArray(self._data[key], dims=self.dims)
        if not isinstance(key, tuple):
            key = (key,)
        return QArray(offsets=self._offsets, diags=self._diags[key], dims=self.dims)

    # ------------------------------------------------------------ arithmetic
    def _binary_dense(self, other, fn) -> QArray:
        return QArray(fn(self.to_numpy(), _as_numpy(other)), dims=self.dims)

    def __mul__(self, y) -> QArray:
        if getattr(y, '_is_timeqarray', False):
            return NotImplemented
        if isinstance(y, QArray):
            if self._data is None and y._data is None:
                return _dia_elmul(self, y)
            return self._binary_dense(y, np.multiply)
        y = np.asarray(y)
        # a scalar or a batch of scalars of shape (..., 1, 1), broadcast against the
        # qarray (is_batched_scalar, dynamiqs/_utils.py:48-57; qarray.py:496-504)
        if not (y.ndim == 0 or y.shape == (1,) or (y.ndim > 1 and y.shape[-2:] == (1, 1))):
            raise NotImplementedError(
                'Element-wise multiplication of two qarrays with the `*` operator is '
                'not supported. For matrix multiplication, use `x @ y`. For '
                'element-wise multiplication, use `x.elmul(y)`.'
            )
        if self._data is not None:
            return QArray(self._data * y, dims=self.dims)
        return QArray(offsets=self._offsets, diags=self._diags * y, dims=self.dims)

    __rmul__ = __mul__

    def elmul(self, y) -> QArray:
        """Element-wise product with a qarray-like (qarray.py ``elmul``)."""
        if isinstance(y, QArray) and self._data is None and y._data is None:
            return _dia_elmul(self, y)
        return self._binary_dense(y, np.multiply)

    def __truediv__(self, y) -> QArray:
        return self * (1 / np.asarray(y))

    def __neg__(self) -> QArray:
        return self * -1

    def __add__(self, y) -> QArray:
        if getattr(y, '_is_timeqarray', False):
            return NotImplemented
        if isinstance(y, QArray):
            if self._data is None and y._data is None:
                return _dia_add(self, y)
            return QArray(self.to_numpy() + y.to_numpy(), dims=self.dims)
        y = np.asarray(y)
        if y.ndim == 0 and y == 0:
            return self
        return QArray(self.to_numpy() + y, dims=self.dims)

    __radd__ = __add__

    def __sub__(self, y) -> QArray:
        if getattr(y, '_is_timeqarray', False):
            return NotImplemented
        return self + (-y if isinstance(y, QArray) else -np.asarray(y))

    def __rsub__(self, y) -> QArray:
        return (-self) + y

    def __matmul__(self, y) -> QArray:
        if not isinstance(y, QArray):
            y = QArray(np.asarray(y), dims=self.dims)
        if self._data is None and y._data is None:
            return _dia_matmul(self, y)
        return QArray(self.to_numpy() @ y.to_numpy(), dims=self.dims)

    def __rmatmul__(self, y) -> QArray:
        return QArray(np.asarray(y) @ self.to_numpy(), dims=self.dims)

    def __pow__(self, k: int) -> QArray:
        out = self
        for _ in range(int(k) - 1):
            out = out @ self
        return out

    def powm(self, k: int) -> QArray:
        return self.__pow__(k)

    def __repr__(self) -> str:
        extra = f', ndiags={self.ndiags}' if self.layout is dia else ''
        return (
            f'QArray: shape={self.shape}, dims={self.dims}, dtype={self.dtype}, layout={self.layout}{extra}\n'
            f'{np.array2string(self.to_numpy(), precision=3, suppress_small=True)}'
        )


def _as_numpy(x) -> np.ndarray:
    return x.to_numpy() if isinstance(x, QArray) else np.asarray(x)


def _dia_add(a: QArray, b: QArray) -> QArray:
    offs = sorted(set(a._offsets) | set(b._offsets))
    batch = np.broadcast_shapes(a._diags.shape[:-2], b._diags.shape[:-2])
    n = a._diags.shape[-1]
    out = np.zeros((*batch, len(offs), n), dtype=np.result_type(a.dtype, b.dtype))
    for i, o in enumerate(offs):
        if o in a._offsets:
            out[..., i, :] += a._diags[..., a._offsets.index(o), :]
        if o in b._offsets:
            out[..., i, :] += b._diags[..., b._offsets.index(o), :]
    return QArray(offsets=offs, diags=out, dims=a.dims)


def _dia_elmul(a: QArray, b: QArray) -> QArray:
    offs = [o for o in a._offsets if o in b._offsets]
    batch = np.broadcast_shapes(a._diags.shape[:-2], b._diags.shape[:-2])
    n = a._diags.shape[-1]
    out = np.zeros((*batch, len(offs), n), dtype=np.result_type(a.dtype, b.dtype))
    for i, o in enumerate(offs):
        out[..., i, :] = a._diags[..., a._offsets.index(o), :] * b._diags[..., b._offsets.index(o), :]
    return QArray(offsets=offs, diags=out, dims=a.dims)


def _dia_matmul(a: QArray, b: QArray) -> QArray:
    # (A B)[r, r+p+q] = A[r, r+p] B[r+p, r+p+q] = a_p[r+p] * b_q[r+p+q]
    n = a._diags.shape[-1]
    batch = np.broadcast_shapes(a._diags.shape[:-2], b._diags.shape[:-2])
    dtype = np.result_type(a.dtype, b.dtype)
    acc: dict[int, np.ndarray] = {}
    for i, p in enumerate(a._offsets):
        for j, q in enumerate(b._offsets):
            o = p + q
            if abs(o) > n - 1:
                continue
            d = acc.setdefault(o, np.zeros((*batch, n), dtype=dtype))
            # column c = r + o; A value at column c - q, B value at column c
            c_lo = max(0, o, q)
            c_hi = min(n, n + o, n + q)
            if c_lo < c_hi:
                d[..., c_lo:c_hi] += a._diags[..., i, c_lo - q : c_hi - q] * b._diags[..., j, c_lo:c_hi]
    offs = sorted(acc)
    if not offs:
        return QArray(offsets=(0,), diags=np.zeros((*batch, 1, n), dtype=dtype), dims=a.dims)
    return QArray(offsets=offs, diags=np.stack([acc[o] for o in offs], axis=-2), dims=a.dims)


def _dia_kron(a: QArray, b: QArray) -> QArray:
    n2 = b._diags.shape[-1]
    n1 = a._diags.shape[-1]
    batch = np.broadcast_shapes(a._diags.shape[:-2], b._diags.shape[:-2])
    offs, diags = [], []
    for i, p in enumerate(a._offsets):
        for j, q in enumerate(b._offsets):
            offs.append(p * n2 + q)
            d = a._diags[..., i, :, None] * b._diags[..., j, None, :]
            diags.append(np.broadcast_to(d, (*batch, n1, n2)).reshape(*batch, n1 * n2))
    order = np.argsort(offs)
    out_offs = [offs[k] for k in order]
    out = np.stack([diags[k] for k in order], axis=-2)
    res = QArray(offsets=out_offs, diags=out, dims=(*a.dims, *b.dims))
    # merge duplicated offsets (cannot happen for |q| < n2, kept for safety)
    if len(set(out_offs)) != len(out_offs):
        res = _dia_add(QArray(offsets=(0,), diags=np.zeros((*batch, 1, n1 * n2), dtype=out.dtype), dims=res.dims), res)
    _zero_oob(res._offsets, res._diags)
    return res


def asqarray(x: Any, dims=None, layout=None) -> QArray:
    """Convert a qarray-like (QArray, array, nested list, torch tensor, list of QArrays)."""
    if isinstance(x, QArray):
        out = x
    elif isinstance(x, (list, tuple)) and len(x) > 0 and all(isinstance(e, QArray) for e in x):
        out = QArray(np.stack([e.to_numpy() for e in x]), dims=x[0].dims)
    else:
        if hasattr(x, 'detach') and hasattr(x, 'cpu'):
            x = x.detach().cpu().numpy()
        arr = np.asarray(x)
        if arr.dtype == object:
            raise TypeError(f'Argument must be a qarray-like, but has type {type(x).__name__}.')
        if arr.ndim < 2:
            raise ValueError(f'Argument must have at least 2 dimensions, but has shape {arr.shape}.')
        if not np.iscomplexobj(arr):
            arr = arr.astype(cdtype())
        elif arr.dtype not in (np.complex64, np.complex128):
            arr = arr.astype(cdtype())
        out = QArray(arr, dims=dims)
    if dims is not None:
        out.dims = tuple(dims)
    if layout is not None:
        out = out.to_layout(layout)
    return out


def isqarraylike(x) -> bool:
    if isinstance(x, QArray):
        return True
    try:
        arr = np.asarray(x)
    except Exception:
        return False
    return arr.dtype != object and np.issubdtype(arr.dtype, np.number)
