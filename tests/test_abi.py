"""CPU checks of the C-ABI library: it loads, exports every symbol declared in
include/dqh.h, and its descriptor layout matches the Python mirror. No compute."""

import ctypes
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_symbols():
    src = open(os.path.join(ROOT, 'include', 'dqh.h')).read()
    return sorted(set(re.findall(r'\b(dqh_[a-z0-9_]+)\s*\(', src)))


def test_library_exports_every_declared_symbol():
    from dynamiqs_amd import _native

    lib = ctypes.CDLL(_native.LIB_PATH)
    names = _declared_symbols()
    assert len(names) >= 7
    for name in names:
        assert hasattr(lib, name), name
    assert set(names) <= set(_native.SYMBOLS), set(names) - set(_native.SYMBOLS)


def test_descriptor_layout_matches_mirror():
    from dynamiqs_amd import _native

    lib = _native.lib()
    assert lib.dqh_sizeof_me_inst() == _native.ME_INST_DTYPE.itemsize
    assert lib.dqh_me_work_elems(64, 2, 0) == 11 * 64 * 64
    src = open(os.path.join(ROOT, 'include', 'dqh.h')).read()
    body = src[src.index('typedef struct dqh_me_inst') : src.index('} dqh_me_inst;')]
    fields = re.findall(r'\b(?:int64_t|double|uint64_t)\s+([a-zA-Z0-9_, ]+);', body)
    names = [f.strip() for group in fields for f in group.split(',')]
    assert names == list(_native.ME_INST_DTYPE.names)


def test_no_fallback_when_library_missing(monkeypatch, tmp_path):
    import pytest

    from dynamiqs_amd import _native

    monkeypatch.setattr(_native, '_lib', None)
    monkeypatch.setattr(_native, 'LIB_PATH', str(tmp_path / 'missing.so'))
    with pytest.raises(RuntimeError, match='missing'):
        _native.lib()
