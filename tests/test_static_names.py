"""Static check of the Python sources (CPU): every name a function reads as a global
is bound at module level (assignment, import, def, class) or is a builtin. Catches
undefined names in code paths that only run on the GPU."""

import builtins
import glob
import os
import symtable

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
MODULE_DUNDERS = {'__file__', '__name__', '__doc__', '__spec__', '__loader__', '__package__', '__builtins__',
                  '__path__', '__annotations__', '__dict__'}
FILES = sorted(glob.glob(os.path.join(ROOT, 'dynamiqs_amd', '**', '*.py'), recursive=True)
               + glob.glob(os.path.join(ROOT, 'oracle', '*.py')) + [os.path.join(ROOT, 'bench.py'),
                                                                   os.path.join(ROOT, '__graft_entry__.py')])


def undefined_globals(path: str) -> list:
    src = open(path).read()
    top = symtable.symtable(src, path, 'exec')
    bound = {s.get_name() for s in top.get_symbols()
             if s.is_assigned() or s.is_imported() or s.is_namespace() or s.is_global() and s.is_declared_global()}
    if 'import *' in src:
        return []
    out = []

    def walk(t):
        for s in t.get_symbols():
            name = s.get_name()
            implicit = s.is_global() and not s.is_declared_global() or t.get_type() == 'module'
            if implicit and s.is_referenced() and name not in bound and not hasattr(builtins, name) \
                    and name not in MODULE_DUNDERS:
                out.append(f'{t.get_name()}: {name}')
        for c in t.get_children():
            walk(c)

    walk(top)
    return out


@pytest.mark.parametrize('path', FILES, ids=[os.path.relpath(f, ROOT) for f in FILES])
def test_no_undefined_globals(path):
    assert undefined_globals(path) == []
