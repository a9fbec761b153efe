"""Build ``libdqh.so`` (HIP kernels + C-ABI) in-tree for gfx950 with hipcc."""

from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, 'csrc')
INCLUDE = os.path.join(os.path.dirname(HERE), 'include')
LIB = os.path.join(HERE, 'libdqh.so')

HIP_SOURCES = ['dqh_dense.hip', 'dqh_dense_grid.hip', 'dqh_dia.hip', 'dqh_dia_rhs.hip', 'dqh_dia_tsit5.hip', 'dqh_adjoint.hip',
               'dqh_sesolve.hip', 'dqh_expm.hip', 'dqh_rouchon_adaptive.hip']
CXX_SOURCES = ['dqh_abi.cpp']

# Packed FP32 VALU forms are disabled for every HIP translation unit.
HIP_FLAGS = [
    '--offload-arch=gfx950',
    '-O3',
    '-std=c++17',
    '-fPIC',
    '-Xclang',
    '-target-feature',
    '-Xclang',
    '-packed-fp32-ops',
    '-munsafe-fp-atomics',
]


def hipcc() -> str:
    return os.environ.get('HIPCC', '/opt/rocm/bin/hipcc')


def build(verbose: bool = False, jobs: int = 4) -> str:
    """Compile every source to an object in ``csrc/_build`` and link ``libdqh.so``."""
    out_dir = os.path.join(CSRC, '_build')
    os.makedirs(out_dir, exist_ok=True)
    cmds = []
    objs = []
    for src in HIP_SOURCES + CXX_SOURCES:
        obj = os.path.join(out_dir, os.path.splitext(src)[0] + '.o')
        objs.append(obj)
        cmd = [hipcc(), *HIP_FLAGS, '-I', INCLUDE, '-I', CSRC, '-c', os.path.join(CSRC, src), '-o', obj]
        if src.endswith('.cpp'):
            cmd = [hipcc(), '-O2', '-std=c++17', '-fPIC', '-I', INCLUDE, '-I', CSRC, '-c', os.path.join(CSRC, src), '-o', obj]
        cmds.append(cmd)
    procs = []
    for cmd in cmds:
        if verbose:
            print(' '.join(cmd), flush=True)
        procs.append(subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT))
        while len([p for p in procs if p.poll() is None]) >= jobs:
            procs[0].wait()
    for cmd, p in zip(cmds, procs):
        out, _ = p.communicate()
        if p.returncode != 0:
            sys.stderr.write(out.decode(errors='replace'))
            raise RuntimeError(f'hipcc failed: {" ".join(cmd)}')
        if verbose and out:
            sys.stderr.write(out.decode(errors='replace'))
    tmp = LIB + '.tmp'
    link = [hipcc(), '--offload-arch=gfx950', '-shared', '-fPIC', '-o', tmp, *objs]
    if verbose:
        print(' '.join(link), flush=True)
    subprocess.run(link, check=True)
    os.replace(tmp, LIB)
    return LIB


if __name__ == '__main__':
    print(build(verbose=True))
