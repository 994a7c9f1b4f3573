"""Build ``libsawfem_hip.so`` (gfx950) in-tree with hipcc.

Usage: ``python -m sawsim_amd.build``.  Always recompiles.
"""
from __future__ import annotations

import os
import subprocess
import sys
import tempfile
from pathlib import Path

HERE = Path(__file__).resolve().parent
SOURCES = ['csrc/assemble.hip', 'csrc/mf.hip', 'csrc/adjoint.hip', 'csrc/h27.hip', 'csrc/refine.hip']
# host-only C++ (no GPU code), compiled by the host compiler with floating-point contraction off
HOST_SOURCES = ['csrc/analyse.cpp']
OUT = HERE / 'libsawfem_hip.so'
HOST_FLAGS = ['-O2', '-std=c++17', '-fPIC', '-ffp-contract=off', '-Wall']

FLAGS = ['--offload-arch=gfx950', '-O3', '-std=c++17', '-fPIC', '-shared',
         # no packed-FP32 instructions in any generated code
         '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops',
         '-Wno-unused-command-line-argument']


def hipcc():
    return os.environ.get('HIPCC', '/opt/rocm/bin/hipcc')


def cxx():
    return os.environ.get('CXX', 'g++')


def _run(cmd, what, verbose):
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError(f'{what} failed for libsawfem_hip.so')
    if verbose:
        sys.stderr.write(res.stderr)


def build(verbose=False) -> Path:
    tmp = OUT.with_suffix('.so.tmp')
    with tempfile.TemporaryDirectory() as d:
        objs = []
        for s in HOST_SOURCES:
            o = str(Path(d) / (Path(s).stem + '.o'))
            _run([cxx(), *HOST_FLAGS, '-c', str(HERE / s), '-o', o], cxx(), verbose)
            objs.append(o)
        cmd = [hipcc(), *FLAGS, *[str(HERE / s) for s in SOURCES], '-x', 'none', *objs, '-I', str(HERE / 'csrc'), '-o', str(tmp)]
        _run(cmd, 'hipcc', verbose)
    os.replace(tmp, OUT)
    return OUT


if __name__ == '__main__':
    print(build(verbose=True))
