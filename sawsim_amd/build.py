"""Build ``libsawfem_hip.so`` (gfx950) in-tree with hipcc.

Usage: ``python -m sawsim_amd.build``.  Always recompiles.
"""
from __future__ import annotations

import os
import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
SOURCES = ['csrc/assemble.hip', 'csrc/mf.hip', 'csrc/adjoint.hip', 'csrc/h27.hip', 'csrc/refine.hip']
OUT = HERE / 'libsawfem_hip.so'

FLAGS = ['--offload-arch=gfx950', '-O3', '-std=c++17', '-fPIC', '-shared',
         # no packed-FP32 instructions in any generated code
         '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops',
         '-Wno-unused-command-line-argument']


def hipcc():
    return os.environ.get('HIPCC', '/opt/rocm/bin/hipcc')


def build(verbose=False) -> Path:
    tmp = OUT.with_suffix('.so.tmp')
    cmd = [hipcc(), *FLAGS, *[str(HERE / s) for s in SOURCES], '-I', str(HERE / 'csrc'), '-o', str(tmp)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError('hipcc failed for libsawfem_hip.so')
    if verbose:
        sys.stderr.write(res.stderr)
    os.replace(tmp, OUT)
    return OUT


if __name__ == '__main__':
    print(build(verbose=True))
