"""No HIP kernel of the library uses scratch (private) memory.

Kernels that spill need the runtime to back scratch for every wave in flight; with a large front
workspace resident that backing can fail and abort the queue.  The check compiles every source
for gfx950 (device code only) with the compiler's resource-usage remarks and reads ScratchSize.
"""
import os
import re
import shutil
import subprocess
from pathlib import Path

import pytest

from sawsim_amd.build import FLAGS, HERE, SOURCES, hipcc

pytestmark = pytest.mark.skipif(shutil.which(hipcc()) is None and not os.path.exists(hipcc()),
                                reason='hipcc is not available')


@pytest.mark.parametrize('src', SOURCES)
def test_no_kernel_uses_scratch(src, tmp_path):
    flags = [f for f in FLAGS if f not in ('-shared', '-fPIC')]
    cmd = [hipcc(), *flags, '--offload-device-only', '-c', '-Rpass-analysis=kernel-resource-usage',
           str(HERE / src), '-I', str(HERE / 'csrc'), '-o', str(tmp_path / 'k.o')]
    res = subprocess.run(cmd, capture_output=True, text=True, timeout=600)
    assert res.returncode == 0, res.stderr[-2000:]
    names = re.findall(r'Function Name: (\S+)', res.stderr)
    scratch = [int(v) for v in re.findall(r'ScratchSize \[bytes/lane\]: (\d+)', res.stderr)]
    assert names and len(names) == len(scratch)
    spilling = [(n, s) for n, s in zip(names, scratch) if s]
    assert not spilling, spilling
