"""Every user body the GPU tests compile goes through hipRTC on the CPU
first (cbm_compile_check, no GPU needed), in a child process so that a
compiler abort fails this test instead of ending the run: once with the
HIP runtime and hipRTC of the ROCm install, and once after importing torch,
which loads the runtime and hipRTC it bundles (an older ROCm) and to which
the library then binds, as it does in the GPU test processes.  The older
compiler's instruction selection aborted the process on a medium skeleton
that indexed kernel-argument arrays at run time."""
import os
import subprocess
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.mark.parametrize("torch_first", [False, True], ids=["rocm", "torch-bundled"])
def test_every_test_body_compiles(torch_first):
    cmd = [sys.executable, os.path.join(HERE, "rtc_bodies_check.py")] + (["--torch"] if torch_first else [])
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, f"exit {r.returncode}\n{r.stdout[-4000:]}\n{r.stderr[-4000:]}"
    assert " compiled" in r.stdout
