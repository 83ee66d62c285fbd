"""The package loaded before torch: libkmst must share torch's HIP runtime
(torch ships its own under the same soname), or torch finds no GPU after the
library has started the system one."""
import os
import subprocess
import sys

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_before_torch():
    code = ("import kamsta_amd as k\n"
            "ctx = k.Context(0)\n"
            "s, d, w = k.gen_gnm_host(1024, 8192, 254, 1)\n"
            "import torch\n"
            "t = torch.empty(16, dtype=torch.int32, device='cuda')\n"
            "t.fill_(1)\n"
            "torch.cuda.synchronize()\n"
            "print('ok', int(t.sum()))\n")
    r = subprocess.run([sys.executable, "-c", code], cwd=ROOT, capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0 and "ok 16" in r.stdout, r.stderr[-2000:]
