"""Build check (CPU): no kernel of the library keeps a private (scratch)
segment.  Kernels that use the Montgomery multiply's inline assembly are only
validated without register spills (a spilling kernel computed wrong values on
gfx950), so every kernel must fit its registers; hipcc's device assembly
states each kernel's private segment size."""
import os
import re
import shutil
import subprocess
import tempfile
from concurrent.futures import ThreadPoolExecutor

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "era-bellman-cuda_amd", "csrc")
HIPCC = "/opt/rocm/bin/hipcc"
FLAGS = ["-O3", "-std=c++17", "--offload-arch=gfx950", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "--cuda-device-only", "-S"]


def private_segments(src, out_dir):
    out = os.path.join(out_dir, os.path.basename(src) + ".s")
    subprocess.run([HIPCC] + FLAGS + ["-o", out, src], check=True, capture_output=True)
    text = open(out).read()
    kernels = re.findall(r"\.amdhsa_kernel (\S+)(.*?)\.end_amdhsa_kernel", text, flags=re.S)
    return [(name, int(re.search(r"\.amdhsa_private_segment_fixed_size (\d+)", body).group(1))) for name, body in kernels]


@pytest.mark.skipif(not os.path.exists(HIPCC) or shutil.which("ld.lld") is None and not os.path.exists("/opt/rocm/llvm/bin/ld.lld"),
                    reason="hipcc not available")
def test_no_kernel_spills_to_scratch():
    sources = sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".hip"))
    with tempfile.TemporaryDirectory() as d, ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 1)) as pool:
        results = list(pool.map(lambda s: private_segments(s, d), sources))
    kernels = [k for r in results for k in r]
    assert len(kernels) > 40
    spilling = [(name, size) for name, size in kernels if size > 0]
    assert not spilling, spilling
