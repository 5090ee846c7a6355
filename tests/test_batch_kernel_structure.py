"""Compiles the batched step kernel to gfx950 assembly and checks the shape of
its persistent loop: one outermost loop per kernel, entered with every lane
active (no exec-mask split in its header), and no out-of-line calls.  A
build in which the thread-0 claim and the loop's barriers end up in
different loops (lane 0 leaving an inner loop that starts with a barrier)
re-runs a stale item forever; this test keeps that shape out of the
library."""
import re
import shutil
import subprocess
from pathlib import Path

import pytest

CSRC = Path(__file__).resolve().parents[1] / "opensayal_amd" / "csrc"
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"


@pytest.mark.skipif(not Path(HIPCC).exists(), reason="hipcc not installed")
def test_persistent_loop_is_a_single_uniform_loop(tmp_path):
    out = tmp_path / "kb.s"
    cmd = [HIPCC, "-O3", "-std=c++17", "-ffp-contract=off", "-fno-fast-math",
           "--offload-arch=gfx950", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
           "-fno-gpu-rdc", "-x", "hip", "-S", "--cuda-device-only",
           str(CSRC / "kernels_batch.hip"), "-o", str(out)]
    subprocess.run(cmd, check=True, capture_output=True, cwd=tmp_path)
    lines = out.read_text().split("\n")
    starts = [i for i, l in enumerate(lines) if re.match(r"^_Z\w*k_batch_step\w*:", l)]
    assert len(starts) == 18
    for a in starts:
        end = next(i for i in range(a, len(lines)) if "s_endpgm" in lines[i])
        body = lines[a:end]
        headers = [i for i, l in enumerate(body) if "This Loop Header: Depth=1" in l]
        assert len(headers) == 1, body[0]
        block = []
        for l in body[headers[0] + 1:]:
            if re.match(r"^(\.LBB|; %bb)", l):
                break
            block.append(l)
        # no write to the exec mask in the loop header
        assert not any(re.search(r"saveexec|^\s+s_\w+\s+exec,", l) for l in block), body[0]
        assert not any("s_swappc_b64" in l for l in body), body[0]
