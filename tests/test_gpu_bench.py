"""bench.py end to end on a small workload with fixed launch shapes: the JSON
line states the shapes in effect and no counter traffic (none was measured
for this workload)."""
import json
import subprocess
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]


@pytest.mark.gpu
@pytest.mark.parametrize("extra", [["--no-autotune"], ["--split-k-patches", "24", "--split-k-group", "2"],
                                   ["--no-autotune", "--persistent", "3"]])
def test_bench_line_with_fixed_launch_shapes(extra):
    cmd = [sys.executable, str(ROOT / "bench.py"), "--workload", "h2o8", "--steps", "1", "--warmup", "0",
           "--modes", "fp64,bf16x3k", "--no-cpu-baseline"] + extra
    print("[case]", " ".join(cmd[2:]), flush=True)
    out = subprocess.run(cmd, capture_output=True, text=True, timeout=240, cwd=ROOT)
    assert out.returncode == 0, out.stderr[-3000:]
    line = json.loads(out.stdout.strip().splitlines()[-1])
    t = line["modes"]["bf16x3k"]["tuning"]
    assert t["split_k_patches"] == 24
    assert t["split_k_group"] == (2 if "--split-k-group" in extra else 4)
    assert t["persistent"] == (3 if "--persistent" in extra else 2)
    assert line["roofline"]["traffic"] is None and "traffic_note" in line["roofline"]
    assert line["modes"]["bf16x3k"]["abs_dE"] == 0.0
    assert line["modes"]["bf16x3k"]["max_abs_dK"] <= 1e-7
