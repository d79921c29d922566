"""bench.py's command line without a GPU: more than one GPU and an out-of-range
split K item size are refused with a message (the build does not shard across
GPUs), and the mode tables agree."""
import json
import subprocess
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parents[1]


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 2
    msg = json.loads(r.stdout.strip().splitlines()[-1])
    assert "one GPU" in msg["error"]


def test_split_k_item_size_out_of_range_is_refused():
    r = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--split-k-patches", "40"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "4..32" in json.loads(r.stdout.strip().splitlines()[-1])["error"]


def test_mode_tables_cover_every_mode():
    sys.path.insert(0, str(ROOT))
    import bench

    assert set(bench.MODES) == set(bench.KERNELS) == set(bench.LABELS) == set(bench.DTYPES)
    from proteindf_amd import engine

    assert {bench.MODES[m] for m in bench.MODES} == set(engine.PRECISIONS.values())
    for (workload, mode, kernel), (nbytes, source) in bench.TRAFFIC_MEASURED.items():
        assert workload in bench.WORKLOADS and mode in bench.MODES
        assert kernel in {k for k, _ in bench.KERNELS[mode]} and nbytes > 0 and "profiles/" in source
