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
    for workload, modes, kernel, need, nbytes, source in bench.TRAFFIC_MEASURED:
        assert workload in bench.WORKLOADS and nbytes > 0 and "profiles/" in source
        for mode in modes:
            assert mode in bench.MODES and kernel in {k for k, _ in bench.KERNELS[mode]}
        assert need and all(isinstance(v, int) for v in need.values())


def test_traffic_is_reported_only_for_the_profiled_launch_shapes():
    sys.path.insert(0, str(ROOT))
    import bench

    tune = {"rho_shift": 0, "k_tiles": 96, "split_rho_shift": 0, "split_k_patches": 24, "split_k_group": 2,
            "persistent": 2}
    g2 = bench.measured_traffic("gly30", "bf16x3k", "k_kmat_split", tune)
    g4 = bench.measured_traffic("gly30", "bf16x3k", "k_kmat_split", dict(tune, split_k_group=4))
    assert g2 and g4 and g2[0] > g4[0]  # a wider group scatters fewer atomics
    assert bench.measured_traffic("gly30", "bf16x3k", "k_kmat_split", dict(tune, split_k_group=8)) is None
    assert bench.measured_traffic("gly30", "bf16x3k", "k_rho", dict(tune, rho_shift=1)) is None
    assert bench.measured_traffic("gly10", "bf16x3k", "k_rho", tune) is None
    plain = bench.measured_traffic("gly30", "bf16x3k", "k_rho", tune)
    persistent = bench.measured_traffic("gly30", "bf16x3k", "k_rho", dict(tune, persistent=3))
    assert plain and persistent and persistent[0] < plain[0]


def test_split_k_group_out_of_range_is_refused():
    r = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--split-k-group", "3"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "1, 2, 4 or 8" in json.loads(r.stdout.strip().splitlines()[-1])["error"]


def test_persistent_mask_out_of_range_is_refused():
    r = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--persistent", "4"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "0..3" in json.loads(r.stdout.strip().splitlines()[-1])["error"]


def test_default_workload_is_the_headline_configuration():
    sys.path.insert(0, str(ROOT))
    import bench

    ap_default = [a for a in open(ROOT / "bench.py").read().splitlines() if '"--workload", default=' in a]
    assert ap_default and 'default="gly30"' in ap_default[0]
    wl = bench.WORKLOADS["gly30"]
    assert (wl["n_rad"], wl["n_ang"]) == (99, 590)


def test_batch_cpu_baseline_samples_every_cluster():
    """The batch's CPU baseline draws points from every cluster in proportion
    to its size and reports points/s for the whole batch (oracle on the CPU,
    stand-in contexts that only hold a grid)."""
    sys.path.insert(0, str(ROOT))
    import bench
    import numpy as np
    from oracle import oracle
    from proteindf_amd import calc_grid, guess, molecules

    class Grid:
        def __init__(self, g):
            self.g = g
            self.n_pts = len(g)

        def get_grid(self):
            return self.g

    ctxs, infos, Ps = [], [], []
    for n in (1, 2, 3):
        mol = molecules.water_cluster(n, seed=n)
        info = calc_grid.orbital_info(mol)
        g, _ = oracle.generate_grid(mol.Z, mol.xyz, 20, 26)
        ctxs.append(Grid(g))
        infos.append(info)
        Ps.append(0.5 * guess.density_guess(mol, info, seed=0, eps=0.05))
    out = bench.cpu_baseline_batch(ctxs, infos, Ps, target_s=0.2)
    assert out["unit"] == "points/s" and out["value"] > 0 and out["kind"] == "port"
    assert str(sum(c.n_pts for c in ctxs)) in out["sample"]
    assert np.isfinite(out["value"])
