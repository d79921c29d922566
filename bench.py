"""Benchmark: one closed-shell B3LYP K_xc + E_xc build per step on one MI355X.

Workload (default): alpha-helical H-(Gly)_30-OH, 213 atoms, 6-31G(d) with 5
spherical d (1878 AOs), 99 radial x 590 Lebedev Becke grid (~10.1M points),
128-point grid blocks with per-block AO screening at 1e-12.  Each run times
the FP64 build and the split-precision builds (bf16 slices on MFMA, FP32
accumulation, FP64 elsewhere) on the same grid and P; value is the mode named
by --precision (default bf16x3k: FP64 rho pass, so E_xc is the FP64 build's,
and a hi+mid+lo K pass), and "modes" lists every mode's time, throughput and,
for the split modes, |dE_xc| and max|dK_xc| against the FP64 build.
P is the deterministic synthetic guess of proteindf_amd.guess (the reference's
own SCF is outside the XC path).  Inputs (grid, blocks, P) are resident in HBM
before the timed region; a step is one pdfxc_build_rks_device call.

value = algorithmic GFLOP/s, F = sum_b 4 n_b N_b^2 (N_b = screened AOs of
block b), divided by the wall time of a build.  The roofline entry is for the
dominant kernel, timed with HIP events on the engine's stream.  The CPU
baseline runs the C oracle (the reference algorithm: all AOs at every point)
on a bounded sample of grid points and scales to the whole grid.

    python bench.py [--gpus 1] [--steps 10] [--warmup 2] [--workload gly10|h2o8|gly20|gly30|h2o_batch]
                    [--precision fp64|bf16x2|bf16x3|bf16x3k|bfp|f16k] [--modes fp64,bf16x2,...]
                    [--no-autotune] [--split-k-patches N] [--split-k-group G] [--persistent MASK]
"""
import argparse
import ctypes
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

from proteindf_amd import calc_grid, engine, guess, molecules  # noqa: E402

PEAK_HBM_GBS = 8000.0

WORKLOADS = {
    "gly10": dict(make=lambda: molecules.polyglycine_helix(10, seed=1), n_rad=75, n_ang=302,
                  ao_screen=1e-12, label="(Gly)10 alpha-helix, 6-31G(d) 638 AOs, B3LYP, 75x302 Becke"),
    "h2o8": dict(make=lambda: molecules.water_cluster(8, seed=0), n_rad=75, n_ang=302,
                 ao_screen=0.0, label="(H2O)8, 6-31G(d) 144 AOs, B3LYP, 75x302 Becke"),
    "gly20": dict(make=lambda: molecules.polyglycine_helix(20, seed=4), n_rad=75, n_ang=302,
                  ao_screen=1e-12, label="(Gly)20 alpha-helix, 6-31G(d) 1258 AOs, B3LYP, 75x302 Becke"),
    "gly30": dict(make=lambda: molecules.polyglycine_helix(30, seed=2), n_rad=99, n_ang=590,
                  ao_screen=1e-12, label="(Gly)30 alpha-helix, 6-31G(d) 1878 AOs, B3LYP, 99x590 Becke"),
}


class Hip:
    """Minimal HIP runtime calls for device buffers (no torch needed)."""

    def __init__(self):
        self.L = ctypes.CDLL("libamdhip64.so")
        self.L.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.L.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        self.L.hipMemset.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_size_t]
        self.L.hipFree.argtypes = [ctypes.c_void_p]

    def check(self, rc, what):
        if rc != 0:
            raise RuntimeError(f"{what} failed with HIP error {rc}")

    def malloc(self, nbytes):
        p = ctypes.c_void_p()
        self.check(self.L.hipMalloc(ctypes.byref(p), nbytes), "hipMalloc")
        return p.value

    def h2d(self, dst, arr):
        self.check(self.L.hipMemcpy(dst, arr.ctypes.data, arr.nbytes, 1), "hipMemcpy H2D")

    def d2h(self, arr, src):
        self.check(self.L.hipMemcpy(arr.ctypes.data, src, arr.nbytes, 2), "hipMemcpy D2H")

    def memset(self, p, nbytes):
        self.check(self.L.hipMemset(p, 0, nbytes), "hipMemset")

    def sync(self):
        self.check(self.L.hipDeviceSynchronize(), "hipDeviceSynchronize")

    def free(self, p):
        self.L.hipFree(p)


def cpu_baseline(info, grid, PA, flops_total, target_s=12.0):
    """Oracle (reference algorithm, OpenMP) on a bounded sample of points, scaled."""
    from oracle import oracle

    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    aos = oracle.AoTable.from_basis(info)
    rng = np.random.default_rng(0)
    n = 256
    t = 0.0
    while True:
        idx = rng.choice(len(grid), size=min(n, len(grid)), replace=False)
        t0 = time.perf_counter()
        oracle.calc_xc_rks(aos, grid[idx], PA, n_threads=threads)
        t = time.perf_counter() - t0
        if t > 0.6 * target_s or n >= len(grid):
            break
        n = int(min(len(grid), n * max(2.0, 1.05 * target_s / max(t, 1e-3))))
    t_full = t * len(grid) / len(idx)
    return {"value": flops_total / t_full / 1e9, "unit": "GFLOP/s", "cores": threads, "kind": "port",
            "sample": f"{len(idx)} random grid points of {len(grid)} in {t:.1f} s, "
                      f"scaled to a {t_full:.0f} s whole-grid build"}


MODES = {"fp64": engine.PDFXC_FP64, "bf16x2": engine.PDFXC_SPLIT_BF16, "bf16x3": engine.PDFXC_SPLIT_BF16X3,
         "bf16x3k": engine.PDFXC_SPLIT_K_BF16X3, "bfp": engine.PDFXC_SPLIT_BFP, "f16k": engine.PDFXC_SPLIT_K_F16}
# (rho-pass kernel, peak TFLOP/s), (K-pass kernel, peak) per mode: dense FP64 and bf16 MFMA peaks
KERNELS = {"fp64": (("k_rho", 78.6), ("k_kmat", 78.6)),
           "bf16x2": (("k_rho_split", 2500.0), ("k_kmat_split", 2500.0)),
           "bf16x3": (("k_rho_split", 2500.0), ("k_kmat_split", 2500.0)),
           "bf16x3k": (("k_rho", 78.6), ("k_kmat_split", 2500.0)),
           "bfp": (("k_rho_bfp", 2500.0), ("k_kmat_split", 2500.0)),
           "f16k": (("k_rho", 78.6), ("k_kmat_split", 2500.0))}
LABELS = {"fp64": "FP64", "bf16x2": "split-BF16 hi+lo", "bf16x3": "split-BF16 hi+mid+lo",
          "bf16x3k": "FP64 rho pass + split-BF16 hi+mid+lo K pass",
          "bfp": "block-FP BF16 rho pass (4 slices, exact sums) + split-BF16 hi+mid+lo K pass",
          "f16k": "FP64 rho pass + split-FP16 hi+lo K pass (per-slab column scales)"}
DTYPES = {"fp64": "f64", "bf16x2": "bf16 slices, f32 accumulate, f64 AO/functional/sums",
          "bf16x3": "bf16 slices, f32 accumulate, f64 AO/functional/sums",
          "bf16x3k": "f64 rho pass; K pass bf16 slices, f32 accumulate; f64 AO/functional/sums",
          "bfp": "rho pass bf16 block-FP slices, exact f32 sums; K pass bf16 slices, f32 accumulate; "
                 "f64 AO/functional/sums",
          "f16k": "f64 rho pass; K pass fp16 slices, f32 accumulate; f64 AO/functional/sums"}
E_TARGET, K_TARGET = 1e-6, 1e-7   # split-precision accuracy target vs the FP64 build
# HBM bytes per build of a dominant kernel (the sum over its class launches),
# measured in separate rocprofv3 --pmc runs (FETCH_SIZE x 2.00 and WRITE_SIZE
# x 1.00, the factors a plain copy kernel of known size gives) of this benchmark;
# the counters are not read inside the timed run.  Each entry holds for the
# launch shapes it names (the tuning fields that change that kernel's traffic)
# and is reported only when the run's tuning matches them.
_KIB = 1024.0
TRAFFIC_MEASURED = [
    # (workload, modes, kernel, tuning fields, bytes per build, source)
    # k_rho: the persistent form of the widest class (persistent bit 0) reads less than the plain one
    *[("gly30", ("bf16x3k", "f16k", "fp64"), "k_rho", {"rho_shift": 0, "persistent": pm},
       (510766.1 + 519605.6 + 79248.1) * _KIB * 2.00 + (239687.4 + 60472.0 + 19658.0) * _KIB * 1.00,
       "rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE, 3 class launches, persistent widest class; "
       "profiles/gly30_kernels.md") for pm in (1, 3)],
    *[("gly30", ("bf16x3k", "f16k", "fp64"), "k_rho", {"rho_shift": 0, "persistent": pm},
       (631813.9 + 514037.4 + 79570.8) * _KIB * 2.00 + (235992.0 + 60472.0 + 19658.6) * _KIB * 1.00,
       "rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE, 3 class launches, plain; profiles/gly30_kernels.md")
      for pm in (0, 2)],
    ("gly30", ("bf16x3k", "bfp"), "k_kmat_split", {"split_k_patches": 24, "split_k_group": 2},
     575575.2 * _KIB * 2.00 + 14948334.8 * _KIB * 1.00,
     "rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE, groups of 2 blocks; profiles/gly30_kernels.md"),
    ("gly30", ("bf16x3k", "bfp"), "k_kmat_split", {"split_k_patches": 24, "split_k_group": 4},
     546979.2 * _KIB * 2.00 + 7839451.1 * _KIB * 1.00,
     "rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE, groups of 4 blocks; profiles/gly30_kernels.md"),
    ("gly30", ("bfp",), "k_rho_bfp", {"split_rho_shift": 0},
     (334056.1 + 1030113.3 + 210071.5) * _KIB * 2.00 + (65830.0 + 246764.0 + 3528.0) * _KIB * 1.00,
     "rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE, 3 class launches; profiles/gly30_kernels.md"),
]


def measured_traffic(workload, mode, kernel, tuning):
    """(bytes, source) of the counter run that matches this launch shape, else None."""
    for wl, modes, k, need, nbytes, src in TRAFFIC_MEASURED:
        if wl == workload and mode in modes and k == kernel and all(tuning.get(f) == v for f, v in need.items()):
            return nbytes, src
    return None


def time_mode(hip, info, grid, PA, precision, ao_screen, steps, warmup, autotune=True, k_patches=0, k_group=0,
              persistent=-1):
    """Build a context with the given precision on the same grid, let the
    autotuner pick its launch shapes (set-up, untimed), time steps builds on
    HBM-resident P/K, return (ms per build, stats, E, K dense, tuning)."""
    N = info.n_aos
    ctx = engine.XcContext(info, engine.default_opts(ao_screen=ao_screen, precision=precision))
    ctx.set_grid(grid)
    dP, dK, dE = hip.malloc(PA.nbytes), hip.malloc(PA.nbytes), hip.malloc(8)
    hip.h2d(dP, PA)
    fixed = bool(k_patches or k_group or persistent >= 0)
    pmask = persistent if persistent >= 0 else 2
    if fixed:
        ctx.set_tuning(split_k_patches=k_patches or 24, split_k_group=k_group or 4, persistent=pmask)
    if autotune and not fixed:
        tune = ctx.autotune(dP, dK, dE, reps=2).as_dict()
    else:  # the launch shapes in effect: the library's defaults, with the K item and group sizes given
        tune = dict(rho_shift=0, k_tiles=96, split_rho_shift=0, split_k_patches=k_patches or 24,
                    split_k_group=k_group or 4, persistent=pmask, ms=None)
    hip.memset(dK, PA.nbytes)
    for _ in range(warmup):
        ctx.build_rks_device(dP, dK, dE)
    hip.sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        ctx.build_rks_device(dP, dK, dE)
    hip.sync()
    dt = (time.perf_counter() - t0) / max(steps, 1)
    st = ctx.stats()
    # one more build into a zeroed K for the result
    hip.memset(dK, PA.nbytes)
    ctx.build_rks_device(dP, dK, dE)
    E = np.zeros(1)
    K = np.zeros((N, N))
    hip.d2h(E, dE)
    hip.d2h(K, dK)
    K = np.tril(K)
    K = K + np.tril(K, -1).T
    # the host-buffer entry point gives the same energy
    E_host, _, _ = ctx.build_rks(guess.pack_lower(PA))
    if abs(E[0] - E_host) > 1e-10 * max(1.0, abs(E_host)):
        raise RuntimeError(f"device/host build energies differ: {E[0]} vs {E_host}")
    ctx.close()
    for p in (dP, dK, dE):
        hip.free(p)
    return dt, st, E[0], K, tune


def _tally(tunes):
    """How many contexts the autotuner put on each launch shape."""
    if not tunes:
        return None
    out = {}
    for t in tunes:
        key = f"rho_shift={t['rho_shift']} k_tiles={t['k_tiles']} split_rho_shift={t['split_rho_shift']} " \
              f"split_k_patches={t['split_k_patches']} split_k_group={t['split_k_group']} " \
              f"persistent={t['persistent']}"
        out[key] = out.get(key, 0) + 1
    return out


def cpu_baseline_batch(ctxs, infos, Ps, target_s=12.0):
    """Oracle (reference algorithm, OpenMP) on a uniform random sample of the
    batch's points (each cluster gets a share in proportion to its points, at
    least one), scaled to the whole batch: points/s."""
    from oracle import oracle

    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    sizes = np.array([c.n_pts for c in ctxs], dtype=np.float64)
    total = sizes.sum()
    rng = np.random.default_rng(0)
    n = 2048
    while True:
        t = 0.0
        taken = 0
        for ctx, info, PA, npts in zip(ctxs, infos, Ps, sizes):
            k = int(min(npts, max(1, round(n * npts / total))))
            grid = ctx.get_grid()
            idx = rng.choice(len(grid), size=k, replace=False)
            aos = oracle.AoTable.from_basis(info)
            t0 = time.perf_counter()
            oracle.calc_xc_rks(aos, grid[idx], PA, n_threads=threads)
            t += time.perf_counter() - t0
            taken += k
        if t > 0.6 * target_s or n >= total:
            break
        n = int(min(total, n * max(2.0, 1.05 * target_s / max(t, 1e-3))))
    t_full = t * total / taken
    return {"value": total / t_full, "unit": "points/s", "cores": threads, "kind": "port",
            "sample": f"{taken} random points of the {int(total)} (each cluster's share in proportion to its points) "
                      f"in {t:.1f} s, scaled to a {t_full:.0f} s whole-batch build"}


def bench_batch(args):
    """64 water clusters (H2O)_n, n ~ U{4..64} from one stream (seed 3), 6-31G(d),
    B3LYP, 75 x 302: every cluster's blocks in one graph-captured batch vs one
    build per cluster.  Reports aggregate points/s and per-cluster latency."""
    mols = molecules.water_cluster_batch(64, 4, 64, seed=3)
    prec = MODES[args.precision]
    hip = Hip()
    ctxs, infos, Ps, singles, tunes = [], [], [], [], []
    t0 = time.perf_counter()
    for k, mol in enumerate(mols):
        if k % 8 == 0:
            print(f"[bench] set-up: cluster {k} of {len(mols)}, {time.perf_counter() - t0:.1f} s",
                  file=sys.stderr, flush=True)
        info = calc_grid.orbital_info(mol)
        ctx = engine.XcContext(info, engine.default_opts(ao_screen=1e-12, precision=prec))
        ctx.generate_grid(mol.Z, mol.xyz, 75, 302)
        PA = np.ascontiguousarray(0.5 * guess.density_guess(mol, info, seed=0, eps=0.05))
        dP, dK, dE = hip.malloc(PA.nbytes), hip.malloc(PA.nbytes), hip.malloc(8)
        hip.h2d(dP, PA)
        if not args.no_autotune:
            tunes.append(ctx.autotune(dP, dK, dE, reps=2).as_dict())
        hip.memset(dK, PA.nbytes)
        ctxs.append(ctx)
        infos.append(info)
        Ps.append(PA)
        singles.append((dP, dK, dE))
    t_setup = time.perf_counter() - t0
    n_pts = sum(c.n_pts for c in ctxs)
    batch = engine.XcBatch(ctxs)
    st = batch.stats()
    total = int(batch.dense_off[-1])
    bP, bK, bE = hip.malloc(8 * total), hip.malloc(8 * total), hip.malloc(8 * len(mols))
    for i, PA in enumerate(Ps):
        hip.h2d(bP + 8 * int(batch.dense_off[i]), PA)
    hip.memset(bK, 8 * total)
    print(f"[bench] 64 clusters: {n_pts} points, {st.n_blocks} blocks, N {min(i.n_aos for i in infos)}.."
          f"{max(i.n_aos for i in infos)}, set-up {t_setup:.1f} s", file=sys.stderr, flush=True)
    for _ in range(args.warmup):
        batch.build_rks_device(bP, bK, bE)
    hip.sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        batch.build_rks_device(bP, bK, bE)
    hip.sync()
    t_batch = (time.perf_counter() - t0) / max(args.steps, 1)
    # one build per cluster, each timed on its own
    lat = []
    for _ in range(args.warmup):
        for ctx, (dP, dK, dE) in zip(ctxs, singles):
            ctx.build_rks_device(dP, dK, dE)
    hip.sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        for ctx, (dP, dK, dE) in zip(ctxs, singles):
            t1 = time.perf_counter()
            ctx.build_rks_device(dP, dK, dE)  # synchronises before returning
            lat.append(time.perf_counter() - t1)
    t_single = (time.perf_counter() - t0) / max(args.steps, 1)
    # the batch's energies are the single builds' energies
    Eb = np.zeros(len(mols))
    hip.d2h(Eb, bE)
    Es = np.zeros(len(mols))
    for i, (dP, dK, dE) in enumerate(singles):
        e = np.zeros(1)
        hip.d2h(e, dE)
        Es[i] = e[0]
    if not np.array_equal(Eb, Es):
        raise RuntimeError(f"batched and single energies differ: max {np.abs(Eb - Es).max()}")
    lat_ms = np.array(lat) * 1e3
    label = LABELS[args.precision]
    line = {
        "metric": f"batched K_xc+E_xc build points/s, 64 water clusters, {label} GPU",
        "value": n_pts / t_batch,
        "unit": "points/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": t_batch * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": DTYPES[args.precision],
        "data": "synthetic (64 (H2O)_n clusters, n ~ U{4..64}, seed 3; deterministic density guess)",
        "config": {"workload": "64 (H2O)_n clusters, 6-31G(d), B3LYP, 75x302 Becke, one graph-captured batch",
                   "precision": args.precision, "n_points": n_pts, "n_blocks": int(st.n_blocks),
                   "algorithmic_gflop": st.flops / 1e9, "ao_screen": 1e-12},
        "batch": {"ms": t_batch * 1e3, "points_per_s": n_pts / t_batch, "gflops": st.flops / t_batch / 1e9,
                  "cluster_latency_ms": t_batch * 1e3},
        "one_launch_per_cluster": {"ms": t_single * 1e3, "points_per_s": n_pts / t_single,
                                   "cluster_latency_ms_p50": float(np.percentile(lat_ms, 50)),
                                   "cluster_latency_ms_p99": float(np.percentile(lat_ms, 99))},
        "speedup_batch_vs_single": t_single / t_batch,
        "tuning_chosen": _tally(tunes),
        # the captured graph gives no per-pass times: the whole build against
        # the FP64 MFMA peak of its rho pass (a lower bound of that pass's fraction)
        "roofline": {"kernel": "batch graph (rho and K passes of all 64 clusters)", "bound": "mfma",
                     "achieved": st.flops / 2.0 / t_batch / 1e12, "peak": KERNELS[args.precision][0][1],
                     "unit": "TFLOP/s", "frac": st.flops / 2.0 / t_batch / 1e12 / KERNELS[args.precision][0][1],
                     "traffic": None, "traffic_note": "no counter run for the batch"},
        "cpu_baseline": None if args.no_cpu_baseline else cpu_baseline_batch(ctxs, infos, Ps),
    }
    print(json.dumps(line), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="gly30", choices=sorted(WORKLOADS) + ["h2o_batch"])
    ap.add_argument("--precision", default="bf16x3k", choices=sorted(MODES),
                    help="mode whose throughput is the reported value (the others are measured beside it)")
    ap.add_argument("--modes", default="fp64,bf16x2,bf16x3,bf16x3k,bfp,f16k", help="comma list of modes to measure")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-autotune", action="store_true", help="keep the default launch shapes")
    ap.add_argument("--split-k-patches", type=int, default=0, metavar="N",
                    help="split K pass items of N patches (4..32: 16 waves x 1 up to 16, 12 x 2 up to 24, "
                         "8 x 4 up to 32) instead of the autotuner's choice")
    ap.add_argument("--split-k-group", type=int, default=0, metavar="G",
                    help="split K pass work units of G blocks (1, 2, 4 or 8; 128 G points) instead of the "
                         "autotuner's choice")
    ap.add_argument("--persistent", type=int, default=-1, metavar="MASK",
                    help="persistent launches (bit 0 the rho kernels, bit 1 the K kernels) instead of the "
                         "autotuner's choice")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run the plain copy kernel over MIB MiB (HBM counter calibration)")
    args = ap.parse_args()
    if args.copy_calibration:
        hip = Hip()
        n = args.copy_calibration * (1 << 20) // 8
        src, dst = hip.malloc(8 * n), hip.malloc(8 * n)
        hip.memset(src, 8 * n)
        engine.check(engine.lib().pdfxc_selftest_copy(ctypes.c_void_p(src), ctypes.c_void_p(dst), n))
        print(json.dumps({"copy_calibration_bytes_read": 8 * n, "bytes_written": 8 * n}), flush=True)
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this engine runs on one GPU per process; the build does not "
                                   "shard across GPUs (use --gpus 1)"}), flush=True)
        sys.exit(2)
    if args.split_k_patches and not 4 <= args.split_k_patches <= 32:
        print(json.dumps({"error": "--split-k-patches takes 4..32"}), flush=True)
        sys.exit(2)
    if args.split_k_group and args.split_k_group not in (1, 2, 4, 8):
        print(json.dumps({"error": "--split-k-group takes 1, 2, 4 or 8"}), flush=True)
        sys.exit(2)
    if not -1 <= args.persistent <= 3:
        print(json.dumps({"error": "--persistent takes 0..3"}), flush=True)
        sys.exit(2)
    if args.workload == "h2o_batch":
        bench_batch(args)
        return
    modes = [m for m in args.modes.split(",") if m]
    if "fp64" not in modes:
        modes.insert(0, "fp64")  # the split errors are stated against the FP64 build
    if args.precision not in modes:
        modes.append(args.precision)
    wl = WORKLOADS[args.workload]
    mol = wl["make"]()
    info = calc_grid.orbital_info(mol)
    N = info.n_aos
    gctx = engine.XcContext(info, engine.default_opts(ao_screen=wl["ao_screen"]))
    t0 = time.perf_counter()
    n_pts = gctx.generate_grid(mol.Z, mol.xyz, wl["n_rad"], wl["n_ang"])
    t_grid = time.perf_counter() - t0
    grid = gctx.get_grid()
    st0 = gctx.stats()
    gctx.close()
    print(f"[bench] {wl['label']}: {mol.n_atoms} atoms, {N} AOs, {n_pts} points, {st0.n_blocks} blocks, "
          f"mean N_b {st0.mean_nao:.1f}, max N_b {st0.max_nao}, grid generation {t_grid:.2f} s",
          file=sys.stderr, flush=True)

    PA = np.ascontiguousarray(0.5 * guess.density_guess(mol, info, seed=0, eps=0.05))
    hip = Hip()
    res = {}
    for m in modes:
        dt, st, E, K, tune = time_mode(hip, info, grid, PA, MODES[m], wl["ao_screen"], args.steps, args.warmup,
                                       not args.no_autotune, args.split_k_patches, args.split_k_group,
                                       args.persistent)
        res[m] = dict(dt=dt, st=st, E=E, K=K, k_ms=list(st.last_kernel_ms), tune=tune)
        print(f"[bench] {m:7s} E_xc {E:.10f}  build {dt * 1e3:.2f} ms  kernels ms: "
              + ", ".join(f"{t:.2f}" for t in res[m]["k_ms"][:3]), file=sys.stderr, flush=True)
    ref = res["fp64"]
    summary = {}
    for m in modes:
        r = res[m]
        e = {"ms": r["dt"] * 1e3, "gflops": r["st"].flops / r["dt"] / 1e9, "tuning": r["tune"]}
        if m != "fp64":
            e["abs_dE"] = abs(r["E"] - ref["E"])
            e["max_abs_dK"] = float(np.abs(r["K"] - ref["K"]).max())
            e["speedup_vs_fp64"] = ref["dt"] / r["dt"]
            e["meets_1e-6_1e-7"] = bool(e["abs_dE"] <= E_TARGET and e["max_abs_dK"] <= K_TARGET)
        summary[m] = e
    head = res[args.precision]
    st = head["st"]
    dt = head["dt"]

    # dominant kernel of the reported mode: per-build algorithmic flops of its
    # contraction (half of 4 n_b N_b^2) over its HIP-event time
    k_ms = head["k_ms"]
    dom = int(np.argmax(k_ms[:2]))
    name, peak = KERNELS[args.precision][dom]
    achieved = st.flops / 2.0 / (k_ms[dom] / 1e3) / 1e12
    tr = measured_traffic(args.workload, args.precision, name, head["tune"])
    roof = {"kernel": name, "bound": "mfma", "achieved": achieved, "peak": peak,
            "unit": "TFLOP/s", "frac": achieved / peak, "traffic": tr[0] if tr else None}
    if tr:
        roof["traffic_source"] = tr[1]
    else:
        roof["traffic_note"] = "no counter run for this workload and launch shape"

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(info, grid, PA, st.flops)

    label = LABELS[args.precision]
    line = {
        "metric": f"K_xc+E_xc build GFLOP/s (algorithmic 4 n_b N_b^2), {label} GPU",
        "value": st.flops / dt / 1e9,
        "unit": "GFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": dt * 1e3,
        "build_s": dt,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": DTYPES[args.precision],
        "data": "synthetic (NeRF alpha-helix geometry, deterministic density guess)",
        "config": {"workload": wl["label"], "precision": args.precision, "n_atoms": mol.n_atoms, "n_aos": N,
                   "n_points": n_pts, "n_blocks": int(st.n_blocks), "block_points": 128,
                   "ao_screen": wl["ao_screen"], "mean_block_aos": round(st.mean_nao, 1),
                   "algorithmic_gflop": st.flops / 1e9, "E_xc": head["E"]},
        "modes": summary,
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
