"""Benchmark: batched FP32 STEM multislice on one MI355X.

Workload (default): Au [001], a = 4.078 A, 8x8x20 cells, 512^2 grid, 41
slices, 300 kV probe of 21.4 mrad, 64x64 scan over one unit cell, ADF ring
80-200 mrad, 16 frozen-phonon configurations.  A step is one full run of that
workload: 64*64*16*41 = 2 686 976 probe.slices.

    python bench.py [--gpus 1] [--steps K] [--warmup W]

Prints one JSON line.  value = whole-job probe.slices/s (wall clock of the
timed steps).  roofline: the dominant slice kernel's algorithmic bytes per
launch over its mean duration (HIP events on the engine stream).
cpu_baseline: the CPU oracle (a C port of the reference path, FP32, one
thread per core) on 8x8 probes x 1 configuration of the same workload.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--ns", type=int, default=64, help="scan points per side")
    ap.add_argument("--nconf", type=int, default=16)
    ap.add_argument("--n", type=int, default=512, help="grid size")
    ap.add_argument("--sto", action="store_true",
                    help="SrTiO3 [001] workload of the uneven-size group at grid --n (256: 6x6x10 cells, "
                         "512: 12x12x25, 1024: 24x24x50) instead of Au[001] (for profiling)")
    ap.add_argument("--batch", type=int, default=0, help="waves per batch (0: planner)")
    ap.add_argument("--no-graph", action="store_true")
    ap.add_argument("--wave-pad", type=int, default=-1,
                    help="row pitch pad of the wave batches in elements (-1: the engine's default)")
    ap.add_argument("--variants", default="", help="row,col,col_det launch variants to force (default: tuner)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-ns", type=int, default=24)
    ap.add_argument("--no-group", action="store_true", help="skip the uneven-size group measurement")
    ap.add_argument("--group-batch", type=int, default=0,
                    help="waves per batch of each group member (0: each member's planner and tuner)")
    ap.add_argument("--no-surrogate", action="store_true", help="skip the surrogate measurement")
    ap.add_argument("--surrogate-scan128", action="store_true",
                    help="also time the surrogate on one 128x128-scan specimen (about a minute of GPU time)")
    ap.add_argument("--no-fp64", action="store_true", help="skip the FP64-path measurement")
    ap.add_argument("--c4", action="store_true",
                    help="only time the C4 CBED run (Pt[110] 1024^2, 20 configurations, 325 slices) and exit")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run the 1 GiB device copy used to calibrate HBM counters")
    return ap.parse_args()


def sync():
    """Device-wide barrier on the engine's HIP runtime.  (torch is not imported:
    it bundles a second HIP runtime whose synchronize would not order the
    engine's work, and two runtimes in one process confuse tracers.)"""
    from multem_amd import engine
    engine.synchronize()


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(args):
    """The FP32 C oracle, one thread per core (the transmission build over
    slices, then the probes over threads), on the cores this process may
    use up to the host's CPU share of one GPU (16 cores per GPU on these
    hosts, whatever nproc shows); nproc and the CPU model are recorded."""
    from multem_amd import configs
    from oracle import oracle
    p = configs.au_stem_512(ns=args.cpu_ns, nconf=1, n=args.n)
    cores = oracle.default_threads()
    nsl = len(oracle.spec_slicing(p))
    t0 = time.perf_counter()
    oracle.run(p, "f32", nthreads=cores)
    dt = time.perf_counter() - t0
    ps = args.cpu_ns * args.cpu_ns * nsl
    return {"value": ps / dt, "unit": "probe·slices/s", "cores": cores, "kind": "port",
            "nproc": os.cpu_count(), "cpu_model": cpu_model(),
            "sample": "Au[001] 512^2, %dx%d probes x 1 configuration x %d slices (%d probe.slices), "
                      "FP32 C oracle (own radix-2 FFT over 16 sequences at a time, AVX2 build; FFTW absent), %.1f s" % (args.cpu_ns, args.cpu_ns, nsl, ps, dt)}


def group_measurement(group_batch=0):
    """SrTiO3 [001] at 256^2 / 512^2 / 1024^2 (21 / 51 / 101 slices), 16x16
    probes and 4 phonon configurations each, group_batch probes per batch
    (0: each member's planner and tuner pick its width): one group
    run (one hipGraph launch per round) against the three runs one after the
    other, and each grid's own rate and column-pass roofline fraction from
    its solo run (HIP events around the passes of each configuration's
    first batch)."""
    from multem_amd import configs, engine
    ps = configs.sto_group(ns=16, nconf=4)
    nsl = [len(engine.spec_slicing(p)) for p in ps]
    work_i = [16 * 16 * 4 * n for n in nsl]
    work = sum(work_i)
    engine.set_batch(group_batch, True)
    per_grid = []
    try:
        engine.run_multislice_group(ps)          # warm-up (tuning, graphs)
        sync()
        t0 = time.perf_counter()
        engine.run_multislice_group(ps)
        sync()
        t_group = time.perf_counter() - t0
        plan = engine.last_plan()
        for p in ps:
            engine.run_multislice(p)
        sync()
        t_seq = 0.0
        for p, w in zip(ps, work_i):
            t0 = time.perf_counter()
            engine.run_multislice(p)
            sync()
            dt = time.perf_counter() - t0
            t_seq += dt
            t = engine.last_timing()
            pl = engine.last_plan()
            B = int([kv.split("=")[1] for kv in pl.split() if kv.startswith("batch=")][0])
            N2 = p.nx * p.ny
            col_ms = t["ms_col"] / max(t["n_col"], 1)
            row_ms = t["ms_row"] / max(t["n_row"], 1)
            col_gbs = 16.0 * N2 * B / (col_ms * 1e-3) / 1e9 if col_ms > 0 else None
            per_grid.append({"grid": "%d^2" % p.nx, "slices": nsl[len(per_grid)], "probe_slices": w,
                             "value": w / dt, "unit": "probe·slices/s", "batch": B,
                             "row_pass_ms": row_ms, "col_pass_ms": col_ms,
                             "col_pass_alg_GBs": col_gbs,
                             "col_pass_roofline_frac": col_gbs / HBM_PEAK_GBS if col_gbs else None,
                             "slice_loop_alg_GBs": 40.0 * N2 * t["probe_slices"] / (t["ms_slice"] * 1e-3) / 1e9
                             if t["ms_slice"] > 0 else None,
                             "plan": pl})
    finally:
        engine.set_batch(0, True)
    return {"workload": "sto_group: SrTiO3[001] 256^2/512^2/1024^2, %s slices, 16x16 probes x 4 configs each, "
                        "%s" % ("/".join(str(n) for n in nsl),
                                "%d probes per batch" % group_batch if group_batch > 0 else "planner batch widths"),
            "probe_slices": work, "value": work / t_group, "unit": "probe·slices/s",
            "sequential_value": work / t_seq, "plan": plan, "per_grid": per_grid}


def c4_measurement():
    """C4: Pt[110] CBED at 1024^2, 20 phonon configurations, full depth (325
    slices, thicknesses every 50 A to 450 A), FP32 wave path: wall time of one
    run after a warm-up (tuning, tables), and the transmission build's
    share (HIP events).  One JSON line; not the headline metric."""
    from multem_amd import configs, engine
    p = configs.pt_cbed_1024(nconf=20, n=1024)
    engine.run_multislice(p)
    sync()
    t0 = time.perf_counter()
    engine.run_multislice(p)
    sync()
    dt = time.perf_counter() - t0
    t = engine.last_timing()
    print(json.dumps({"metric": "C4 CBED run time", "value": dt, "unit": "s", "higher_is_better": False,
                      "workload": "pt_cbed_1024: Pt[110] 10x14 cells, 1024^2, 200 kV, 10 mrad, -100 A defocus, "
                                  "20 configurations, 325 slices, 9 thicknesses",
                      "ms_total_device": t["ms_total"], "plan": engine.last_plan()}))
    sys.stdout.flush()


def fp64_measurement():
    """The FP64 wave path (system_conf.precision = 2) on the same Au[001]
    512^2 STEM workload, 32x32 probes x 2 configurations: probe.slices/s and
    the column pass's algorithmic HBM rate (32 N^2 B per wave)."""
    from multem_amd import configs, engine
    p = configs.au_stem_512(ns=32, nconf=2, n=512)
    p.system_conf.precision = 2
    nsl = len(engine.spec_slicing(p))
    engine.run_multislice(p)                   # warm-up (tuning, graphs)
    sync()
    t0 = time.perf_counter()
    engine.run_multislice(p)
    sync()
    wall = time.perf_counter() - t0
    t = engine.last_timing()
    plan = engine.last_plan()
    B = int([kv.split("=")[1] for kv in plan.split() if kv.startswith("batch=")][0])
    col_ms = t["ms_col"] / max(t["n_col"], 1)
    return {"workload": "au_stem_512 FP64: 512^2, %d slices, 32x32 probes x 2 configs" % nsl,
            "value": 32 * 32 * 2 * nsl / wall, "unit": "probe·slices/s", "dtype": "f64", "plan": plan,
            "col_pass_ms": col_ms,
            "col_pass_alg_GBs": 32.0 * 512 * 512 * B / (col_ms * 1e-3) / 1e9 if col_ms > 0 else None}


SURROGATE_WEIGHTS = os.path.join(ROOT, "multem_amd", "data", "surrogate_6x32.npz")


def surrogate_measurement(scan128=False):
    """Frozen-phonon surrogate on 4 held-out random specimens (32x32 scans,
    256^2 grid, illumination tilt up to 1 degree): GPU time of plain
    64-configuration multislice over that of the surrogate path
    (2-configuration run + potential map + network), and relative RMS errors
    against the 64-configuration image and against an independent
    256-configuration image (another phonon seed; the 64-configuration
    image's own error is listed beside), with the errors against the
    converged image estimated by removing the 256-configuration image's own
    noise (surrogate.evaluate); with scan128, also one specimen at a 128x128
    scan (4 x 4 unit cells, the training's 32 points per cell) for the time
    ratio and the error against its 64-configuration image."""
    from multem_amd import surrogate
    model = surrogate.Surrogate.load(SURROGATE_WEIGHTS)
    ev = surrogate.evaluate(model, n_specimens=4, seed=2024, n_ref=256, tilt_max=1.0)
    out = {"metric": "surrogate ADF: plain 64-config GPU time / surrogate GPU time",
            "gpu_time_ratio": ev["gpu_time_ratio"],
            "rel_rms_surrogate_vs_64conf": ev["rel_rms_surrogate_mean"],
            "rel_rms_2conf_vs_64conf": ev["rel_rms_small_mean"],
            "rel_rms_surrogate_vs_256conf": ev["rel_rms_surrogate_ref_mean"],
            "rel_rms_2conf_vs_256conf": ev["rel_rms_small_ref_mean"],
            "rel_rms_64conf_vs_256conf": ev["rel_rms_big_ref_mean"],
            "rel_rms_vs_converged_est": {"surrogate": ev["rel_rms_surrogate_conv_est"],
                                         "2conf": ev["rel_rms_small_conv_est"],
                                         "64conf": ev["rel_rms_big_conv_est"]},
            "ms_plain": sum(r["ms_plain"] for r in ev["specimens"]),
            "ms_surrogate": sum(r["ms_surrogate"] for r in ev["specimens"]),
            "specimens": ["%s %.0f A tilt %.2f deg" % (r["desc"], r["thick"], r["theta"]) for r in ev["specimens"]]}
    if scan128:
        e128 = surrogate.evaluate(model, n_specimens=1, seed=2025, ns=128, tilt_max=1.0)
        r128 = e128["specimens"][0]
        out["scan_128"] = {"gpu_time_ratio": e128["gpu_time_ratio"],
                           "rel_rms_surrogate_vs_64conf": e128["rel_rms_surrogate_mean"],
                           "rel_rms_2conf_vs_64conf": e128["rel_rms_small_mean"],
                           "ms_plain": r128["ms_plain"], "ms_surrogate": r128["ms_surrogate"],
                           "specimen": "%s %.0f A tilt %.2f deg, 128x128 scan over 4x4 cells"
                                       % (r128["desc"], r128["thick"], r128["theta"])}
    return out


def copy_calibration():
    """Device-to-device copy of 1 GiB on the engine's HIP runtime: the known-size
    reference for rocprofv3's FETCH_SIZE / WRITE_SIZE, and the achievable copy
    bandwidth on this GPU."""
    import ctypes as C
    from multem_amd import engine
    engine.lib()
    hip = C.CDLL("libamdhip64.so.7")
    n = 1 << 30
    a, b = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(a), C.c_size_t(n)) == 0 and hip.hipMalloc(C.byref(b), C.c_size_t(n)) == 0
    assert hip.hipMemset(a, 1, C.c_size_t(n)) == 0
    hip.hipDeviceSynchronize()
    reps = 10
    t0 = time.perf_counter()
    for _ in range(reps):
        assert hip.hipMemcpy(b, a, C.c_size_t(n), 3) == 0     # hipMemcpyDeviceToDevice
    hip.hipDeviceSynchronize()
    dt = (time.perf_counter() - t0) / reps
    hip.hipFree(a)
    hip.hipFree(b)
    print(json.dumps({"copy_bytes_per_launch": n, "read_plus_write_GBs": 2 * n / dt / 1e9, "ms": dt * 1e3,
                      "stream_kernel_read_plus_write_GBs": engine.stream_copy_bandwidth(n, reps)}))


def traffic_from_profiles(kernel, workload, n, batch):
    """HBM bytes per launch measured with rocprofv3 --pmc (FETCH_SIZE + WRITE_SIZE,
    scaled by the copy-kernel calibration), if a summary has been recorded:
    the recorded bytes per wave times this run's batch width."""
    path = os.path.join(ROOT, "profiles", "traffic_%s_%s_%d.txt" % (kernel, workload, n))
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("bytes_per_wave"):
            return float(line.split("=")[1].split()[0]) * batch
    return None


def main():
    args = parse()
    if args.copy_calibration:
        copy_calibration()
        return
    if args.c4:
        c4_measurement()
        return
    if args.gpus != 1:
        print("bench.py: the engine runs on one GPU; --gpus %d is not provided (several GPUs would each run "
              "a share of the phonon configurations or probe tiles and reduce the detector images)" % args.gpus,
              file=sys.stderr)
        sys.exit(2)
    from multem_amd import configs, engine
    if args.sto:
        cells = {256: (6, 10), 512: (12, 25), 1024: (24, 50)}[args.n]
        p = configs.sto_stem(n=args.n, ncells=cells[0], nc=cells[1], ns=args.ns, nconf=args.nconf)
    else:
        p = configs.au_stem_512(ns=args.ns, nconf=args.nconf, n=args.n)
    engine.set_batch(args.batch, not args.no_graph)
    engine.set_wave_pad(args.wave_pad)
    if args.variants:
        engine.set_pass_variants(*[int(v) for v in args.variants.split(",")])
    nsl = len(engine.spec_slicing(p))

    for _ in range(args.warmup):
        engine.run_multislice(p)
    sync()
    t0 = time.perf_counter()
    rows = cols = 0.0
    nrow = ncol = 0
    dev_ms = 0.0
    ps_total = 0.0
    for k in range(args.steps):
        engine.run_multislice(p)
        t = engine.last_timing()
        print("bench: step %d done, device %.1f ms" % (k, t["ms_total"]), file=sys.stderr, flush=True)
        rows += t["ms_row"]
        cols += t["ms_col"]
        nrow += t["n_row"]
        ncol += t["n_col"]
        dev_ms += t["ms_slice"]
        ps_total += t["probe_slices"]
    sync()
    wall = time.perf_counter() - t0
    ps_step = args.ns * args.ns * args.nconf * nsl
    value = ps_step * args.steps / wall
    plan = engine.last_plan()
    B = int([kv.split("=")[1] for kv in plan.split() if kv.startswith("batch=")][0])
    N2 = args.n * args.n
    # algorithmic bytes per launch: row pass reads+writes the wave and reads T
    # (24 N^2 B per wave), column pass reads+writes the wave (16 N^2 B)
    row_ms, col_ms = rows / max(nrow, 1), cols / max(ncol, 1)
    row_bytes, col_bytes = 24.0 * N2 * B, 16.0 * N2 * B
    if row_ms * nrow >= col_ms * ncol:
        kern, kb, kms = "row_pass", row_bytes, row_ms
    else:
        kern, kb, kms = "col_pass", col_bytes, col_ms
    achieved = kb / (kms * 1e-3) / 1e9 if kms > 0 else None   # None when engine events are off
    slice_loop_gbs = 40.0 * N2 * ps_total / (dev_ms * 1e-3) / 1e9 if dev_ms > 0 else None
    out = {
        "metric": "STEM probe·slices/s (+ HBM roofline %)",
        "value": value,
        "unit": "probe·slices/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (crystal built in-process, seeded frozen phonons)",
        "config": {
            "workload": ("sto_stem: SrTiO3[001] %d^2 grid, %d slices, 300 kV 25 mrad probe, %dx%d scan, %d phonon "
                         "configs, ADF 80-200 mrad" % (args.n, nsl, args.ns, args.ns, args.nconf)) if args.sto else
                        ("au_stem_512: Au[001] a=4.078 8x8x20 cells, %d^2 grid, %d slices, 300 kV 21.4 mrad probe, "
                         "%dx%d scan, %d phonon configs, ADF 80-200 mrad" % (args.n, nsl, args.ns, args.ns, args.nconf)),
            "probe_slices_per_step": ps_step,
            "batch": B,
            "plan": plan,
            "tune": engine.tune_log(),
        },
        "roofline": {
            "bound": "hbm",
            "kernel": kern,
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS if achieved else None,
            "traffic": traffic_from_profiles(kern, "sto_stem" if args.sto else "au_stem", args.n, B),
            "bytes_alg_per_launch": kb,
            "mean_launch_ms": kms,
            "row_pass_ms": row_ms,
            "col_pass_ms": col_ms,
            "slice_loop_alg_GBs": slice_loop_gbs,
            # the achievable rate on this GPU: a streaming copy kernel (16-byte
            # nontemporal loads/stores, 1 GiB, read + write)
            "copy_ceiling_GBs": engine.stream_copy_bandwidth(1 << 30, 10),
        },
    }
    out["roofline"]["frac_of_copy_ceiling"] = (achieved / out["roofline"]["copy_ceiling_GBs"]
                                               if achieved else None)
    if not args.no_group:
        out["uneven_group"] = group_measurement(args.group_batch)
    if not args.no_fp64:
        out["fp64"] = fp64_measurement()
    if not args.no_surrogate and os.path.exists(SURROGATE_WEIGHTS):
        out["surrogate"] = surrogate_measurement(args.surrogate_scan128)
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(args)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
