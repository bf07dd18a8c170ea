#!/usr/bin/env python3
"""Benchmark of the MI355X SOS solver on the batched uneven workload.

Workload (one step = one solve of the whole batch, inputs resident on the GPU):
  8 bands (412..865 nm) x 1024 geometries = 8192 instances drawn from one
  seed: SZA U(0,80), VZA U(0,70), RAA U(0,180), AOD(865) logU(0.01,0.5),
  Angstrom U(0,2), HG g U(0.6,0.8) with 80 Legendre terms, SSA U(0.85,1),
  wind U(1,15) m/s rough ocean, 50..400 layers, 16..64 streams, Fourier
  modes 0..16.  Headline value: solves/s of the mixed form (bf16 MFMA
  scattering integral, orders >= 2).

Also reported: the FP64 form (batched and one instance per launch), the
error of the mixed form against the FP64 form, the single-instance
Rayleigh + aerosol case (400 layers, 48 streams, modes 0..95, IPSS), the
roofline of the dominant kernel from HIP events over the timed steps, and
the reference CPU solver on a bounded sample of the same workload.

    python bench.py [--steps K] [--warmup W] [--gpus 1]
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "TOA Stokes solves/s at <=1e-4 rel err (mixed vs FP64 form); adjoint/fwd cost"
BF16_PEAK_TFLOPS = 2500.0      # MI355X dense bf16 MFMA
FP64_PEAK_TFLOPS = 78.6        # MI355X fp64 matrix
HBM_PEAK_GBS = 8000.0          # MI355X HBM3E


TRAFFIC_PROFILE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles",
                               "sos_traffic_8192inst_mixed.md")


def measured_traffic(geoms, gemm_tuning, sweep_tuning):
    """HBM bytes per launch of the scattering integral (the classes that ran,
    each with its chosen tile shape, summed) and of the sweep (its chosen
    launch shape), measured with rocprofv3 FETCH_SIZE / WRITE_SIZE
    (copy-calibrated) on this workload and committed with the profile
    summary; None for another workload size or a shape not profiled."""
    try:
        lines = open(TRAFFIC_PROFILE).read().splitlines()
    except OSError:
        return None
    for ln in lines:
        if ln.startswith("bench-traffic-bytes-per-launch:"):
            tok = ln.split(":", 1)[1].split()
            kv = dict(zip(tok[0::2], tok[1::2]))
            if kv.get("workload") != f"geoms={geoms}":
                return None
            try:
                gemm = 0.0
                for k, (tm, tn) in enumerate(gemm_tuning["shape_of_class"]):
                    if gemm_tuning["ms"] is None or max(gemm_tuning["ms"][k]) > 0:
                        gemm += float(kv[f"gemm_c{k}_{tm}x{tn}"])
                sweep = float(kv[f"sweep_{sweep_tuning['unroll']}x{sweep_tuning['waves']}"])
            except KeyError:
                return None
            return {"gemm_bf16": gemm, "sweep": sweep}
    return None


def log(msg):
    sys.stderr.write(f"[bench] {msg}\n")
    sys.stderr.flush()


def timed_steps(plan, steps, warmup, sync):
    for _ in range(warmup):
        plan.execute()
        sync()
    sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        plan.execute()
    sync()
    return (time.perf_counter() - t0) / steps


def _cpu_lib():
    """The reference solver built from its sources (oracle/_ref), else the C
    restatement; both are the CPU baseline, never the measured GPU path."""
    ref = os.path.join(ROOT, "oracle", "_ref", "libocrt_ref.so")
    if os.path.exists(ref):
        return C.CDLL(ref), "ref_solve_batch", "reference"
    path = os.path.join(ROOT, "oracle", "liboracle.so")
    if not os.path.exists(path):
        import subprocess
        subprocess.run(["make", "-C", os.path.join(ROOT, "oracle")], check=True,
                       stdout=subprocess.DEVNULL)
    return C.CDLL(path), "oracle_solve_batch", "port"


def spread_sample(descs, k):
    """k instances spread evenly over the workload (every len/k-th), so the
    sample's mix of layers, streams and bands follows the whole batch."""
    step = max(1, len(descs) // k)
    return [descs[i] for i in range(0, len(descs), step)][:k]


def cpu_baseline(descs, n_sample, threads, n_single):
    """Reference solver on a spread sample of the workload, OpenMP dynamic over
    instances as rt_io_run_batch does (rt_io.c:199); then single-thread ms per
    instance on a smaller spread sample."""
    from ocrt_amd.abi import Result, instance_array
    lib, fn, kind = _cpu_lib()
    sample = spread_sample(descs, n_sample)
    arr = instance_array(sample)
    out = (Result * len(sample))()
    t0 = time.perf_counter()
    failed = getattr(lib, fn)(arr, len(sample), out, threads)
    dt = time.perf_counter() - t0
    one = spread_sample(descs, n_single)
    arr1 = instance_array(one)
    out1 = (Result * len(one))()
    t0 = time.perf_counter()
    getattr(lib, fn)(arr1, len(one), out1, 1)
    dt1 = time.perf_counter() - t0
    return {"value": len(sample) / dt, "unit": "solves/s", "cores": threads, "kind": kind,
            "nproc": os.cpu_count(),
            "single_thread_ms_per_instance": dt1 / len(one) * 1e3,
            "sample": f"{len(sample)} instances spread over the workload (every {len(descs) // len(sample)}-th; "
                      f"{failed} failed), OpenMP dynamic over instances on {threads} threads "
                      f"(the host's CPU share for one GPU; nproc {os.cpu_count()}), {dt:.1f} s wall; "
                      f"single thread: {len(one)} spread instances, {dt1:.1f} s"}


def survey_work(descs, orders, n_ring):
    """Algorithmic work of the orders >= 2 (the ones the scattering-integral
    and sweep kernels run), by the per-unit figures of the survey:
      F = 72 n_out n_in (nt + 1) flop per order and mode (n_out = ring size,
          n_in = Gauss nodes of non-zero weight);
      B = (2 * 3 * 2 n (nt + 1) + 2 n nt) * 4 bytes per order and mode (fp32
          source read and field written for I, Q, U in both hemispheres, plus
          the transmission table)."""
    F = B = 0.0
    for i, d in enumerate(descs):
        n, nin, nt = int(n_ring[i]), d.n_mu, d.n_layers
        runs = float(np.maximum(orders[i, : d.m_max + 1] - 1, 0).sum())
        if d.L_max > 0:
            F += runs * 72.0 * n * nin * (nt + 1)
        B += runs * (2 * 3 * 2 * n * (nt + 1) + 2 * n * nt) * 4.0
    return F, B


def adjoint_leg(sync):
    """Config-3 Jacobian timing: 256 instances, 200 layers, 32 streams,
    modes 0..16, checkpoints every 4 orders."""
    import ocrt_amd
    from ocrt_amd import synthetic as S
    log("adjoint workload: 256 instances, 200 layers, 32 streams, modes 0..16")
    d3, k3 = S.config3()
    g3 = np.tile(np.array([1.0, 0.5, -0.5]), (len(d3), 1))

    def jacobian_times(prec):
        """(forward with checkpoints every 4 orders + backward, backward
        alone) in seconds: the Jacobian of one new input, and of one more
        output weighting on the same forward."""
        pa = ocrt_amd.Plan(d3, precision=prec, adjoint=True, ckpt_sparse=True)
        pa.execute()
        pa.adjoint(g3)
        sync()
        t0 = time.perf_counter()
        for _ in range(3):
            pa.execute()
            pa.adjoint(g3)
        sync()
        t_full = (time.perf_counter() - t0) / 3
        t0 = time.perf_counter()
        for _ in range(3):
            pa.adjoint(g3)
        sync()
        t_bwd = (time.perf_counter() - t0) / 3
        pa.close()
        return t_full, t_bwd

    t_j64, t_b64 = jacobian_times("fp64")
    t_jx3, t_bx3 = jacobian_times("mixed")
    pf = ocrt_amd.Plan(d3, precision="fp64")
    t_fwd = timed_steps(pf, 3, 1, sync)
    pf.close()
    pm3 = ocrt_amd.Plan(d3, precision="mixed")
    t_mix = timed_steps(pm3, 3, 1, sync)
    pm3.close()
    adjoint = {"instances": len(d3), "checkpoint_every": 4,
               "fp64_forward_ms": t_fwd * 1e3, "mixed_forward_ms": t_mix * 1e3,
               "jacobian_fp64_ms": t_j64 * 1e3, "backward_only_fp64_ms": t_b64 * 1e3,
               "jacobian_mixed_ms": t_jx3 * 1e3, "backward_only_mixed_ms": t_bx3 * 1e3,
               "jacobian_fp64_over_mixed_forward": t_j64 / t_mix,
               "jacobian_fp64_over_fp64_forward": t_j64 / t_fwd,
               "jacobian_mixed_over_mixed_forward": t_jx3 / t_mix,
               "jacobian_rows_per_s_mixed": len(d3) / t_jx3,
               "note": "jacobian = checkpointing forward + backward recursion (recomputing the orders "
                       "between checkpoints) for one new input; mixed: the mixed order loop with every field "
                       "kept as its bf16 pair, fp32 adjoint fields, bf16 MFMA operator products; fp64: "
                       "everything fp64"}
    log(f"adjoint: {adjoint}")
    return adjoint


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--geoms", type=int, default=1024, help="geometries per band (workload size)")
    ap.add_argument("--plain-sample", type=int, default=64)
    ap.add_argument("--cpu-sample", type=int, default=64)
    ap.add_argument("--cpu-single", type=int, default=8)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-single", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--no-ocean", action="store_true")
    ap.add_argument("--mixed-only", action="store_true",
                    help="time only the mixed form (for counter-collection runs)")
    ap.add_argument("--adjoint-only", action="store_true",
                    help="time only the Jacobian workload (for profiling runs of the adjoint kernels)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="first run a plain device copy of MIB MiB (3 times): the reference "
                         "for scaling FETCH_SIZE / WRITE_SIZE counters")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU only (instances are independent; a "
              "multi-GPU run would shard them with no exchange). Use --gpus 1.", file=sys.stderr)
        sys.exit(2)

    import torch
    import ocrt_amd
    from ocrt_amd import synthetic as S

    sync = torch.cuda.synchronize
    if args.copy_calibration > 0:
        nbytes = args.copy_calibration << 20
        src = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        sync()
        log(f"copy calibration: 3 copies of {nbytes} bytes (read {nbytes}, write {nbytes} each)")
        del src, dst
    if args.adjoint_only:
        print(json.dumps({"adjoint": adjoint_leg(sync)}))
        return
    log("building workload")
    descs, keep = S.config2(n_geom=args.geoms)
    n = len(descs)

    log(f"mixed plan over {n} instances (sweep launch shape autotuned at the first execute)")
    pm = ocrt_amd.Plan(descs, precision="mixed", autotune=True)
    pm.execute()            # autotuning: one warm-up and one execute per sweep shape
    sync()
    tuning = pm.tuning()
    gemm_tuning = pm.gemm_tuning()
    log(f"GEMM tile shape per operator class: {gemm_tuning}")
    log(f"autotune: {tuning}")
    pm.profile(True)
    ms_mixed = timed_steps(pm, args.steps, args.warmup, sync) * 1e3
    kt = pm.kernel_times()
    res_m = pm.results(modes=True)
    work = pm.work()
    log(f"mixed: {ms_mixed:.1f} ms/step, kernels {kt}")
    pg = ocrt_amd.Plan(descs, precision="mixed", graph=True)
    pg.set_tuning(tuning["unroll"], tuning["waves"])        # the eager plan's autotuned shapes
    for k, shape in enumerate(gemm_tuning["shape_of_class"]):
        pg.set_gemm_shape(k, ocrt_amd.Plan.GEMM_SHAPES.index(tuple(shape)))
    ms_graph = timed_steps(pg, args.steps, 1, sync) * 1e3
    same_bits = all(np.array_equal(pg.results()[f], res_m[f]) for f in ("rho_I", "rho_Q", "rho_U"))
    pg.close()
    log(f"mixed, hipGraph replay: {ms_graph:.1f} ms/step (results equal to eager: {same_bits})")

    if args.mixed_only:
        print(json.dumps({"mixed_ms_per_step": ms_mixed, "instances": n, "kernels": kt}))
        pm.close()
        return

    log("fp64 batched plan")
    pf = ocrt_amd.Plan(descs, precision="fp64")
    ms_fp64 = timed_steps(pf, 1, 1, sync) * 1e3
    res_f = pf.results(modes=True)
    pf.close()
    log(f"fp64 batched: {ms_fp64:.1f} ms/step")

    ok = (res_m["status"] == 0) & (res_f["status"] == 0)
    rel_I = np.abs(res_m["rho_I"] - res_f["rho_I"]) / np.abs(res_f["rho_I"])
    abs_QU = np.maximum(np.abs(res_m["rho_Q"] - res_f["rho_Q"]), np.abs(res_m["rho_U"] - res_f["rho_U"]))
    err_I = float(np.max(rel_I[ok]))
    err_QU = float(np.max(abs_QU[ok]))
    worst = int(np.argmax(np.where(ok, rel_I, 0.0)))
    log(f"mixed vs fp64: max rel I {err_I:.3e} (instance {worst}), max abs Q/U {err_QU:.3e}")

    log(f"fp64 one instance per launch on {args.plain_sample} instances spread over the workload")
    plain = spread_sample(descs, args.plain_sample)
    pp = ocrt_amd.Plan(plain, precision="fp64", one_per_launch=True)
    ms_plain = timed_steps(pp, 1, 1, sync) * 1e3
    pp.close()
    plain_rate = len(plain) / (ms_plain / 1e3)

    adjoint = None if args.no_adjoint else adjoint_leg(sync)

    ocean_leg = None
    if not args.no_ocean:
        from ocrt_amd import ocean
        log("coupled atmosphere-ocean workload: 512 instances, 100 wavelengths, 64 streams, modes 0..127")
        d4, w4, meta4 = S.config4()
        ocean.coupled_solve(d4[:8], w4[:8], precision="mixed")   # warm-up
        sync()
        st_m, st_f = {}, {}
        t0 = time.perf_counter()
        om = ocean.coupled_solve(d4, w4, precision="mixed", stage_times=st_m)
        sync()
        t_m = time.perf_counter() - t0
        t0 = time.perf_counter()
        of = ocean.coupled_solve(d4, w4, precision="fp64", stage_times=st_f)
        sync()
        t_f = time.perf_counter() - t0
        ei4 = np.abs(om["rho_I"] - of["rho_I"]) / np.abs(of["rho_I"])
        eq4 = np.maximum(np.abs(om["rho_Q"] - of["rho_Q"]), np.abs(om["rho_U"] - of["rho_U"]))
        nj = 64

        def make(c, a):
            d, w, _ = S.config4(n_inst=nj, chl=c, aod_scale=a)
            return d, w
        t0 = time.perf_counter()
        ocean.coupled_jacobian(make, meta4["chl"][:nj], rel=1e-3)
        sync()
        t_j = time.perf_counter() - t0
        ocean_leg = {"instances": len(d4), "mixed_s": t_m, "fp64_s": t_f,
                     "mixed_solves_per_s": len(d4) / t_m, "fp64_solves_per_s": len(d4) / t_f,
                     "stages_mixed_s": st_m, "stages_fp64_s": st_f,
                     "max_rel_err_I_mixed_vs_fp64": float(ei4.max()), "max_abs_err_QU_mixed_vs_fp64": float(eq4.max()),
                     "water_orders_max": int(of["orders_water"].max()),
                     "jacobian_instances": nj, "jacobian_fp64_s": t_j,
                     "jacobian_note": "d(rho_I,Q,U)/d ln Chl and d ln AOD by central differences (5 coupled solves "
                                      "per instance in one batch, fp64)"}
        log(f"coupled ocean: {ocean_leg}")

    single = None
    if not args.no_single:
        log("single instance: Rayleigh + aerosol, 400 layers, 48 streams, modes 0..95, IPSS")
        d1, k1 = S.config1()
        single = {}
        for prec in ("mixed", "fp64"):
            p1 = ocrt_amd.Plan(d1, precision=prec)
            t = timed_steps(p1, 3, 1, sync)
            single[prec] = {"ms": t * 1e3, "solves_per_s": 1.0 / t, "rho_I": float(p1.results()["rho_I"][0])}
            p1.close()
        single["rel_err_I_mixed_vs_fp64"] = abs(single["mixed"]["rho_I"] - single["fp64"]["rho_I"]) / abs(
            single["fp64"]["rho_I"])

    value = n / (ms_mixed / 1e3)
    gemm_s = kt["gemm_ms"] / 1e3
    sweep_s = kt["sweep_ms"] / 1e3
    F_alg, B_alg = survey_work(descs, res_m["orders_per_m"], res_m["n_ring"])
    gemm_tf = F_alg / gemm_s / 1e12 if gemm_s > 0 else 0.0
    sweep_gbs = B_alg / sweep_s / 1e9 if sweep_s > 0 else 0.0
    traffic = measured_traffic(args.geoms, gemm_tuning, tuning)
    gl, sl = max(kt["gemm_launches"], 1), max(kt["sweep_launches"], 1)
    rl_gemm = {"kernel": "gemm_bf16", "bound": "mfma", "achieved": gemm_tf, "peak": BF16_PEAK_TFLOPS,
               "unit": "TFLOP/s", "frac": gemm_tf / BF16_PEAK_TFLOPS,
               "traffic": traffic["gemm_bf16"] if traffic else None,
               "algorithmic_per_launch": F_alg / gl, "avg_launch_ms": kt["gemm_ms"] / gl,
               "ms_total": kt["gemm_ms"], "launches": kt["gemm_launches"],
               "issued_model_tflops": work["gemm_flops"] / gemm_s / 1e12 if gemm_s > 0 else 0.0}
    rl_sweep = {"kernel": "sweep<float,true>", "bound": "hbm", "achieved": sweep_gbs, "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": sweep_gbs / HBM_PEAK_GBS,
                "traffic": traffic["sweep"] if traffic else None,
                "algorithmic_per_launch": B_alg / sl, "avg_launch_ms": kt["sweep_ms"] / sl,
                "ms_total": kt["sweep_ms"], "launches": kt["sweep_launches"],
                "moved_model_gbs": work["sweep_bytes"] / sweep_s / 1e9 if sweep_s > 0 else 0.0}
    roofline = rl_gemm if kt["gemm_ms"] >= kt["sweep_ms"] else rl_sweep
    om = res_m["orders_per_m"].astype(np.float64)
    of = res_f["orders_per_m"].astype(np.float64)
    orders_per_mode = {"mixed_mean": [round(v, 3) for v in om.mean(axis=0).tolist()],
                       "fp64_mean": [round(v, 3) for v in of.mean(axis=0).tolist()],
                       "max_abs_diff": int(np.abs(om - of).max())}

    cpu = None
    if not args.no_cpu:
        threads = min(os.cpu_count() or 1, 16)
        log(f"cpu baseline on {args.cpu_sample} spread instances, {threads} threads (nproc {os.cpu_count()})")
        cpu = cpu_baseline(descs, args.cpu_sample, threads, args.cpu_single)

    line = {
        "metric": METRIC,
        "value": value,
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_mixed,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "bf16",
        "data": "synthetic",
        "config": {
            "workload": f"batched uneven SOS solves: 8 bands x {args.geoms} geometries = {n} instances, "
                        "rough ocean, 50-400 layers, 16-64 streams, modes 0..16, HG aerosol L=80",
            "instances": n,
            "form": "mixed (split-bf16 MFMA scattering integral, fp32 sweeps for orders >= 2, fp64 elsewhere)",
        },
        "fp64_batched_solves_per_s": n / (ms_fp64 / 1e3),
        "fp64_plain_solves_per_s": plain_rate,
        "mixed_over_plain_fp64": value / plain_rate,
        "mixed_over_fp64_batched": ms_fp64 / ms_mixed,
        "max_rel_err_I_vs_fp64": err_I,
        "max_abs_err_QU_vs_fp64": err_QU,
        "worst_instance_I": worst,
        "orders_per_mode": orders_per_mode,
        "adjoint_fwd_ratio": adjoint["jacobian_mixed_over_mixed_forward"] if adjoint else None,
        "adjoint": adjoint,
        "single_instance_config1": single,
        "coupled_ocean": ocean_leg,
        "roofline": roofline,
        "roofline_other": rl_sweep if roofline is rl_gemm else rl_gemm,
        "cpu_baseline": cpu,
        "device_bytes": pm.device_bytes(),
        "sweep_autotune": tuning,
        "gemm_autotune": gemm_tuning,
        "graph_replay_ms_per_step": ms_graph,
    }
    pm.close()
    print(json.dumps(line))


if __name__ == "__main__":
    main()
