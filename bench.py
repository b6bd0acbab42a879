"""Benchmark of the simulation step on one MI355X.

Headline (value): the plain FP32 step at 256 x 256 x 128 voxels (x, z
horizontal; 128 levels vertical), dx = 100 m, dt = 0.5 s, the randomised
sounding (seed 20261203, three warm bubbles, surface theta 296-304 K, RH
60-95 %, 0-25 m/s shear). One step is one environmentGPU::updateGPU over the
whole domain with the reference's solver settings (20 diffusion sweeps,
reference-mode PCG: absolute tolerance 1e-5, at most 200 iterations, which
this workload always reaches). Inputs are resident in HBM before the timed
region; the timed region is K steps bracketed by device synchronisation.

Beside it, each in its own key (skip any with --skip):
  standard  the same step with the textbook Jacobi-PCG at a relative residual
            of 1e-6, K steps as the headline;
  mixed     the mixed-precision form at 512 x 512 x 256 (hydrometeors in BF16
            and in FP16 with power-of-two scales, multigrid-preconditioned CG
            to 1e-6 with two FP64 refinement passes) next to the FP32 step at
            that size with the Jacobi and with the multigrid preconditioner;
  ensemble  64 members drawn from {64^3, 128^3, 192x192x96, 128x128x160},
            1-4 bubbles each, in the mixed form, one launch per phase (HIP
            graph), against the same members run one after the other as plain
            FP32 steps (Jacobi-PCG to 1e-6) and as mixed-form steps; also 64
            members of 64^3 (the launch-bound case);
  serving   Poisson arrivals at 2, 5, 10 and 20 requests/s of 'simulate
            sounding S for T s' jobs (grids of the ensemble plus
            256x256x128, T uniform 300-1800 s) with T and the 60 s report
            interval scaled by SERVE_TIME_SCALE so the leg runs in about a
            minute: throughput and p50/p99 latency per rate;
  rows      sounding ingestion, serving diagnostics readback, terrain;
  cpu       the CPU baseline: the oracle (FP32 C++ restatement) on the host
            cores, 2 timed steps of the headline workload after a warm-up
            step with all threads the box grants, plus a 1-thread figure on
            one 64^3 step; one GPU step is checked against it.

Prints one JSON line. The roofline object is for the dominant kernel class
of the headline step: algorithmic bytes per launch / average launch
duration from HIP events on the env's stream during the timed region
(attached to every tenth launch of that class; no other launch carries
events there), with the HBM traffic per launch from the committed rocprofv3
counter summary when one exists for it.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

NX, NY, NZ, DX, DT = 256, 128, 256, 100.0, 0.5
SEED = 20261203
HBM_PEAK_GBS = 8000.0
TIMING_STRIDE = 10
KEYS = ("potTemps", "velField", "Qv", "groundTemp", "groundPres", "pressures", "smallPressure")
STANDARD = dict(pcg_mode=1, pcg_tol=1e-6, pcg_max_iter=1000)
# the mixed-precision form: 16-bit hydrometeors, multigrid-preconditioned CG
# to a relative residual of 1e-6 with two FP64 refinement passes
# (at most 20 iterations per solve: an ensemble step holds that many per
# pass; a multigrid solve takes about 10 per pass here)
MIXED = {"bf16": dict(pcg_mode=1, pcg_tol=1e-6, pcg_max_iter=20, tracer_format=1, pcg_refine=2,
                      pcg_refine_tol=1e-3, pcg_precon=1),
         "f16": dict(pcg_mode=1, pcg_tol=1e-6, pcg_max_iter=20, tracer_format=2, pcg_refine=2,
                     pcg_refine_tol=1e-3, pcg_precon=1)}
ENSEMBLE_GRIDS = [(64, 64, 64), (128, 128, 128), (192, 96, 192), (128, 160, 128)]  # nx, ny vertical, nz
SERVE_RATES = (2.0, 5.0, 10.0, 20.0)
SERVE_REQUESTS = 16
SERVE_TIME_SCALE = 1.0 / 60.0

# Fields one launch of a batched class moves, on average over a step: the
# diffusion launches sweep 1, 3 and 6 fields (20 sweeps each), the PPM passes
# 1 (theta), 1, 1, 1 (u, v, w), 1 (vapour) and 5 fields (5 passes each), the
# sedimentation launch 3 fields.
FIELDS_PER_LAUNCH = {"diffusion_sweep": 10.0 / 3.0, "ppm_pass": 10.0 / 6.0, "precip_fall": 3.0}


def algorithmic_bytes(cls, n_sky, tb=4):
    """Compulsory HBM bytes of one launch of a kernel class per sky cell, with
    the five hydrometeors stored in tb bytes (4 FP32, 2 BF16/FP16):
    diffusion sweep 8 (read x, write x; a 16-bit field is read by the first
    of the 20 sweeps and written by the last, the sweeps between run on FP32
    buffers); PPM pass 12 (read q and the face velocity, write q; a 16-bit
    field is read by the first pass and written by the last); PCG spmv 20
    (read z, s_old and the density-derived operator as 4 B of density, write
    s and As); PCG update 32 (read p, s, r, As, 1/diag; write p, r, z);
    microphysics 24 + 10 tb (9 fields in, 7 out, five of each hydrometeors);
    sedimentation 12 + 2 tb (read q, theta, p, vapour; write q)."""
    sweep_q = (tb + 4 + 8 * 18 + 4 + tb) / 20.0  # a hydrometeor over one diffusion's 20 sweeps
    pass_q = (tb + 8 + 12 * 3 + 8 + tb) / 5.0     # and over the five PPM passes
    per = {"diffusion_sweep": (5 * 8 + 5 * sweep_q) / 10.0, "ppm_pass": (5 * 12 + 5 * pass_q) / 10.0,
           "pcg_spmv": 20, "pcg_update": 32, "microphysics": 24 + 10 * tb, "precip_fall": 12 + 2 * tb}
    return per.get(cls, 0) * n_sky * FIELDS_PER_LAUNCH.get(cls, 1.0)


SKY_FIELDS = ("QV", "QW", "QC", "QR", "QS", "QI", "POTTEMP", "VELX", "VELY", "VELZ", "PRESSURE", "DENSITY")


def _sync():
    import skewtcloudsim_amd as skt
    skt.load_library().skt_device_synchronize(0)


def cpu_info():
    model = ""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        allowed = len(os.sched_getaffinity(0))
    except AttributeError:
        allowed = os.cpu_count()
    return {"nproc": os.cpu_count(), "affinity": allowed, "model": model}


def cpu_baseline(snd_args, threads, steps):
    """The oracle's step timed on the host (one untimed warm-up step, then
    'steps' timed ones) with 'threads' OpenMP threads, and a 1-thread figure
    from one 64^3 step after a warm-up step. One GPU run of the same steps
    from the same initial state is checked against the oracle (relative L2
    over the sky cells per field, PCG iteration counts)."""
    import numpy as np
    import skewtcloudsim_amd as skt
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle_bind import Oracle
    o = Oracle(NX, NY, NZ, DX, threads=threads)
    o.init(*snd_args)
    o.step(DT, 1.0)
    t0 = time.perf_counter()
    its = []
    for k in range(steps):
        o.step(DT, 1.0)
        its.append(o.stats()["pcg_iterations"])
        print(f"cpu baseline step {k + 1}/{steps}: {time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)
    t1 = time.perf_counter()
    env = skt.EnvironmentGPU(NX, NY, NZ, DX)
    env.init(*snd_args)
    for _ in range(steps + 1):
        env.updateGPU(DT, 1.0)
    sky = (np.arange(NX * NY * NZ) // NX) % NY >= 1
    rel = {}
    for f in SKY_FIELDS:
        a, b = env.get(f)[sky].astype(np.float64), o.get(f)[sky].astype(np.float64)
        rel[f] = float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))
    check = {"steps": steps + 1, "rel_l2": rel, "max_rel_l2": max(rel.values()),
             "pcg_iterations_last": {"gpu": int(env.stats().pcg_iterations), "oracle": int(its[-1])}}
    env.close()
    o.close()
    # 1-thread figure, bounded sample
    n1 = 64
    s1 = skt.make_sounding(n1, n1, n1, DX, seed=SEED, n_bubbles=3)
    o1 = Oracle(n1, n1, n1, DX, threads=1)
    o1.init(*[s1[k] for k in KEYS])
    o1.step(DT, 1.0)
    t2 = time.perf_counter()
    o1.step(DT, 1.0)
    t3 = time.perf_counter()
    o1.close()
    one = n1 ** 3 / (t3 - t2)
    return (NX * NY * NZ) * steps / (t1 - t0), (t1 - t0) / steps, its, one, t3 - t2, check


def cpu_single_bubble(threads, steps=200):
    """The plain method's CPU-runnable case: 64^3, the fixed single-bubble
    sounding, Jacobi-PCG to a relative residual of 1e-6, 'steps' steps of
    dt 0.5 s, on the oracle with 'threads' threads and on the GPU."""
    import skewtcloudsim_amd as skt
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle_bind import Oracle
    n = 64
    snd = skt.make_sounding(n, n, n, DX, seed=SEED, n_bubbles=0)
    args = [snd[k] for k in KEYS]
    o = Oracle(n, n, n, DX, threads=threads, **{k: v for k, v in STANDARD.items()})
    o.init(*args)
    its = []
    t0 = time.perf_counter()
    for k in range(steps):
        o.step(DT, 1.0)
        its.append(o.stats()["pcg_iterations"])
        if k % 20 == 19:
            print(f"cpu 64^3 step {k + 1}/{steps}: {time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)
    cpu_s = time.perf_counter() - t0
    o.close()
    env = skt.EnvironmentGPU(n, n, n, DX, **STANDARD)
    env.init(*args)
    env.device_synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        env.updateGPU(DT, 1.0)
    env.device_synchronize()
    gpu_s = time.perf_counter() - t0
    env.close()
    cells = n ** 3 * steps
    return {"grid": [n, n, n], "steps": steps, "threads": threads, "cpu_s": cpu_s, "cpu_cell_steps_per_s": cells / cpu_s,
            "gpu_s": gpu_s, "gpu_cell_steps_per_s": cells / gpu_s, "gpu_speedup": cpu_s / gpu_s,
            "pcg_iterations_oracle": {"min": min(its), "max": max(its), "mean": sum(its) / len(its)}}


def traffic_from_profile(kernel_class):
    """HBM bytes per launch of a kernel class from the committed counter
    summary (rocprofv3 FETCH_SIZE / WRITE_SIZE runs, each scaled by the same
    counter's reading for a copy kernel of known size). None if absent."""
    path = os.path.join(ROOT, "profiles", "traffic_256x256x128.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        f = line.split()
        if len(f) >= 2 and f[0] == kernel_class:
            try:
                return float(f[1])
            except ValueError:
                return None
    return None


def run_standard(snd_args, steps, warmup):
    """The step with textbook Jacobi-PCG to a relative residual of 1e-6."""
    import skewtcloudsim_amd as skt
    env = skt.EnvironmentGPU(NX, NY, NZ, DX, **STANDARD)
    env.init(*snd_args)
    for _ in range(warmup):
        env.updateGPU(DT, 1.0)
    env.device_synchronize()
    ks = []
    t0 = time.perf_counter()
    for _ in range(steps):
        env.updateGPU(DT, 1.0)
        ks.append(env.stats().pcg_iterations)
    env.device_synchronize()
    wall = time.perf_counter() - t0
    env.close()
    return {"value": NX * NY * NZ * steps / wall, "unit": "cell-steps/s", "ms_per_step": wall / steps * 1e3,
            "steps": steps, "pcg": "textbook Jacobi-PCG, relative residual 1e-6, max 1000",
            "pcg_iterations": ks}


def _timed_env(env, steps, warmup, tb):
    """Warm-up, one instrumented step (per-class breakdown), then 'steps'
    timed steps; returns the figures of the run."""
    for _ in range(warmup):
        env.updateGPU(DT, 1.0)
    env.device_synchronize()
    env.set_timing(True)
    env.updateGPU(DT, 1.0)
    env.device_synchronize()
    br = env.timing()
    env.set_timing(False)
    t0 = time.perf_counter()
    st = []
    for _ in range(steps):
        env.updateGPU(DT, 1.0)
        s = env.stats()
        st.append((s.pcg_iterations_total, s.pcg_passes, s.pcg_residual64 / s.pcg_rhs64 if s.pcg_rhs64 else None))
    env.device_synchronize()
    wall = time.perf_counter() - t0
    n_sky = env.nx * (env.ny - 1) * env.nz
    cls = {c: v for c, v in br.items() if v[0] > 0}
    dom = max(cls, key=lambda c: cls[c][1])
    l, ms = cls[dom]
    alg = algorithmic_bytes(dom, n_sky, tb)
    return {"ms_per_step": wall / steps * 1e3, "value": env.N * steps / wall, "steps": steps,
            "pcg_iterations_passes_residual64": st,
            "dominant_class": dom, "dominant_GBps_algorithmic": alg / (ms / l / 1e3) / 1e9 if alg else None,
            "classes_ms_one_step": {c: round(v[1], 3) for c, v in cls.items()}}


def run_mixed(steps, warmup):
    """512 x 512 x 256 (x, z 512; 256 vertical): FP32 (Jacobi-PCG to 1e-6)
    and the mixed forms (16-bit hydrometeors, FP64 refinement)."""
    import skewtcloudsim_amd as skt
    nx, ny, nz = 512, 256, 512
    snd = skt.make_sounding(nx, ny, nz, DX, seed=SEED, n_bubbles=3)
    args = [snd[k] for k in KEYS]
    out = {"grid": [nx, ny, nz], "cells": nx * ny * nz}
    legs = [("fp32", dict(STANDARD, pcg_max_iter=1000), 4), ("fp32_multigrid", dict(STANDARD, pcg_precon=1), 4)]
    for name, cfg, tb in legs + [(k, v, 2) for k, v in MIXED.items()]:
        env = skt.EnvironmentGPU(nx, ny, nz, DX, **cfg)
        env.init(*args)
        r = _timed_env(env, steps, warmup, tb)
        r["config"] = cfg
        r["tracer_bytes"] = tb
        out[name] = r
        env.close()
        print(f"mixed leg {name}: {r['ms_per_step']:.1f} ms/step", file=sys.stderr, flush=True)
    out["speedup_bf16_vs_fp32"] = out["fp32"]["ms_per_step"] / out["bf16"]["ms_per_step"]
    out["speedup_f16_vs_fp32"] = out["fp32"]["ms_per_step"] / out["f16"]["ms_per_step"]
    out["note"] = ("per-field error against the FP64 oracle: tests/test_gpu_mixed.py "
                   "(test_mixed_steps_against_the_fp64_oracle)")
    return out


def _members(shapes, cfg, rng, seed0):
    import skewtcloudsim_amd as skt
    envs = []
    for m, (nx, ny, nz) in enumerate(shapes):
        snd = skt.make_sounding(nx, ny, nz, DX, seed=seed0 + 7919 * m, n_bubbles=int(rng.integers(1, 5)))
        e = skt.EnvironmentGPU(nx, ny, nz, DX, **cfg)
        e.init(*[snd[k] for k in KEYS])
        envs.append(e)
    return envs


def run_ensemble(steps, members=64):
    """Ensemble members/s: mixed-precision members in one ensemble against the
    same members run one after the other as plain FP32 steps (Jacobi-PCG to
    1e-6, the same stop rule), for the configs' size mix and for 64^3 only."""
    import numpy as np
    import skewtcloudsim_amd as skt
    out = {}
    for case, grids in (("size_mix", ENSEMBLE_GRIDS), ("all_64cubed", [(64, 64, 64)])):
        rng = np.random.default_rng(SEED)
        shapes = [grids[int(i)] for i in rng.integers(0, len(grids), members)]
        cells = sum(a * b * c for a, b, c in shapes)
        plain = _members(shapes, dict(STANDARD, pcg_max_iter=400), np.random.default_rng(SEED + 1), SEED)
        for e in plain:
            e.updateGPU(DT, 1.0)
        _sync()
        t0 = time.perf_counter()
        for e in plain:
            for _ in range(steps):
                e.updateGPU(DT, 1.0)
        _sync()
        seq = (time.perf_counter() - t0) / steps
        for e in plain:
            e.close()
        mixed = _members(shapes, MIXED["bf16"], np.random.default_rng(SEED + 1), SEED)
        for e in mixed:
            e.updateGPU(DT, 1.0)
        _sync()
        t0 = time.perf_counter()
        for e in mixed:
            for _ in range(steps):
                e.updateGPU(DT, 1.0)
        _sync()
        seq_mixed = (time.perf_counter() - t0) / steps
        ens = skt.EnsembleGPU()
        for e in mixed:
            ens.add(e)
        tun = ens.autotune(DT, 1.0)
        ens.step(DT, 1.0)
        ens.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            ens.step(DT, 1.0)
        ens.synchronize()
        bat = (time.perf_counter() - t0) / steps
        prof = ens.profile(DT, 1.0)  # one more step, timed launch by launch
        top = sorted(prof.items(), key=lambda kv: -kv[1][0])[:12]
        ens.close()
        for e in mixed:
            e.close()
        out[case] = {"members": members, "cells": cells, "steps": steps,
                     "sequential_plain_fp32_ms_per_round": seq * 1e3,
                     "sequential_mixed_ms_per_round": seq_mixed * 1e3, "ensemble_mixed_ms_per_step": bat * 1e3,
                     "member_steps_per_s": members / bat, "member_steps_per_s_sequential": members / seq,
                     "members_per_s_600s_runs": members / (1200 * bat),
                     "members_per_s_600s_runs_sequential": members / (1200 * seq),
                     "speedup": seq / bat, "speedup_batching_alone": seq_mixed / bat, "autotune": tun,
                     "bodies_ms_one_step": {k: round(v[0], 3) for k, v in top}}
        print(f"ensemble leg {case}: sequential {seq*1e3:.1f} ms, ensemble {bat*1e3:.1f} ms", file=sys.stderr,
              flush=True)
    return out


def run_serving():
    import skewtcloudsim_amd as skt
    grids = ENSEMBLE_GRIDS + [(NX, NY, NZ)]
    res = []
    for k, rate in enumerate(SERVE_RATES):
        r = skt.run_poisson(rate, SERVE_REQUESTS, seed=SEED + k, grids=grids, time_scale=SERVE_TIME_SCALE,
                            member_config=MIXED["bf16"])
        r.pop("events")
        res.append(r)
        print(f"serving {rate}/s: p50 {r['latency_p50_s']:.2f} s p99 {r['latency_p99_s']:.2f} s", file=sys.stderr,
              flush=True)
    return {"time_scale": SERVE_TIME_SCALE, "requests_per_rate": SERVE_REQUESTS,
            "job": "grid from the ensemble sizes plus 256x256x128, random sounding with 1-4 bubbles, "
                   "T uniform 300-1800 s and reports every 60 s, both times time_scale; BF16 hydrometeors, "
                   "multigrid-preconditioned CG to 1e-6 with two FP64 refinement passes",
            "rates": res}


def measure_next_rows(reps=20):
    """Wall times at the bench grid of the paths around the step: sounding
    ingestion (DWD fixture archive -> levels on the host -> device broadcast,
    against the host-array init over PCIe), the serving diagnostics readback
    and terrain generation. Kernel times of the device parts come from
    rocprofv3 (k_sounding_fill, k_diag, k_terrain)."""
    import numpy as np
    import skewtcloudsim_amd as skt
    fixture = os.path.join(ROOT, "tests", "golden", "sekundenwerte_aero_01303_fixture.zip")
    out = {}
    r = skt.ReadTable()
    t0 = time.perf_counter()
    for _ in range(reps):
        r.readDWDFile(fixture)
    out["dwd_read_ms"] = (time.perf_counter() - t0) / reps * 1e3
    out["dwd_rows"] = int(r.data.dataSize)
    env = skt.EnvironmentGPU(NX, NY, NZ, DX)
    r.initEnvironment(env)  # first call: allocations and code paging outside the timing
    env.device_synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        r.initEnvironment(env)
    env.device_synchronize()
    out["init_sounding_ms"] = (time.perf_counter() - t0) / reps * 1e3
    dev = {f: env.get(f) for f in ("POTTEMP", "VELX", "QV", "PRESSURE")}
    t0 = time.perf_counter()
    host = r.init_fields(NX, NY, NZ, DX)
    out["init_fields_host_ms"] = (time.perf_counter() - t0) * 1e3
    t0 = time.perf_counter()
    env.init(*[host[k] for k in KEYS])
    env.device_synchronize()
    out["init_host_arrays_ms"] = (time.perf_counter() - t0) * 1e3
    out["device_equals_host_path"] = all(np.array_equal(dev[f], env.get(f)) for f in dev)
    env.updateGPU(DT, 1.0)
    env.diagnostics()
    env.device_synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        d = env.diagnostics()
    out["diagnostics_ms"] = (time.perf_counter() - t0) / reps * 1e3
    out["diagnostics"] = d
    env.close()
    skt.terrain(NX, NZ, seed=100, octaves=12, max_height=6.0)
    t0 = time.perf_counter()
    for _ in range(reps):
        skt.terrain(NX, NZ, seed=100, octaves=12, max_height=6.0)
    out["terrain_ms"] = (time.perf_counter() - t0) / reps * 1e3
    return out


LEGS = ("standard", "mixed", "ensemble", "serving", "rows", "cpu")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--skip", default="", help="comma-separated legs to skip: " + ",".join(LEGS))
    ap.add_argument("--cpu-threads", type=int, default=0, help="OpenMP threads of the CPU baseline "
                    "(default: the CPUs this process may use, at most 16)")
    ap.add_argument("--cpu-steps", type=int, default=2)
    ap.add_argument("--cpu-single-bubble", action="store_true",
                    help="also run the 64^3 single-bubble case for 200 steps on the CPU (several minutes)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); the step is single-domain, one GPU per "
              "process (on several GPUs the domain would split into z slabs, one per GPU).", file=sys.stderr)
        sys.exit(2)
    skip = {s for s in args.skip.split(",") if s}
    bad = skip - set(LEGS)
    if bad:
        ap.error(f"unknown legs {sorted(bad)}")

    import skewtcloudsim_amd as skt
    snd = skt.make_sounding(NX, NY, NZ, DX, seed=SEED, n_bubbles=3)
    snd_args = [snd[k] for k in KEYS]
    env = skt.EnvironmentGPU(NX, NY, NZ, DX)
    env.init(*snd_args)
    ks = []  # PCG iterations of the untimed steps and of the last timed one
    for _ in range(args.warmup):
        env.updateGPU(DT, 1.0)
        ks.append(env.stats().pcg_iterations)
    env.device_synchronize()

    # one fully instrumented step: per-class breakdown and the dominant class
    env.set_timing(True)
    env.updateGPU(DT, 1.0)
    env.device_synchronize()
    breakdown = env.timing()
    classes = {c: v for c, v in breakdown.items() if v[0] > 0}
    dom = max(classes, key=lambda c: classes[c][1])

    # timed region: HIP events on a sample of the dominant class's launches
    # (every TIMING_STRIDE-th; each timed launch idles the device for about 3 us)
    env.set_timing(True, classes=[dom], stride=TIMING_STRIDE)
    env.device_synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        env.updateGPU(DT, 1.0)  # no host reads in the loop: stats wait for the device
    env.device_synchronize()
    t1 = time.perf_counter()
    timing = env.timing()
    env.set_timing(False)
    st = env.stats()
    ks.append(st.pcg_iterations)
    copy_ms, copy_bytes = env.copy_probe(20)
    env.close()
    wall = t1 - t0
    n = NX * NY * NZ
    n_sky = NX * (NY - 1) * NZ
    value = n * args.steps / wall
    print(f"headline: {wall / args.steps * 1e3:.2f} ms/step", file=sys.stderr, flush=True)

    launches, total_ms = timing[dom]
    avg_s = total_ms / launches / 1e3
    alg = algorithmic_bytes(dom, n_sky)
    achieved = alg / avg_s / 1e9 if alg else None
    traffic = traffic_from_profile(dom)

    standard = None if "standard" in skip else run_standard(snd_args, args.steps, 2)
    mixed = None if "mixed" in skip else run_mixed(2, 1)
    ensemble = None if "ensemble" in skip else run_ensemble(2)
    serving = None if "serving" in skip else run_serving()
    rows = None if "rows" in skip else measure_next_rows()

    cpu = None
    parity = None
    info = cpu_info()
    if "cpu" not in skip:
        threads = args.cpu_threads or min(16, info["affinity"] or 1)
        cps, secs, its, one, one_s, parity = cpu_baseline(snd_args, threads, args.cpu_steps)
        cpu = {"value": cps, "unit": "cell-steps/s", "cores": threads, "kind": "port",
               "sample": f"{args.cpu_steps} steps of the same {NX}x{NY}x{NZ} workload after init and one "
                         f"untimed step ({secs:.1f} s per step, PCG iterations {its}), oracle/skt_oracle.cpp "
                         f"FP32, OpenMP {threads} threads",
               "host": info, "one_thread": {"value": one, "unit": "cell-steps/s",
                                            "sample": f"one 64^3 step after a warm-up step ({one_s:.1f} s)"},
               "gpu_speedup": value / cps}
        if args.cpu_single_bubble:
            cpu["single_bubble_64cubed_200_steps"] = cpu_single_bubble(threads)

    out = {
        "metric": "cell-steps/s at 256x256x128 (FP32, reference solver: 200 PCG iterations per step)",
        "value": value,
        "unit": "cell-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic (seeded sounding, 3 warm bubbles)",
        "config": {"workload": "plain FP32 step, 256x256x128 voxels (x,z 256; 128 vertical), dx 100 m, dt 0.5 s",
                   "nx": NX, "ny_vertical": NY, "nz": NZ, "pcg": "reference mode, abs tol 1e-5, max 200",
                   "pcg_iterations_per_step": sorted(set(ks)),
                   "pcg_iterations_sampled": "warm-up steps and the last timed step", "diffusion_sweeps": 20,
                   "advection_substeps_last_step": list(st.substeps)[:st.n_advect]},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": (achieved / HBM_PEAK_GBS) if achieved else None, "traffic": traffic,
                     "avg_launch_us": avg_s * 1e6, "launches_timed": launches, "timed_every": TIMING_STRIDE,
                     "algorithmic_bytes_per_launch": alg},
        "kernel_classes_one_step": {c: {"launches": v[0], "total_ms": v[1],
                                        "avg_us": (v[1] / v[0] * 1e3) if v[0] else 0.0,
                                        "GBps_algorithmic": (algorithmic_bytes(c, n_sky) / (v[1] / v[0] / 1e3) / 1e9)
                                        if v[0] and algorithmic_bytes(c, n_sky) else None}
                                    for c, v in breakdown.items()},
        "pcg_standard_1e-6": standard,
        "mixed_512x512x256": mixed,
        "ensemble": ensemble,
        "serving": serving,
        "copy_probe": {"GBps": copy_bytes / (copy_ms / 1e3) / 1e9, "bytes": copy_bytes, "ms": copy_ms},
        "cpu_baseline": cpu,
        "full_size_parity": parity,
        "ingestion_diagnostics_terrain": rows,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
