"""Benchmark of the plain FP32 simulation step on one MI355X.

Workload: 256 x 256 x 128 voxels (x, z horizontal; 128 levels vertical) at
dx = 100 m, dt = 0.5 s, the randomised sounding (seed 20261203, three warm
bubbles, surface theta 296-304 K, RH 60-95 %, 0-25 m/s shear). One step is one
environmentGPU::updateGPU over the whole domain with the reference's solver
settings (20 diffusion sweeps, reference-mode PCG: absolute tolerance 1e-5,
at most 200 iterations, which this workload always reaches). Inputs are
resident in HBM before the timed region; the timed region is K steps
bracketed by device synchronisation.

Also measured: the same step with the textbook Jacobi-PCG at a relative
residual of 1e-6 (the solver the workload description names), and the CPU
baseline (the oracle, one step of the same workload on the host cores), whose
result is compared with one GPU step from the same initial state.

Also timed (key ingestion_diagnostics_terrain, off with --no-rows): reading
the DWD fixture archive, initialising the grid from it on the device against
the host-array init, the serving-diagnostics readback and terrain generation.

Prints one JSON line: cell-steps/s (value), the roofline of the dominant
kernel class (algorithmic bytes per launch / average launch duration from HIP
events on the env's stream during the timed region, attached to every tenth
launch of that class; no other launch carries events there), the per-class breakdown of one extra fully instrumented step,
and the HBM traffic per launch of that kernel from the committed rocprofv3
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


# Fields one launch of a batched class moves, on average over a step: the
# diffusion launches sweep 1, 3 and 6 fields (20 sweeps each), the PPM passes
# 1 (theta), 1, 1, 1 (u, v, w), 1 (vapour) and 5 fields (5 passes each), the
# sedimentation launch 3 fields.
FIELDS_PER_LAUNCH = {"diffusion_sweep": 10.0 / 3.0, "ppm_pass": 10.0 / 6.0, "precip_fall": 3.0}


def algorithmic_bytes(cls, n_sky):
    """Compulsory HBM bytes of one launch of a kernel class (FP32, per sky cell):
    diffusion sweep 8 (read x, write x); PPM pass 12 (read q and the face
    velocity, write q; the unit-density field is computed, not stored);
    PCG spmv 20 (read z, s_old and the density-derived operator as 4 B of
    density, write s and As); PCG update 32 (read p, s, r, As, 1/diag; write
    p, r, z); microphysics 64 (9 fields in, 7 out)."""
    per = {"diffusion_sweep": 8, "ppm_pass": 12, "pcg_spmv": 20, "pcg_update": 32, "microphysics": 64,
           "precip_fall": 20}
    return per.get(cls, 0) * n_sky * FIELDS_PER_LAUNCH.get(cls, 1.0)


SKY_FIELDS = ("QV", "QW", "QC", "QR", "QS", "QI", "POTTEMP", "VELX", "VELY", "VELZ", "PRESSURE", "DENSITY")


def cpu_baseline(snd_args, threads):
    """The oracle's step timed on the host; its result also checks one GPU step
    of the same workload from the same initial state (relative L2 over the sky
    cells per field, PCG iteration counts)."""
    import numpy as np
    import skewtcloudsim_amd as skt
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle_bind import Oracle
    o = Oracle(NX, NY, NZ, DX, threads=threads)
    o.init(*snd_args)
    t0 = time.perf_counter()
    o.step(DT, 1.0)
    t1 = time.perf_counter()
    it = o.stats()["pcg_iterations"]
    env = skt.EnvironmentGPU(NX, NY, NZ, DX)
    env.init(*snd_args)
    env.updateGPU(DT, 1.0)
    sky = (np.arange(NX * NY * NZ) // NX) % NY >= 1
    rel = {}
    for f in SKY_FIELDS:
        a, b = env.get(f)[sky].astype(np.float64), o.get(f)[sky].astype(np.float64)
        rel[f] = float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))
    check = {"rel_l2": rel, "max_rel_l2": max(rel.values()),
             "pcg_iterations": {"gpu": int(env.stats().pcg_iterations), "oracle": int(it)}}
    env.close()
    o.close()
    return (NX * NY * NZ) / (t1 - t0), t1 - t0, it, check


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
    env = skt.EnvironmentGPU(NX, NY, NZ, DX, pcg_mode=skt.PCG_STANDARD, pcg_tol=1e-6, pcg_max_iter=1000)
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
    return {"value": NX * NY * NZ * steps / wall, "ms_per_step": wall / steps * 1e3, "steps": steps,
            "pcg": "textbook Jacobi-PCG, relative residual 1e-6, max 1000", "pcg_iterations": ks}


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
    keys = ("potTemps", "velField", "Qv", "groundTemp", "groundPres", "pressures", "smallPressure")
    t0 = time.perf_counter()
    env.init(*[host[k] for k in keys])
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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-standard", action="store_true", help="skip the relative-1e-6 PCG measurement")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-rows", action="store_true", help="skip the ingestion/diagnostics/terrain timings")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); the step is single-domain, one GPU per "
              "process (on several GPUs the domain would split into z slabs, one per GPU).", file=sys.stderr)
        sys.exit(2)

    import skewtcloudsim_amd as skt
    snd = skt.make_sounding(NX, NY, NZ, DX, seed=SEED, n_bubbles=3)
    keys = ("potTemps", "velField", "Qv", "groundTemp", "groundPres", "pressures", "smallPressure")
    snd_args = [snd[k] for k in keys]
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

    launches, total_ms = timing[dom]
    avg_s = total_ms / launches / 1e3
    alg = algorithmic_bytes(dom, n_sky)
    achieved = alg / avg_s / 1e9 if alg else None
    traffic = traffic_from_profile(dom)

    standard = None if args.no_standard else run_standard(snd_args, min(args.steps, 5), 2)
    rows = None if args.no_rows else measure_next_rows()

    cpu = None
    parity = None
    if not args.no_cpu:
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        cps, secs, it, parity = cpu_baseline(snd_args, threads)
        cpu = {"value": cps, "unit": "cell-steps/s", "cores": threads, "kind": "port",
               "sample": f"one step of the same {NX}x{NY}x{NZ} workload after init ({secs:.1f} s, "
                         f"{it} PCG iterations), oracle/skt_oracle.cpp FP32, OpenMP {threads} threads"}

    out = {
        "metric": "cell-steps/s at 256x256x128 (FP32)",
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
        "copy_probe": {"GBps": copy_bytes / (copy_ms / 1e3) / 1e9, "bytes": copy_bytes, "ms": copy_ms},
        "cpu_baseline": cpu,
        "full_size_parity_one_step": parity,
        "ingestion_diagnostics_terrain": rows,
    }
    env.close()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
