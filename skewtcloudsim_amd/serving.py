"""Serving: 'simulate sounding S for T seconds' jobs on one GPU (skt_server_*).

Server wraps the native server: jobs queue, are admitted by HBM footprint
into free slots, run as members of one ensemble (mixed precision by
default), report max updraft, cloud-top height and ground rain every
report interval of simulated time, and leave when done.

run_poisson drives a server with Poisson arrivals on the wall clock and
returns throughput and latency percentiles.
"""
import ctypes
import time

import numpy as np

from . import _lib
from .environment import SktError, default_config

# grid sizes of the request mix (nx, ny vertical, nz): the ensemble sizes plus
# the 256 x 256 x 128 workload
REQUEST_GRIDS = [(64, 64, 64), (128, 128, 128), (192, 96, 192), (128, 160, 128), (256, 128, 256)]


def job_footprint(nx, ny, nz, dx=100.0, **config):
    """HBM bytes an environment of this grid and configuration allocates."""
    lib = _lib.load_library()
    cfg = default_config(nx, ny, nz, dx, **config)
    b = ctypes.c_int64()
    rc = lib.skt_env_footprint(ctypes.byref(cfg), ctypes.byref(b))
    if rc:
        raise SktError(f"skt_env_footprint failed ({rc})")
    return b.value


class Server:
    def __init__(self, device=0, dt=0.5, hbm_budget_bytes=0, max_active=64, cloud_threshold=1e-5,
                 member_config=None, use_graph=True, advection_substep_cap=4, launch_mode=0):
        self.lib = _lib.load_library()
        o = _lib.ServerOptions()
        self.lib.skt_server_options_default(ctypes.byref(o))
        o.device, o.dt, o.hbm_budget_bytes, o.max_active = device, dt, int(hbm_budget_bytes), max_active
        o.cloud_threshold = cloud_threshold
        for k, v in (member_config or {}).items():
            if isinstance(v, (list, tuple)):
                arr = getattr(o.member_config, k)
                for i, x in enumerate(v):
                    arr[i] = x
            else:
                setattr(o.member_config, k, v)
        o.batch = _lib.BatchOptions(int(bool(use_graph)), int(advection_substep_cap), int(launch_mode))
        h = ctypes.c_void_p()
        rc = self.lib.skt_server_create(ctypes.byref(o), ctypes.byref(h))
        if rc:
            raise SktError(f"skt_server_create failed ({rc})")
        self.h = h
        self.options = o

    def _chk(self, rc, what):
        if rc:
            msg = self.lib.skt_server_last_error(self.h)
            raise SktError(f"{what}: status {rc}: {msg.decode() if msg else ''}")

    def submit(self, nx, ny, nz, sim_seconds, dx=100.0, seed=1, n_bubbles=2, report_every=60.0):
        """Queues 'simulate the seeded sounding for sim_seconds'; returns the job id."""
        r = _lib.JobRequest()
        r.nx, r.ny, r.nz, r.dx = nx, ny, nz, dx
        rc = self.lib.skt_sounding_spec_random(ctypes.byref(r.sounding), int(seed), nx, ny, nz, dx, int(n_bubbles))
        if rc:
            raise SktError(f"skt_sounding_spec_random failed ({rc})")
        r.sim_seconds, r.report_every = sim_seconds, report_every
        jid = ctypes.c_int64()
        self._chk(self.lib.skt_server_submit(self.h, ctypes.byref(r), ctypes.byref(jid)), "submit")
        return jid.value

    def step(self):
        self._chk(self.lib.skt_server_step(self.h), "step")

    def poll(self):
        out, evs = (_lib.JobEvent * 256)(), []
        n = ctypes.c_int32()
        while True:
            self._chk(self.lib.skt_server_poll(self.h, out, 256, ctypes.byref(n)), "poll")
            for e in out[:n.value]:
                evs.append({"job": e.job, "kind": _lib.EVENT_KINDS[e.kind], "steps": e.steps,
                            "sim_time": e.sim_time, "wall_s": e.wall_s, "submitted_s": e.submitted_s,
                            "max_updraft": e.diag.max_updraft, "cloud_top_level": e.diag.cloud_top_level,
                            "cloud_top_height": e.diag.cloud_top_height, "surface_rain": e.diag.surface_rain})
            if n.value < 256:
                return evs

    def status(self):
        s = _lib.ServerStatus()
        self._chk(self.lib.skt_server_status_get(self.h, ctypes.byref(s)), "status")
        return {k: getattr(s, k) for k, _ in s._fields_}

    def close(self):
        if self.h:
            self.lib.skt_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def run_poisson(rate, n_requests, seed=1, grids=REQUEST_GRIDS, t_range=(300.0, 1800.0), report_every=60.0,
                time_scale=1.0, **server_kw):
    """Poisson arrivals at 'rate' requests/s (wall clock) of n_requests jobs,
    each a random grid from 'grids', a random sounding with 1-4 bubbles and T
    uniform in t_range, both T and the report interval multiplied by
    time_scale. The server steps whenever it has work; returns latency
    percentiles (submission to the last report), throughput and the events."""
    rng = np.random.default_rng(seed)
    gaps = rng.exponential(1.0 / rate, n_requests)
    arrive = np.cumsum(gaps) - gaps[0]
    jobs = [(grids[int(rng.integers(len(grids)))], float(rng.uniform(*t_range)) * time_scale,
             int(rng.integers(1, 2 ** 31)), int(rng.integers(1, 5))) for _ in range(n_requests)]
    srv = Server(**server_kw)
    events, done, nxt = [], {}, 0
    t0 = time.perf_counter()
    try:
        while len(done) < n_requests:
            now = time.perf_counter() - t0
            while nxt < n_requests and arrive[nxt] <= now:
                (nx, ny, nz), T, sd, nb = jobs[nxt]
                srv.submit(nx, ny, nz, T, seed=sd, n_bubbles=nb, report_every=report_every * time_scale)
                nxt += 1
            st = srv.status()
            if st["active"] or st["queued"]:
                srv.step()
            elif nxt < n_requests:
                time.sleep(max(0.0, min(arrive[nxt] - (time.perf_counter() - t0), 0.05)))
            for e in srv.poll():
                events.append(e)
                if e["kind"] == "done":
                    done[e["job"]] = e
        wall = time.perf_counter() - t0
        st = srv.status()
    finally:
        srv.close()
    lat = np.array([e["wall_s"] - e["submitted_s"] for e in done.values()])
    first = min(e["submitted_s"] for e in done.values())
    last = max(e["wall_s"] for e in done.values())
    dt = server_kw.get("dt", 0.5)
    cells = sum(nx * ny * nz * max(1, round(T / dt)) for (nx, ny, nz), T, _, _ in jobs)
    return {"rate_per_s": rate, "requests": n_requests, "completed": len(done),
            "throughput_jobs_per_s": len(done) / (last - first), "cell_steps_per_s": st["cell_steps"] / wall,
            "latency_p50_s": float(np.percentile(lat, 50)), "latency_p99_s": float(np.percentile(lat, 99)),
            "latency_max_s": float(lat.max()), "ensemble_steps": st["steps"], "busy_s": st["busy_s"],
            "admit_s": st["admit_s"], "replan_s": st["replan_s"],
            "wall_s": wall, "offered_cell_steps": cells, "events": events}
