"""Ensembles (skt_batch_*): members of uneven shapes advanced by one launch per
phase give, member by member, the bits of the plain step run alone.

Each test builds every member twice from the same initial state: once alone
(EnvironmentGPU.updateGPU) and once inside an ensemble (EnsembleGPU.step),
then compares every field of every member bit for bit, and the PCG
iteration counts.
"""
import numpy as np
import pytest

import skewtcloudsim_amd as skt

pytestmark = pytest.mark.gpu

KEYS = ("potTemps", "velField", "Qv", "groundTemp", "groundPres", "pressures", "smallPressure")
COMPARED = ("QV", "QW", "QC", "QR", "QS", "QI", "POTTEMP", "VELX", "VELY", "VELZ", "PRESSURE", "DENSITY",
            "PRESSURE_PROJ", "QRS", "QGR", "QGS", "QGI", "GROUND_T", "GROUND_TIME", "CLOUD_COVER")


def _terrain(nx, nz, kind, ny, seed=3):
    rng = np.random.default_rng(seed)
    if kind == "flat":
        return None
    if kind == "random":
        return rng.integers(0, 5, nx * nz).astype(np.int32)
    gh = rng.integers(0, ny - 6, nx * nz)
    gh[rng.uniform(size=nx * nz) < 0.5] = 1
    return gh.astype(np.int32)


def _make(shape, seed, kind="flat", bubbles=2, **cfg):
    nx, ny, nz = shape
    snd = skt.make_sounding(nx, ny, nz, 100.0, seed=seed, n_bubbles=bubbles)
    env = skt.EnvironmentGPU(nx, ny, nz, 100.0, **cfg)
    gh = _terrain(nx, nz, kind, ny, seed)
    if gh is not None:
        env.set_ground_height(gh)
    env.init(*[snd[k] for k in KEYS])
    return env


def _wet(env, seed):
    """Clouds, rain, snow and ice and a cold upper half, so every branch runs."""
    rng = np.random.default_rng(seed)
    n = env.N
    for f, hi in (("QW", 2e-3), ("QC", 1e-3), ("QR", 1.5e-3), ("QS", 1e-3), ("QI", 1e-3)):
        v = rng.uniform(0, hi, n).astype(np.float32)
        v[rng.uniform(size=n) < 0.3] = 0.0
        env.set(f, v)
    th = env.get("POTTEMP").reshape(env.nz, env.ny, env.nx)
    th[:, env.ny // 2:, :] -= 25.0
    env.set("POTTEMP", th.ravel())
    env.initDensity()


MEMBERS = [((24, 20, 16), 11, "flat"), ((17, 13, 11), 12, "random"), ((40, 48, 24), 13, "tall"),
           ((33, 18, 21), 14, "random"), ((64, 32, 48), 15, "flat"), ((8, 2, 3), 16, "flat")]


def _pairs(members, wet=True, **cfg):
    solo, mix = [], []
    for shape, seed, kind in members:
        a, b = _make(shape, seed, kind, **cfg), _make(shape, seed, kind, **cfg)
        if wet:
            _wet(a, seed)
            _wet(b, seed)
        solo.append(a)
        mix.append(b)
    return solo, mix


def _assert_same(a, b, tag):
    for f in COMPARED:
        x, y = a.get(f), b.get(f)
        nd = int((x.view(np.uint32) != y.view(np.uint32)).sum())
        assert nd == 0, (tag, f, nd, "cells differ")
    assert a.stats().pcg_iterations == b.stats().pcg_iterations, tag


@pytest.mark.parametrize("use_graph,mode", [(True, 0), (False, 0), (True, 1), (True, 2)])
def test_members_alone_equal_members_in_the_mix(use_graph, mode):
    solo, mix = _pairs(MEMBERS)
    ens = skt.EnsembleGPU(use_graph=use_graph, launch_mode=mode)
    for env in mix:
        ens.add(env)
    for s in range(3):
        print(f"ensemble step {s} graph={use_graph} mode={mode}", flush=True)
        for env in solo:
            env.updateGPU(0.5, 1.0)
        ens.step(0.5, 1.0)
    ens.synchronize()
    for k, (a, b) in enumerate(zip(solo, mix)):
        _assert_same(a, b, f"member {k} {MEMBERS[k][0]}")
        assert a.params().time == b.params().time
    ens.close()
    for env in solo + mix:
        env.close()


def test_members_join_and_leave_between_steps():
    """Continuous batching: a member leaves after two steps, another joins
    after one; each still follows its own plain run."""
    solo, mix = _pairs(MEMBERS[:4])
    ens = skt.EnsembleGPU()
    ids = [ens.add(env) for env in mix[:3]]
    steps = [0, 0, 0, 0]
    for s in range(4):
        print(f"join/leave step {s}", flush=True)
        if s == 1:
            ids.append(ens.add(mix[3]))
        if s == 2:
            ens.remove(ids[0])
        active = [k for k in range(4) if (k != 0 or s < 2) and (k != 3 or s >= 1)]
        for k in active:
            solo[k].updateGPU(0.5, 1.0)
            steps[k] += 1
        ens.step(0.5, 1.0)
    ens.synchronize()
    assert steps == [2, 4, 4, 3]
    for k in range(4):
        _assert_same(solo[k], mix[k], f"member {k}")
    # the removed member runs on its own again, on its own stream
    solo[0].updateGPU(0.5, 1.0)
    mix[0].updateGPU(0.5, 1.0)
    _assert_same(solo[0], mix[0], "member 0 after leaving")
    ens.close()
    for env in solo + mix:
        env.close()


def test_autotuned_widths_change_no_result():
    solo, mix = _pairs(MEMBERS[:5])
    ens = skt.EnsembleGPU()
    for env in mix:
        ens.add(env)
    tun = ens.autotune(0.5, 1.0)  # three ordinary steps, one per launch width
    print("autotune", tun, flush=True)
    assert tun["bodies"] > 10 and all(t > 0 for t in tun["step_ms_by_width"])
    assert tun["tuned_step_ms"] <= min(tun["step_ms_by_width"]) + 1e-9
    for _ in range(3):
        for env in solo:
            env.updateGPU(0.5, 1.0)
    ens.step(0.5, 1.0)
    for env in solo:
        env.updateGPU(0.5, 1.0)
    ens.synchronize()
    for k, (a, b) in enumerate(zip(solo, mix)):
        _assert_same(a, b, f"tuned member {k}")
    ens.close()
    for env in solo + mix:
        env.close()


def test_standard_solver_members():
    """Members with the textbook solver to a relative residual of 1e-6: their
    solves stop at different iteration counts inside one ensemble launch."""
    cfg = dict(pcg_mode=1, pcg_tol=1e-6, pcg_max_iter=400)
    solo, mix = _pairs(MEMBERS[:4], **cfg)
    ens = skt.EnsembleGPU()
    for env in mix:
        ens.add(env)
    its = set()
    for _ in range(2):
        for env in solo:
            env.updateGPU(0.5, 1.0)
            its.add(env.stats().pcg_iterations)
        ens.step(0.5, 1.0)
    ens.synchronize()
    assert len(its) > 1
    for k, (a, b) in enumerate(zip(solo, mix)):
        _assert_same(a, b, f"standard-solver member {k}")
    ens.close()
    for env in solo + mix:
        env.close()


def test_substep_cap_overflow_is_reported():
    """A member whose advection needs more substep passes than the ensemble
    holds (Courant number above the cap) is reported, not silently wrong."""
    env = _make((24, 20, 16), 21)
    u = env.get("VELX")
    u[:] = 450.0  # Courant 450 * 0.5 / 100 = 2.25: three substeps
    env.set("VELX", u)
    ens = skt.EnsembleGPU(advection_substep_cap=2)
    ens.add(env)
    ens.step(0.5, 1.0)
    with pytest.raises(skt.SktError, match="substeps"):
        ens.synchronize()
    ens.close()
    env.close()


def test_members_must_share_solver_settings():
    a = _make((24, 20, 16), 31)
    b = _make((24, 20, 16), 32, pcg_mode=1, pcg_tol=1e-6, pcg_max_iter=400)
    ens = skt.EnsembleGPU()
    ens.add(a)
    with pytest.raises(skt.SktError, match="share"):
        ens.add(b)
    ens.close()
    a.close()
    b.close()
