"""The multigrid preconditioner (pcg_precon = SKT_PRECON_MG).

* the preconditioned solve reaches the solution of the Jacobi-preconditioned
  one: from the same state, both solved to a relative residual of 1e-9, the
  pressure corrections agree to about the solves' tolerance, on flat ground
  and on tall random terrain with odd grid sizes;
* its iteration count is a small fraction of Jacobi's (a fifth or less on
  flat ground, half or less on tall random terrain);
* with FP64 refinement, the final FP64 residual meets the stop value;
* an ensemble of members of uneven sizes (their level counts differ) gives
  each member's bits alone.
"""
import numpy as np
import pytest

import oracle_bind as ob
from test_gpu_parity import KEYS, rel_l2, rich_state, sky_mask

pytestmark = pytest.mark.gpu

MG = 1


def _terrain(nx, nz, ny, kind, seed=3):
    if kind == "flat":
        return None
    rng = np.random.default_rng(seed)
    gh = rng.integers(0, ny - 6, nx * nz)
    gh[rng.uniform(size=nx * nz) < 0.5] = 1
    return gh.astype(np.int32)


def _state(shape, kind, seed, cfgs):
    import skewtcloudsim_amd as skt
    nx, ny, nz = shape
    snd = skt.make_sounding(nx, ny, nz, 100.0, seed=seed, n_bubbles=3)
    gh = _terrain(nx, nz, ny, kind, seed)
    o32 = ob.Oracle(nx, ny, nz, 100.0, pcg_mode=1, pcg_tol=1e-6, pcg_max_iter=600)
    if gh is not None:
        o32.set_ground_height(gh)
    o32.init(*[snd[k] for k in KEYS])
    envs = []
    for cfg in cfgs:
        e = skt.EnvironmentGPU(nx, ny, nz, 100.0, **cfg)
        if gh is not None:
            e.set_ground_height(gh)
        e.init(*[snd[k] for k in KEYS])
        envs.append(e)
    rich_state(envs[0], o32, seed=seed)
    for e in envs[1:]:
        for f in ob.FIELDS:
            e.set(f, o32.get(f))
    m = sky_mask(nx, ny, nz, gh)
    o32.close()
    return envs, m


@pytest.mark.parametrize("shape,kind", [((64, 48, 40), "flat"), ((45, 37, 29), "tall"), ((128, 64, 96), "flat")])
def test_multigrid_solve_reaches_the_jacobi_solution(shape, kind):
    tight = dict(pcg_mode=1, pcg_tol=1e-9, pcg_max_iter=4000)
    envs, m = _state(shape, kind, 41, [tight, dict(tight, pcg_precon=MG)])
    its = []
    for e in envs:
        e.calculatePressureProject(0.5)
        its.append(e.stats().pcg_iterations)
    a, b = envs[0].get("PRESSURE_PROJ")[m], envs[1].get("PRESSURE_PROJ")[m]
    d = rel_l2(b, a)
    print(f"{shape} {kind}: iterations jacobi {its[0]}, multigrid {its[1]}; p' rel L2 {d:.2e}", flush=True)
    assert its[0] < 4000 and its[1] < 4000
    # tall random terrain (half the columns 1 to ny - 7 cells high) leaves
    # coarse cells that are partly ground, where the coarse operator fits less
    assert its[1] * (5 if kind == "flat" else 2) <= its[0], its
    assert d <= 1e-5, d
    for e in envs:
        e.close()


def test_multigrid_steps_track_jacobi_steps():
    """Five whole steps at the bench's solver setting (relative 1e-6): every
    field stays within the distance two solves at that tolerance allow."""
    cfg = dict(pcg_mode=1, pcg_tol=1e-6, pcg_max_iter=2000)
    envs, m = _state((64, 48, 40), "flat", 42, [cfg, dict(cfg, pcg_precon=MG)])
    for s in range(5):
        for e in envs:
            e.updateGPU(0.5, 1.0)
        print(f"step {s}: iterations jacobi {envs[0].stats().pcg_iterations}, "
              f"multigrid {envs[1].stats().pcg_iterations}", flush=True)
    for f in ("POTTEMP", "PRESSURE", "QV", "VELX", "VELY", "VELZ", "QW", "QC", "QR", "QS", "QI"):
        d = rel_l2(envs[1].get(f)[m], envs[0].get(f)[m])
        print(f"{f}: {d:.2e}", flush=True)
        assert d <= 2e-3, (f, d)
    for e in envs:
        e.close()


def test_multigrid_with_refinement_meets_the_fp64_stop_value():
    cfg = dict(pcg_mode=1, pcg_tol=1e-10, pcg_max_iter=400, pcg_refine=3, pcg_refine_tol=1e-4, pcg_precon=MG,
               tracer_format=1)
    envs, m = _state((64, 48, 40), "tall", 43, [cfg])
    env = envs[0]
    for s in range(2):
        env.updateGPU(0.5, 1.0)
        st = env.stats()
        rel = st.pcg_residual64 / st.pcg_rhs64
        print(f"step {s}: passes {st.pcg_passes}, iterations {st.pcg_iterations_total}, FP64 residual {rel:.2e}",
              flush=True)
        assert rel <= 3e-10, rel
    env.close()


def test_multigrid_members_alone_equal_members_in_the_mix():
    import skewtcloudsim_amd as skt
    cfg = dict(pcg_mode=1, pcg_tol=1e-6, pcg_max_iter=300, pcg_precon=MG, tracer_format=1, pcg_refine=2,
               pcg_refine_tol=1e-3)
    shapes = [((24, 20, 16), 81), ((8, 2, 3), 82), ((64, 48, 40), 83), ((33, 18, 21), 84), ((128, 64, 96), 85)]
    solo, mix = [], []
    for shape, seed in shapes:
        snd = skt.make_sounding(*shape, 100.0, seed=seed, n_bubbles=2)
        for lst in (solo, mix):
            e = skt.EnvironmentGPU(*shape, 100.0, **cfg)
            e.init(*[snd[k] for k in KEYS])
            lst.append(e)
    ens = skt.EnsembleGPU()
    for e in mix:
        ens.add(e)
    for s in range(3):
        print(f"multigrid ensemble step {s}", flush=True)
        for e in solo:
            e.updateGPU(0.5, 1.0)
        ens.step(0.5, 1.0)
    ens.synchronize()
    for k, (a, b) in enumerate(zip(solo, mix)):
        for f in ("QV", "QW", "POTTEMP", "VELX", "VELY", "VELZ", "PRESSURE", "PRESSURE_PROJ", "DENSITY"):
            x, y = a.get(f), b.get(f)
            nd = int((x.view(np.uint32) != y.view(np.uint32)).sum())
            assert nd == 0, (k, f, nd)
        assert a.stats().pcg_iterations_total == b.stats().pcg_iterations_total
    ens.close()
    for e in solo + mix:
        e.close()
