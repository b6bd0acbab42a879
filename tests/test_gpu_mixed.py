"""The mixed-precision form: hydrometeors stored in 16 bits with per-field
power-of-two scales, and FP64 residual refinement of the pressure solve.

* storage: the host copies of a 16-bit field equal a numpy restatement of
  round-to-nearest-even BF16 / FP16 conversion of value * scale, bit for bit;
* refinement: the FP64 residual of the final solution meets the requested
  stop value, down to 1e-10 of max|b| -- which FP32 arithmetic cannot reach,
  and which needs the FP64 residual operator to equal the FP32 solver's
  (the refinement converges only as fast as the two agree);
* the mixed step against the FP64 oracle, field by field, next to the FP32
  step's own distance to it;
* an ensemble of mixed members gives each member's bits alone.
"""
import numpy as np
import pytest

import oracle_bind as ob
from test_gpu_parity import KEYS, rel_l2, rich_state, sky_mask

pytestmark = pytest.mark.gpu

TRACERS = ("QW", "QC", "QR", "QS", "QI")
FMT = {"bf16": 1, "f16": 2}


def _bf16_rne(v, scale):
    """value * scale rounded to bfloat16 (nearest even), returned as float32 / scale."""
    s = (np.asarray(v, np.float32) * np.float32(scale)).astype(np.float32)
    u = s.view(np.uint32).astype(np.uint64)
    h = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint32)
    nan = (u & 0x7FFFFFFF) > 0x7F800000
    h[nan] = ((u[nan] >> 16) | 0x40).astype(np.uint32)
    return ((h << 16).astype(np.uint32).view(np.float32) * np.float32(1.0 / scale)).astype(np.float32)


def _f16_rne(v, scale):
    s = (np.asarray(v, np.float32) * np.float32(scale)).astype(np.float32)
    return (s.astype(np.float16).astype(np.float32) * np.float32(1.0 / scale)).astype(np.float32)


def _env(shape, seed=5, **cfg):
    import skewtcloudsim_amd as skt
    nx, ny, nz = shape
    snd = skt.make_sounding(nx, ny, nz, 100.0, seed=seed, n_bubbles=2)
    env = skt.EnvironmentGPU(nx, ny, nz, 100.0, **cfg)
    env.init(*[snd[k] for k in KEYS])
    return env, snd


@pytest.mark.parametrize("fmt", ["bf16", "f16"])
def test_16bit_storage_round_trip(fmt):
    """Values over 1e-8..5e-2 kg/kg (both signs, zeros, halfway cases) go
    through the device's storage exactly as the numpy restatement rounds them;
    every other field stays float."""
    scales = [16, 12, 0, 20, -3]
    env, _ = _env((24, 20, 16), tracer_format=FMT[fmt], tracer_scale_log2=scales)
    rng = np.random.default_rng(7)
    n = env.N
    conv = _bf16_rne if fmt == "bf16" else _f16_rne
    for k, f in enumerate(TRACERS):
        v = (10.0 ** rng.uniform(-8, np.log10(5e-2), n)).astype(np.float32)
        v[rng.uniform(size=n) < 0.1] = 0.0
        v[rng.uniform(size=n) < 0.05] *= -1
        # scaled values exactly halfway between two 16-bit neighbours: the
        # bits below BF16's 7 (FP16's 10) mantissa bits are 1000...0
        if fmt == "bf16":
            u = (rng.integers(0x3C000000, 0x3E000000, 64).astype(np.uint32) & np.uint32(0xFFFF0000)) | 0x8000
        else:
            u = (rng.integers(0x3F800000, 0x40000000, 64).astype(np.uint32) & np.uint32(0xFFFFE000)) | 0x1000
        v[:64] = u.view(np.float32) / np.float32(2.0 ** scales[k])
        env.set(f, v)
        got = env.get(f)
        want = conv(v, 2.0 ** scales[k])
        nd = int((got.view(np.uint32) != want.view(np.uint32)).sum())
        assert nd == 0, (fmt, f, nd)
        assert env.field_format(f) == (fmt, 2.0 ** scales[k])
    for f in ("QV", "POTTEMP", "VELX", "PRESSURE", "QGR"):
        assert env.field_format(f) == ("f32", 1.0)
    env.close()


def _refine_cfg(tol, refine, rtol, **kw):
    return dict(pcg_mode=1, pcg_tol=tol, pcg_max_iter=600, pcg_refine=refine, pcg_refine_tol=rtol, **kw)


@pytest.mark.parametrize("tol,refine,rtol", [(1e-6, 2, 1e-3), (1e-10, 3, 1e-4)])
def test_refinement_meets_the_fp64_stop_value(tol, refine, rtol):
    """After a step, max|b - A x| computed in FP64 for the final x is at most
    tol * max|b| (a small slack covers the last solve stopping on its FP32
    recurrence). At 1e-10 this lies far below what FP32 arithmetic reaches,
    so the passes must converge, which they do only as fast as the FP64
    residual operator matches the FP32 solver's."""
    shape = (40, 32, 24)
    env, snd = _env(shape, **_refine_cfg(tol, refine, rtol))
    o32 = ob.Oracle(*shape, 100.0, pcg_mode=1, pcg_tol=1e-6, pcg_max_iter=600)
    o32.init(*[snd[k] for k in KEYS])
    rich_state(env, o32, seed=61)
    for s in range(2):
        env.updateGPU(0.5, 1.0)
        st = env.stats()
        rel = st.pcg_residual64 / st.pcg_rhs64
        print(f"step {s}: passes {st.pcg_passes}, iterations {st.pcg_iterations_total}, "
              f"FP64 residual / max|b| {rel:.2e}", flush=True)
        assert st.pcg_rhs64 > 0 and st.pcg_passes >= 2, (st.pcg_passes, st.pcg_rhs64)
        assert rel <= 3.0 * tol, rel
    env.close()
    o32.close()


def test_refined_solution_matches_a_tight_solution():
    """The default refinement (two passes of 1e-3 to a 1e-6 stop value) and
    the plain FP32 solve at 1e-6, each against a solve refined to 1e-11 from
    the same state: the refined pressure correction is at least as close."""
    shape = (40, 32, 24)
    base = dict(pcg_mode=1, pcg_max_iter=600)
    cfgs = {"plain": dict(base, pcg_tol=1e-6), "refined": _refine_cfg(1e-6, 2, 1e-3),
            "tight": _refine_cfg(1e-11, 4, 1e-4)}
    envs = {}
    for k, c in cfgs.items():
        envs[k], snd = _env(shape, **c)
    o32 = ob.Oracle(*shape, 100.0, pcg_mode=1, pcg_tol=1e-6, pcg_max_iter=600)
    o32.init(*[snd[k] for k in KEYS])
    rich_state(envs["plain"], o32, seed=62)
    for k in ("refined", "tight"):
        for f in ob.FIELDS:
            envs[k].set(f, o32.get(f))
    m = sky_mask(*shape)
    for e in envs.values():
        e.updateGPU(0.5, 1.0)
    t = envs["tight"].get("PRESSURE_PROJ")[m]
    e_ref = rel_l2(envs["refined"].get("PRESSURE_PROJ")[m], t)
    e_pl = rel_l2(envs["plain"].get("PRESSURE_PROJ")[m], t)
    print(f"p' vs the 1e-11 solve: refined {e_ref:.2e}, plain FP32 {e_pl:.2e}", flush=True)
    assert e_ref <= e_pl * 1.5 + 1e-7, (e_ref, e_pl)
    for e in envs.values():
        e.close()
    o32.close()


def test_mixed_steps_against_the_fp64_oracle():
    """20 steps from a state with every hydrometeor, a cold upper half and
    winds: the FP32 step, the BF16 form and the FP16 form (both with FP64
    refinement), and the BF16 form with the multigrid preconditioner,
    against the FP64 oracle. Per field, relative L2 over the sky
    cells. Each mixed form stays within twice the FP32 step's distance plus
    a storage term: half a 16-bit unit in the last place (2^-9 BF16, 2^-11
    FP16) times the number of rounded stores a value takes (about ten per
    step) for the hydrometeors, and a tenth of that for the fields they feed
    (theta, vapour, velocities, pressure)."""
    shape = (40, 32, 24)
    cfg = dict(pcg_mode=1, pcg_tol=1e-6, pcg_max_iter=600)
    env32, snd = _env(shape, **cfg)
    args = [snd[k] for k in KEYS]
    o32 = ob.Oracle(*shape, 100.0, **cfg)
    o32.init(*args)
    rich_state(env32, o32, seed=63)
    o64 = ob.Oracle(*shape, 100.0, fp64=True, **cfg)
    o64.init(*args)
    for f in ob.FIELDS:
        o64.set(f, o32.get(f))
    mixed = {}
    variants = [(name, dict(tracer_format=fmt)) for name, fmt in FMT.items()]
    # the benchmarked form: BF16 with the multigrid preconditioner
    variants.append(("bf16+mg", dict(tracer_format=1, pcg_precon=1)))
    for name, extra in variants:
        e, _ = _env(shape, pcg_refine=2, pcg_refine_tol=1e-3, **extra, **cfg)
        for f in ob.FIELDS:
            e.set(f, o32.get(f))
        mixed[name] = e
    steps = 20
    for _ in range(steps):
        env32.updateGPU(0.5, 1.0)
        o64.step(0.5, 1.0)
        for e in mixed.values():
            e.updateGPU(0.5, 1.0)
    m = sky_mask(*shape)
    ulp = {"bf16": 2.0 ** -9, "f16": 2.0 ** -11, "bf16+mg": 2.0 ** -9}
    rows = []
    for f in ("POTTEMP", "QV", "PRESSURE", "VELX", "VELY", "VELZ") + TRACERS:
        d = o64.get64(f)[m]
        e32 = rel_l2(env32.get(f)[m], d)
        row = [f, e32]
        for name, e in mixed.items():
            em = rel_l2(e.get(f)[m], d)
            assert np.isfinite(e.get(f)).all(), (name, f)
            storage = 10 * steps * ulp[name] * (1.0 if f in TRACERS else 0.1)
            assert em <= 2.0 * e32 + storage, (name, f, em, e32)
            row.append(em)
        rows.append(row)
    print("field      fp32-fp64  " + "  ".join(f"{n + '-fp64':>12}" for n in mixed), flush=True)
    for r in rows:
        print(f"{r[0]:<10} {r[1]:9.2e}  " + "  ".join(f"{v:12.2e}" for v in r[2:]), flush=True)
    for e in [env32, *mixed.values()]:
        e.close()
    o32.close()
    o64.close()


def test_mixed_members_alone_equal_members_in_the_mix():
    """BF16 members with refinement, of uneven shapes, in one ensemble."""
    import skewtcloudsim_amd as skt
    cfg = dict(tracer_format=1, pcg_mode=1, pcg_tol=1e-6, pcg_max_iter=400, pcg_refine=2, pcg_refine_tol=1e-3)
    shapes = [((24, 20, 16), 71), ((33, 18, 21), 72), ((40, 32, 24), 73)]
    solo, mix = [], []
    for shape, seed in shapes:
        for lst in (solo, mix):
            e, snd = _env(shape, seed=seed, **cfg)
            rng = np.random.default_rng(seed)
            for f, hi in (("QW", 2e-3), ("QC", 1e-3), ("QR", 1.5e-3), ("QS", 1e-3), ("QI", 1e-3)):
                v = rng.uniform(0, hi, e.N).astype(np.float32)
                v[rng.uniform(size=e.N) < 0.3] = 0.0
                e.set(f, v)
            e.initDensity()
            lst.append(e)
    ens = skt.EnsembleGPU()
    for e in mix:
        ens.add(e)
    for s in range(3):
        print(f"mixed ensemble step {s}", flush=True)
        for e in solo:
            e.updateGPU(0.5, 1.0)
        ens.step(0.5, 1.0)
    ens.synchronize()
    for k, (a, b) in enumerate(zip(solo, mix)):
        for f in ("QV", "QW", "QC", "QR", "QS", "QI", "POTTEMP", "VELX", "VELY", "VELZ", "PRESSURE",
                  "PRESSURE_PROJ", "DENSITY"):
            x, y = a.get(f), b.get(f)
            nd = int((x.view(np.uint32) != y.view(np.uint32)).sum())
            assert nd == 0, (k, f, nd)
        sa, sb = a.stats(), b.stats()
        assert (sa.pcg_iterations_total, sa.pcg_passes, sa.pcg_residual64) == \
            (sb.pcg_iterations_total, sb.pcg_passes, sb.pcg_residual64), k
    ens.close()
    for e in solo + mix:
        e.close()
