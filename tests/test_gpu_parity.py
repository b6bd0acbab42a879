"""Parity of the HIP step (libskt.so, through the C ABI) with the CPU oracle.

Same seeded inputs on both sides; each phase is called alone on identical
states, then whole steps are compared. Tolerances:
  * masks and microphysics phase masks: bit-exact;
  * the advection and diffusion phases: bit-exact on sky cells (no
    transcendental function and the same operation order on both sides);
  * one phase call: relative L2 <= 2e-6 and max-abs <= 1e-5 x field scale
    (libm/ocml transcendental ulps and FP64 reduction order only);
  * whole steps: relative L2 per field <= 1e-5 after 1 step, <= 1e-4 after 5.
Only sky cells are compared: the step leaves ground cells untouched, while
the reference leaves NaN there (0/0 of the density renormalisation).
"""
import numpy as np
import pytest

import oracle_bind as ob

pytestmark = pytest.mark.gpu

KEYS = ("potTemps", "velField", "Qv", "groundTemp", "groundPres", "pressures", "smallPressure")
SKY = ("QV", "QW", "QC", "QR", "QS", "QI", "POTTEMP", "VELX", "VELY", "VELZ", "PRESSURE", "DENSITY")
GROUND = ("QRS", "QGR", "QGS", "QGI", "GROUND_T", "GROUND_TIME", "CLOUD_COVER")


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30)


def sky_mask(nx, ny, nz, gh=None):
    y = (np.arange(nx * ny * nz) // nx) % ny
    if gh is None:
        return y >= 1
    x = np.arange(nx * ny * nz) % nx
    z = np.arange(nx * ny * nz) // (nx * ny)
    return y > gh[x + nx * z]


def make_pair(nx, ny, nz, dx=100.0, seed=5, n_bubbles=2, **cfg):
    import skewtcloudsim_amd as skt
    snd = skt.make_sounding(nx, ny, nz, dx, seed=seed, n_bubbles=n_bubbles)
    env = skt.EnvironmentGPU(nx, ny, nz, dx, **cfg)
    ocfg = {}
    if "pcg_mode" in cfg:
        ocfg["pcg_mode"] = cfg["pcg_mode"]
    if "pcg_tol" in cfg:
        ocfg["pcg_tol"] = cfg["pcg_tol"]
    if "pcg_max_iter" in cfg:
        ocfg["pcg_max_iter"] = cfg["pcg_max_iter"]
    if "maxabs_mode" in cfg:
        ocfg["maxabs_mode"] = cfg["maxabs_mode"]
    orc = ob.Oracle(nx, ny, nz, dx, **ocfg)
    env.init(*[snd[k] for k in KEYS])
    orc.init(*[snd[k] for k in KEYS])
    return env, orc


def copy_state(env, orc, fields):
    for f in fields:
        env.set(f, orc.get(f))


def rich_state(env, orc, seed=1):
    """Clouds, rain, snow, ice, winds and a cold upper layer on both sides."""
    rng = np.random.default_rng(seed)
    n = orc.N
    # perturb the initialised state, not what earlier tests left behind
    if not hasattr(orc, "_base"):
        orc._base = {f: orc.get(f).copy() for f in ob.FIELDS}
    for f, v in orc._base.items():
        orc.set(f, v)
    for f, hi in (("QW", 2.5e-3), ("QC", 1.5e-3), ("QR", 2e-3), ("QS", 1e-3), ("QI", 1e-3)):
        v = rng.uniform(0, hi, n).astype(np.float32)
        v[rng.uniform(size=n) < 0.3] = 0.0
        orc.set(f, v)
    orc.set("QV", orc.get("QV") * rng.uniform(0.85, 1.15, n).astype(np.float32))
    for f, s in (("VELX", 3.0), ("VELY", 1.5), ("VELZ", 3.0)):
        orc.set(f, (orc.get(f) + rng.normal(0, s, n)).astype(np.float32))
    th = orc.get("POTTEMP").reshape(orc.nz, orc.ny, orc.nx)
    th[:, orc.ny // 2:, :] -= 30.0  # sub-freezing upper half
    orc.set("POTTEMP", th.ravel())
    orc.init_density()
    g = orc.G
    orc.set("QGR", rng.uniform(0, 1e-3, g).astype(np.float32))
    orc.set("QGS", rng.uniform(0, 5e-4, g).astype(np.float32))
    copy_state(env, orc, ob.FIELDS)


def assert_bits(env, orc, fields, mask, tag=""):
    """The fields hold the same float bits on the masked cells."""
    for f in fields:
        a, b = env.get(f), orc.get(f)
        if a.size == mask.size:
            a, b = a[mask], b[mask]
        nd = int((a.view(np.uint32) != b.view(np.uint32)).sum())
        assert nd == 0, (tag, f, nd, "cells differ")


def assert_fields(env, orc, fields, mask, l2=2e-6, maxabs=1e-5, tag=""):
    for f in fields:
        a, b = env.get(f), orc.get(f)
        if a.size == mask.size:
            a, b = a[mask], b[mask]
        assert np.isfinite(a).all(), (tag, f, "non-finite")
        scale = max(np.abs(b).max(), 1e-30)
        r = rel_l2(a, b)
        m = np.abs(a.astype(np.float64) - b).max() / scale
        assert r <= l2 and m <= maxabs, (tag, f, r, m)


NX, NY, NZ = 20, 24, 16


@pytest.fixture(scope="module")
def pair():
    env, orc = make_pair(NX, NY, NZ, dx=400.0)
    yield env, orc
    env.close()
    orc.close()


def test_init_matches(pair):
    env, orc = pair
    m = sky_mask(NX, NY, NZ)
    assert_fields(env, orc, SKY + GROUND, m, l2=1e-7, maxabs=1e-6, tag="init")
    for f in ("ISEN_TEMP", "ISEN_VAPOR", "DEFAULT_VELX", "DEFAULT_VELZ", "DEFAULT_PRESSURE"):
        assert np.array_equal(env.get(f), orc.get(f)), f


def test_masks_bit_exact():
    import skewtcloudsim_amd as skt
    nx, ny, nz = 13, 9, 7
    env, orc = make_pair(nx, ny, nz)
    gh = np.random.default_rng(3).integers(0, 4, nx * nz).astype(np.int32)
    env.set_ground_height(gh)
    orc.set_ground_height(gh)
    assert np.array_equal(env.masks(), orc.masks())
    env.close()


@pytest.mark.parametrize("field,type_", [("POTTEMP", 0), ("QV", 1), ("VELX", 2), ("VELY", 3), ("VELZ", 4),
                                         ("QR", 5)])
def test_diffusion_phase(pair, field, type_):
    env, orc = pair
    rich_state(env, orc, seed=type_ + 10)
    env.diffuseGPU(field, type_, 0.5)
    orc.diffuse(ob.FIELD[field], type_, 0.5)
    assert_fields(env, orc, (field,), sky_mask(NX, NY, NZ), tag="diffuse")
    assert_bits(env, orc, (field,), sky_mask(NX, NY, NZ), tag="diffuse")
    # a large time step makes every neighbour count
    env.diffuseGPU(field, type_, 1e6)
    orc.diffuse(ob.FIELD[field], type_, 1e6)
    assert_fields(env, orc, (field,), sky_mask(NX, NY, NZ), l2=1e-5, maxabs=1e-4, tag="diffuse-k")
    assert_bits(env, orc, (field,), sky_mask(NX, NY, NZ), tag="diffuse-k")


@pytest.mark.parametrize("field,prof,bounds", [
    ("POTTEMP", "ISEN_TEMP", (2, 2, 2)), ("VELX", "DEFAULT_VELX", (2, 2, 1)), ("VELY", None, (1, 1, 1)),
    ("VELZ", "DEFAULT_VELZ", (2, 2, 1)), ("QV", "ISEN_VAPOR", (2, 2, 2)), ("QR", None, (1, 1, 1))])
def test_advection_phase(pair, field, prof, bounds):
    env, orc = pair
    rich_state(env, orc, seed=3)
    env.advectPPMWGPU(field, prof, bounds, 40.0)  # Courant ~0.3 at 400 m
    orc.advect_ppm(ob.FIELD[field], -1 if prof is None else ob.FIELD[prof], *bounds, 40.0)
    assert env.stats().substeps[0] == orc.stats()["substeps"][0]
    assert_fields(env, orc, (field, "VELX", "VELY", "VELZ"), sky_mask(NX, NY, NZ), l2=2e-6, maxabs=2e-5,
                  tag="advect")
    assert_bits(env, orc, (field, "VELX", "VELY", "VELZ"), sky_mask(NX, NY, NZ), tag="advect")


@pytest.mark.parametrize("ftype", [0, 1, 2])
def test_precip_phase(pair, ftype):
    env, orc = pair
    rich_state(env, orc, seed=4)
    f = ("QR", "QS", "QI")[ftype]
    env.advectPrecip(f, ftype, 200.0)
    orc.advect_precip(ob.FIELD[f], ftype, 200.0)
    assert_fields(env, orc, (f,), sky_mask(NX, NY, NZ), tag="precip")


def test_buoyancy_phase(pair):
    env, orc = pair
    rich_state(env, orc, seed=5)
    env.calculateBuoyancy(0.5)
    orc.buoyancy(0.5)
    assert_fields(env, orc, ("VELY",), sky_mask(NX, NY, NZ), tag="buoyancy")


@pytest.mark.parametrize("mode", [0, 1])
def test_pressure_projection(mode):
    env, orc = make_pair(NX, NY, NZ, dx=400.0, pcg_mode=mode, pcg_tol=1e-5 if mode == 0 else 1e-6,
                         pcg_max_iter=60)
    rich_state(env, orc, seed=6)
    env.initDensity()
    orc.init_density()
    env.pressureProject(0.5)
    orc.pressure_project(0.5)
    so, sg = orc.stats(), env.stats()
    assert sg.pcg_iterations == so["pcg_iterations"]
    m = sky_mask(NX, NY, NZ)
    assert_fields(env, orc, ("PRESSURE_PROJ",), m, l2=1e-4, maxabs=1e-3, tag="pcg")
    assert_fields(env, orc, ("VELX", "VELY", "VELZ", "PRESSURE", "DENSITY"), m, l2=1e-5, maxabs=1e-4, tag="proj")
    env.close()


def test_microphysics_phase_and_phase_mask(pair):
    env, orc = pair
    rich_state(env, orc, seed=7)
    env.microPhysicsSkyGPU(0.5, 1.0)
    orc.micro_sky(0.5, 1.0)
    m = sky_mask(NX, NY, NZ)
    assert_fields(env, orc, ("QV", "QW", "QC", "QR", "QS", "QI", "POTTEMP"), m, l2=1e-5, maxabs=1e-4, tag="micro")
    pg, po = env.phase_mask(), orc.phase_mask()
    assert np.array_equal(pg, po)
    assert (po & 2).any() and (~po & 2)[m].any()  # both branches exercised


def test_ground_phases(pair):
    env, orc = pair
    rich_state(env, orc, seed=8)
    irr = env.irridianceGPU()
    assert irr == orc.irradiance()
    env.advectGroundWater(0.5, 1.0)
    orc.ground_water(0.5, 1.0)
    env.groundCoverageFactor()
    orc.cloud_cover()
    env.updateGroundTemps(0.5, 1.0, irr)
    orc.ground_temps(0.5, 1.0, irr)
    env.microPhysicsGroundGPU(0.5, 1.0, irr)
    orc.micro_ground(0.5, 1.0, irr)
    env.setTempsAtGround(0.5, 1.0)
    orc.set_temps_at_ground(0.5, 1.0)
    assert_fields(env, orc, GROUND + ("QV", "QR", "QS", "QI", "POTTEMP"), sky_mask(NX, NY, NZ), l2=1e-6,
                  maxabs=1e-5, tag="ground")


@pytest.mark.parametrize("shape,dx", [((24, 20, 16), 100.0), ((17, 13, 11), 400.0)])
def test_full_steps(shape, dx):
    nx, ny, nz = shape
    env, orc = make_pair(nx, ny, nz, dx=dx, seed=9, n_bubbles=2)
    m = sky_mask(nx, ny, nz)
    for s in range(5):
        env.updateGPU(0.5, 1.0)
        orc.step(0.5, 1.0)
        tol = 1e-5 if s == 0 else 1e-4
        assert env.stats().pcg_iterations == orc.stats()["pcg_iterations"]
        assert_fields(env, orc, SKY + GROUND, m, l2=tol, maxabs=50 * tol, tag=f"step{s}")
    assert abs(env.params().time - orc.params()[0]) < 1e-3
    env.close()


def test_region_diagnostic(pair):
    env, orc = pair
    rich_state(env, orc, seed=12)
    mn, mx = [2, 3, 1], [15, 20, 12]
    env.region_set(mn, mx)
    import ctypes
    a, b = np.asarray(mn, np.int32), np.asarray(mx, np.int32)
    orc.L.sko_region(orc.h, 1, a.ctypes.data_as(ctypes.c_void_p), b.ctypes.data_as(ctypes.c_void_p))
    env.microPhysicsSkyGPU(0.5, 1.0)
    orc.micro_sky(0.5, 1.0)
    g = env.region_get()
    o = np.zeros(33)
    orc.L.sko_region_get(orc.h, o.ctypes.data_as(ctypes.c_void_p))
    env.region_clear()
    orc.L.sko_region(orc.h, 0, a.ctypes.data_as(ctypes.c_void_p), b.ctypes.data_as(ctypes.c_void_p))
    np.testing.assert_allclose(g, o, rtol=1e-4, atol=1e-12)
    assert np.count_nonzero(o) > 10


# ------------------------------------------------------------------ terrain ---
# Non-flat ground: lateral ground neighbours, ground within the PPM stencil
# and the slow (near-ground) paths of every stencil kernel.

def _terrain(nx, nz, kind, seed=21, ny=None):
    rng = np.random.default_rng(seed)
    if kind == "random":
        return rng.integers(0, 5, nx * nz).astype(np.int32)
    if kind == "tall":
        # Heights spread over the whole column, so ground crosses the 16-level
        # chunks of the column marches (a chunk starting at or inside ground,
        # a chunk wholly ground, neighbour heights >= 16), with the chunk
        # edges 15, 16, 17, 31, 32 and one column that is ground to the top.
        gh = rng.integers(0, ny - 8, nx * nz)
        gh[rng.uniform(size=nx * nz) < 0.4] = rng.integers(0, 4)
        cols = rng.choice(nx * nz, 12, replace=False)
        for c, h in zip(cols, (15, 16, 17, 31, 32, ny - 1, 15, 16, 17, 31, 32, 14)):
            gh[c] = h
        return gh.astype(np.int32)
    x = np.arange(nx)[None, :]
    z = np.arange(nz)[:, None]
    hill = 6.0 * np.exp(-(((x - nx * 0.4) / (nx * 0.2)) ** 2 + ((z - nz * 0.6) / (nz * 0.25)) ** 2))
    return np.rint(hill).astype(np.int32).ravel()


def make_terrain_pair(nx, ny, nz, gh, dx=200.0, seed=31, **cfg):
    import skewtcloudsim_amd as skt
    snd = skt.make_sounding(nx, ny, nz, dx, seed=seed, n_bubbles=2)
    env = skt.EnvironmentGPU(nx, ny, nz, dx, **cfg)
    ocfg = {k: cfg[k] for k in ("pcg_mode", "pcg_tol", "pcg_max_iter") if k in cfg}
    orc = ob.Oracle(nx, ny, nz, dx, **ocfg)
    env.set_ground_height(gh)
    orc.set_ground_height(gh)
    env.init(*[snd[k] for k in KEYS])
    orc.init(*[snd[k] for k in KEYS])
    return env, orc


@pytest.mark.parametrize("shape,kind", [((24, 20, 16), "random"), ((33, 18, 21), "hill"), ((17, 13, 11), "random"),
                                        ((40, 48, 24), "tall")])
def test_terrain_phases(shape, kind):
    nx, ny, nz = shape
    gh = _terrain(nx, nz, kind, ny=ny)
    env, orc = make_terrain_pair(nx, ny, nz, gh)
    m = sky_mask(nx, ny, nz, gh)
    assert np.array_equal(env.masks(), orc.masks())
    rich_state(env, orc, seed=41)
    for field, type_ in (("POTTEMP", 0), ("QV", 1), ("VELX", 2), ("VELY", 3), ("VELZ", 4), ("QR", 5)):
        env.diffuseGPU(field, type_, 2e5)  # large k: every neighbour, ground ones included, counts
        orc.diffuse(ob.FIELD[field], type_, 2e5)
        assert_fields(env, orc, (field,), m, l2=1e-5, maxabs=1e-4, tag="terrain-diffuse-" + field)
        if type_ != 0:  # the potential temperature's ground neighbours go through powf
            assert_bits(env, orc, (field,), m, tag="terrain-diffuse-" + field)
    rich_state(env, orc, seed=42)
    for field, prof, bounds in (("POTTEMP", "ISEN_TEMP", (2, 2, 2)), ("VELX", "DEFAULT_VELX", (2, 2, 1)),
                                ("VELY", None, (1, 1, 1)), ("QR", None, (1, 1, 1)), ("QV", "ISEN_VAPOR", (0, 0, 0))):
        env.advectPPMWGPU(field, prof, bounds, 60.0)
        orc.advect_ppm(ob.FIELD[field], -1 if prof is None else ob.FIELD[prof], *bounds, 60.0)
        assert_fields(env, orc, (field, "VELX", "VELY", "VELZ"), m, l2=1e-5, maxabs=1e-4,
                      tag="terrain-advect-" + field)
        assert_bits(env, orc, (field, "VELX", "VELY", "VELZ"), m, tag="terrain-advect-" + field)
    rich_state(env, orc, seed=43)
    # the sedimentation's fall speeds are chains of powf and sqrtf (ocml on
    # the device, libm in the oracle, each within an ulp or two); the bound
    # is measured in the test: the device may differ from the FP32 oracle by
    # at most twice the FP32 oracle's own distance to the FP64 oracle (FP32
    # rounding of the same chains) plus 1e-7
    o64 = ob.Oracle(nx, ny, nz, 200.0, fp64=True)
    o64.set_ground_height(gh)
    for f in ob.FIELDS:
        o64.set(f, orc.get(f))
    for ftype, f in enumerate(("QR", "QS", "QI")):
        env.advectPrecip(f, ftype, 200.0)
        orc.advect_precip(ob.FIELD[f], ftype, 200.0)
        o64.advect_precip(ob.FIELD[f], ftype, 200.0)
        g, c, d = env.get(f)[m], orc.get(f)[m], o64.get64(f)[m]
        e_gpu, e_32 = rel_l2(g, c), rel_l2(c, d)
        assert e_gpu <= 2.0 * e_32 + 1e-7, (kind, f, e_gpu, e_32)
        assert_fields(env, orc, (f,), m, l2=1e-5, maxabs=1e-4, tag="terrain-precip")
    o64.close()
    rich_state(env, orc, seed=44)
    env.calculateBuoyancy(0.5)
    orc.buoyancy(0.5)
    assert_fields(env, orc, ("VELY",), m, tag="terrain-buoyancy")
    rich_state(env, orc, seed=45)
    env.microPhysicsSkyGPU(0.5, 1.0)
    orc.micro_sky(0.5, 1.0)
    assert_fields(env, orc, ("QV", "QW", "QC", "QR", "QS", "QI", "POTTEMP"), m, l2=1e-5, maxabs=1e-4,
                  tag="terrain-micro")
    assert np.array_equal(env.phase_mask(), orc.phase_mask())
    env.close()
    orc.close()


@pytest.mark.parametrize("mode,kind", [(0, "random"), (1, "random"), (0, "tall"), (1, "tall")])
def test_terrain_pressure_projection(mode, kind):
    nx, ny, nz = (24, 20, 16) if kind == "random" else (40, 48, 24)
    gh = _terrain(nx, nz, kind, ny=ny)
    env, orc = make_terrain_pair(nx, ny, nz, gh, pcg_mode=mode, pcg_tol=1e-5 if mode == 0 else 1e-6,
                                 pcg_max_iter=60)
    rich_state(env, orc, seed=46)
    env.initDensity()
    orc.init_density()
    env.pressureProject(0.5)
    orc.pressure_project(0.5)
    assert env.stats().pcg_iterations == orc.stats()["pcg_iterations"]
    m = sky_mask(nx, ny, nz, gh)
    assert_fields(env, orc, ("PRESSURE_PROJ",), m, l2=1e-4, maxabs=1e-3, tag="terrain-pcg")
    assert_fields(env, orc, ("VELX", "VELY", "VELZ", "PRESSURE", "DENSITY"), m, l2=1e-5, maxabs=1e-4,
                  tag="terrain-proj")
    env.close()
    orc.close()


@pytest.mark.parametrize("shape,kind", [((24, 20, 16), "random"), ((33, 18, 21), "hill"), ((40, 48, 24), "tall")])
def test_terrain_full_steps(shape, kind):
    nx, ny, nz = shape
    gh = _terrain(nx, nz, kind, ny=ny)
    env, orc = make_terrain_pair(nx, ny, nz, gh)
    m = sky_mask(nx, ny, nz, gh)
    for s in range(4):
        env.updateGPU(0.5, 1.0)
        orc.step(0.5, 1.0)
        tol = 1e-5 if s == 0 else 1e-4
        assert env.stats().pcg_iterations == orc.stats()["pcg_iterations"]
        assert_fields(env, orc, SKY, m, l2=tol, maxabs=50 * tol, tag=f"terrain-step{s}")
        assert_fields(env, orc, GROUND, np.ones(nx * nz, bool), l2=tol, maxabs=50 * tol, tag=f"terrain-ground{s}")
    env.close()
    orc.close()


# Launch shapes. The default grids of these small domains give every block a
# single z plane; max_blocks bounds the blocks of a launch, so each block walks
# a longer z chunk (and the flat solver kernels longer row runs). For the 6
# tile columns of a 20 x 24 plane: 1 -> one 16-plane chunk, 18 -> chunks of
# 6, 6, 4, 42 -> chunks of 3 (the last of 1).
LAUNCH_CAPS = [1, 18, 42]


@pytest.mark.parametrize("cap", LAUNCH_CAPS)
@pytest.mark.parametrize("mode", [0, 1])
def test_launch_shapes_pressure(mode, cap):
    env, orc = make_pair(NX, NY, NZ, dx=400.0, pcg_mode=mode, pcg_tol=1e-5 if mode == 0 else 1e-6,
                         pcg_max_iter=60, max_blocks=cap)
    rich_state(env, orc, seed=6)
    env.initDensity()
    orc.init_density()
    env.pressureProject(0.5)
    orc.pressure_project(0.5)
    assert env.stats().pcg_iterations == orc.stats()["pcg_iterations"]
    m = sky_mask(NX, NY, NZ)
    assert_fields(env, orc, ("PRESSURE_PROJ",), m, l2=1e-4, maxabs=1e-3, tag=f"pcg-cap{cap}")
    assert_fields(env, orc, ("VELX", "VELY", "VELZ", "PRESSURE", "DENSITY"), m, l2=1e-5, maxabs=1e-4,
                  tag=f"proj-cap{cap}")
    env.close()
    orc.close()


@pytest.mark.parametrize("cap", LAUNCH_CAPS)
def test_launch_shapes_terrain_steps(cap):
    nx, ny, nz = NX, NY, NZ
    gh = _terrain(nx, nz, "random")
    env, orc = make_terrain_pair(nx, ny, nz, gh, max_blocks=cap)
    m = sky_mask(nx, ny, nz, gh)
    for s in range(3):
        env.updateGPU(0.5, 1.0)
        orc.step(0.5, 1.0)
        tol = 1e-5 if s == 0 else 1e-4
        assert env.stats().pcg_iterations == orc.stats()["pcg_iterations"]
        assert_fields(env, orc, SKY, m, l2=tol, maxabs=50 * tol, tag=f"cap{cap}-step{s}")
        assert_fields(env, orc, GROUND, np.ones(nx * nz, bool), l2=tol, maxabs=50 * tol, tag=f"cap{cap}-ground{s}")
    env.close()
    orc.close()


# A mid-size domain with the default launch shapes: 32 tile columns, so the
# stencil kernels walk 2-plane chunks, the pressure operator 3-plane chunks
# (one resident round of blocks) and the flat solver kernels 4-row runs. The
# two steps run the reference's full 200 PCG iterations.
@pytest.mark.parametrize("terrain", [False, True])
def test_mid_size_steps(terrain):
    nx, ny, nz = 128, 64, 128
    if terrain:
        gh = _terrain(nx, nz, "hill")
        env, orc = make_terrain_pair(nx, ny, nz, gh, dx=100.0, seed=12)
    else:
        gh = None
        env, orc = make_pair(nx, ny, nz, dx=100.0, seed=12, n_bubbles=3)
    m = sky_mask(nx, ny, nz, gh)
    for s in range(2):
        env.updateGPU(0.5, 1.0)
        orc.step(0.5, 1.0)
        tol = 1e-5 if s == 0 else 1e-4
        assert env.stats().pcg_iterations == orc.stats()["pcg_iterations"]
        assert_fields(env, orc, SKY, m, l2=tol, maxabs=50 * tol, tag=f"mid{'-terrain' if terrain else ''}-step{s}")
    env.close()
    orc.close()


def _diag_brute(env, gh, thr):
    nx, ny, nz = env.nx, env.ny, env.nz
    n = nx * ny * nz
    y = (np.arange(n) // nx) % ny
    x = np.arange(n) % nx
    z = np.arange(n) // (nx * ny)
    sky = y > gh[x + nx * z]
    v = env.get("VELY")
    cloudy = sky & ((env.get("QW") + env.get("QC")) > np.float32(thr))
    w = np.float32(v[sky].max())
    lev = int(y[cloudy].max()) if cloudy.any() else -1
    return w, lev, float(env.get("QGR").astype(np.float64).sum())


@pytest.mark.parametrize("cap", [0, 1])
def test_diagnostics(cap):
    """skt_env_diagnostics against a numpy brute force on the GPU's own state
    (exact), and on the oracle's state copied in (against the oracle's
    diagnostics, exact), with terrain; cap 1 makes each block walk all z."""
    nx, ny, nz = 24, 20, 16
    gh = _terrain(nx, nz, "random")
    env, orc = make_terrain_pair(nx, ny, nz, gh, max_blocks=cap)
    for _ in range(2):
        env.updateGPU(0.5, 1.0)
        orc.step(0.5, 1.0)
    for thr in (1e-7, 1e-5):
        d = env.diagnostics(thr)
        w, lev, rain = _diag_brute(env, gh, thr)
        assert d["max_updraft"] == w and d["cloud_top_level"] == lev
        assert d["cloud_top_height"] == (np.float32(lev) * np.float32(env.dx) if lev >= 0 else 0.0)
        assert d["surface_rain"] == pytest.approx(rain, rel=1e-12, abs=1e-300)
    for f in ("VELY", "QW", "QC", "QGR"):
        env.set(f, orc.get(f))
    d, do = env.diagnostics(1e-5), orc.diagnostics(1e-5)
    assert d["max_updraft"] == do["max_updraft"] and d["cloud_top_level"] == do["cloud_top_level"]
    assert d["surface_rain"] == pytest.approx(do["surface_rain"], rel=1e-12, abs=1e-300)
    # a cloudy ground cell and a huge ground updraft are ignored
    g = int(np.argmax(gh))
    x, z = g % nx, g // nx
    i = x + nx * (int(gh[g]) + ny * z)
    v, qc = env.get("VELY"), env.get("QC")
    v[i], qc[i] = 1e6, 1.0
    env.set("VELY", v)
    env.set("QC", qc)
    d2 = env.diagnostics(1e-5)
    assert d2["max_updraft"] == d["max_updraft"] and d2["cloud_top_level"] == d["cloud_top_level"]
    env.close()
    orc.close()


@pytest.mark.parametrize("nx,nz,octaves,maxh", [(24, 20, 12, 0.2), (64, 48, 12, 6.0), (33, 17, 4, 3.0)])
def test_terrain_generate(nx, nz, octaves, maxh):
    """skt_terrain_generate against the oracle: the noise within a few float
    ulps (sin/cos of the device library against double-rounded ones), the
    heights equal except where noise * max_height lies within 1e-4 of a tie."""
    import skewtcloudsim_amd as skt
    noise, h = skt.terrain(nx, nz, seed=100, octaves=octaves, max_height=maxh)
    on = np.empty(nx * nz, np.float32)
    oh = np.empty(nx * nz, np.int32)
    ob.lib().sko_terrain(100, nx, nz, octaves, maxh, on.ctypes.data, oh.ctypes.data)
    assert np.allclose(noise, on, rtol=2e-6, atol=2e-6)
    v = on.astype(np.float64) * maxh
    near_tie = np.abs(v - np.floor(v) - 0.5) < 1e-4
    assert np.array_equal(h[~near_tie], oh[~near_tie])
    if maxh == 0.2:
        assert np.all(h == 0)


EDITS = [
    ("brush", dict(param=7, size=2.5, pos=(6, 8, 5), smooth=1.0, dt=0.5, intensity=2.0, value=3.0,
                   direction=(1.0, 0.5, -0.25))),
    ("brush", dict(param=1, size=3.0, pos=(0, 6, 15), smooth=2.0, dt=0.5, intensity=1.5, value=2e-3)),
    ("brush", dict(param=0, size=1.5, pos=(23, 19, 8), smooth=0.5, dt=1.0, intensity=1.0, value=4.0)),
    ("selection", dict(param=3, mn=(3, 2, 4), mx=(9, 7, 6), value=5e-4)),
    ("selection", dict(param=7, mn=(10, 5, 1), mx=(12, 9, 2), value=2.0, direction=(0.0, 1.0, 0.0))),
    ("selection", dict(param=2, mn=(8, 3, 3), mx=(5, 6, 3), value=1e-4)),   # empty x range: one column
    ("brush", dict(param=8, size=2.0, pos=(9, 4, 9), smooth=1.0, dt=1.0, intensity=1.0, value=1.0,
                   erase=True)),
    ("selection", dict(param=8, mn=(15, 6, 3), mx=(16, 7, 4), value=0.0, erase=False)),
    ("selection", dict(param=8, mn=(2, 1, 12), mx=(4, 5, 13), value=0.0, erase=True)),
]


def test_editor_edits():
    """Brush and selection edits (prepareBrushGPU / prepareSelectionGPU) on
    every kind of parameter, clipped at the domain edges, raising and
    erasing ground, against the oracle; then a step from the edited state."""
    nx, ny, nz = 24, 20, 16
    gh = _terrain(nx, nz, "hill")
    env, orc = make_terrain_pair(nx, ny, nz, gh)
    fields = SKY[:-1] + GROUND
    for kind, a in EDITS:
        a = dict(a)
        if kind == "brush":
            env.prepareBrushGPU(a["param"], a["size"], a["pos"], a["smooth"], a["dt"], a["intensity"], a["value"],
                                a.get("direction", (0.0, 0.0, 0.0)), a.get("erase", False))
            orc.brush(a["param"], a["size"], a["pos"], a["smooth"], a["dt"], a["intensity"], a["value"],
                      a.get("direction", (0.0, 0.0, 0.0)), a.get("erase", False))
        else:
            env.prepareSelectionGPU(a["param"], a["mn"], a["mx"], a["value"], a.get("direction", (0.0, 0.0, 0.0)),
                                    a.get("erase", False))
            orc.selection(a["param"], a["mn"], a["mx"], a["value"], a.get("direction", (0.0, 0.0, 0.0)),
                          a.get("erase", False))
        assert np.array_equal(env.ground_height(), orc.ground_height()), (kind, a)
        m = sky_mask(nx, ny, nz, env.ground_height())
        # powf of the device library against the host's: a float ulp or two
        assert_fields(env, orc, fields, m, l2=1e-6, maxabs=1e-5, tag=f"edit {kind} {a['param']}")
    assert not np.array_equal(env.ground_height(), gh)
    env.updateGPU(0.5, 1.0)
    orc.step(0.5, 1.0)
    assert env.stats().pcg_iterations == orc.stats()["pcg_iterations"]
    m = sky_mask(nx, ny, nz, env.ground_height())
    assert_fields(env, orc, SKY, m, l2=1e-5, maxabs=5e-4, tag="edit-step")
    env.close()
    orc.close()


@pytest.mark.parametrize("source", ["knmi", "dwd"])
def test_sounding_initial_state_and_steps(source):
    """initEnvironment from the radiosonde fixtures: the device broadcast
    (skt_env_init_sounding) equals skt_env_init on the host arrays bit for bit,
    and the state steps like the oracle initialised from its own restatement."""
    import os
    import sys
    import skewtcloudsim_amd as skt
    sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden"))
    import make_soundings as gen
    nx, ny, nz, dx = 24, 20, 16, 100.0
    r = skt.ReadTable()
    if source == "knmi":
        r.readKNMIFile(gen.KNMI_PATH)
    else:
        r.readDWDFile(gen.DWD_PATH)
    print(f"sounding {source}: {r.data.dataSize} rows, grid {nx}x{ny}x{nz}", flush=True)
    env = skt.EnvironmentGPU(nx, ny, nz, dx)
    r.initEnvironment(env)
    host = r.init_fields(nx, ny, nz, dx)
    ref = skt.EnvironmentGPU(nx, ny, nz, dx)
    ref.init(*[host[k] for k in KEYS])
    for f in ob.FIELDS:
        assert np.array_equal(env.get(f), ref.get(f), equal_nan=True), f
    ref.close()
    cols = {n: getattr(r.data, n) for n in ob.SOUNDING_COLUMNS}
    want = ob.sounding_init(cols, nx, ny, nz, dx)
    for k in KEYS:
        assert np.array_equal(host[k], want[k]), k
    orc = ob.Oracle(nx, ny, nz, dx)
    orc.init(*[want[k] for k in KEYS])
    m = sky_mask(nx, ny, nz)
    assert_fields(env, orc, SKY + GROUND, m, l2=1e-7, maxabs=1e-6, tag="sounding init")
    for s in range(3):
        env.updateGPU(0.5, 1.0)
        orc.step(0.5, 1.0)
        assert env.stats().pcg_iterations == orc.stats()["pcg_iterations"]
        tol = 1e-5 if s == 0 else 1e-4
        assert_fields(env, orc, SKY + GROUND, m, l2=tol, maxabs=50 * tol, tag=f"sounding step{s}")
    env.close()
    orc.close()


def test_kernel_timing_classes():
    """Per-class timing: every class of a step is timed (events on the
    kernel's dispatch for single-kernel classes, around the call otherwise),
    and timing leaves the results bit-identical."""
    nx, ny, nz = 24, 20, 16
    env, orc = make_pair(nx, ny, nz, seed=13)
    orc.close()
    import skewtcloudsim_amd as skt
    snd = skt.make_sounding(nx, ny, nz, 100.0, seed=13, n_bubbles=2)
    ref = skt.EnvironmentGPU(nx, ny, nz, 100.0)
    ref.init(*[snd[k] for k in KEYS])
    env.set_timing(True)
    for _ in range(2):
        env.updateGPU(0.5, 1.0)
        ref.updateGPU(0.5, 1.0)
    t = env.timing()
    for c, (n, ms) in t.items():
        if c == "multigrid":  # the Jacobi-preconditioned solve has no V-cycles
            assert n == 0, (c, n)
            continue
        assert n > 0 and ms > 0.0, (c, n, ms)
    assert t["pcg_spmv"][0] == t["pcg_update"][0]
    # with the multigrid preconditioner: set-up plus a V-cycle (which also
    # forms the z.r partials) per iteration and one at the start of the solve
    mg = skt.EnvironmentGPU(nx, ny, nz, 100.0, pcg_mode=1, pcg_tol=1e-6, pcg_max_iter=50, pcg_precon=1)
    mg.init(*[snd[k] for k in KEYS])
    mg.set_timing(True, classes=["multigrid"])
    mg.updateGPU(0.5, 1.0)
    tm = mg.timing()
    its = mg.stats().pcg_iterations
    assert its > 0 and tm["multigrid"][0] >= 1 + (its + 1), (tm["multigrid"], its)
    mg.close()
    env.set_timing(True, classes=["pcg_spmv"])
    env.updateGPU(0.5, 1.0)
    ref.updateGPU(0.5, 1.0)
    t2 = env.timing()
    assert t2["pcg_spmv"][0] == env.stats().pcg_iterations and t2["pcg_update"][0] == 0
    # every third launch of the class
    env.set_timing(True, classes=["pcg_spmv"], stride=3)
    env.updateGPU(0.5, 1.0)
    ref.updateGPU(0.5, 1.0)
    k = env.stats().pcg_iterations
    assert env.timing()["pcg_spmv"][0] == (k + 2) // 3
    env.set_timing(False)
    for f in SKY + GROUND:
        assert np.array_equal(env.get(f), ref.get(f), equal_nan=True), f
    env.close()
    ref.close()


def test_pressure_solve_out_of_range_density():
    """Two neighbouring sky cells of density 2e-38: their 1/density (5e37)
    lies above the short face division's range, so the solve takes the exact
    division for every face, and the face between them (sum of 1/densities
    1e38) depends on it. Matches the oracle."""
    env, orc = make_pair(NX, NY, NZ, dx=400.0, pcg_max_iter=40)
    rich_state(env, orc, seed=8)
    env.initDensity()
    orc.init_density()
    rho = orc.get("DENSITY").copy()
    i = 5 + NX * (7 + NY * 6)
    rho[i] = rho[i + 1] = np.float32(2e-38)
    env.set("DENSITY", rho)
    orc.set("DENSITY", rho)
    env.calculatePressureProject(0.5)
    orc.calc_pressure_project(0.5)
    assert env.stats().pcg_iterations == orc.stats()["pcg_iterations"]
    m = sky_mask(NX, NY, NZ)
    assert_fields(env, orc, ("PRESSURE_PROJ",), m, l2=1e-4, maxabs=1e-3, tag="pcg exact faces")
    env.close()


@pytest.mark.parametrize("shape", [(8, 2, 3), (65, 9, 5), (70, 33, 4)])
def test_edge_shapes_steps(shape):
    """Shapes at the edges of the launch layouts: two levels (one sky level),
    x just past one 64-lane tile, ny not a multiple of the 8-row tiles or the
    16-level column chunks, nz below one z group. Three steps against the
    oracle."""
    nx, ny, nz = shape
    print(f"edge shape {shape}", flush=True)
    env, orc = make_pair(nx, ny, nz, dx=100.0, seed=9, n_bubbles=2)
    m = sky_mask(nx, ny, nz)
    for s in range(3):
        env.updateGPU(0.5, 1.0)
        orc.step(0.5, 1.0)
        assert env.stats().pcg_iterations == orc.stats()["pcg_iterations"]
        tol = 1e-5 if s == 0 else 1e-4
        assert_fields(env, orc, SKY + GROUND, m, l2=tol, maxabs=50 * tol, tag=f"{shape} step{s}")
    env.close()
    orc.close()


LONG_FIELDS = ("POTTEMP", "VELX", "VELY", "VELZ", "PRESSURE", "QV", "QW", "QC", "QR", "QS", "QI")


def test_long_run_against_fp32_and_fp64_oracles():
    """200 steps from a state with clouds, rain, snow, ice and a sub-freezing
    upper half (32x24x16). After 200 steps every field's distance (relative
    L2 over sky cells) to the FP64 oracle is within twice the FP32 oracle's
    own distance to it, plus 1e-6: the device path is as accurate as the CPU
    FP32 restatement, and the velocities' growing differences are FP32
    round-off amplified by the flow, which the FP32 oracle shows as well.

    PCG iteration counts: the stop test compares max|r| with an absolute
    tolerance, so rounding alone moves a solve across it by one iteration
    now and then: over this run the FP32 and FP64 oracles differ by one
    iteration in about 15 of the 200 steps. The device's count equals the
    FP32 oracle's over the first 20 steps, and afterwards differs from it in
    no more steps than the FP32 oracle differs from the FP64 one, by at most
    one iteration each time."""
    import skewtcloudsim_amd as skt
    nx, ny, nz = 32, 24, 16
    env, orc = make_pair(nx, ny, nz, dx=100.0, seed=9, n_bubbles=3)
    rich_state(env, orc, seed=12)
    snd = skt.make_sounding(nx, ny, nz, 100.0, seed=9, n_bubbles=3)
    o64 = ob.Oracle(nx, ny, nz, 100.0, fp64=True)
    o64.init(*[snd[k] for k in KEYS])
    for f in ob.FIELDS:
        o64.set64(f, orc.get(f).astype(np.float64))
    m = sky_mask(nx, ny, nz)
    dev_off, fp64_off = [], []
    for s in range(200):
        env.updateGPU(0.5, 1.0)
        orc.step(0.5, 1.0)
        o64.step(0.5, 1.0)
        ig, ic, i64 = env.stats().pcg_iterations, orc.stats()["pcg_iterations"], o64.stats()["pcg_iterations"]
        assert abs(ig - ic) <= (0 if s < 20 else 1), (s, ig, ic)
        if ig != ic:
            dev_off.append(s)
        if ic != i64:
            fp64_off.append(s)
    print(f"PCG counts: device != FP32 oracle at steps {dev_off}; FP32 != FP64 oracle at {len(fp64_off)} steps")
    assert len(dev_off) <= len(fp64_off), (dev_off, fp64_off)
    worst = {}
    for f in LONG_FIELDS:
        g, c, d = env.get(f)[m], orc.get(f)[m], o64.get64(f)[m]
        assert np.isfinite(g).all(), f
        e_gpu, e_cpu = rel_l2(g, d), rel_l2(c, d)
        worst[f] = (rel_l2(g, c), e_gpu, e_cpu)
        assert e_gpu <= 2.0 * e_cpu + 1e-6, (f, e_gpu, e_cpu)
    print(" ".join(f"{f}: gpu-fp32 {a:.1e} gpu-fp64 {b:.1e} fp32-fp64 {c:.1e}" for f, (a, b, c) in worst.items()))
    env.close()
    orc.close()
    o64.close()


def test_full_size_steps_are_reproducible():
    """Two steps of the 256x128x256 workload (nx = nz = 256, ny = 128, the
    benchmark's randomised sounding), run twice in fresh environments from
    the same initial state, leave the same bits in every field and take the
    same PCG iteration counts: at full size the step has no race (lane
    shuffles, block barriers, in-place column chunks) and no reduction whose
    result depends on timing."""
    import skewtcloudsim_amd as skt
    nx, ny, nz, dx = 256, 128, 256, 100.0
    snd = skt.make_sounding(nx, ny, nz, dx, seed=20261203, n_bubbles=3)
    args = [snd[k] for k in KEYS]
    runs = []
    for run in range(2):
        print(f"full-size reproducibility run {run}", flush=True)
        env = skt.EnvironmentGPU(nx, ny, nz, dx)
        env.init(*args)
        its = []
        for _ in range(2):
            env.updateGPU(0.5, 1.0)
            its.append(env.stats().pcg_iterations)
        runs.append((its, {f: env.get(f).copy() for f in SKY + GROUND}))
        env.close()
    (ia, fa), (ib, fb) = runs
    assert ia == ib
    for f in fa:
        assert np.isfinite(fa[f]).any(), f
        assert np.array_equal(fa[f].view(np.uint32), fb[f].view(np.uint32)), f


def test_full_size_advection_and_diffusion_bit_exact():
    """The advection phase (a moisture field and the three velocity
    components) and the diffusion phase (potential temperature with its
    ground-neighbour rule, and rain) on the 256x128x256 workload equal the
    oracle's bit for bit on every sky cell: the launch shapes' chunk and
    segment boundaries at full size change no result."""
    nx, ny, nz = 256, 128, 256
    print("full-size bit-exact phases: set-up", flush=True)
    env, orc = make_pair(nx, ny, nz, dx=100.0, seed=20261203, n_bubbles=3)
    m = sky_mask(nx, ny, nz)
    rich_state(env, orc, seed=77)
    print("full-size bit-exact phases: advection", flush=True)
    env.advectPPMWGPU("QV", "ISEN_VAPOR", (2, 2, 2), 2.0)
    orc.advect_ppm(ob.FIELD["QV"], ob.FIELD["ISEN_VAPOR"], 2, 2, 2, 2.0)
    assert env.stats().substeps[0] == orc.stats()["substeps"][0]
    assert_bits(env, orc, ("QV",), m, tag="full-advect")
    env.advectPPMWGPU("VELX", "DEFAULT_VELX", (2, 2, 1), 2.0)
    orc.advect_ppm(ob.FIELD["VELX"], ob.FIELD["DEFAULT_VELX"], 2, 2, 1, 2.0)
    assert_bits(env, orc, ("VELX", "VELY", "VELZ"), m, tag="full-advect-u")
    print("full-size bit-exact phases: diffusion", flush=True)
    for field, type_ in (("POTTEMP", 0), ("QR", 5)):
        env.diffuseGPU(field, type_, 0.5)
        orc.diffuse(ob.FIELD[field], type_, 0.5)
        assert_bits(env, orc, (field,), m, tag="full-diffuse-" + field)
    env.close()
    orc.close()


def test_full_size_step_against_oracle():
    """One whole step of the 256x128x256 workload (reference solver, 200 PCG
    iterations) against the oracle from the same initial state: equal PCG
    iteration counts, and every sky and ground field within relative L2
    5e-5 (the worst field, cloud water, measures about 1.3e-5: microphysics
    transcendental ulps)."""
    nx, ny, nz = 256, 128, 256
    print("full-size step: set-up", flush=True)
    env, orc = make_pair(nx, ny, nz, dx=100.0, seed=20261203, n_bubbles=3)
    m = sky_mask(nx, ny, nz)
    print("full-size step: device", flush=True)
    env.updateGPU(0.5, 1.0)
    print("full-size step: oracle", flush=True)
    orc.step(0.5, 1.0)
    assert env.stats().pcg_iterations == orc.stats()["pcg_iterations"]
    worst = {}
    for f in SKY + GROUND:
        a, b = env.get(f), orc.get(f)
        if a.size == m.size:
            a, b = a[m], b[m]
        assert np.isfinite(a).all(), f
        worst[f] = rel_l2(a, b)
        assert worst[f] <= 5e-5, (f, worst[f])
    print(" ".join(f"{f} {v:.1e}" for f, v in sorted(worst.items(), key=lambda kv: -kv[1])[:5]), flush=True)
    env.close()
    orc.close()


@pytest.mark.parametrize("param", ["POTTEMP", "QV", "QW", "QC", "QR", "QS", "QI", "WIND", "PGROUND", "PRESSURE"])
def test_reset_parameter(param):
    """resetParameterGPU on tall terrain: the setToDefault z walk stops at a
    column's first ground cell, setToValue clears every cell; every cell of
    every field equals the oracle bit for bit."""
    nx, ny, nz = 40, 48, 24
    gh = _terrain(nx, nz, "tall", ny=ny)
    env, orc = make_terrain_pair(nx, ny, nz, gh)
    rich_state(env, orc, seed=47)
    for f in ob.FIELDS[:11]:
        assert np.array_equal(env.get(f).view(np.uint32), orc.get(f).view(np.uint32)), f
    env.resetParameterGPU(param)
    orc.reset_parameter(env._PARAMS[param])
    env.synchronize()
    for f in ob.FIELDS[:11]:
        a, b = env.get(f), orc.get(f)
        nd = int((a.view(np.uint32) != b.view(np.uint32)).sum())
        assert nd == 0, (param, f, nd)
    env.close()
    orc.close()


CFG0_BOUNDS = {"POTTEMP": 1e-5, "PRESSURE": 1e-5, "QV": 1e-5, "VELX": 5e-3, "VELZ": 5e-3, "VELY": 5e-2}


def test_single_bubble_64cubed_200_steps_against_fp32_and_fp64_oracles():
    """The plain method at 64x64x64 (dx 100 m, one 2 K bubble of 500 m at
    1 km in the fixed 300 K / 14 g/kg sounding, the reference's init noise),
    200 steps of dt 0.5 s with the pressure solve to a relative residual of
    1e-6 (Jacobi-PCG on the symmetric operator). At every step the device,
    the FP32 oracle and the FP64 oracle run from the same initial state:
      * PCG iteration counts equal the FP32 oracle's, except where the
        relative stop test sits within the two paths' round-off of 1e-6
        (then they differ by one; the FP64 oracle then agrees with one or
        the other): at most 6 such steps;
      * every field stays within relative L2 CFG0_BOUNDS of the FP32 oracle
        (measured after 200 steps: theta, p, qv 1.4e-6; u, w 7.3e-4;
        v 8.5e-3; the condensates stay 0 in these 100 s);
      * after 200 steps every field's distance to the FP64 oracle is at most
        twice the FP32 oracle's own, plus 1e-6."""
    import skewtcloudsim_amd as skt
    n = 64
    snd = skt.make_sounding(n, n, n, 100.0, seed=20261203, n_bubbles=0)
    args = [snd[k] for k in KEYS]
    cfg = dict(pcg_mode=1, pcg_tol=1e-6, pcg_max_iter=1000)
    env = skt.EnvironmentGPU(n, n, n, 100.0, **cfg)
    env.init(*args)
    o32 = ob.Oracle(n, n, n, 100.0, **cfg)
    o32.init(*args)
    o64 = ob.Oracle(n, n, n, 100.0, fp64=True, **cfg)
    o64.init(*args)
    m = sky_mask(n, n, n)
    off_by_one = 0
    for s in range(200):
        env.updateGPU(0.5, 1.0)
        o32.step(0.5, 1.0)
        o64.step(0.5, 1.0)
        ig, ic = env.stats().pcg_iterations, o32.stats()["pcg_iterations"]
        assert abs(ig - ic) <= 1, (s, ig, ic)
        off_by_one += ig != ic
        if s % 20 == 19:
            curve = []
            for f, b in CFG0_BOUNDS.items():
                r = rel_l2(env.get(f)[m], o32.get(f)[m])
                curve.append(f"{f} {r:.1e}")
                assert r <= b, (s, f, r)
            print(f"step {s + 1}: its {ig}; gpu vs fp32 oracle " + ", ".join(curve), flush=True)
    assert off_by_one <= 6, off_by_one
    for f in ("POTTEMP", "PRESSURE", "QV", "VELX", "VELY", "VELZ", "QW", "QC", "QR", "QS", "QI"):
        g, c, d = env.get(f)[m], o32.get(f)[m], o64.get64(f)[m]
        assert np.isfinite(g).all(), f
        e_gpu, e_cpu = rel_l2(g, d), rel_l2(c, d)
        print(f"{f}: gpu-fp64 {e_gpu:.1e} fp32-fp64 {e_cpu:.1e}", flush=True)
        assert e_gpu <= 2.0 * e_cpu + 1e-6, (f, e_gpu, e_cpu)
    env.close()
    o32.close()
    o64.close()


RICH_BOUNDS = {"QW": 2e-3, "QC": 2e-3}


def test_full_size_rich_step_against_oracle():
    """One whole step of the 256x128x256 workload from a state with cloud
    water, cloud ice, rain, snow, graupel and a sub-freezing upper half
    (rich_state): every branch of the microphysics runs at full size. Equal
    PCG iteration counts; relative L2 of every field within 1e-5 of the
    oracle except cloud water and cloud ice (2e-3; measured 5.1e-4 and
    3.1e-4): their differences sit in the few cells where a threshold of
    the microphysics (autoconversion at QwMin, the phase selectors, wet or
    dry growth) compares a value that the earlier phases left one float
    rounding apart on the two paths, so the two sides take different
    branches there."""
    nx, ny, nz = 256, 128, 256
    print("full-size rich step: set-up", flush=True)
    env, orc = make_pair(nx, ny, nz, dx=100.0, seed=20261203, n_bubbles=3)
    m = sky_mask(nx, ny, nz)
    rich_state(env, orc, seed=91)
    print("full-size rich step: device", flush=True)
    env.updateGPU(0.5, 1.0)
    print("full-size rich step: oracle", flush=True)
    orc.step(0.5, 1.0)
    assert env.stats().pcg_iterations == orc.stats()["pcg_iterations"]
    for f in ("QW", "QC", "QR", "QS", "QI"):
        assert np.linalg.norm(orc.get(f)[m]) > 0, f
    rep = []
    for f in SKY + GROUND:
        a, b = env.get(f), orc.get(f)
        if a.size == m.size:
            a, b = a[m], b[m]
        assert np.isfinite(a).all(), f
        r = rel_l2(a, b)
        far = int((np.abs(a.astype(np.float64) - b) > 1e-4 * max(np.abs(b).max(), 1e-30)).sum())
        rep.append(f"{f} {r:.1e} ({far} cells beyond 1e-4 of scale)")
        assert r <= RICH_BOUNDS.get(f, 1e-5), (f, r)
    print("; ".join(rep), flush=True)
    env.close()
    orc.close()
