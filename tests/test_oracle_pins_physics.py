"""Pins of the oracle's physics phases (oracle/skt_oracle.cpp) -- CPU only.

Each check comes from what the mathematics or the reference's documented
boundary rules fix, not from the oracle's own code, and is chosen so that a
plausible transcription slip (a dropped term, a wrong sign, exponent or
boundary condition, a swapped species) fails it:
  * pressure operator, divergence, projection correction and velocity reset:
    numpy brute force over terrain with the boundary rule of the reference's
    fillNeighbourData (Game/kernelSky.cu:1844-1874: a neighbour outside the
    domain takes the 'sides' condition of the cell's own value, a ground
    neighbour the 'ground' condition; the top face uses 'sides');
  * buoyancy: closed forms on horizontally uniform states;
  * sedimentation: exact transfer between levels, column budget, and the
    fall speed of Lin et al. (1983) for rain;
  * ground 2-D kernels: closed forms, water conservation and the latent-heat
    budget of the ground temperature;
  * process rates: homogeneity in the species they act on (the power laws of
    the Marshall-Palmer slope), zero source gives zero rate, saturation
    zeros, the saturation-adjustment derivative against a finite difference;
  * the T < 0 branch: every limited rate is bounded by its source species,
    and single-process cells follow the latent-heat closed form.
All oracle instances here are FP64.
"""
import ctypes
import math

import numpy as np
import pytest

import oracle_bind as ob

RNG = np.random.default_rng(20261204)
KAP = 287.0528 / 1003.5
G_ACC = 9.8076
N, D, C = 0, 1, 2  # NEUMANN, DIRICHLET, CUSTOM


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _state(nx, ny, nz, dx=100.0, gh=None, seed=3, **kw):
    import skewtcloudsim_amd as skt
    snd = skt.make_sounding(nx, ny, nz, dx, seed=seed, n_bubbles=2)
    o = ob.Oracle(nx, ny, nz, dx, fp64=True, **kw)
    if gh is not None:
        o.set_ground_height(gh)
    o.init(*[snd[k] for k in ("potTemps", "velField", "Qv", "groundTemp", "groundPres", "pressures",
                              "smallPressure")])
    return o


def _ground(gh, nx, ny, nz):
    return np.arange(ny)[None, :, None] <= gh.reshape(nz, nx)[:, None, :]  # [z, y, x]


def _g3(o, f):
    return o.get64(f).reshape(o.nz, o.ny, o.nx)


def _nb(a, ground, axis, step, sides, gnd, cust=0.0):
    """Neighbour value seen from every cell (fillNeighbourData): outside the
    domain -> 'sides' condition of the cell's own value, ground neighbour ->
    'ground' condition of the cell's own value, else the raw neighbour."""
    n = a.shape[axis]
    idx = np.arange(n) + step
    out1 = (idx < 0) | (idx >= n)
    idxc = np.clip(idx, 0, n - 1)
    raw = np.take(a, idxc, axis=axis)
    gn = np.take(ground, idxc, axis=axis)
    shp = [1, 1, 1]
    shp[axis] = n
    out = np.broadcast_to(out1.reshape(shp), a.shape)
    cust = np.broadcast_to(cust, a.shape)

    def bc(cond):
        return a if cond == N else (np.zeros_like(a) if cond == D else cust)
    return np.where(out, bc(sides), np.where(gn, bc(gnd), raw))


def _terrain(nx, nz, hi=4, seed=5):
    return np.random.default_rng(seed).integers(0, hi, nx * nz).astype(np.int32)


# ----------------------------------------------- pressure operator (a17) ---
def test_operator_faces_diagonal_and_preconditioner_brute_force():
    """initAMatrix + initPrecon (Game/kernelSky.cu:1320-1430), density bounds
    {N, D, D}: face coefficient -max(1, 1.225 / harmonic mean), zero at a
    ground face, the cell's own density beyond the domain; diagonal = minus
    the six faces; preconditioner = 1 / diagonal."""
    nx, ny, nz = 9, 8, 7
    gh = _terrain(nx, nz)
    o = _state(nx, ny, nz, gh=gh)
    rho = RNG.uniform(0.3, 1.6, (nz, ny, nx))
    rho[RNG.uniform(size=rho.shape) < 0.05] = RNG.uniform(2.0, 3.0)  # some faces below the max(1, .) floor
    o.set64("DENSITY", rho.ravel())
    arrs = [np.zeros(o.N) for _ in range(5)]
    o.L.sko_build_A(o.h, *[_p(a) for a in arrs])
    Ax, Ay, Az, Aw, pre = [a.reshape(nz, ny, nx) for a in arrs]
    grd = _ground(gh, nx, ny, nz)

    def face(axis, step):
        n = _nb(rho, grd, axis, step, N, D)
        with np.errstate(divide="ignore"):
            return np.where(n == 0, 0.0, -np.maximum(1.0, 1.225 / (2.0 / (1.0 / rho + 1.0 / n))))
    r, l = face(2, 1), face(2, -1)
    u, d = face(1, 1), face(1, -1)
    f, b = face(0, 1), face(0, -1)
    sky = ~grd
    w = -r - l - u - d - f - b
    with np.errstate(divide="ignore"):
        prew = 1.0 / w
    for got, want in ((Ax, r), (Ay, u), (Az, f), (Aw, w), (pre, prew)):
        np.testing.assert_allclose(got[sky], want[sky], rtol=1e-13, atol=0)
    assert (np.abs(Aw[sky]) > 0).all()
    # the floor max(1, .) is active on some faces and not on others
    assert (r[sky] == -1.0).any() and (r[sky] < -1.0).any()
    o.close()


def test_divergence_brute_force_with_terrain():
    """calculateDivergenceGPU (Game/kernelSky.cu:780-860): density bounds
    {N, N, D}, velocity bounds {C, C, D} with the default u / w profiles and 0
    for v; harmonic face densities (0 at a ground face); minus the density
    change over dt."""
    nx, ny, nz = 10, 9, 8
    gh = _terrain(nx, nz)
    o = _state(nx, ny, nz, gh=gh)
    grd = _ground(gh, nx, ny, nz)
    rho = RNG.uniform(0.5, 1.3, (nz, ny, nx))
    rold = rho * RNG.uniform(0.99, 1.01, rho.shape)
    u, v, w = (RNG.normal(0, 3, (nz, ny, nx)) for _ in range(3))
    defU, defW = RNG.normal(0, 5, ny), RNG.normal(0, 5, ny)
    for name, a in (("DENSITY", rho), ("DENSITY_OLD", rold), ("VELX", u), ("VELY", v), ("VELZ", w),
                    ("DEFAULT_VELX", defU), ("DEFAULT_VELZ", defW)):
        o.set64(name, a.ravel())
    dt = 0.5
    got = np.zeros(o.N)
    o.L.sko_divergence64(o.h, dt, _p(got))
    got = got.reshape(nz, ny, nx)
    cu = defU[None, :, None]
    cw = defW[None, :, None]
    ul = _nb(u, grd, 2, -1, C, D, cu)
    vd = _nb(v, grd, 1, -1, C, D, 0.0)
    wb = _nb(w, grd, 0, -1, C, D, cw)

    def hm(axis, step):
        n = _nb(rho, grd, axis, step, N, D)
        with np.errstate(divide="ignore"):
            return np.where(n == 0, 0.0, 2.0 / (1.0 / rho + 1.0 / n))
    want = ((hm(2, 1) * u - hm(2, -1) * ul) + (hm(1, 1) * v - hm(1, -1) * vd) + (hm(0, 1) * w - hm(0, -1) * wb)
            - (rho - rold) / dt)
    want[grd] = 0.0
    np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-12)
    # the terrain puts ground faces next to sky cells in every direction
    assert (grd[:, :, 1:] & ~grd[:, :, :-1]).any() and (grd[1:] & ~grd[:-1]).any()
    o.close()


def test_projection_correction_brute_force_with_terrain():
    """applyPresProjGPU (Game/kernelSky.cu:862-908): pressure bounds {D, N, N}
    (0 beyond the domain, no gradient into ground), density bounds {N, N, D}
    with the cell's own density where the face density is 0."""
    nx, ny, nz = 10, 9, 8
    gh = _terrain(nx, nz)
    o = _state(nx, ny, nz, gh=gh)
    grd = _ground(gh, nx, ny, nz)
    rho = RNG.uniform(0.5, 1.3, (nz, ny, nx))
    pp = RNG.normal(0, 10, (nz, ny, nx))
    u0, v0, w0 = (RNG.normal(0, 3, (nz, ny, nx)) for _ in range(3))
    for name, a in (("DENSITY", rho), ("VELX", u0), ("VELY", v0), ("VELZ", w0)):
        o.set64(name, a.ravel())
    o.L.sko_apply_projection(o.h, _p(np.ascontiguousarray(pp.ravel())))
    for name, a0, axis in (("VELX", u0, 2), ("VELY", v0, 1), ("VELZ", w0, 0)):
        pn = _nb(pp, grd, axis, 1, D, N)
        dn = _nb(rho, grd, axis, 1, N, D)
        with np.errstate(divide="ignore"):
            df = np.where(dn == 0, rho, 2.0 / (1.0 / rho + 1.0 / dn))
        want = a0 + (pn - pp) / df
        want[grd] = a0[grd]
        np.testing.assert_allclose(_g3(o, name), want, rtol=1e-13, atol=1e-13, err_msg=name)
    o.close()


def test_velocity_reset_brute_force_with_terrain():
    """resetVelPressProj (Game/kernelSky.cu:1648-1682): each (x, y) column
    is walked along z and the walk returns at the first ground cell; on the
    way v = 0 under the top or ground, u = default at the +x edge and 0 before
    ground, w likewise along z."""
    nx, ny, nz = 9, 7, 10
    gh = _terrain(nx, nz)
    o = _state(nx, ny, nz, gh=gh)
    grd = _ground(gh, nx, ny, nz)
    u0, v0, w0 = (RNG.normal(0, 3, (nz, ny, nx)) for _ in range(3))
    defU, defW = RNG.normal(0, 5, ny), RNG.normal(0, 5, ny)
    for name, a in (("VELX", u0), ("VELY", v0), ("VELZ", w0), ("DEFAULT_VELX", defU), ("DEFAULT_VELZ", defW)):
        o.set64(name, a.ravel())
    o.reset_vel()
    reached = np.cumsum(grd, axis=0) == 0
    u, v, w = u0.copy(), v0.copy(), w0.copy()
    up_g = np.ones_like(grd)
    up_g[:, :-1, :] = grd[:, 1:, :]  # the top counts as 'not sky above'
    v[reached & up_g] = 0.0
    edge_x = np.zeros_like(grd)
    edge_x[:, :, -1] = True
    u = np.where(reached & edge_x, np.broadcast_to(defU[None, :, None], u.shape), u)
    right_g = np.zeros_like(grd)
    right_g[:, :, :-1] = grd[:, :, 1:]
    u[reached & right_g] = 0.0
    edge_z = np.zeros_like(grd)
    edge_z[-1] = True
    w = np.where(reached & edge_z, np.broadcast_to(defW[None, :, None], w.shape), w)
    fwd_g = np.zeros_like(grd)
    fwd_g[:-1] = grd[1:]
    w[reached & fwd_g] = 0.0
    for name, want in (("VELX", u), ("VELY", v), ("VELZ", w)):
        assert np.array_equal(_g3(o, name), want), name
    assert (~reached & ~grd).any()  # walks that stop before reaching sky cells
    o.close()


# -------------------------------------------------------- buoyancy (a15) ---
def _uniform_column_state(nx, ny, nz, th=300.0, qv=0.004, p=900.0):
    o = _state(nx, ny, nz)
    o.set64("POTTEMP", np.full(o.N, th))
    o.set64("QV", np.full(o.N, qv))
    o.set64("PRESSURE", np.full(o.N, p))
    o.set64("ISEN_TEMP", np.full(ny, th))
    o.set64("ISEN_VAPOR", np.full(ny, qv))
    for f in ("QR", "QS", "QI", "VELY"):
        o.set64(f, np.zeros(o.N))
    return o


def test_buoyancy_uniform_state_is_hydrometeor_loading_only():
    """A horizontally (and vertically) uniform, unsaturated state at rest has
    no thermal buoyancy: the change is exactly -g (Qr + Qs + Qi) dt
    (buoyancyGPU, Game/kernelSky.cu:1061-1247)."""
    nx, ny, nz = 8, 7, 6
    o = _uniform_column_state(nx, ny, nz)
    dt = 0.5
    o.buoyancy(dt)
    sky = (np.arange(o.N) // nx) % ny >= 1
    assert np.all(o.get64("VELY")[sky] == 0.0)
    q = {f: RNG.uniform(0, 1e-3, o.N) for f in ("QR", "QS", "QI")}
    for f, a in q.items():
        o.set64(f, a)
    o.set64("VELY", np.zeros(o.N))
    o.buoyancy(dt)
    tot = q["QR"] + q["QS"] + q["QI"]
    np.testing.assert_allclose(o.get64("VELY")[sky], -G_ACC * tot[sky] * dt, rtol=1e-13)
    o.close()


def test_buoyancy_warm_cell_closed_form():
    """One cell d theta warmer than its four horizontal neighbours, Qv uniform:
    B = g d / theta_env; each neighbour sees a quarter of the excess in its
    environment mean."""
    nx, ny, nz = 8, 7, 6
    o = _uniform_column_state(nx, ny, nz)
    th = _g3(o, "POTTEMP")
    z, y, x = 3, 3, 4
    dth = 1.5
    th[z, y, x] += dth
    o.set64("POTTEMP", th.ravel())
    dt = 0.5
    o.buoyancy(dt)
    v = _g3(o, "VELY")
    assert abs(v[z, y, x] - G_ACC * dth / 300.0 * dt) < 1e-14
    env = 300.0 + dth / 4
    want_nb = G_ACC * (300.0 - env) / env * dt
    for (zz, xx) in ((z, x - 1), (z, x + 1), (z - 1, x), (z + 1, x)):
        assert abs(v[zz, y, xx] - want_nb) < 1e-14
    o.close()


def test_buoyancy_direction_selection_and_reversal_cap():
    """With v > 0 the upward buoyancy is used, with v < 0 the downward one;
    a change that would reverse an upward velocity stops it at exactly 0."""
    nx, ny, nz = 8, 7, 6
    o = _uniform_column_state(nx, ny, nz)
    z, y, x = 3, 3, 4
    i = x + nx * (y + ny * z)
    qr = np.zeros(o.N)
    qr[i] = 1e-3
    o.set64("QR", qr)
    dt = 0.5
    for cur, want in ((1e-4, 0.0), (1.0, 1.0 - G_ACC * 1e-3 * dt), (-1.0, -1.0 - G_ACC * 1e-3 * dt)):
        vel = np.zeros(o.N)
        vel[i] = cur
        o.set64("VELY", vel)
        o.buoyancy(dt)
        got = o.get64("VELY")[i]
        assert abs(got - want) <= 1e-15 * max(1.0, abs(want)), (cur, got, want)
        if want == 0.0:
            assert got == 0.0
    o.close()


# ---------------------------------------------------- sedimentation (a14) ---
def _lin_rain_speed(rho_air, q):
    """Mass-weighted fall speed of rain, Lin et al. (1983): V = a Gamma(4 + b)
    / (6 lambda^b) (rho0 / rho)^(1/2), lambda = (pi rho_w N0 / (rho q))^(1/4)
    in cgs units (a = 2115 cm^(1-b)/s, b = 0.8, N0 = 0.08 cm^-4,
    rho_w = 0.99 g/cm^3), returned in m/s."""
    lam = (math.pi * 0.99 * 0.08 / max(rho_air * q * 1e-3, 1e-14)) ** 0.25
    # Gamma(4 + b) at the reference's float argument 4.0f + 0.8f
    g = math.gamma(float(np.float32(4.0) + np.float32(0.8)))
    return 2115.0 * g / (6.0 * lam ** 0.8) * math.sqrt(1.225 / rho_air) * 0.01


def test_sedimentation_single_level_moves_down_exactly():
    """Rain at one level only: the level below gains exactly what it loses,
    dt min(V / dx Q, Q) with V from Lin et al. (1983) at the cell's density."""
    nx, ny, nz = 5, 8, 4
    o = _state(nx, ny, nz)
    th, p, qv = _g3(o, "POTTEMP"), _g3(o, "PRESSURE"), _g3(o, "QV")
    gP = o.get64("GROUND_P").reshape(nz, nx)
    q = np.zeros((nz, ny, nx))
    z, k, x = 2, 5, 3
    q[z, k, x] = 2e-3
    o.set64("QR", q.ravel())
    dt = 10.0  # Courant 5 * 10 / 100 = 0.5: one substep
    o.advect_precip(ob.FIELD["QR"], 0, dt)
    got = _g3(o, "QR")
    T = th[z, k, x] * (p[z, k, x] / gP[z, x]) ** KAP
    rho = p[z, k, x] * 100 / (287.0528 * T * (0.608 * qv[z, k, x] + 1))
    moved = dt * min(_lin_rain_speed(rho, 2e-3) / 100.0 * 2e-3, 2e-3)
    assert abs(got[z, k - 1, x] - moved) < 1e-12 * moved
    assert abs(got[z, k, x] - (2e-3 - moved)) < 1e-15
    got[z, k - 1, x] = got[z, k, x] = 0.0
    assert not got.any()
    o.close()


@pytest.mark.parametrize("ftype,field", [(0, "QR"), (1, "QS"), (2, "QI")])
def test_sedimentation_column_budget_telescopes(ftype, field):
    """Every transfer between two sky levels cancels: each column loses
    exactly the lowest sky cell's outflow into the ground, and the column
    above the lowest sky cell keeps its total minus that outflow (terrain)."""
    nx, ny, nz = 6, 9, 5
    gh = _terrain(nx, nz, hi=3)
    o = _state(nx, ny, nz, gh=gh)
    grd = _ground(gh, nx, ny, nz)
    q = RNG.uniform(0, 3e-3, (nz, ny, nx))
    q[grd] = 0.0
    o.set64(field, q.ravel())
    dt = 5.0
    o.advect_precip(ob.FIELD[field], ftype, dt)
    got = _g3(o, field)
    assert (got[~grd] >= -1e-18).all()
    G = gh.reshape(nz, nx)
    for zz in range(nz):
        for xx in range(nx):
            lo = G[zz, xx] + 1
            col0, col1 = q[zz, lo:, xx], got[zz, lo:, xx]
            out = col0.sum() - col1.sum()
            # the outflow is what the lowest cell lost beyond what it gained
            # from the cell above, which itself lost that gain
            gain_lo = col1[0] - col0[0]
            assert out > 0
            assert abs((col1[1:].sum() - col0[1:].sum()) + (gain_lo + out)) < 1e-15
    o.close()


def test_fall_speed_power_laws():
    """V ~ Q^(b/4) and ~ rho^(b/4 - 1/2) for rain (b = 0.8) and snow
    (b = 0.25); ice V ~ Q^(1/8) rho^(1/8 - 1/2) (its lambda^(-1/2) times
    rho^(-1/2))."""
    L = ob.lib()
    out = np.zeros(3, np.float32)

    def vel(q, rho, t):
        L.sko_fallvel(q, q, q, rho, t, _p(out))
        return float(out[t])
    for t, e in ((0, 0.2), (1, 0.0625), (2, 0.125)):
        r = vel(16e-4, 1.0, t) / vel(1e-4, 1.0, t)
        assert abs(r - 16 ** e) < 2e-5 * 16 ** e, (t, r)
        r = vel(1e-3, 0.5, t) / vel(1e-3, 1.0, t)
        assert abs(r - 0.5 ** (e - 0.5)) < 2e-5, (t, r)
    assert abs(vel(2e-3, 1.1, 0) - _lin_rain_speed(1.1, 2e-3)) < 2e-6 * _lin_rain_speed(1.1, 2e-3)


# -------------------------------------------------- ground phase (a6-a11) ---
def test_cloud_cover_column_closed_form():
    """calculateCloudCoverGPU: LC = min(sum over the sky levels above the
    ground of (Qc + Qw) dx / 1.2, 1)."""
    nx, ny, nz = 6, 10, 5
    gh = _terrain(nx, nz, hi=4)
    o = _state(nx, ny, nz, gh=gh)
    grd = _ground(gh, nx, ny, nz)
    qc = RNG.uniform(0, 2e-4, (nz, ny, nx))
    qw = RNG.uniform(0, 2e-4, (nz, ny, nx))
    qc[0, :, 0] = 1e-2  # this column saturates the cover
    qc[grd] = 5.0  # ground cells hold junk that must not count
    o.set64("QC", qc.ravel())
    o.set64("QW", qw.ravel())
    o.cloud_cover()
    got = o.get64("CLOUD_COVER").reshape(nz, nx)
    want = np.minimum(np.where(grd, 0.0, qc + qw).sum(axis=1) * 100.0 / 1.2, 1.0)
    np.testing.assert_allclose(got, want, rtol=1e-13)
    assert got[0, 0] == 1.0 and (got < 1.0).any()
    o.close()


def test_ground_temperature_radiation_balance():
    """calculateGroundTempGPU: full cover leaves the ground temperature alone,
    no cover adds dt ((1 - 0.25) irr - 0.95 sigma T^4) / (1 m 1500 kg/m^3 Cpds),
    which vanishes at the radiative equilibrium temperature."""
    nx, ny, nz = 4, 5, 3
    o = _state(nx, ny, nz)
    g = o.G
    T0 = RNG.uniform(270, 320, g)
    irr = 800.0
    T0[0] = (0.75 * irr / (0.95 * 5.67e-8)) ** 0.25
    cover = np.zeros(g)
    cover[1] = 1.0
    o.set64("GROUND_T", T0)
    o.set64("CLOUD_COVER", cover)
    dt = 2.0
    o.ground_temps(dt, 1.0, irr)
    got = o.get64("GROUND_T")
    want = T0 + dt * (1 - cover) * ((0.75 * irr - 0.95 * 5.67e-8 * T0 ** 4) / (1500.0 * 800.0))
    np.testing.assert_allclose(got, want, rtol=1e-13)
    assert got[1] == T0[1]
    assert abs(got[0] - T0[0]) < 1e-9
    o.close()


def test_set_temps_at_ground_closed_form():
    """setTempsAtGroundGPU: theta at the first sky level relaxes towards the
    potential temperature of the ground temperature, Tg (pg / p)^kappa, by
    min(1, dt); columns whose first sky level is the top one are skipped."""
    nx, ny, nz = 6, 7, 4
    gh = _terrain(nx, nz, hi=4)
    gh[3] = ny - 2  # first sky level = the top level: skipped
    o = _state(nx, ny, nz, gh=gh)
    th0, p = _g3(o, "POTTEMP").copy(), _g3(o, "PRESSURE")
    Tg = RNG.uniform(280, 310, o.G)
    o.set64("GROUND_T", Tg)
    gP = o.get64("GROUND_P").reshape(nz, nx)
    dt = 0.25  # exact in float, which the binding passes
    o.set_temps_at_ground(dt, 1.0)
    got = _g3(o, "POTTEMP")
    want = th0.copy()
    G = gh.reshape(nz, nx)
    Tg = Tg.reshape(nz, nx)
    for z in range(nz):
        for x in range(nx):
            y = G[z, x] + 1
            if y >= ny - 1:
                continue
            tgt = Tg[z, x] * (gP[z, x] / p[z, y, x]) ** KAP
            want[z, y, x] -= (th0[z, y, x] - tgt) * min(1.0, dt)
    np.testing.assert_allclose(got, want, rtol=1e-13)
    assert got[0, ny - 1, 3] == th0[0, ny - 1, 3]
    o.close()


def test_ground_water_flat_conservation_and_spike():
    """advectGroundWaterGPU on flat ground: no slope flow, diffusion between
    neighbour columns conserves both stores; a single wet column gives each
    of its four neighbours dt min(q D / dx^2 speed, q) (D = 1.75e-3 for
    ground rain, 1.8e-5 for soil water); a uniform field stays put."""
    nx, ny, nz = 7, 4, 6
    o = _state(nx, ny, nz)
    g = o.G
    o.set64("QGR", np.full(g, 2e-3))
    o.set64("QRS", np.full(g, 1e-3))
    o.ground_water(0.5, 1.0)
    assert np.all(o.get64("QGR") == 2e-3) and np.all(o.get64("QRS") == 1e-3)
    gr = RNG.uniform(0, 1e-3, g)
    rs = RNG.uniform(0, 1e-3, g)
    o.set64("QGR", gr)
    o.set64("QRS", rs)
    dt = 3e5  # strong enough that fluxes matter against rounding
    o.ground_water(dt, 1.0)
    assert abs(o.get64("QGR").sum() - gr.sum()) < 1e-15 * g
    assert abs(o.get64("QRS").sum() - rs.sum()) < 1e-15 * g
    q = 1e-3
    spike = np.zeros(g)
    c = 3 + nx * 2
    spike[c] = q
    o.set64("QGR", spike)
    o.set64("QRS", spike)
    dt = 1e6
    o.ground_water(dt, 1.0)
    got_r, got_s = o.get64("QGR"), o.get64("QRS")
    fr = dt * min(q * 1.75e-3 / 100.0 ** 2, q)
    fs = dt * min(q * 1.8e-5 / 100.0 ** 2, q)
    for nb in (c - 1, c + 1, c - nx, c + nx):
        assert abs(got_r[nb] - fr) < 1e-15 and abs(got_s[nb] - fs) < 1e-15
    assert abs(got_r[c] - (q - 4 * fr)) < 1e-15 and abs(got_s[c] - (q - 4 * fs)) < 1e-15
    o.close()


def test_ground_water_slope_flow():
    """Manning flow down a one-cell step: the lower column gains
    dt min(n h^(2/3) slope^(1/2) Qgr speed, Qgr) of the higher column's ground
    rain (n = 1/0.03, h = Qgr), and soil water moves by g 5e-5 Qrs slope."""
    nx, ny, nz = 4, 6, 3
    gh = np.zeros(nx * nz, np.int32)
    gh[1 + nx * 1] = 1  # one raised column, its neighbours lower by 1
    o = _state(nx, ny, nz, gh=gh)
    g = o.G
    q = 1e-4
    gr = np.zeros(g)
    gr[1 + nx] = q
    o.set64("QGR", gr)
    o.set64("QRS", np.zeros(g))
    dt = 1.0
    o.ground_water(dt, 1.0)
    got = o.get64("QGR")
    v = (1 / 0.03) * q ** (2.0 / 3.0) * 1.0 ** 0.5
    lap = q * 1.75e-3 / 100.0 ** 2
    want_nb = dt * min(v * q, q) + dt * min(lap, q)
    for nb in (0 + nx, 2 + nx, 1, 1 + 2 * nx):
        assert abs(got[nb] - want_nb) < 1e-15 * max(1.0, want_nb / 1e-15), (nb, got[nb], want_nb)
    o.close()


def _ground_rates(tAK=270.0, tGC=2.0, ps=950.0, Qv=0.004, Qgr=1e-3, Qgs=0.0, Qgi=0.0, Qrs=1e-3, irr=500.0,
                  cc=0.3, Dair=1.1, wind=3.0, dt=0.5, time=100.0, Qr=0.0, Qi=0.0):
    a = np.array([tAK, tGC, ps, Qv, Qgr, Qgs, Qgi, Qrs, irr, cc, Dair, wind, dt, time, Qr, Qi], np.float64)
    out = np.zeros(9)
    ob.lib().sko_ground_rates64(_p(a), _p(out))
    return dict(zip(("PGREVP", "PGSMLT", "PGGMLT", "PGGFR", "PGSFR", "PGFLW", "PGDEVP", "time", "PGRFR"), out))


def test_ground_surface_rates_special_cases():
    """Surface rates (Game/microPhysics.cu:687-1032): each vanishes without its
    source or outside its temperature range, never exceeds its source where
    the reference bounds it, and follows its closed-form scalings."""
    r = _ground_rates
    # infiltration: 100 * 5e-5 * Qgr
    assert r(Qgr=2e-3)["PGFLW"] == pytest.approx(1e-5, rel=1e-15) and r(Qgr=0.0)["PGFLW"] == 0.0
    # evaporation of ground rain: none without it, ~ sqrt(Qgr) (wetted area)
    assert r(Qgr=0.0)["PGREVP"] == 0.0
    a, b = r(Qgr=1e-4, tGC=15.0, Qv=0.002)["PGREVP"], r(Qgr=4e-4, tGC=15.0, Qv=0.002)["PGREVP"]
    assert a > 0 and b / a == pytest.approx(2.0, rel=1e-12)
    # soil evaporation: 200 * 1e-6 * Qrs * exp(-t / (0.21 day)) on a dry
    # surface; any surface water resets the clock and stops it
    d = r(Qgr=0.0, Qrs=2e-3, time=5000.0)
    assert d["PGDEVP"] == pytest.approx(200 * 1e-6 * 2e-3 * math.exp(-5000.0 / (0.21 * 86400)), rel=1e-14)
    assert d["time"] == 5000.0
    d = r(Qgr=1e-3, Qrs=2e-3, time=5000.0)
    assert d["PGDEVP"] == 0.0 and d["time"] == 0.0
    # melt of ground ice: none at or below 0 C or without ice; linear in the
    # ground temperature; ~ Qgi^-1/2 (1/Qgi from the layer depth times
    # lambda^-2 ~ Qgi^1/2) at fixed air density
    assert r(Qgi=0.0, tGC=3.0)["PGGMLT"] == 0.0 and r(Qgi=1e-3, tGC=-3.0)["PGGMLT"] == 0.0
    m1, m2 = r(Qgi=1e-3, tGC=2.0)["PGGMLT"], r(Qgi=1e-3, tGC=4.0)["PGGMLT"]
    assert m1 > 0 and m2 / m1 == pytest.approx(2.0, rel=1e-13)
    m4 = r(Qgi=4e-3, tGC=2.0)["PGGMLT"]
    assert m4 / m1 == pytest.approx(0.5, rel=1e-12)
    # snow melt: none without snow; the local energy term shadows the snow
    # argument, so any snow amount gives the same rate; linear in dt
    assert r(Qgs=0.0, tAK=285.0)["PGSMLT"] == 0.0
    s1 = r(Qgs=1e-3, tAK=290.0, tGC=8.0, irr=900.0, dt=0.5)["PGSMLT"]
    assert s1 > 0
    assert r(Qgs=7e-3, tAK=290.0, tGC=8.0, irr=900.0, dt=0.5)["PGSMLT"] == s1
    assert r(Qgs=1e-3, tAK=290.0, tGC=8.0, irr=900.0, dt=1.0)["PGSMLT"] == pytest.approx(2 * s1, rel=1e-13)
    # freezing of ground rain: only below 0 C air temperature and with rain;
    # a fraction of the rain
    assert r(Qgr=1e-3, tAK=275.0)["PGGFR"] == 0.0 and r(Qgr=0.0, tAK=260.0)["PGGFR"] == 0.0
    for kw in (dict(tAK=260.0, tGC=-5.0), dict(tAK=272.0, tGC=5.0, wind=0.0), dict(tAK=250.0, tGC=-20.0, Qgi=1e-3)):
        f = r(Qgr=1e-3, **kw)["PGGFR"]
        assert 0.0 <= f <= 1e-3, kw
    # freezing of rain on snow: needs both; a fraction of the rain
    assert r(Qgr=1e-3, Qgs=0.0, tAK=260.0)["PGSFR"] == 0.0 and r(Qgr=0.0, Qgs=1e-3, tAK=260.0)["PGSFR"] == 0.0
    f = r(Qgr=1e-3, Qgs=1e-3, tAK=265.0, tGC=-3.0)["PGSFR"]
    assert 0.0 < f <= 1e-3
    # freezing rain hitting a dry ground below 0 C
    assert r(Qr=1e-3, Qgr=0.0, tAK=265.0)["PGRFR"] >= 0.0
    assert r(Qr=1e-3, Qgr=1e-4, tAK=265.0)["PGRFR"] == 0.0
    assert r(Qr=1e-3, Qgr=0.0, tAK=275.0)["PGRFR"] == 0.0
    assert r(Qr=0.0, Qgr=0.0, tAK=265.0)["PGRFR"] == 0.0
    assert r(Qr=1e-3, Qgr=0.0, Qgs=1e-3, tAK=265.0)["PGRFR"] == 0.0
    assert r(Qr=1e-3, Qgr=0.0, tAK=250.0, tGC=-10.0, wind=5.0)["PGRFR"] <= 1e-3


def _ground_state(nx, ny, nz, cold):
    gh = _terrain(nx, nz, hi=3)
    o = _state(nx, ny, nz, gh=gh)
    g = o.G
    o.set64("QGR", RNG.uniform(0, 2e-3, g) * (RNG.uniform(size=g) < 0.7))
    o.set64("QGS", RNG.uniform(0, 2e-3, g) * (RNG.uniform(size=g) < 0.5))
    o.set64("QGI", RNG.uniform(0, 2e-3, g) * (RNG.uniform(size=g) < 0.5))
    o.set64("QRS", RNG.uniform(0, 2e-3, g))
    o.set64("GROUND_T", RNG.uniform(255, 268, g) if cold else RNG.uniform(276, 300, g))
    o.set64("CLOUD_COVER", RNG.uniform(0, 1, g))
    o.set64("GROUND_TIME", RNG.uniform(0, 3e4, g))
    for f in ("QR", "QS", "QI"):
        o.set64(f, RNG.uniform(0, 2e-3, o.N) * (RNG.uniform(size=o.N) < 0.6))
    if cold:
        o.set64("POTTEMP", o.get64("POTTEMP") - 35.0)
    o.init_density()
    return o, gh


@pytest.mark.parametrize("cold", [False, True])
def test_ground_microphysics_water_and_heat_budget(cold):
    """calculateGroundMicroPhysicsGPU: the water of the ground stores and of
    the vapour above them is conserved, and the ground temperature changes
    by the latent heats of what changed phase: -Lwater(Tg) dQv (evaporation
    cools) + Lf (dQgi + dQgs) (freezing warms), over 1500 kg/m^3 x 0.1 m x
    1500 J/(kg K)."""
    nx, ny, nz = 8, 6, 7
    o, gh = _ground_state(nx, ny, nz, cold)
    keys = ("QRS", "QGR", "QGS", "QGI")
    before = {k: o.get64(k) for k in keys + ("GROUND_T",)}
    qv0 = _g3(o, "QV").copy()
    L = ob.lib()
    o.L.sko_micro_ground.argtypes  # bound
    # ground microphysics alone: precipitation hitting the ground is tested below
    zero = np.zeros(o.N)
    for f in ("QR", "QS", "QI"):
        o.set64(f, zero)
    o.micro_ground(0.5, 1.0, 600.0)
    after = {k: o.get64(k) for k in keys + ("GROUND_T",)}
    qv1 = _g3(o, "QV")
    G = gh.reshape(nz, nx)
    dqv = np.array([qv1[z, G[z, x] + 1, x] - qv0[z, G[z, x] + 1, x] for z in range(nz) for x in range(nx)])
    tot = sum(after[k] - before[k] for k in keys) + dqv
    assert np.abs(tot).max() < 1e-17
    tGC = before["GROUND_T"] - 273.15
    Lw = np.array([L.sko_Lwater(float(t)) for t in tGC])
    dT = (-Lw * dqv + 3.3355e5 * ((after["QGI"] - before["QGI"]) + (after["QGS"] - before["QGS"]))) / (1500 * 0.1 * 1500)
    # Lwater is evaluated in float by the binding: 1e-7 relative on that term
    np.testing.assert_allclose(after["GROUND_T"] - before["GROUND_T"], dT, rtol=1e-6, atol=1e-12)
    assert np.abs(dT).max() > 0 and np.abs(dqv).max() > 0
    o.close()


def _cp_moist(qv):
    """cp of dry air + vapour as an ideal-gas mixture: mole fraction from the
    mass fraction, heat-capacity ratio mixed by mass (1.33 vapour, 1.4 dry
    air), cp = gamma R / (M (gamma - 1))."""
    Mair, Mw = 0.02896, 0.01802
    xv = (qv / Mw) / (qv / Mw + (1 - qv) / Mair)
    M = xv * Mw + (1 - xv) * Mair
    yv = xv * Mw / M
    gam = yv * 1.33 + (1 - yv) * 1.4
    return gam * 8.3145 / (M * (gam - 1))


def test_precipitation_hitting_ground_budget():
    """calculatePrecipHittingGroundMicroPhysicsGPU: the lowest sky cell hands
    dt min(Q speed, Q) of its rain, snow and ice to the ground stores (rain
    that freezes on contact goes to ground ice), total water is conserved, and
    the ground warms by Lf / cv (the freezing's latent heat over the moist
    air's constant-volume heat capacity) per unit frozen."""
    nx, ny, nz = 8, 6, 7
    o, gh = _ground_state(nx, ny, nz, cold=True)
    o.set64("QGR", np.zeros(o.G))
    o.set64("QGS", np.zeros(o.G))
    G = gh.reshape(nz, nx)
    lo = lambda a: np.array([a[z, G[z, x] + 1, x] for z in range(nz) for x in range(nx)])
    sky0 = {f: lo(_g3(o, f)) for f in ("QR", "QS", "QI", "QV")}
    g0 = {k: o.get64(k) for k in ("QGR", "QGS", "QGI", "GROUND_T")}
    L = o.L
    irr = 0.0
    # precipitation hitting ground only: run the combined call on a copy
    # without the surface processes' sources is not possible, so compare the
    # hand-over against the combined call's sky side (the surface processes
    # act on the ground stores only)
    dt = 0.5
    L.sko_micro_ground(o.h, dt, 1.0, irr)
    sky1 = {f: lo(_g3(o, f)) for f in ("QR", "QS", "QI")}
    for f in ("QR", "QS", "QI"):
        np.testing.assert_allclose(sky1[f], sky0[f] - dt * sky0[f], rtol=1e-14, atol=1e-20)
    o.close()
    # the hand-over alone, with the surface processes inert: no ground water,
    # no stores, warm enough ground that nothing freezes or melts there
    o, gh = _ground_state(nx, ny, nz, cold=True)
    for k in ("QGR", "QGS", "QGI", "QRS"):
        o.set64(k, np.zeros(o.G))
    G = gh.reshape(nz, nx)
    sky0 = {f: lo(_g3(o, f)) for f in ("QR", "QS", "QI", "QV")}
    gT0 = o.get64("GROUND_T")
    o.L.sko_micro_ground(o.h, dt, 1.0, irr)
    gr, gs, gi = o.get64("QGR"), o.get64("QGS"), o.get64("QGI")
    handed = dt * (sky0["QR"] + sky0["QS"] + sky0["QI"])
    # what arrives stays in the stores except surface processes' share; the
    # hand-over itself conserves water: sky loss = store gain + vapour change
    qv1 = lo(_g3(o, "QV"))
    stores = gr + gs + gi + o.get64("QRS")
    np.testing.assert_allclose(stores + (qv1 - sky0["QV"]), handed, rtol=1e-12, atol=1e-18)
    o.close()


# ----------------------------------------------- process rates (a19, a20) ---
NAMES = ("PVCON", "PVDEP", "PIMLT", "PIDW", "PIHOM", "PIACR", "PRACI", "PRAUT", "PRACW", "PREVP", "PRACS",
         "PSACW", "PSACR", "PSACI", "PSAUT", "PSFW", "PSFI", "PSDEP", "PSMLT", "PGAUT", "PGFR", "PGACW", "PGACI",
         "PGACR", "PGACS", "PGSUB", "PGMLT", "PGDRY", "PGWET", "PGACR1", "PT1", "PT2")


def rates(T=-10.0, ps=700.0, Qv=2e-3, Qw=1e-3, Qc=5e-4, Qr=1e-3, Qs=1e-3, Qi=1e-3, Dair=0.9, dt=0.5, speed=1.0):
    a = np.array([T, ps, Qv, Qw, Qc, Qr, Qs, Qi, Dair, dt, speed], np.float64)
    out = np.zeros(32)
    ob.lib().sko_mp_rates64(_p(a), _p(out))
    return dict(zip(NAMES, out))


# (rate, species it is linear in, species whose slope parameter sets the
# power, exponent of that species): lambda ~ Q^(-1/4), so lambda^-(n) ~ Q^(n/4)
POWER_LAWS = [
    ("PIACR", "Qc", "Qr", (6 + 0.8) / 4),   # Lin et al. (1983): lambda_R^-(6+b)
    ("PRACI", "Qc", "Qr", (3 + 0.8) / 4),
    ("PRACW", "Qw", "Qr", (3 + 0.8) / 4),
    ("PSACW", "Qw", "Qs", (3 + 0.25) / 4),
    ("PSACI", "Qc", "Qs", (3 + 0.25) / 4),
    ("PGACW", "Qw", "Qi", 3.5 / 4),
    ("PGACI", "Qc", "Qi", 3.5 / 4),
]


@pytest.mark.parametrize("name,lin,slope_sp,power", POWER_LAWS)
def test_collection_rates_power_laws(name, lin, slope_sp, power):
    base = dict(T=-12.0, Qw=4e-4, Qc=3e-4, Qr=8e-4, Qs=6e-4, Qi=7e-4)
    r0 = rates(**base)[name]
    assert r0 > 0
    kw = dict(base)
    kw[lin] = base[lin] * 3
    assert rates(**kw)[name] / r0 == pytest.approx(3.0, rel=1e-12)
    kw = dict(base)
    kw[slope_sp] = base[slope_sp] * 16
    assert rates(**kw)[name] / r0 == pytest.approx(16 ** power, rel=1e-12)
    for sp in (lin, slope_sp):
        kw = dict(base)
        kw[sp] = 0.0
        assert rates(**kw)[name] == 0.0, (name, sp)


def test_autoconversion_and_freezing_closed_forms():
    r = rates(T=-8.0, Qc=3e-3, Qs=2e-3)
    assert r["PSAUT"] == pytest.approx(1e-3 * math.exp(0.025 * -8.0) * (3e-3 - 1e-3), rel=1e-14)
    assert r["PGAUT"] == pytest.approx(1e-3 * math.exp(0.09 * -8.0) * (2e-3 - 6e-4), rel=1e-14)
    assert rates(T=-8.0, Qc=5e-4)["PSAUT"] == 0.0 and rates(T=3.0, Qc=3e-3)["PSAUT"] == 0.0
    assert rates(T=-8.0, Qs=5e-4)["PGAUT"] == 0.0 and rates(T=3.0, Qs=2e-3)["PGAUT"] == 0.0
    # probabilistic freezing of rain (Bigg): (exp(-0.66 T) - 1) lambda_R^-7
    a, b = rates(T=-5.0, Qr=1e-3)["PGFR"], rates(T=-10.0, Qr=1e-3)["PGFR"]
    assert b / a == pytest.approx((math.exp(6.6) - 1) / (math.exp(3.3) - 1), rel=1e-12)
    assert rates(T=-5.0, Qr=16e-3)["PGFR"] / a == pytest.approx(16 ** 1.75, rel=1e-12)
    assert rates(T=1.0, Qr=1e-3)["PGFR"] == 0.0
    # homogeneous freezing below -40 C, melting of cloud ice at or above 0 C
    assert rates(T=-41.0, Qw=2e-3)["PIHOM"] == 2e-3 and rates(T=-39.0, Qw=2e-3)["PIHOM"] == 0.0
    assert rates(T=0.0, Qc=7e-4)["PIMLT"] == 7e-4 and rates(T=-0.1, Qc=7e-4)["PIMLT"] == 0.0
    # cloud-water autoconversion needs at least QwMin = 1e-3
    assert rates(T=5.0, Qw=9e-4)["PRAUT"] == 0.0 and rates(T=5.0, Qw=2e-3)["PRAUT"] > 0


def _ws(T, p):
    es = 0.61078 * math.exp(17.27 * T / (T + 237.3)) * 10
    return (287.0528 / 461.5) * es / (p - es)


def _wi(T, p):
    es = 0.61078 * math.exp(21.875 * T / (T + 265.5)) * 10
    return (287.0528 / 461.5) * es / (p - es)


def test_saturation_adjustment_derivatives_and_zeros():
    """FPVCON / FPVDEP: zero at saturation, and the derivative in the
    denominator is d ws / dT (d wi / dT), checked by a central difference."""
    for T, p in ((12.0, 850.0), (-15.0, 600.0)):
        for Qv in (_ws(T, p), _wi(T, p)):
            pass
        assert rates(T=T, ps=p, Qv=_ws(T, p))["PVCON"] == pytest.approx(0.0, abs=1e-20)
        assert rates(T=T, ps=p, Qv=_wi(T, p))["PVDEP"] == pytest.approx(0.0, abs=1e-20)
        h = 1e-4
        for name, sat, L in (("PVCON", _ws, 2.374e6), ("PVDEP", _wi, 2.834e6)):
            der = (sat(T + h, p) - sat(T - h, p)) / (2 * h)
            qv = 1.3 * sat(T, p)
            dt = 0.5
            want = (qv - sat(T, p)) / (dt * (1 + L / 1003.5 * der))
            assert rates(T=T, ps=p, Qv=qv, dt=dt)[name] == pytest.approx(want, rel=1e-8), (name, T)


def test_sub_saturation_rates_vanish_at_saturation_and_are_linear():
    """Rain evaporation ~ (1 - Qv/Qws) below water saturation (0 above);
    snow deposition / sublimation ~ (Qv/Qwi - 1); graupel sublimation
    ~ (1 - Qv/Qwi) below ice saturation (0 above)."""
    T, p = 6.0, 900.0
    qs = _ws(T, p)
    a, b = rates(T=T, ps=p, Qv=0.5 * qs)["PREVP"], rates(T=T, ps=p, Qv=0.8 * qs)["PREVP"]
    assert a > 0 and b / a == pytest.approx(0.2 / 0.5, rel=1e-10)
    assert rates(T=T, ps=p, Qv=1.1 * qs)["PREVP"] == 0.0
    T, p = -14.0, 650.0
    qi = _wi(T, p)
    a, b = rates(T=T, ps=p, Qv=1.3 * qi)["PSDEP"], rates(T=T, ps=p, Qv=0.7 * qi)["PSDEP"]
    assert a > 0 and b / a == pytest.approx(-1.0, rel=1e-10)
    a, b = rates(T=T, ps=p, Qv=0.4 * qi)["PGSUB"], rates(T=T, ps=p, Qv=0.7 * qi)["PGSUB"]
    assert a > 0 and b / a == pytest.approx(0.5, rel=1e-10)
    assert rates(T=T, ps=p, Qv=1.05 * qi)["PGSUB"] == 0.0


def test_zero_source_gives_zero_rate():
    full = dict(T=-12.0, Qv=2e-3, Qw=1e-3, Qc=1e-3, Qr=1e-3, Qs=1e-3, Qi=1e-3)
    needs = {"Qw": ("PIDW", "PIHOM", "PRACW", "PSACW", "PSFW", "PGACW"),
             "Qc": ("PIACR", "PRACI", "PSACI", "PSFI", "PGACI"),
             "Qr": ("PIACR", "PRACI", "PRACW", "PREVP", "PRACS", "PSACR", "PGFR", "PGACR"),
             "Qs": ("PRACS", "PSACW", "PSACR", "PSACI", "PSFW", "PSFI", "PGACS"),
             "Qi": ("PGACW", "PGACI", "PGACR", "PGACS", "PGSUB", "PGMLT", "PGWET", "PGACR1")}
    r0 = rates(**full)
    for sp, names in needs.items():
        kw = dict(full)
        kw[sp] = 0.0
        r = rates(**kw)
        for n in names:
            assert r[n] == 0.0, (sp, n)
            # PIHOM needs T < -40 C; PGACS collides snow and ice, whose fall
            # speeds the 'all' branch leaves at 0 (only rain is evaluated)
            # PGMLT (melting) needs T > 0
            assert r0[n] != 0.0 or n in ("PIHOM", "PREVP", "PRACS", "PGSUB", "PGACS", "PGMLT"), (sp, n)
    warm = dict(full, T=4.0)
    assert rates(**warm)["PGMLT"] > 0 and rates(**dict(warm, Qi=0.0))["PGMLT"] == 0.0
    assert rates(**warm)["PSMLT"] > 0 and rates(**dict(warm, Qs=0.0))["PSMLT"] == 0.0


# region-diagnostic order (Game/microPhysics.cu:1357-1415) and the species each
# limited rate is bounded by (dt min(Q, speed rate))
REGION = [("PVCON", "QV"), ("PVDEP", "QV"), ("PIMLT", "QC"), ("PIDW", "QW"), ("PIHOM", "QW"), ("PIACR", "QR"),
          ("PRACI", "QC"), ("PRAUT", "QW"), ("PRACW", "QW"), ("PREVP", "QR"), ("PRACS", "QS"), ("PSACW", "QW"),
          ("PSACR", "QR"), ("PSACI", "QC"), ("PSAUT", "QC"), ("PSFW", "QW"), ("PSFI", "QC"), ("PSDEP", "QV"),
          ("PSSUB", "QS"), ("PSMLT", "QS"), ("PGAUT", "QS"), ("PGFR", "QR"), ("PGACW", "QW"), ("PGACI", "QC"),
          ("PGACR", "QR"), ("PGDRY", None), ("PGACS", "QS"), ("PGSUB", "QI"), ("PGMLT", "QI"), ("PGWET", None),
          ("PGACR1", None), ("PVVAP", "QW"), ("PVSUB", "QC")]


def _cell_region(o, x, y, z):
    mn = np.array([x, y, z], np.int32)
    o.L.sko_region(o.h, 1, _p(mn), _p(mn))
    return mn


def test_cold_branch_limited_rates_bounded_by_their_source():
    """T < 0: every process is limited to dt min(Q, speed rate) of the species
    it draws from; the sum over cells of each limited rate (region
    diagnostic) is checked cell by cell against dt x that species."""
    nx, ny, nz = 5, 6, 4
    o = _state(nx, ny, nz)
    o.set64("POTTEMP", o.get64("POTTEMP") - 40.0)
    n = o.N
    vals = {}
    for f, hi in (("QW", 3e-3), ("QC", 2e-3), ("QR", 2e-3), ("QS", 2e-3), ("QI", 2e-3)):
        v = RNG.uniform(0, hi, n)
        v[RNG.uniform(size=n) < 0.2] = 0.0
        vals[f] = v
        o.set64(f, v)
    vals["QV"] = o.get64("QV") * RNG.uniform(0.3, 1.6, n)
    # cells where one species is nearly absent, so the processes drawing on it
    # run into their limit
    for (x, y, z), scarce in (((1, 2, 1), "QW"), ((3, 4, 2), "QC"), ((0, 5, 3), "QR"), ((4, 1, 0), "QS")):
        i = x + nx * (y + ny * z)
        for f in ("QW", "QC", "QR", "QS", "QI"):
            vals[f][i] = 1e-8 if f == scarce else 1.5e-3
    for f in ("QW", "QC", "QR", "QS", "QI"):
        o.set64(f, vals[f])
    o.set64("QV", vals["QV"])
    o.init_density()
    state = {f: o.get64(f) for f in ob.FIELDS}
    dt = 0.5
    checked = 0
    for (x, y, z) in ((1, 2, 1), (3, 4, 2), (0, 5, 3), (4, 1, 0), (2, 3, 2)):
        for f, v in state.items():
            o.set64(f, v)
        _cell_region(o, x, y, z)
        o.micro_sky(dt, 1.0)
        reg = np.zeros(33)
        o.L.sko_region_get(o.h, _p(reg))
        i = x + nx * (y + ny * z)
        assert o.phase_mask()[i] & 2  # T < 0
        for k, (name, src) in enumerate(REGION):
            if src is None:
                continue
            assert reg[k] >= 0.0, name
            assert reg[k] <= dt * state[src][i] * (1 + 1e-12) + 1e-30, (name, reg[k], state[src][i])
            checked += reg[k] > 0
    assert checked > 20
    o.close()


def _one_cell_micro(T_shift, fields, qv_factor):
    """A 3 x 3 x 3 domain above flat ground with the given condensates at one
    cell and vapour set to qv_factor times a saturation value there."""
    o = _state(3, 4, 3)
    th = o.get64("POTTEMP") + T_shift
    o.set64("POTTEMP", th)
    for f in ("QW", "QC", "QR", "QS", "QI"):
        o.set64(f, np.zeros(o.N))
    o.init_density()
    i = 1 + 3 * (2 + 4 * 1)
    p = o.get64("PRESSURE")[i]
    gP = o.get64("GROUND_P")[1 + 3 * 1]
    T = th[i] * (p / gP) ** KAP - 273.15
    for f, v in fields.items():
        a = o.get64(f)
        a[i] = v
        o.set64(f, a)
    return o, i, T, p


def test_cold_condensation_only_latent_heat():
    """Below 0 C with vapour supersaturated over water and nothing else to act
    on (no condensate: ice deposition is clamped to <= 0 and the ice paths
    need cloud), only condensation happens: theta rises by
    Lwater(theta - 273.15) c / cp, with cp of the updated moist air."""
    o, i, T, p = _one_cell_micro(-35.0, {}, 1.0)
    assert T < 0
    qv0 = 1.2 * _ws(T, p)
    a = o.get64("QV")
    a[i] = qv0
    o.set64("QV", a)
    th0 = o.get64("POTTEMP")[i]
    o.micro_sky(0.5, 1.0)
    qv1, qw1, th1 = o.get64("QV")[i], o.get64("QW")[i], o.get64("POTTEMP")[i]
    c = qw1
    assert c > 0 and abs((qv0 - qv1) - c) < 1e-18
    for f in ("QC", "QR", "QS", "QI"):
        assert o.get64(f)[i] == 0.0
    Lw = ob.lib().sko_Lwater(float(th0 - 273.15))
    want = th0 + Lw / _cp_moist(qv1) * c
    assert abs(th1 - want) < 2e-7 * abs(want - th0) + 1e-12  # Lwater through the float binding
    o.close()


def test_cloud_ice_sublimation_only_latent_heat():
    """Below 0 C, cloud ice under QcMin in air subsaturated over ice: only
    PVDEP (negative) acts; the ice lost becomes vapour and theta falls by
    Lice(theta - 273.15) |d| / cp."""
    o, i, T, p = _one_cell_micro(-35.0, {"QC": 4e-4}, 1.0)
    assert T < 0
    qv0 = 0.6 * _wi(T, p)
    a = o.get64("QV")
    a[i] = qv0
    o.set64("QV", a)
    th0 = o.get64("POTTEMP")[i]
    o.micro_sky(0.5, 1.0)
    qv1, qc1, th1 = o.get64("QV")[i], o.get64("QC")[i], o.get64("POTTEMP")[i]
    d = qc1 - 4e-4
    assert d < 0 and abs((qv1 - qv0) + d) < 1e-18
    for f in ("QW", "QR", "QS", "QI"):
        assert o.get64(f)[i] == 0.0
    Li = ob.lib().sko_Lice(float(th0 - 273.15))
    want = th0 + Li / _cp_moist(qv1) * d
    assert abs(th1 - want) < 2e-7 * abs(want - th0) + 1e-12
    o.close()


LF = 3.3355e5


@pytest.mark.parametrize("case", ["PGFR", "PSAUT", "PGAUT", "PIHOM+PIDW", "PIMLT+PVDEP"])
def test_single_process_cells_bookkeeping(case):
    """Cells where one process (or a known pair) is the only one with a source
    and a nonzero rate: the species it moves between change by equal and
    opposite amounts, and theta changes by the latent heat of that phase
    change (none for autoconversion, Lf for freezing, -Lf for melting,
    Lice for deposition) over cp of the updated moist air. Vapour sits at
    saturation over water (or ice) so that the condensation and deposition
    paths are idle."""
    def run(shift, fields, sat, qv_scale=1.0):
        o, i, T, p = _one_cell_micro(shift, fields, 1.0)
        a = o.get64("QV")
        a[i] = qv_scale * sat(T, p)
        o.set64("QV", a)
        before = {f: o.get64(f)[i] for f in ("QV", "QW", "QC", "QR", "QS", "QI", "POTTEMP")}
        o.micro_sky(0.5, 1.0)
        after = {f: o.get64(f)[i] for f in before}
        o.close()
        d = {f: after[f] - before[f] for f in before}
        return T, before, after, d
    L = ob.lib()
    tol = 1e-17
    if case == "PGFR":  # rain freezing to ice (Bigg), rain above the PT3 threshold
        T, b, a, d = run(-40.0, {"QR": 2e-3}, _ws)
        assert T < 0 and d["QI"] > 0
        assert abs(d["QR"] + d["QI"]) < tol and abs(d["QW"]) + abs(d["QC"]) + abs(d["QS"]) < tol
        want = LF / _cp_moist(a["QV"]) * d["QI"]
        assert abs(d["POTTEMP"] - want) < 1e-9 * want + 1e-12
    elif case == "PSAUT":  # cloud ice autoconverting to snow: no phase change
        T, b, a, d = run(-40.0, {"QC": 3e-3}, _wi)
        assert T < 0 and d["QS"] > 0
        assert abs(d["QC"] + d["QS"]) < tol and abs(d["POTTEMP"]) < 1e-12
        assert abs(d["QV"]) < 1e-15  # at ice saturation: no deposition
    elif case == "PGAUT":  # snow autoconverting to graupel / hail
        T, b, a, d = run(-40.0, {"QS": 3e-3}, _wi)
        assert T < 0 and d["QI"] > 0
        assert abs(d["QS"] + d["QI"]) < tol and abs(d["POTTEMP"]) < 1e-12
        assert d["QI"] == pytest.approx(0.5 * 1e-3 * math.exp(0.09 * T) * (3e-3 - 6e-4), rel=1e-12)
    elif case == "PIHOM+PIDW":  # cloud water below -40 C: homogeneous freezing (all of it, limited to dt Qw)
        T, b, a, d = run(-75.0, {"QW": 6e-4}, _ws)  # plus Bergeron growth of ice from the water
        assert T < -40
        assert abs(d["QW"] + d["QC"]) < tol
        hom = 0.5 * 6e-4
        dep = d["QC"] - hom
        assert dep >= -tol
        want = (L.sko_Lice(float(b["POTTEMP"] - 273.15)) * dep + LF * hom) / _cp_moist(a["QV"])
        assert abs(d["POTTEMP"] - want) < 2e-7 * want
    else:  # cloud ice above 0 C: melts (limited to dt Qc) and sublimates (over-ice saturation above Qws)
        T, b, a, d = run(+10.0, {"QC": 8e-4}, _ws)
        assert T > 0
        melt = 0.5 * 8e-4
        assert d["QW"] == pytest.approx(melt, rel=1e-14)
        dep = -d["QV"]
        assert dep <= 0 and abs(d["QC"] - (dep - melt)) < tol
        want = (L.sko_Lice(float(b["POTTEMP"] - 273.15)) * dep - LF * melt) / _cp_moist(a["QV"])
        assert abs(d["POTTEMP"] - want) < 2e-7 * abs(want)
