"""ctypes binding of the oracle (oracle/liboracle.so) — test infrastructure.

Only tests, __graft_entry__.smoke() and bench.py's CPU baseline load this.
"""
import ctypes
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
# SKT_ORACLE_SO points the tests at another build of the oracle (the mutation
# tests use it to run the pins against oracles with planted slips).
ORACLE_SO = os.environ.get("SKT_ORACLE_SO") or os.path.join(ROOT, "oracle", "liboracle.so")
REF_SO = os.path.join(ROOT, "oracle", "_ref", "libref_meteo.so")

FIELDS = [
    "QV", "QW", "QC", "QR", "QS", "QI", "POTTEMP", "VELX", "VELY", "VELZ", "PRESSURE",
    "DENSITY", "DENSITY_OLD", "PRESSURE_PROJ",
    "QRS", "QGR", "QGS", "QGI", "GROUND_P", "GROUND_TIME", "GROUND_T", "CLOUD_COVER",
    "ISEN_TEMP", "ISEN_VAPOR", "DEFAULT_VELX", "DEFAULT_VELZ", "DEFAULT_PRESSURE",
]
FIELD = {n: i for i, n in enumerate(FIELDS)}


class Opts(ctypes.Structure):
    _fields_ = [("pcg_mode", ctypes.c_int), ("pcg_tol", ctypes.c_float), ("pcg_max_iter", ctypes.c_int),
                ("sweeps", ctypes.c_int), ("max_substeps", ctypes.c_int), ("maxabs_mode", ctypes.c_int),
                ("noise_seed", ctypes.c_uint), ("init_noise", ctypes.c_int), ("threads", ctypes.c_int)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            raise RuntimeError(f"{ORACLE_SO} missing: run 'make -C oracle'")
        L = ctypes.CDLL(ORACLE_SO)
        L.sko_create.restype = ctypes.c_void_p
        L.sko_create.argtypes = [ctypes.c_int] * 3 + [ctypes.c_float, ctypes.c_int, ctypes.POINTER(Opts)]
        f, vp, i = ctypes.c_float, ctypes.c_void_p, ctypes.c_int
        sig = {
            "sko_destroy": [vp], "sko_set_ground_height": [vp, vp], "sko_init": [vp] * 8,
            "sko_step": [vp, f, f], "sko_get": [vp, i, vp], "sko_set": [vp, i, vp], "sko_get64": [vp, i, vp],
            "sko_set64": [vp, i, vp], "sko_set_params": [vp, f, i, f, f, i, i], "sko_get_params": [vp, vp, vp],
            "sko_diffuse": [vp, i, i, f], "sko_advect_ppm": [vp, i, i, i, i, i, f],
            "sko_ppm_pass": [vp, i, i, i, i, i, i, i, i, f], "sko_advect_precip": [vp, i, i, f],
            "sko_buoyancy": [vp, f], "sko_init_density": [vp], "sko_new_pressure": [vp],
            "sko_pressure_project": [vp, f], "sko_calc_pressure_project": [vp, f], "sko_divergence": [vp, f, vp],
            "sko_build_A": [vp] * 6, "sko_apply_A": [vp, vp, vp], "sko_reset_vel": [vp],
            "sko_micro_sky": [vp, f, f], "sko_micro_ground": [vp, f, f, f], "sko_ground_water": [vp, f, f],
            "sko_cloud_cover": [vp], "sko_ground_temps": [vp, f, f, f], "sko_set_temps_at_ground": [vp, f, f],
            "sko_editor_data": [vp], "sko_masks": [vp, vp], "sko_phase": [vp, vp],
            "sko_stats": [vp, vp, vp, vp, vp, vp], "sko_region": [vp, i, vp, vp], "sko_region_get": [vp, vp],
            "sko_set_threads": [i], "sko_fallvel": [f, f, f, f, i, vp], "sko_gammas": [vp],
            "sko_diagnostics": [vp, f, vp, vp, vp], "sko_terrain": [i, i, i, i, f, vp, vp],
            "sko_apply_brush": [vp, i, f, vp, f, f, f, f, vp, i], "sko_apply_selection": [vp, i, vp, vp, f, vp, i],
            "sko_get_ground_height": [vp, vp], "sko_reset_parameter": [vp, i],
        }
        for name, args in sig.items():
            getattr(L, name).argtypes = args
            getattr(L, name).restype = None
        L.sko_set_threads.restype = ctypes.c_int
        L.sko_irradiance.restype = f
        L.sko_irradiance.argtypes = [vp]
        for name, n in [("sko_es", 1), ("sko_ei", 1), ("sko_ws", 2), ("sko_wi", 2), ("sko_qs", 2), ("sko_Ni", 1),
                        ("sko_pwater", 1), ("sko_Lwater", 1), ("sko_Lice", 1), ("sko_DQVair", 2),
                        ("sko_ViscAir", 1), ("sko_MLR", 2), ("sko_cvd", 3), ("sko_potTemp", 3)]:
            getattr(L, name).restype = f
            getattr(L, name).argtypes = [f] * n
        L.sko_msvc_rand_first.restype = i
        L.sko_msvc_rand_first.argtypes = [i]
        L.sko_perlin.restype = f
        L.sko_perlin.argtypes = [i, f, f, i]
        L.sko_slopePrecip.restype = f
        L.sko_slopePrecip.argtypes = [f, f, i]
        L.sko_potTemp64.restype = ctypes.c_double
        L.sko_potTemp64.argtypes = [ctypes.c_double] * 3
        for name in ("sko_mp_rates64", "sko_ground_rates64"):
            getattr(L, name).restype = None
            getattr(L, name).argtypes = [vp, vp]
        L.sko_divergence64.restype = None
        L.sko_divergence64.argtypes = [vp, ctypes.c_double, vp]
        L.sko_apply_projection.restype = None
        L.sko_apply_projection.argtypes = [vp, vp]
        L.sko_hash.restype = ctypes.c_uint32
        L.sko_hash.argtypes = [ctypes.c_uint32]
        L.sko_ppm_flux.restype = f
        L.sko_ppm_flux.argtypes = [f] * 6
        L.sko_ppm_flux64.restype = ctypes.c_double
        L.sko_ppm_flux64.argtypes = [ctypes.c_double] * 6
        cp, lg = ctypes.c_char_p, ctypes.c_long
        L.sko_sounding_knmi.restype = i
        L.sko_sounding_knmi.argtypes = [cp, lg, i, vp, vp]
        L.sko_sounding_dwd.restype = i
        L.sko_sounding_dwd.argtypes = [cp, lg, cp, lg, i, i, vp, vp]
        L.sko_sounding_init.restype = i
        L.sko_sounding_init.argtypes = [i, vp, i, i, i, f] + [vp] * 7
        _lib = L
    return _lib


SOUNDING_COLUMNS = ["altitude", "pressure", "dew_point", "temperature", "wind_speed", "wind_dir"]


def _sounding_out(call, cap=8192):
    rows = np.zeros(1, np.int32)
    out = np.zeros((6, cap), np.float32)
    rc = call(cap, _p(rows), _p(out))
    if rc == -2:
        return _sounding_out(call, int(rows[0]))
    if rc:
        return None
    return {n: out[k, :rows[0]].copy() for k, n in enumerate(SOUNDING_COLUMNS)}


def sounding_knmi(text):
    """readKNMIFile's parser on the XML text: dict of columns, or None where the reference fails."""
    b = text.encode() if isinstance(text, str) else bytes(text)
    return _sounding_out(lambda cap, r, o: lib().sko_sounding_knmi(b, len(b), cap, r, o))


def sounding_dwd(meta, product, date_index=161):
    """readDWDFile's parsers on the member texts: dict of columns, or None where the reference fails."""
    m = meta.encode() if isinstance(meta, str) else bytes(meta)
    p = product.encode() if isinstance(product, str) else bytes(product)
    return _sounding_out(lambda cap, r, o: lib().sko_sounding_dwd(m, len(m), p, len(p), date_index, cap, r, o))


def sounding_init(cols, nx, ny, nz, dx):
    """initEnvironment's init() arrays from sounding columns, or None where the reference fails."""
    rows = len(cols["altitude"])
    c = np.ascontiguousarray(np.stack([np.asarray(cols[n], np.float32) for n in SOUNDING_COLUMNS]))
    n, g = nx * ny * nz, nx * nz
    out = dict(potTemps=np.zeros(n, np.float32), velField=np.zeros(3 * n, np.float32), Qv=np.zeros(n, np.float32),
               groundTemp=np.zeros(g, np.float32), groundPres=np.zeros(g, np.float32),
               pressures=np.zeros(n, np.float32), smallPressure=np.zeros(ny, np.float32))
    rc = lib().sko_sounding_init(rows, _p(c), nx, ny, nz, dx, *[_p(out[k]) for k in (
        "potTemps", "velField", "Qv", "groundTemp", "groundPres", "pressures", "smallPressure")])
    return None if rc else out


def ref_lib():
    """The reference's own CPU thermodynamics, if oracle/_ref was built here."""
    if not os.path.exists(REF_SO):
        return None
    L = ctypes.CDLL(REF_SO)
    f = ctypes.c_float
    for name, n in [("ref_es", 1), ("ref_ei", 1), ("ref_ws", 2), ("ref_wi", 2), ("ref_qs", 2), ("ref_Ni", 1),
                    ("ref_pwater", 1), ("ref_Lwater", 1), ("ref_Lice", 1), ("ref_DQVair", 2), ("ref_ViscAir", 1),
                    ("ref_MLR", 2), ("ref_cvd", 3), ("ref_potentialTemp", 3), ("ref_gamma", 1)]:
        getattr(L, name).restype = f
        getattr(L, name).argtypes = [f] * n
    L.ref_slopePrecip.restype = f
    L.ref_slopePrecip.argtypes = [f, f, ctypes.c_int]
    return L


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


class Oracle:
    """One CPU domain of the restated step (FP32 or FP64)."""

    def __init__(self, nx, ny, nz, dx=100.0, fp64=False, pcg_mode=0, pcg_tol=1e-5, pcg_max_iter=200, sweeps=20,
                 max_substeps=100, maxabs_mode=0, noise_seed=100, init_noise=1, threads=0):
        self.L = lib()
        self.nx, self.ny, self.nz, self.dx = nx, ny, nz, dx
        self.N, self.G = nx * ny * nz, nx * nz
        o = Opts(pcg_mode, pcg_tol, pcg_max_iter, sweeps, max_substeps, maxabs_mode, noise_seed, init_noise, threads)
        self.h = ctypes.c_void_p(self.L.sko_create(nx, ny, nz, dx, 1 if fp64 else 0, ctypes.byref(o)))
        if threads:
            self.L.sko_set_threads(threads)

    def close(self):
        if self.h:
            self.L.sko_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def field_len(self, field):
        i = FIELD[field]
        return self.N if i <= FIELD["PRESSURE_PROJ"] else (self.G if i <= FIELD["CLOUD_COVER"] else self.ny)

    def init(self, potTemps, velField, Qv, groundTemp, groundPres, pressures, smallPressure):
        arrs = [np.ascontiguousarray(a, np.float32) for a in
                (potTemps, velField, Qv, groundTemp, groundPres, pressures, smallPressure)]
        self.L.sko_init(self.h, *[_p(a) for a in arrs])

    def set_ground_height(self, gh):
        gh = np.ascontiguousarray(gh, np.int32)
        self.L.sko_set_ground_height(self.h, _p(gh))

    def step(self, dt, speed=1.0):
        self.L.sko_step(self.h, dt, speed)

    def get(self, field):
        a = np.empty(self.field_len(field), np.float32)
        self.L.sko_get(self.h, FIELD[field], _p(a))
        return a

    def get64(self, field):
        a = np.empty(self.field_len(field), np.float64)
        self.L.sko_get64(self.h, FIELD[field], _p(a))
        return a

    def set(self, field, v):
        a = np.ascontiguousarray(v, np.float32).ravel()
        assert a.size == self.field_len(field)
        self.L.sko_set(self.h, FIELD[field], _p(a))

    def set64(self, field, v):
        a = np.ascontiguousarray(v, np.float64).ravel()
        assert a.size == self.field_len(field)
        self.L.sko_set64(self.h, FIELD[field], _p(a))

    def set_params(self, time=43200.0, day=130, lon=52.37, sun=1.0, pause=0, ground_changed=0):
        self.L.sko_set_params(self.h, time, day, lon, sun, pause, ground_changed)

    def params(self):
        t, g = ctypes.c_float(), ctypes.c_int()
        self.L.sko_get_params(self.h, ctypes.byref(t), ctypes.byref(g))
        return t.value, g.value

    def stats(self):
        it, res, r0, nc = ctypes.c_int(), ctypes.c_float(), ctypes.c_float(), ctypes.c_int()
        ns = (ctypes.c_int * 16)()
        self.L.sko_stats(self.h, ctypes.byref(it), ctypes.byref(res), ctypes.byref(r0), ns, ctypes.byref(nc))
        return dict(pcg_iterations=it.value, pcg_residual=res.value, pcg_residual0=r0.value,
                    substeps=list(ns)[:min(nc.value, 16)])

    def masks(self):
        m = np.empty(12 * self.N, np.uint8)
        self.L.sko_masks(self.h, _p(m))
        return m

    def phase_mask(self):
        m = np.empty(self.N, np.uint8)
        self.L.sko_phase(self.h, _p(m))
        return m

    def irradiance(self):
        return self.L.sko_irradiance(self.h)

    def ground_height(self):
        gh = np.empty(self.nx * self.nz, np.int32)
        self.L.sko_get_ground_height(self.h, _p(gh))
        return gh

    def brush(self, param, size, pos, smooth, dt, intensity, value, direction=(0.0, 0.0, 0.0), erase=False):
        p, d = np.asarray(pos, np.int32), np.asarray(direction, np.float32)
        self.L.sko_apply_brush(self.h, param, size, _p(p), smooth, dt, intensity, value, _p(d), int(bool(erase)))

    def selection(self, param, mn, mx, value, direction=(0.0, 0.0, 0.0), erase=False):
        a, b, d = np.asarray(mn, np.int32), np.asarray(mx, np.int32), np.asarray(direction, np.float32)
        self.L.sko_apply_selection(self.h, param, _p(a), _p(b), value, _p(d), int(bool(erase)))

    def diagnostics(self, thr=1e-5):
        w, lev, rain = ctypes.c_float(), ctypes.c_int(), ctypes.c_double()
        self.L.sko_diagnostics(self.h, ctypes.c_float(thr), ctypes.byref(w), ctypes.byref(lev), ctypes.byref(rain))
        return {"max_updraft": w.value, "cloud_top_level": lev.value, "surface_rain": rain.value}

    def __getattr__(self, name):
        # per-phase entry points: oracle.diffuse(...), oracle.micro_sky(...), ...
        fn = getattr(self.__dict__.get("L") or lib(), "sko_" + name)
        return lambda *a: fn(self.h, *a)
