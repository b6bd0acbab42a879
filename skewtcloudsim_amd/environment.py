"""Python host mirror of the reference's environmentGPU step API
(Game/environment.cuh:39-98) over the C ABI in include/skt.h.

Method names, argument meaning and call order follow the reference; errors
raise SktError instead of breaking into a debugger. Arrays are numpy float32
in the reference's layout (x fastest, then y vertical, then z).
"""
import ctypes

import numpy as np

from . import _lib
from ._lib import FIELD, SOUNDING_COLUMNS, Bounds, Config, Params, SoundingSpec, SoundingTable, Stats


class SktError(RuntimeError):
    pass


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def default_config(nx, ny, nz, dx=100.0, **overrides):
    lib = _lib.load_library()
    cfg = Config()
    lib.skt_config_default(ctypes.byref(cfg), nx, ny, nz, dx)
    for k, v in overrides.items():
        if isinstance(v, (list, tuple)):
            arr = getattr(cfg, k)
            if len(v) != len(arr):
                raise SktError(f"config {k}: expected {len(arr)} values")
            for i, x in enumerate(v):
                arr[i] = x
        else:
            setattr(cfg, k, v)
    return cfg


def terrain(nx, nz, seed=100, octaves=12, max_height=0.2, device=0):
    """Terrain noise and ground heights of environmentGPU::init
    (skt_terrain_generate: PNoise2D, then round(noise * max_height)).
    Returns (noise float32[nz*nx], heights int32[nz*nx]), column x + nx*z."""
    lib = _lib.load_library()
    noise = np.empty(nx * nz, np.float32)
    heights = np.empty(nx * nz, np.int32)
    rc = lib.skt_terrain_generate(device, seed, nx, nz, octaves, max_height, _ptr(noise), _ptr(heights))
    if rc:
        raise SktError(f"skt_terrain_generate failed ({rc})")
    return noise, heights


def make_sounding(nx, ny, nz, dx=100.0, seed=None, n_bubbles=0):
    """Init arrays of the synthetic sounding (skt_sounding_generate).

    seed=None / n_bubbles<=0: the fixed single-bubble case; otherwise the
    randomised generator drawn from seed with n_bubbles bubbles.
    Returns a dict with the seven init() arguments.
    """
    lib = _lib.load_library()
    spec = SoundingSpec()
    rc = lib.skt_sounding_spec_random(ctypes.byref(spec), 0 if seed is None else int(seed), nx, ny, nz, dx,
                                      0 if seed is None else int(n_bubbles))
    if rc:
        raise SktError(f"skt_sounding_spec_random failed ({rc})")
    n, g = nx * ny * nz, nx * nz
    out = dict(potTemps=np.empty(n, np.float32), velField=np.empty(3 * n, np.float32), Qv=np.empty(n, np.float32),
               groundTemp=np.empty(g, np.float32), groundPres=np.empty(g, np.float32),
               pressures=np.empty(n, np.float32), smallPressure=np.empty(ny, np.float32))
    rc = lib.skt_sounding_generate(ctypes.byref(spec), nx, ny, nz, dx, *[_ptr(out[k]) for k in (
        "potTemps", "velField", "Qv", "groundTemp", "groundPres", "pressures", "smallPressure")])
    if rc:
        raise SktError(f"skt_sounding_generate failed ({rc})")
    return out


class _Sounding:
    """Column arrays of one radiosonde profile (radioSondeData, Game/readTable.h:8-47)."""

    def __init__(self, rows=0):
        for n in SOUNDING_COLUMNS:
            setattr(self, n, np.zeros(rows, np.float32))

    @property
    def dataSize(self):
        return len(self.altitude)

    def table(self):
        """An skt_sounding_table over these arrays (kept alive by self)."""
        cols = [np.ascontiguousarray(getattr(self, n), np.float32) for n in SOUNDING_COLUMNS]
        rows = len(cols[0])
        if any(len(c) != rows for c in cols):
            raise SktError("sounding columns differ in length")
        for n, c in zip(SOUNDING_COLUMNS, cols):
            setattr(self, n, c)
        t = SoundingTable(rows, rows, *[c.ctypes.data_as(ctypes.POINTER(ctypes.c_float)) for c in cols])
        return t


def _read_table(call):
    """Runs a reader taking an skt_sounding_table, growing the arrays once if needed."""
    lib = _lib.load_library()
    cap = 8192
    for _ in range(2):
        d = _Sounding(cap)
        t = SoundingTable(0, cap, *[getattr(d, n).ctypes.data_as(ctypes.POINTER(ctypes.c_float))
                                     for n in SOUNDING_COLUMNS])
        rc = call(lib, t)
        if rc == _lib.SKT_ERR_ARG and t.rows > cap:
            cap = t.rows
            continue
        if rc:
            raise SktError(f"sounding input: status {rc}: {lib.skt_input_last_error().decode()}")
        for n in SOUNDING_COLUMNS:
            setattr(d, n, getattr(d, n)[:t.rows].copy())
        return d
    raise SktError("sounding input: table did not fit after growing")


class ReadTable:
    """Host mirror of readTable (Game/readTable.h:49-88): reads a radiosonde
    profile into .skewTData.data and builds the initial state from it.
    Readers raise SktError where the reference would print an error and
    return (or crash)."""

    def __init__(self):
        self.data = _Sounding(0)

    @property
    def skewTData(self):
        return self

    def readKNMIFile(self, path):
        """readKNMIFile (Game/readTable.cpp:28-129): a KNMI .mwx archive."""
        self.data = _read_table(lambda lib, t: lib.skt_sounding_read_knmi(str(path).encode(), ctypes.byref(t)))

    def readDWDFile(self, path, date_index=_lib.DWD_REFERENCE_DATE_INDEX):
        """readDWDFile (Game/readTable.cpp:134-354): a DWD sekundenwerte_aero
        zip; the reference reads launch index 161 (the default here)."""
        self.data = _read_table(
            lambda lib, t: lib.skt_sounding_read_dwd(str(path).encode(), int(date_index), ctypes.byref(t)))

    def parseKNMI(self, text):
        """The KNMI row parser on the text of SynchronizedSoundingData.xml."""
        b = text.encode() if isinstance(text, str) else bytes(text)
        self.data = _read_table(lambda lib, t: lib.skt_sounding_parse_knmi(b, len(b), ctypes.byref(t)))

    def parseDWD(self, metadata, product, date_index=_lib.DWD_REFERENCE_DATE_INDEX):
        """The DWD metadata and product parsers on the member texts."""
        m = metadata.encode() if isinstance(metadata, str) else bytes(metadata)
        p = product.encode() if isinstance(product, str) else bytes(product)
        self.data = _read_table(
            lambda lib, t: lib.skt_sounding_parse_dwd(m, len(m), p, len(p), int(date_index), ctypes.byref(t)))

    def set_data(self, **cols):
        """Sets the profile from column arrays (altitude, pressure, dew_point,
        temperature, wind_speed, wind_dir)."""
        d = _Sounding(0)
        for n in SOUNDING_COLUMNS:
            setattr(d, n, np.ascontiguousarray(cols[n], np.float32))
        d.table()  # length check
        self.data = d

    def init_fields(self, nx, ny, nz, dx=100.0):
        """initEnvironment's init() arrays on the host (skt_sounding_init_fields)."""
        lib = _lib.load_library()
        n, g = nx * ny * nz, nx * nz
        out = dict(potTemps=np.empty(n, np.float32), velField=np.empty(3 * n, np.float32),
                   Qv=np.empty(n, np.float32), groundTemp=np.empty(g, np.float32),
                   groundPres=np.empty(g, np.float32), pressures=np.empty(n, np.float32),
                   smallPressure=np.empty(ny, np.float32))
        t = self.data.table()
        rc = lib.skt_sounding_init_fields(ctypes.byref(t), nx, ny, nz, dx, *[_ptr(out[k]) for k in (
            "potTemps", "velField", "Qv", "groundTemp", "groundPres", "pressures", "smallPressure")])
        if rc:
            raise SktError(f"skt_sounding_init_fields: status {rc}: {lib.skt_input_last_error().decode()}")
        return out

    def initEnvironment(self, env):
        """initEnvironment (Game/readTable.cpp:364-474) into env: per-level
        values on the host, broadcast over the grid on the device."""
        t = self.data.table()
        env._chk(env.lib.skt_env_init_sounding(env.h, ctypes.byref(t)), "initEnvironment")


class EnvironmentGPU:
    """One simulation domain resident on one GPU (environmentGPU)."""

    def __init__(self, nx, ny, nz, dx=100.0, **config):
        self.lib = _lib.load_library()
        self.cfg = default_config(nx, ny, nz, dx, **config)
        self.nx, self.ny, self.nz, self.dx = nx, ny, nz, dx
        self.N, self.G = nx * ny * nz, nx * nz
        h = ctypes.c_void_p()
        rc = self.lib.skt_env_create(ctypes.byref(self.cfg), ctypes.byref(h))
        if rc:
            raise SktError(f"skt_env_create failed ({rc}) for {nx}x{ny}x{nz}")
        self.h = h

    def _chk(self, rc, what):
        if rc:
            msg = self.lib.skt_env_last_error(self.h)
            raise SktError(f"{what}: status {rc}: {msg.decode() if msg else ''}")

    def close(self):
        if self.h:
            self.lib.skt_env_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- environmentGPU API ----
    def init(self, potTemps, velField, Qv, groundTemp, groundPres, pressures, smallPressure):
        arrs = [np.ascontiguousarray(a, np.float32) for a in
                (potTemps, velField, Qv, groundTemp, groundPres, pressures, smallPressure)]
        self._chk(self.lib.skt_env_init(self.h, *[_ptr(a) for a in arrs]), "init")

    def set_ground_height(self, gh):
        gh = np.ascontiguousarray(gh, np.int32)
        self._chk(self.lib.skt_env_set_ground_height(self.h, _ptr(gh)), "set_ground_height")

    def updateGPU(self, dt, speed=1.0):
        self._chk(self.lib.skt_env_update(self.h, dt, speed), "updateGPU")

    def microPhysicsGroundGPU(self, dt, speed, irradiance):
        self._chk(self.lib.skt_env_micro_physics_ground(self.h, dt, speed, irradiance), "microPhysicsGroundGPU")

    def microPhysicsSkyGPU(self, dt, speed):
        self._chk(self.lib.skt_env_micro_physics_sky(self.h, dt, speed), "microPhysicsSkyGPU")

    def diffuseGPU(self, field, type_, dt):
        self._chk(self.lib.skt_env_diffuse(self.h, FIELD[field], type_, dt), "diffuseGPU")

    def advectGroundWater(self, dt, speed):
        self._chk(self.lib.skt_env_advect_ground_water(self.h, dt, speed), "advectGroundWater")

    def setTempsAtGround(self, dt, speed):
        self._chk(self.lib.skt_env_set_temps_at_ground(self.h, dt, speed), "setTempsAtGround")

    def advectPPMWGPU(self, field, profile, bounds, dt):
        prof = -1 if profile is None else FIELD[profile]
        self._chk(self.lib.skt_env_advect_ppm(self.h, FIELD[field], prof, Bounds(*bounds), dt), "advectPPMWGPU")

    def advectPrecip(self, field, fall_type, dt):
        self._chk(self.lib.skt_env_advect_precip(self.h, FIELD[field], fall_type, dt), "advectPrecip")

    def pressureProject(self, dt):
        self._chk(self.lib.skt_env_pressure_project(self.h, dt), "pressureProject")

    def calculatePressureProject(self, dt):
        self._chk(self.lib.skt_env_calculate_pressure_project(self.h, dt), "calculatePressureProject")

    def editorDataGPU(self):
        self._chk(self.lib.skt_env_editor_data(self.h), "editorDataGPU")

    def irridianceGPU(self):
        v = ctypes.c_float()
        self._chk(self.lib.skt_env_irradiance(self.h, ctypes.byref(v)), "irridianceGPU")
        return v.value

    def groundCoverageFactor(self):
        self._chk(self.lib.skt_env_ground_coverage_factor(self.h), "groundCoverageFactor")

    def updateGroundTemps(self, dt, speed, irradiance):
        self._chk(self.lib.skt_env_update_ground_temps(self.h, dt, speed, irradiance), "updateGroundTemps")

    def calculateBuoyancy(self, dt):
        self._chk(self.lib.skt_env_calculate_buoyancy(self.h, dt), "calculateBuoyancy")

    def initDensity(self):
        self._chk(self.lib.skt_env_init_density(self.h), "initDensity")

    def calculateNewPressure(self):
        self._chk(self.lib.skt_env_calculate_new_pressure(self.h), "calculateNewPressure")

    def getParamArray(self, parameter, direction="RIGHT"):
        p = ctypes.c_void_p()
        self._chk(self.lib.skt_env_get_param_array(self.h, _lib.PARAM[parameter], _lib.DIRECTION[direction],
                                                   ctypes.byref(p)), "getParamArray")
        return p.value

    # ---- state access ----
    def field_len(self, field):
        i = FIELD[field]
        return self.N if i <= FIELD["PRESSURE_PROJ"] else (self.G if i <= FIELD["CLOUD_COVER"] else self.ny)

    def get(self, field):
        a = np.empty(self.field_len(field), np.float32)
        self._chk(self.lib.skt_env_copy_field_to_host(self.h, FIELD[field], _ptr(a)), f"get {field}")
        return a

    def set(self, field, values):
        a = np.ascontiguousarray(values, np.float32).ravel()
        if a.size != self.field_len(field):
            raise SktError(f"set {field}: expected {self.field_len(field)} values, got {a.size}")
        self._chk(self.lib.skt_env_copy_field_from_host(self.h, FIELD[field], _ptr(a)), f"set {field}")

    FORMATS = {0: "f32", 1: "bf16", 2: "f16"}

    def field_format(self, field):
        """(storage, scale) of a field: ("f32", 1.0), or for a hydrometeor of
        the mixed form ("bf16" | "f16", 2^k) -- the device holds value * scale."""
        f, sc = ctypes.c_int32(), ctypes.c_float()
        self._chk(self.lib.skt_env_get_field_format(self.h, FIELD[field], ctypes.byref(f), ctypes.byref(sc)),
                  "field_format")
        return self.FORMATS[f.value], sc.value

    def device_ptr(self, field):
        p, n = ctypes.c_void_p(), ctypes.c_int64()
        self._chk(self.lib.skt_env_get_field_ptr(self.h, FIELD[field], ctypes.byref(p), ctypes.byref(n)), "ptr")
        return p.value, n.value

    def params(self):
        p = Params()
        self._chk(self.lib.skt_env_get_params(self.h, ctypes.byref(p)), "get_params")
        return p

    def set_params(self, **kw):
        p = self.params()
        for k, v in kw.items():
            setattr(p, k, v)
        self._chk(self.lib.skt_env_set_params(self.h, ctypes.byref(p)), "set_params")

    def stats(self):
        s = Stats()
        self._chk(self.lib.skt_env_get_stats(self.h, ctypes.byref(s)), "stats")
        return s

    _PARAMS = {"POTTEMP": 0, "QV": 1, "QW": 2, "QC": 3, "QR": 4, "QS": 5, "QI": 6, "WIND": 7, "PGROUND": 8,
               "PRESSURE": 9}

    def prepareBrushGPU(self, param, brushSize, mousePos, brushSmoothness, dt, brushIntensity, applyValue,
                        valueDir=(0.0, 0.0, 0.0), groundErase=False):
        """environmentGPU::prepareBrushGPU (skt_env_apply_brush)."""
        p = self._PARAMS.get(param, param)
        pos = np.asarray(mousePos, np.int32)
        d = np.asarray(valueDir, np.float32)
        self._chk(self.lib.skt_env_apply_brush(self.h, p, brushSize, _ptr(pos), brushSmoothness, dt, brushIntensity,
                                               applyValue, _ptr(d), int(bool(groundErase))), "prepareBrushGPU")

    def prepareSelectionGPU(self, param, minPos, maxPos, applyValue, valueDir=(0.0, 0.0, 0.0), groundErase=False):
        """environmentGPU::prepareSelectionGPU (skt_env_apply_selection)."""
        p = self._PARAMS.get(param, param)
        mn, mx = np.asarray(minPos, np.int32), np.asarray(maxPos, np.int32)
        d = np.asarray(valueDir, np.float32)
        self._chk(self.lib.skt_env_apply_selection(self.h, p, _ptr(mn), _ptr(mx), applyValue, _ptr(d),
                                                   int(bool(groundErase))), "prepareSelectionGPU")

    def resetParameterGPU(self, param):
        """environmentGPU::resetParameterGPU (skt_env_reset_parameter)."""
        p = self._PARAMS.get(param, param)
        self._chk(self.lib.skt_env_reset_parameter(self.h, p), "resetParameterGPU")

    def ground_height(self):
        """The ground heights in use, int32[nz*nx] (column x + nx*z)."""
        gh = np.empty(self.G, np.int32)
        self._chk(self.lib.skt_env_get_ground_height(self.h, _ptr(gh)), "ground_height")
        return gh

    def diagnostics(self, cloud_threshold=1e-5):
        """Maximum updraft, cloud-top level and height, total ground rain
        (skt_env_diagnostics), as a dict."""
        d = _lib.Diagnostics()
        self._chk(self.lib.skt_env_diagnostics(self.h, cloud_threshold, ctypes.byref(d)), "diagnostics")
        return {"max_updraft": d.max_updraft, "cloud_top_level": d.cloud_top_level,
                "cloud_top_height": d.cloud_top_height, "surface_rain": d.surface_rain}

    def masks(self):
        m = np.empty(12 * self.N, np.uint8)
        self._chk(self.lib.skt_env_compute_masks(self.h, _ptr(m)), "masks")
        return m

    def phase_mask(self):
        m = np.empty(self.N, np.uint8)
        self._chk(self.lib.skt_env_phase_mask(self.h, _ptr(m)), "phase_mask")
        return m

    def region_set(self, mn, mx):
        a, b = np.asarray(mn, np.int32), np.asarray(mx, np.int32)
        self._chk(self.lib.skt_mphys_region_set(self.h, _ptr(a), _ptr(b)), "region_set")

    def region_clear(self):
        self._chk(self.lib.skt_mphys_region_clear(self.h), "region_clear")

    def region_get(self):
        out = np.empty(33, np.float32)
        self._chk(self.lib.skt_mphys_region_get(self.h, _ptr(out)), "region_get")
        return out

    def set_timing(self, enable=True, classes=None, stride=1):
        """Per-class kernel timing with HIP events; classes: names from
        TIMING_CLASSES to time (default all), stride: time every stride-th
        launch of a class; enable=False switches it off."""
        mask = 0
        if enable:
            mask = 0x1FF if classes is None else sum(1 << _lib.TIMING_CLASSES.index(c) for c in classes)
        self._chk(self.lib.skt_env_set_timing_stride(self.h, int(stride)), "set_timing_stride")
        self._chk(self.lib.skt_env_set_timing(self.h, mask), "set_timing")

    def timing(self):
        t = _lib.Timing()
        self._chk(self.lib.skt_env_get_timing(self.h, ctypes.byref(t)), "timing")
        return {c: (int(t.launches[k]), float(t.total_ms[k])) for k, c in enumerate(_lib.TIMING_CLASSES)}

    def copy_probe(self, repeats=20):
        ms, nb = ctypes.c_float(), ctypes.c_double()
        self._chk(self.lib.skt_env_copy_probe(self.h, repeats, ctypes.byref(ms), ctypes.byref(nb)), "copy_probe")
        return ms.value, nb.value

    def device_synchronize(self):
        self._chk(self.lib.skt_device_synchronize(self.cfg.device), "device_synchronize")

    def synchronize(self):
        self._chk(self.lib.skt_env_synchronize(self.h), "synchronize")


class EnsembleGPU:
    """An ensemble of EnvironmentGPU members advanced by one launch per phase
    (skt_batch_*). The members are borrowed: keep them alive while they are
    in the ensemble. Each member's results equal its own updateGPU's."""

    def __init__(self, device=0, use_graph=True, advection_substep_cap=4, launch_mode=0):
        self.lib = _lib.load_library()
        h = ctypes.c_void_p()
        rc = self.lib.skt_batch_create(device, ctypes.byref(h))
        if rc:
            raise SktError(f"skt_batch_create failed ({rc})")
        self.h = h
        self.members = {}
        self.set_options(use_graph=use_graph, advection_substep_cap=advection_substep_cap, launch_mode=launch_mode)

    def _chk(self, rc, what):
        if rc:
            msg = self.lib.skt_batch_last_error(self.h)
            raise SktError(f"{what}: status {rc}: {msg.decode() if msg else ''}")

    def set_options(self, use_graph=True, advection_substep_cap=4, launch_mode=0):
        o = _lib.BatchOptions(int(bool(use_graph)), int(advection_substep_cap), int(launch_mode))
        self._chk(self.lib.skt_batch_set_options(self.h, ctypes.byref(o)), "set_options")

    def add(self, env):
        mid = ctypes.c_int32()
        self._chk(self.lib.skt_batch_add(self.h, env.h, ctypes.byref(mid)), "add")
        self.members[mid.value] = env
        return mid.value

    def remove(self, member_id):
        self._chk(self.lib.skt_batch_remove(self.h, int(member_id)), "remove")
        self.members.pop(member_id, None)

    def __len__(self):
        n = ctypes.c_int32()
        self._chk(self.lib.skt_batch_size(self.h, ctypes.byref(n)), "size")
        return n.value

    def step(self, dt, speed=1.0):
        self._chk(self.lib.skt_batch_step(self.h, dt, speed), "step")

    def synchronize(self):
        self._chk(self.lib.skt_batch_synchronize(self.h), "synchronize")

    def autotune(self, dt, speed=1.0):
        t = _lib.BatchTuning()
        self._chk(self.lib.skt_batch_autotune(self.h, dt, speed, ctypes.byref(t)), "autotune")
        return {"step_ms_by_width": list(t.step_ms), "tuned_step_ms": t.tuned_step_ms, "bodies": t.bodies,
                "persistent_bodies": t.persistent_bodies}

    def profile(self, dt, speed=1.0):
        """One ordinary ensemble step timed launch by launch: {body: (ms, launches)}."""
        buf = ctypes.create_string_buffer(1 << 16)
        self._chk(self.lib.skt_batch_profile(self.h, dt, speed, buf, len(buf)), "profile")
        out = {}
        for line in buf.value.decode().splitlines():
            name, ms, n = line.split("\t")
            out[name] = (float(ms), int(n))
        return out

    def close(self):
        if self.h:
            self.lib.skt_batch_destroy(self.h)
            self.h = None
            self.members = {}

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
