"""Sounding ingestion (readTable, Game/readTable.cpp) — CPU tests.

The readers and skt_sounding_init_fields are host code in libskt.so, so these
run without a GPU. Three kinds of check:
* fixtures (tests/golden/make_soundings.py): the committed archives are read
  and compared with the rows the generator wrote, rebuilt here from its
  strings (by construction, not from either implementation);
* the oracle's readers are the reference's control flow on std::stringstream,
  so the product's stream model (getline/seekg/tellg state rules) is compared
  with it on fixed edge cases and on randomised product files whose launches
  straddle the 100000-byte probe stride;
* initEnvironment's arithmetic is pinned by closed forms: a level at a
  sounding row's height takes that row's values, a level between rows takes
  the reference's interpolation of the previous and current level pressures,
  and wind directions map to the documented axes.
"""
import ctypes
import os
import random
import sys
import zipfile

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden"))
import make_soundings as gen  # noqa: E402
import oracle_bind as ob  # noqa: E402
import skewtcloudsim_amd as skt  # noqa: E402

F32 = np.float32


def f32(s):
    return F32(float(s))


def cols_of(rt):
    return {n: getattr(rt.data, n) for n in ob.SOUNDING_COLUMNS}


def product_knmi(text):
    r = skt.ReadTable()
    r.parseKNMI(text)
    return cols_of(r)


def product_dwd(meta, prod, idx):
    r = skt.ReadTable()
    try:
        r.parseDWD(meta, prod, idx)
    except skt.SktError:
        return None
    return cols_of(r)


def assert_same(a, b):
    assert (a is None) == (b is None), (a is None, b is None)
    if a is None:
        return
    for n in ob.SOUNDING_COLUMNS:
        np.testing.assert_array_equal(a[n], b[n], err_msg=n)


# ------------------------------------------------------------------ fixtures ---

def test_fixture_archives_match_generator():
    with zipfile.ZipFile(gen.KNMI_PATH) as z:
        assert z.read("SynchronizedSoundingData.xml").decode() == gen.knmi_text()
    with zipfile.ZipFile(gen.DWD_PATH) as z:
        assert z.read(f"Metadaten_Sekunde_Aero_{gen.DWD_STATION}.txt").decode() == gen.dwd_metadata_text()
        assert z.read(gen.dwd_product_name()).decode() == gen.dwd_product_text()


def knmi_expected(rows):
    p = np.array([f32(r[0]) for r in rows], F32)
    T = np.array([f32(r[1]) - F32(273.15) for r in rows], F32)
    rh = np.array([f32(r[2]) for r in rows], F32)
    dew = (T - ((F32(100) - rh) * F32(0.2)) * ((T + F32(273.15)) * F32(0.01))).astype(F32)
    return dict(altitude=np.array([f32(r[5]) for r in rows], F32), pressure=p, dew_point=dew, temperature=T,
                wind_speed=np.array([f32(r[4]) for r in rows], F32), wind_dir=np.array([f32(r[3]) for r in rows], F32))


def test_knmi_fixture_by_construction():
    r = skt.ReadTable()
    r.readKNMIFile(gen.KNMI_PATH)
    got = cols_of(r)
    assert r.data.dataSize == gen.KNMI_ROWS
    assert_same(got, knmi_expected(gen.knmi_rows()))
    assert_same(got, ob.sounding_knmi(gen.knmi_text()))


def dwd_expected(rows):
    c = list(zip(*rows))
    return dict(altitude=np.array([f32(v) for v in c[0]], F32), pressure=np.array([f32(v) for v in c[1]], F32),
                temperature=np.array([f32(v) for v in c[2]], F32), dew_point=np.array([f32(v) for v in c[3]], F32),
                wind_speed=np.array([f32(v) for v in c[4]], F32), wind_dir=np.array([f32(v) for v in c[5]], F32))


def test_dwd_fixture_by_construction():
    r = skt.ReadTable()
    r.readDWDFile(gen.DWD_PATH)  # the reference's launch index 161
    got = cols_of(r)
    want = gen.dwd_launches()[gen.dwd_dates()[gen.DWD_TARGET]]
    assert r.data.dataSize == len(want) == 1400
    assert_same(got, dwd_expected(want))
    assert_same(got, ob.sounding_dwd(gen.dwd_metadata_text(), gen.dwd_product_text(), gen.DWD_TARGET))


def test_dwd_first_launch_loses_two_rows():
    # The first probe reads the first data row; it matches launch 0, the
    # 100000-byte step back fails at the file start, and the reader resumes
    # after the probe row and the row it then skips: rows 2.. are returned.
    meta, prod = gen.dwd_metadata_text(), gen.dwd_product_text()
    want = gen.dwd_launches()[gen.dwd_dates()[0]][2:]
    got = product_dwd(meta, prod, 0)
    assert_same(got, dwd_expected(want))
    assert_same(got, ob.sounding_dwd(meta, prod, 0))


@pytest.mark.parametrize("idx", [1, 5, 100, 160, 162, 169])
def test_dwd_other_launches_match_stream_semantics(idx):
    meta, prod = gen.dwd_metadata_text(), gen.dwd_product_text()
    assert_same(product_dwd(meta, prod, idx), ob.sounding_dwd(meta, prod, idx))


def test_dwd_archive_errors(tmp_path):
    lib = skt.load_library()
    with pytest.raises(skt.SktError, match="date index"):
        skt.ReadTable().readDWDFile(gen.DWD_PATH, 170)
    with pytest.raises(skt.SktError, match="cannot open"):
        skt.ReadTable().readDWDFile(str(tmp_path / "sekundenwerte_aero_01303_x.zip"))
    # station number taken from the file name: a different station has no metadata member
    bad = tmp_path / "sekundenwerte_aero_09999_fixture.zip"
    bad.write_bytes(open(gen.DWD_PATH, "rb").read())
    with pytest.raises(skt.SktError, match="Metadaten_Sekunde_Aero_09999"):
        skt.ReadTable().readDWDFile(str(bad))
    notzip = tmp_path / "knmi.mwx"
    notzip.write_bytes(b"not a zip archive at all" * 3)
    with pytest.raises(skt.SktError, match="zip"):
        skt.ReadTable().readKNMIFile(str(notzip))
    # stored (uncompressed) members read as well as deflated ones
    stored = tmp_path / "stored.mwx"
    with zipfile.ZipFile(stored, "w", compression=zipfile.ZIP_STORED) as z:
        z.writestr("SynchronizedSoundingData.xml", gen.knmi_text())
    r = skt.ReadTable()
    r.readKNMIFile(str(stored))
    assert r.data.dataSize == gen.KNMI_ROWS
    assert lib.skt_input_last_error() is not None
    # a deflated member whose directory entry claims more than deflate can
    # expand to is rejected before anything is allocated for it
    raw = bytearray(open(gen.KNMI_PATH, "rb").read())
    cd = raw.find(b"PK\x01\x02")
    assert cd > 0 and raw[cd + 46:cd + 46 + 28] == b"SynchronizedSoundingData.xml"
    raw[cd + 24:cd + 28] = (0xFFFFFF00).to_bytes(4, "little")
    huge = tmp_path / "huge.mwx"
    huge.write_bytes(bytes(raw))
    with pytest.raises(skt.SktError, match="corrupt zip member size"):
        skt.ReadTable().readKNMIFile(str(huge))


def test_table_capacity_reports_rows():
    lib = skt.load_library()
    cap = 10
    arrs = [np.zeros(cap, F32) for _ in range(6)]
    t = skt._lib.SoundingTable(0, cap, *[a.ctypes.data_as(ctypes.POINTER(ctypes.c_float)) for a in arrs])
    rc = lib.skt_sounding_read_knmi(gen.KNMI_PATH.encode(), ctypes.byref(t))
    assert rc == skt._lib.SKT_ERR_ARG and t.rows == gen.KNMI_ROWS


# -------------------------------------------------------------------- KNMI ---

def knmi_line(k, p="1000.5", tk="290.15", rh="80.0", wd="270.0", ws="5.5", alt=None):
    alt = alt if alt is not None else f"{10.0 * k:.1f}"
    return (f'<Row DataSrvTime="t{k}" SondeTime="{k}" RadioRxTimePk="1" Pressure="{p}" Temperature="{tk}" '
            f'Humidity="{rh}" WindDir="{wd}" WindSpeed="{ws}" Altitude="{alt}" Dropping="0" />')


def test_knmi_quirks():
    rows = [knmi_line(k, rh="100.0") for k in range(5)]
    full = "head\n" + "\n".join(rows) + "\n</end>\n"
    a = product_knmi(full)
    # RH 100 %: the dew point equals the temperature exactly
    np.testing.assert_array_equal(a["dew_point"], a["temperature"])
    assert len(a["altitude"]) == 5
    assert_same(a, ob.sounding_knmi(full))
    # without the closing line the last data row is dropped (row count - 1)
    cut = "head\n" + "\n".join(rows) + "\n"
    b = product_knmi(cut)
    assert len(b["altitude"]) == 4
    assert_same(b, ob.sounding_knmi(cut))
    # CRLF line ends and tabs are whitespace for the tokenizer
    crlf = full.replace("\n", "\r\n").replace(" ", "\t")
    assert_same(product_knmi(crlf), a)
    assert_same(ob.sounding_knmi(crlf), a)


@pytest.mark.parametrize("bad", [
    knmi_line(1, p="x"),            # no number in Pressure
    knmi_line(1, tk=""),            # token shorter than its key
    '<Row a="1" b="2" c="3" Pressure="1000"',  # row without temperature/humidity columns
    knmi_line(1).replace('Altitude="10.0"', 'Altitude'),
])
def test_knmi_malformed_rows_fail(bad):
    text = "head\n" + knmi_line(0) + "\n" + bad + "\n" + knmi_line(2) + "\n</end>\n"
    assert ob.sounding_knmi(text) is None
    with pytest.raises(skt.SktError):
        product_knmi(text)


# --------------------------------------------------------------------- DWD ---

def dwd_meta(dates):
    return "head\n" + "".join(f"1;{d}12;a;b;c;d;e;f;{7.5};eor\n" for d in dates)


def dwd_product(launches, pad=0, rng=None, trailing_newline=True):
    lines = ["STATIONS_ID;MESS_DATUM;QN"]
    for d, n in launches:
        for k in range(n):
            extra = "x" * (rng.randint(0, pad) if rng else pad)
            lines.append(f"1;{d}12;{extra};{k};0;0;{10.0 + k};{1000 - 0.1 * k:.2f};{15 - 0.01 * k:.2f};50;"
                         f"{5 - 0.01 * k:.2f};{3 + 0.01 * k:.2f};{(200 + k) % 360};eor")
    return "\n".join(lines) + ("\n" if trailing_newline else "")


def test_dwd_quirks():
    dates = [f"202401{d:02d}" for d in range(1, 9)]
    meta = dwd_meta(dates)
    # long launches: launch 3 is found and read whole
    prod = dwd_product([(d, 1500) for d in dates], pad=8)
    got = product_dwd(meta, prod, 3)
    assert got is not None and len(got["altitude"]) == 1500
    assert_same(got, ob.sounding_dwd(meta, prod, 3))
    # last launch at the end of a file without its final newline: the reader fails
    prod2 = dwd_product([(d, 1500) for d in dates], pad=8, trailing_newline=False)
    assert product_dwd(meta, prod2, 7) is None and ob.sounding_dwd(meta, prod2, 7) is None
    # a short file: the first probe misses and the 100000-byte step runs past the end
    short = dwd_product([(d, 20) for d in dates])
    assert product_dwd(meta, short, 4) is None and ob.sounding_dwd(meta, short, 4) is None
    # ... while the first launch is read without its first two rows
    got = product_dwd(meta, short, 0)
    assert got is not None and len(got["altitude"]) == 18
    np.testing.assert_array_equal(got["altitude"], np.arange(12, 30, dtype=F32))
    # a date absent from the product file
    meta2 = dwd_meta(dates[:3] + ["20240920"])
    assert product_dwd(meta2, prod, 3) is None and ob.sounding_dwd(meta2, prod, 3) is None
    # a non-numeric value in a parsed column
    broken = prod.replace("1;2024010412;", "1;2024010412;", 1)
    i = broken.index("1;2024010412;")
    j = broken.index("\n", i)
    fields = broken[i:j].split(";")
    fields[7] = "n/a"
    broken = broken[:i] + ";".join(fields) + broken[j:]
    assert product_dwd(meta, broken, 3) is None and ob.sounding_dwd(meta, broken, 3) is None


def test_dwd_stream_walk_randomised():
    """The product's cursor model against the std::stringstream walk on
    product files whose launch sizes straddle the probe stride."""
    rng = random.Random(20261203)
    cases = 0
    for trial in range(60):
        nd = rng.randint(2, 9)
        dates = [f"2023{m:02d}{d:02d}" for m, d in sorted(rng.sample([(m, d) for m in range(1, 13)
                                                                      for d in range(1, 29)], nd))]
        launches = [(d, rng.choice([1, 2, 3, 40, 300, 900, 1100, 1300, 2500])) for d in dates]
        prod = dwd_product(launches, pad=rng.choice([0, 10, 40]), rng=rng,
                           trailing_newline=rng.random() < 0.85)
        if rng.random() < 0.2:
            prod = prod.replace("\n", "\r\n")
        meta = dwd_meta(dates)
        for idx in range(nd):
            a, b = product_dwd(meta, prod, idx), ob.sounding_dwd(meta, prod, idx)
            assert_same(a, b)
            cases += 1
    assert cases > 100


# ---------------------------------------------------------- initEnvironment ---

def table(**cols):
    r = skt.ReadTable()
    r.set_data(**cols)
    return r


SIMPLE = dict(altitude=[0, 1000, 2000, 4000], pressure=[1000, 900, 800, 600], dew_point=[10, 5, 0, -10],
              temperature=[20, 13, 6, -8], wind_speed=[5, 10, 15, 20], wind_dir=[270, 270, 250, 200])


def levels(out, nx, ny):
    return {k: out[k].reshape(-1, ny, nx)[0, :, 0] for k in ("potTemps", "Qv", "pressures")}


def test_init_fields_match_oracle_on_fixtures():
    for path, reader in ((gen.KNMI_PATH, "readKNMIFile"), (gen.DWD_PATH, "readDWDFile")):
        r = skt.ReadTable()
        getattr(r, reader)(path)
        for (nx, ny, nz, dx) in ((8, 24, 5, 100.0), (16, 64, 4, 64.0), (5, 40, 7, 250.0)):
            got = r.init_fields(nx, ny, nz, dx)
            want = ob.sounding_init(cols_of(r), nx, ny, nz, dx)
            for k in got:
                np.testing.assert_array_equal(got[k], want[k], err_msg=f"{path} {k} {nx}x{ny}x{nz}")


def test_init_fields_closed_forms():
    nx, ny, nz, dx = 3, 21, 2, 100.0
    out = table(**SIMPLE).init_fields(nx, ny, nz, dx)
    lv = levels(out, nx, ny)
    pS = out["smallPressure"]
    # pressure profile: linear between rows (level j at height 100 j)
    np.testing.assert_allclose(pS[:11], 1000 - 10 * np.arange(11), rtol=0, atol=1e-4)
    np.testing.assert_allclose(pS[10:21], 900 - 10 * np.arange(11), rtol=0, atol=1e-4)
    # level pressure: previous and current level values with the row heights as interval
    assert lv["pressures"][0] == pS[0]
    t1 = F32(100) / F32(1000)
    assert lv["pressures"][1] == F32(pS[0] + t1 * (pS[1] - pS[0]))  # 999, not 990
    np.testing.assert_allclose(lv["pressures"][1], 999.0, atol=1e-4)
    # a level at a row's height takes that row's values
    for j, row in ((0, 0), (10, 1), (20, 2)):
        assert pS[j] == F32(SIMPLE["pressure"][row])
        T, P, P0 = SIMPLE["temperature"][row], SIMPLE["pressure"][row], SIMPLE["pressure"][0]
        theta = (T + 273.15) * (P0 / P) ** (287.0528 / 1003.5)
        np.testing.assert_allclose(lv["potTemps"][j], theta, rtol=2e-7 * 4)
    # ground: first row's temperature in K and its pressure
    assert np.all(out["groundTemp"] == F32(20) + F32(273.15)) and np.all(out["groundPres"] == F32(1000))
    # wind: from 270 deg (west) blows toward +x; uniform per level; no vertical component
    vel = out["velField"].reshape(nz, ny, nx, 3)
    np.testing.assert_allclose(vel[:, 0, :, 0], 5.0, rtol=1e-6)
    np.testing.assert_allclose(vel[:, 0, :, 2], 0.0, atol=1e-6)
    assert np.all(vel[..., 1] == 0)
    # levels 1..10 take the wind of row 1 (the first row at or above them, not interpolated)
    np.testing.assert_allclose(vel[:, 1:11, :, 0], 10.0, rtol=1e-6)
    # levels 11..20 take row 2 (from 250 deg, west-south-west): toward +x (sin 70 deg) and +z (cos 70 deg)
    a = np.deg2rad(70.0)
    np.testing.assert_allclose(vel[:, 11:21, :, 0], 15 * np.sin(a), rtol=1e-5)
    np.testing.assert_allclose(vel[:, 11:21, :, 2], 15 * np.cos(a), rtol=1e-5)


def test_init_fields_wind_axes():
    for wd, ex, ez in ((180, 0.0, 1.0), (90, -1.0, 0.0), (0, 0.0, -1.0), (270, 1.0, 0.0)):
        c = dict(SIMPLE, wind_dir=[wd] * 4, wind_speed=[4] * 4)
        vel = table(**c).init_fields(2, 5, 2, 100.0)["velField"].reshape(-1, 3)
        np.testing.assert_allclose(vel[:, 0], 4 * ex, atol=2e-6)
        np.testing.assert_allclose(vel[:, 2], 4 * ez, atol=2e-6)


def test_init_fields_vapour_matches_reference_thermo():
    ref = ob.ref_lib()
    out = table(**SIMPLE).init_fields(2, 21, 2, 100.0)
    lv = levels(out, 2, 21)
    for j, row in ((0, 0), (10, 1), (20, 2)):
        if ref is not None:
            # the reference's own ws and potentialTemp (Engine/source/math/meteoformulas.cpp)
            assert lv["Qv"][j] == F32(ref.ref_ws(SIMPLE["dew_point"][row], lv["pressures"][j]))
            th = ref.ref_potentialTemp(SIMPLE["temperature"][row], SIMPLE["pressure"][row], SIMPLE["pressure"][0])
            assert lv["potTemps"][j] == F32(F32(th) + F32(273.15))
        es = 0.61078 * np.exp(17.27 * SIMPLE["dew_point"][row] / (SIMPLE["dew_point"][row] + 237.3)) * 10
        ws = (287.0528 / 461.5) * es / (lv["pressures"][j] - es)
        np.testing.assert_allclose(lv["Qv"][j], ws, rtol=2e-6)


def test_init_fields_errors():
    with pytest.raises(skt.SktError, match="ends below"):
        table(**SIMPLE).init_fields(2, 50, 2, 100.0)  # 4900 m > 4000 m
    ok = table(**SIMPLE).init_fields(2, 41, 2, 100.0)  # exactly 4000 m
    assert np.isfinite(ok["potTemps"]).all()
    assert ob.sounding_init(SIMPLE, 2, 50, 2, 100.0) is None
    # a spacing whose float accumulation stays below ny*dx after ny steps (33.3 summed 10 times < 333):
    # the reference would write an eleventh level past its arrays
    with pytest.raises(skt.SktError, match="accumulates"):
        table(**SIMPLE).init_fields(2, 10, 2, 33.3)
    assert ob.sounding_init(SIMPLE, 2, 10, 2, 33.3) is None
