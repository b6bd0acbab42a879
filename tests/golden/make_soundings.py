"""Writes the radiosonde fixture archives of the sounding-ingestion tests.

The reference's input archives are not distributed with it, so these are
synthetic files in the two layouts its readers parse (Game/readTable.cpp):

* knmi_fixture.mwx — a zip holding SynchronizedSoundingData.xml: one header
  line, one <Row .../> element per line whose whitespace tokens 4-8 are
  Pressure, Temperature (K), Humidity (%), WindDir, WindSpeed and which carries
  an Altitude attribute further on, then a closing line
  (Game/readTable.cpp:43-116).
* sekundenwerte_aero_01303_fixture.zip — Metadaten_Sekunde_Aero_01303.txt
  (';'-separated, launch date and hour in column 1, ground temperature in
  column 8) and produkt_sec_aero_<first>_<last>_01303.txt (';'-separated,
  date and hour in column 1; altitude, pressure, temperature in columns 6-8,
  dew point, wind speed, wind direction in columns 10-12), with 170 launches
  so that the reference's launch index 161 exists
  (Game/readTable.cpp:134-340). Launch 161's profile is long enough
  (more than 100000 bytes) for the reader's 100000-byte probing to land
  inside it.

Every value is written with at most two decimals. The rows are deterministic
functions of their index; knmi_rows() and dwd_launches() return them as the
strings written, so a test can rebuild the expected tables without the files.

Usage: python tests/golden/make_soundings.py   (rewrites the two archives)
"""
import math
import os
import zipfile

HERE = os.path.dirname(os.path.abspath(__file__))
KNMI_PATH = os.path.join(HERE, "knmi_fixture.mwx")
DWD_PATH = os.path.join(HERE, "sekundenwerte_aero_01303_fixture.zip")
DWD_STATION = "01303"
DWD_LAUNCHES = 170
DWD_TARGET = 161
KNMI_ROWS = 800


def _profile(h, phase):
    """Temperature (C), dew point (C), RH (%), wind speed, wind direction and
    pressure (hPa) at height h (m) of a conditionally unstable sounding."""
    t = 24.0 - 0.0068 * h if h < 11500 else 24.0 - 0.0068 * 11500
    t += 0.6 * math.sin(h / 730.0 + phase)
    rh = max(5.0, min(100.0, 85.0 - 0.004 * h + 10.0 * math.sin(h / 1900.0 + phase)))
    dew = t - (100.0 - rh) * 0.2
    ws = 3.0 + 0.0022 * h + 2.0 * math.sin(h / 1300.0 + 2 * phase)
    wd = (225.0 + 0.006 * h + 40.0 * math.sin(h / 2500.0 + phase)) % 360.0
    p = 1013.25 * (1.0 - 2.25577e-5 * h) ** 5.25588
    return t, dew, rh, ws, wd, p


def knmi_rows():
    """Rows of the KNMI fixture: (pressure, temperature K, humidity, wind dir,
    wind speed, altitude) as written."""
    rows = []
    for k in range(KNMI_ROWS):
        h = 2.0 + 20.0 * k
        t, _, rh, ws, wd, p = _profile(h, 0.3)
        rows.append((f"{p:.2f}", f"{t + 273.15:.2f}", f"{rh:.1f}", f"{wd:.1f}", f"{ws:.2f}", f"{h:.1f}"))
    return rows


def knmi_text():
    lines = ['<SynchronizedSoundingData Source="fixture" Station="06260">']
    for k, (p, tk, rh, wd, ws, alt) in enumerate(knmi_rows()):
        lines.append(f'<Row DataSrvTime="2025-08-07T11:{31 + k // 60:02d}:{k % 60:02d}" SondeTime="{k}" '
                     f'RadioRxTimePk="{k + 1}" Pressure="{p}" Temperature="{tk}" Humidity="{rh}" '
                     f'WindDir="{wd}" WindSpeed="{ws}" Height="{alt}" Altitude="{alt}" Dropping="0" />')
    lines.append("</SynchronizedSoundingData>")
    return "\n".join(lines) + "\n"


def dwd_dates():
    """Launch dates (YYYYMMDD) of the metadata: one a day from 2024-10-14."""
    import datetime
    d0 = datetime.date(2024, 10, 14)
    return [(d0 + datetime.timedelta(days=i)).strftime("%Y%m%d") for i in range(DWD_LAUNCHES)]


def dwd_launches():
    """{date: [(altitude, pressure, temperature, dew point, wind speed, wind dir)]} as written:
    1400 samples for the launch at DWD_TARGET, 8 for every other launch."""
    out = {}
    for i, date in enumerate(dwd_dates()):
        n = 1400 if i == DWD_TARGET else 8
        rows = []
        for k in range(n):
            h = 51.0 + 11.5 * k
            t, dew, _, ws, wd, p = _profile(h - 51.0, 0.05 * i)
            rows.append((f"{h:.1f}", f"{p:.2f}", f"{t:.2f}", f"{dew:.2f}", f"{ws:.2f}", f"{wd:.1f}"))
        out[date] = rows
    return out


def dwd_metadata_text():
    lines = ["Stations_ID;Datum_Start;Sondentyp;Gewicht;Ballontyp;Fuellung;Abbruch;Hoehe_Abbruch;Temp_Boden;eor"]
    for i, date in enumerate(dwd_dates()):
        lines.append(f"{DWD_STATION};{date}12;RS41;{350 + i % 3};TX;H2;platzer;{30000 + 17 * i};"
                     f"{8.0 + 0.05 * i:.2f};eor")
    return "\n".join(lines) + "\n"


def dwd_product_name():
    d = dwd_dates()
    return f"produkt_sec_aero_{d[0]}_{d[-1]}_{DWD_STATION}.txt"


def dwd_product_text():
    lines = ["STATIONS_ID;MESS_DATUM;QN;MESS_DATUM_SEK;GEO_BREITE;GEO_LAENGE;HOEHE;LUFTDRUCK;TEMPERATUR;"
             "REL_FEUCHTE;TAUPUNKT;WINDGESCHW;WINDRICHTUNG;eor"]
    for date, rows in dwd_launches().items():
        for k, (alt, p, t, dew, ws, wd) in enumerate(rows):
            lines.append(f"{DWD_STATION};{date}12;3;{k};52.2153;14.1197;{alt};{p};{t};77;{dew};{ws};{wd};eor")
    return "\n".join(lines) + "\n"


def _write_zip(path, members):
    with zipfile.ZipFile(path, "w", compression=zipfile.ZIP_DEFLATED, compresslevel=9) as z:
        for name, text in members:
            info = zipfile.ZipInfo(name, date_time=(2025, 1, 1, 0, 0, 0))
            info.compress_type = zipfile.ZIP_DEFLATED
            z.writestr(info, text.encode())


def main():
    _write_zip(KNMI_PATH, [("SynchronizedSoundingData.xml", knmi_text()), ("Info.txt", "fixture\n")])
    _write_zip(DWD_PATH, [(f"Metadaten_Sekunde_Aero_{DWD_STATION}.txt", dwd_metadata_text()),
                          (dwd_product_name(), dwd_product_text())])
    for p in (KNMI_PATH, DWD_PATH):
        print(p, os.path.getsize(p))


if __name__ == "__main__":
    main()
