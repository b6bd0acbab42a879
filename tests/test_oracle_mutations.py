"""The oracle pins catch planted slips -- CPU only.

Each case plants one plausible transcription slip (a wrong sign, index,
species, constant, exponent or boundary condition) in a copy of
oracle/skt_oracle.cpp, builds that copy, and runs the pins that should
notice it against the mutant (tests/oracle_bind.py loads the library named
by SKT_ORACLE_SO). A case passes when those pins fail. The unmutated oracle
passes the same pins in the ordinary test run.
"""
import concurrent.futures as cf
import os
import subprocess
import sys
import tempfile

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "oracle", "skt_oracle.cpp")
PHYS = "tests/test_oracle_pins_physics.py"
PINS = "tests/test_oracle_pins.py"

# (name, text in the oracle, replacement, pin file, -k selection)
MUTATIONS = [
    ("buoyancy loading sign", "B = C::g * ((VTP - VTE) / VTE) - totalQ * C::g;",
     "B = C::g * ((VTP - VTE) / VTE) + totalQ * C::g;", PHYS, "buoyancy_uniform"),
    ("buoyancy reversal cap", "change = -cur;", "change = R(0);", PHYS, "reversal_cap"),
    ("sedimentation inflow speed", "if (up) q += dts * MN((fvu / e.dx) * Qu, Qu);",
     "if (up) q += dts * MN((fv / e.dx) * Qu, Qu);", PHYS, "sedimentation"),
    ("divergence harmonic mean", "R(2) / (R(1) / cur + R(1) / n); };", "R(2) / (R(1) / cur + R(2) / n); };",
     PHYS, "divergence"),
    ("divergence back-face profile", "ub = velnb(e, w, bv, e.f[F_DEF_W][y]", "ub = velnb(e, w, bv, e.f[F_DEF_U][y]",
     PHYS, "divergence"),
    ("projection pressure bounds", "const Bounds bp{DIRICHLET, NEUMANN, NEUMANN}",
     "const Bounds bp{NEUMANN, NEUMANN, NEUMANN}", PHYS, "projection_correction"),
    ("operator diagonal drops a face", "w -= f; w -= bb;", "w -= f;", PHYS, "operator_faces"),
    ("velocity reset walk", "                if (e.ground(x, y, z)) break;\n                const long i = e.id(x, y, z);\n"
     "                if (y == e.Y - 1",
     "                if (e.ground(x, y, z)) continue;\n                const long i = e.id(x, y, z);\n"
     "                if (y == e.Y - 1", PHYS, "velocity_reset"),
    ("cloud cover counts the ground cell", "for (int y = e.GH[g] + 1; y < e.Y; y++) tot",
     "for (int y = e.GH[g]; y < e.Y; y++) tot", PHYS, "cloud_cover"),
    ("ground temperature albedo", "((R(1) - R(0.25)) * irr)", "((R(0.25)) * irr)", PHYS, "radiation_balance"),
    ("set temps pressure order", "potTemp(e.f[F_GT][g] - R(273.15), e.f[F_P][i], e.f[F_GP][g])",
     "potTemp(e.f[F_GT][g] - R(273.15), e.f[F_GP][g], e.f[F_P][i])", PHYS, "set_temps"),
    ("ground snow store sign", "e.f[F_QGS][g] += PGSFR - PGSMLT;", "e.f[F_QGS][g] += PGSFR + PGSMLT;", PHYS,
     "ground_microphysics"),
    ("ground freezing heat", "sum += C::Lf / (ms * cps) * freeze;", "sum += C::Ls / (ms * cps) * freeze;", PHYS,
     "ground_microphysics"),
    ("ground ice melt exponent", "return mr * R(4e-4) * P(slopeI, R(-2))", "return mr * R(4e-4) * P(slopeI, R(-1))",
     PHYS, "surface_rates"),
    ("ground water diffusivity", "const R DG = R(1.75e-3)", "const R DG = R(1.75e-2)", PHYS, "ground_water_flat"),
    ("precipitation onto the snow store", "e.f[F_QGS][g] += GS;", "e.f[F_QGS][g] += GI;", PHYS,
     "hitting_ground"),
    ("PIACR slope exponent", "(R(24) * R(4.19e-10) * P(sR, R(6) + R(0.8)))", "(R(24) * R(4.19e-10) * P(sR, R(6)))",
     PHYS, "power_laws and PIACR"),
    ("PSACW collects snow", "R(1) * R(3e-2) * R(152.93) * Qw * gm.gS", "R(1) * R(3e-2) * R(152.93) * Qs * gm.gS",
     PHYS, "power_laws and PSACW"),
    ("PGFR temperature sign", "(EX(R(0.66) * (-T)) - R(1))", "(EX(R(0.66) * (T)) - R(1))", PHYS,
     "closed_forms"),
    ("condensation saturation slope", "(ES * (R(17.27) * R(237.3) /", "(ES * (R(17.27) * R(273.3) /", PHYS,
     "derivatives"),
    ("condensation latent heat", "sum += Lwater(th - R(273.15)) / cpth * cond;",
     "sum += Lice(th - R(273.15)) / cpth * cond;", PHYS, "condensation_only"),
    ("sublimation latent heat", "sum += Lice(th - R(273.15)) / cpth * dep;",
     "sum += Lwater(th - R(273.15)) / cpth * dep;", PHYS, "sublimation_only"),
    ("ice growth limited by the wrong species", "PIDW = dt * MN(Qw, speed * PIDW);",
     "PIDW = dt * MN(Qc, speed * PIDW);", PHYS, "bounded_by_their_source"),
    ("graupel gain drops rain freezing", "Qi_ += PGAUT + PGFR + PGDRY", "Qi_ += PGAUT + PGDRY", PHYS,
     "bookkeeping and PGFR"),
    ("rain freezing heat", "PGFR = dt * MN(Qr, speed * PGFR); frz += PGFR;", "PGFR = dt * MN(Qr, speed * PGFR);",
     PHYS, "bookkeeping and PGFR"),
    ("snow gain drops autoconversion", "Qs_ += PSAUT + PSACI", "Qs_ += PSACI", PHYS, "bookkeeping and PSAUT"),
    ("homogeneous freezing heat", "PIHOM = dt * MN(Qw, speed * PIHOM); frz += PIHOM;",
     "PIHOM = dt * MN(Qw, speed * PIHOM); dep += PIHOM;", PHYS, "bookkeeping and PIHOM"),
    ("cloud ice melt sign", "Qw_ += PVCON + PIMLT - PRAUT", "Qw_ += PVCON - PIMLT - PRAUT", PHYS,
     "bookkeeping and PIMLT"),
    ("reset parameter walk", "                    if (e.ground(x, y, z)) break;\n"
     "                    e.f[fld][e.id(x, y, z)] = e.f[prof][y];",
     "                    if (e.ground(x, y, z)) continue;\n"
     "                    e.f[fld][e.id(x, y, z)] = e.f[prof][y];", PINS, "reset_parameter"),
]

FLAGS = ["-std=c++17", "-O1", "-march=x86-64-v3", "-ffp-contract=off", "-fPIC", "-fopenmp", "-shared", "-w"]


@pytest.fixture(scope="module")
def mutants():
    src = open(SRC).read()
    tmp = tempfile.TemporaryDirectory(prefix="sko_mut_")

    def build(k):
        name, old, new = MUTATIONS[k][:3]
        assert src.count(old) == 1, f"{name}: anchor not unique in the oracle"
        cpp = os.path.join(tmp.name, f"m{k}.cpp")
        so = os.path.join(tmp.name, f"m{k}.so")
        with open(cpp, "w") as f:
            f.write(src.replace(old, new))
        r = subprocess.run(["g++"] + FLAGS + ["-o", so, cpp], capture_output=True, text=True)
        assert r.returncode == 0, r.stderr[-2000:]
        return so
    with cf.ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 1)) as ex:
        sos = list(ex.map(build, range(len(MUTATIONS))))
    yield sos
    tmp.cleanup()


def _run_pins(so, pins, sel):
    env = dict(os.environ, SKT_ORACLE_SO=so)
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-x", "-p", "no:cacheprovider", pins, "-k", sel],
                       cwd=ROOT, env=env, capture_output=True, text=True, timeout=300)
    return r


def test_every_planted_slip_fails_its_pins(mutants):
    def check(k):
        name, _, _, pins, sel = MUTATIONS[k]
        r = _run_pins(mutants[k], pins, sel)
        # exit code 1: tests ran and failed (5 would mean nothing was selected)
        return name, r.returncode, r.stdout[-400:]
    with cf.ThreadPoolExecutor(max_workers=4) as ex:
        results = list(ex.map(check, range(len(MUTATIONS))))
    missed = [(n, rc, out) for n, rc, out in results if rc != 1]
    assert not missed, missed
