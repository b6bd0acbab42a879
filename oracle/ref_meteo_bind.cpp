// ORACLE — TEST INFRASTRUCTURE ONLY.
// C entry points over the reference's own CPU thermodynamics
// (Engine/source/math/meteoformulas.cpp, class meteoformulas in
// Engine/include/math/meteoformulas.h), compiled from the reference sources
// where they lie by oracle/Makefile into oracle/_ref/libref_meteo.so. Tests use
// it to pin the oracle's thermo functions; nothing else loads it.
#include "math/meteoformulas.h"

extern "C" {
float ref_es(float T) { return meteoformulas::es(T); }
float ref_ei(float T) { return meteoformulas::ei(T); }
float ref_ws(float T, float P) { return meteoformulas::ws(T, P); }
float ref_wi(float T, float P) { return meteoformulas::wi(T, P); }
float ref_qs(float T, float P) { return meteoformulas::qs(T, P); }
float ref_Ni(float T) { return meteoformulas::Ni(T); }
float ref_pwater(float T) { return meteoformulas::pwater(T); }
float ref_Lwater(float T) { return meteoformulas::Lwater(T); }
float ref_Lice(float T) { return meteoformulas::Lice(T); }
float ref_DQVair(float T, float P) { return meteoformulas::DQVair(T, P); }
float ref_ViscAir(float T) { return meteoformulas::ViscAir(T); }
float ref_slopePrecip(float D, float Q, int t) { return meteoformulas::slopePrecip(D, Q, t); }
float ref_MLR(float T, float P) { return meteoformulas::MLR(T, P); }
float ref_cvd(float T, float P, float d) { return meteoformulas::cvd(T, P, d); }
float ref_potentialTemp(float T, float Pk, float Pt) { return meteoformulas::potentialTemp(T, Pk, Pt); }
float ref_gamma(float x) { return meteoformulas::gamma(x, 0.0f); }
}
