// Internal interface of the sounding ingestion (skt_ingest.cpp).
#pragma once
#include <vector>

#include "../../include/skt.h"

namespace skt {

// One value per vertical level of the initial state readTable::initEnvironment
// builds (Game/readTable.cpp:364-474); every cell of a level takes that value.
struct SoundingLevels {
    std::vector<float> theta;   // potential temperature, K
    std::vector<float> velX;    // MAC face wind, m/s
    std::vector<float> velZ;
    std::vector<float> qv;      // vapour mixing ratio
    std::vector<float> p;       // pressure per cell of the level, hPa
    std::vector<float> pSmall;  // pressure profile by level, hPa
    float groundT = 0.0f;       // K
    float groundP = 0.0f;       // hPa
};

// Levels of an ny-level grid with spacing dx; SKT_OK or an skt_status with the
// text in skt_input_last_error().
int soundingLevels(const skt_sounding_table* t, int ny, float dx, SoundingLevels& out);

}  // namespace skt
