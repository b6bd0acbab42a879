// Sounding ingestion: the radiosonde readers and the sounding -> initial-state
// conversion of readTable (Game/readTable.cpp:28-129, :134-354, :356-474).
//
// Host code by design: a sounding is a few thousand rows and the conversion
// produces one value per vertical level; the per-cell broadcast of those levels
// runs on the device (skt_env_init_sounding, k_sounding_fill).
//
// Reader contracts, restated from the reference:
//  * KNMI .mwx archive, member SynchronizedSoundingData.xml: whitespace tokens
//    of each line after the first; token 4 Pressure="hPa", 5 Temperature="K",
//    6 Humidity="%" (dew point from the reference's linear approximation),
//    7 WindDir="deg", 8 WindSpeed="m/s", and any other token that starts with
//    "Altitude" (Altitude="m"). The row count is the number of lines after the
//    first minus one (the closing tag).
//  * DWD zip: station number = third '_'-separated token of the archive's file
//    name; member Metadaten_Sekunde_Aero_<station>.txt lists the launch dates
//    (';' column 1 without its two-digit hour); the profile of date index k is
//    read from produkt_sec_aero_<first>_<last>_<station>.txt, rows whose column 1
//    equals that date followed by "12": altitude col 6, pressure 7,
//    temperature 8, dew point 10, wind speed 11, wind direction 12. The row
//    search reproduces the reference's stream walk (probe one line every
//    100000 bytes, step back 100000 bytes on a hit), including the rows it
//    drops and the inputs it rejects; Cursor models std::istream's getline,
//    seekg and tellg state rules for that walk.
#include <zlib.h>

#include <cerrno>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/skt.h"
#include "skt_ingest.h"

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
    g_err = msg;
    return code;
}

// ------------------------------------------------------------ zip reader ---
// The reference links miniz (Game/zip_file.hpp) for zip_file::has_file/read.
// This reader handles what the radiosonde archives hold: a single-disk archive
// whose members are stored (method 0) or deflated (method 8), without ZIP64.

uint32_t rd16(const unsigned char* p) { return (uint32_t)p[0] | ((uint32_t)p[1] << 8); }
uint32_t rd32(const unsigned char* p) { return rd16(p) | (rd16(p + 2) << 16); }

int readFile(const char* path, std::string& out) {
    FILE* f = std::fopen(path, "rb");
    if (!f) return fail(SKT_ERR_INPUT, std::string("cannot open ") + path);
    out.clear();
    char buf[1 << 16];
    size_t n;
    while ((n = std::fread(buf, 1, sizeof buf, f)) > 0) out.append(buf, n);
    const bool err = std::ferror(f) != 0;
    std::fclose(f);
    if (err) return fail(SKT_ERR_INPUT, std::string("read error on ") + path);
    return SKT_OK;
}

struct ZipMember {
    std::string name;
    uint32_t method, crc, csize, usize, local;
};

int zipDirectory(const std::string& z, std::vector<ZipMember>& dir) {
    const auto* b = reinterpret_cast<const unsigned char*>(z.data());
    const size_t n = z.size();
    if (n < 22) return fail(SKT_ERR_INPUT, "not a zip archive (too short)");
    // End-of-central-directory record: last 22 bytes plus up to 65535 bytes of comment.
    size_t eocd = SIZE_MAX;
    const size_t lo = n - 22 > 65535 ? n - 22 - 65535 : 0;
    for (size_t p = n - 22 + 1; p-- > lo;)
        if (rd32(b + p) == 0x06054b50u) {
            eocd = p;
            break;
        }
    if (eocd == SIZE_MAX) return fail(SKT_ERR_INPUT, "not a zip archive (no end record)");
    const uint32_t count = rd16(b + eocd + 10), cdSize = rd32(b + eocd + 12), cdOff = rd32(b + eocd + 16);
    if (count == 0xFFFF || cdOff == 0xFFFFFFFFu) return fail(SKT_ERR_INPUT, "ZIP64 archives are not supported");
    if ((uint64_t)cdOff + cdSize > eocd) return fail(SKT_ERR_INPUT, "corrupt zip central directory");
    size_t p = cdOff;
    dir.clear();
    for (uint32_t k = 0; k < count; k++) {
        if (p + 46 > eocd || rd32(b + p) != 0x02014b50u) return fail(SKT_ERR_INPUT, "corrupt zip directory entry");
        ZipMember m;
        m.method = rd16(b + p + 10);
        m.crc = rd32(b + p + 16);
        m.csize = rd32(b + p + 20);
        m.usize = rd32(b + p + 24);
        const uint32_t nl = rd16(b + p + 28), el = rd16(b + p + 30), cl = rd16(b + p + 32);
        m.local = rd32(b + p + 42);
        if (p + 46 + nl > eocd) return fail(SKT_ERR_INPUT, "corrupt zip directory entry");
        m.name.assign(reinterpret_cast<const char*>(b + p + 46), nl);
        dir.push_back(m);
        p += 46 + (size_t)nl + el + cl;
    }
    return SKT_OK;
}

int zipExtract(const std::string& z, const ZipMember& m, std::string& out) {
    const auto* b = reinterpret_cast<const unsigned char*>(z.data());
    const size_t n = z.size();
    if ((uint64_t)m.local + 30 > n || rd32(b + m.local) != 0x04034b50u)
        return fail(SKT_ERR_INPUT, "corrupt zip local header for " + m.name);
    const size_t data = (size_t)m.local + 30 + rd16(b + m.local + 26) + rd16(b + m.local + 28);
    if ((uint64_t)data + m.csize > n) return fail(SKT_ERR_INPUT, "truncated zip member " + m.name);
    // deflate expands at most about 1032:1, so a larger stated size is corrupt
    // (and would only allocate memory that inflate can never fill)
    if (m.method == 8 && (uint64_t)m.usize > (uint64_t)m.csize * 1032u + 1024u)
        return fail(SKT_ERR_INPUT, "corrupt zip member size for " + m.name);
    out.assign(m.usize, '\0');
    if (m.method == 0) {
        if (m.csize != m.usize) return fail(SKT_ERR_INPUT, "corrupt stored zip member " + m.name);
        if (m.usize) std::memcpy(&out[0], b + data, m.usize);
    } else if (m.method == 8) {
        z_stream s;
        std::memset(&s, 0, sizeof s);
        if (inflateInit2(&s, -MAX_WBITS) != Z_OK) return fail(SKT_ERR_INPUT, "zlib init failed");
        s.next_in = const_cast<Bytef*>(b + data);
        s.avail_in = m.csize;
        s.next_out = reinterpret_cast<Bytef*>(out.empty() ? nullptr : &out[0]);
        s.avail_out = m.usize;
        const int r = inflate(&s, Z_FINISH);
        const uLong got = s.total_out;
        inflateEnd(&s);
        if (r != Z_STREAM_END || got != m.usize) return fail(SKT_ERR_INPUT, "cannot inflate zip member " + m.name);
    } else {
        return fail(SKT_ERR_INPUT, "unsupported zip compression method for " + m.name);
    }
    const uLong crc = crc32(crc32(0L, Z_NULL, 0), reinterpret_cast<const Bytef*>(out.data()), (uInt)out.size());
    if ((uint32_t)crc != m.crc) return fail(SKT_ERR_INPUT, "CRC mismatch in zip member " + m.name);
    return SKT_OK;
}

// Reads member 'name' (exact match, as zip_file::has_file) into out.
int zipMember(const std::string& z, const std::string& name, std::string& out) {
    std::vector<ZipMember> dir;
    if (int rc = zipDirectory(z, dir)) return rc;
    for (const ZipMember& m : dir)
        if (m.name == name) return zipExtract(z, m, out);
    return fail(SKT_ERR_INPUT, "archive has no member " + name);
}

// ---------------------------------------------------------- text helpers ---

// std::stof: leading blanks skipped, at least one character converted,
// trailing characters ignored, out-of-range results rejected.
bool parseFloat(const std::string& s, float& v) {
    const char* c = s.c_str();
    char* end = nullptr;
    errno = 0;
    v = std::strtof(c, &end);
    return end != c && errno != ERANGE;
}

// Fields of 'line' split at 'delim' as repeated std::getline(ss, word, delim)
// yields them: an empty line has none, a trailing delimiter adds none.
void splitFields(const std::string& line, char delim, std::vector<std::string>& out) {
    out.clear();
    size_t p = 0;
    while (p < line.size()) {
        size_t q = line.find(delim, p);
        if (q == std::string::npos) q = line.size();
        out.emplace_back(line, p, q - p);
        p = q + 1;
    }
}

// Whitespace tokens (operator>> on std::string in the "C" locale).
void splitWords(const std::string& line, std::vector<std::string>& out) {
    out.clear();
    size_t p = 0;
    auto blank = [](char c) { return c == ' ' || c == '\t' || c == '\n' || c == '\v' || c == '\f' || c == '\r'; };
    while (p < line.size()) {
        while (p < line.size() && blank(line[p])) p++;
        size_t q = p;
        while (q < line.size() && !blank(line[q])) q++;
        if (q > p) out.emplace_back(line, p, q - p);
        p = q;
    }
}

// Input cursor over an in-memory text with the state rules of std::istream that
// the DWD row search depends on:
//  getline  fails (failbit, line untouched) unless the stream is good; else
//           clears the line and extracts up to and including '\n'; reaching the
//           end sets eofbit, extracting nothing at all also sets failbit.
//  seekg    clears eofbit first, does nothing while failbit is set, and sets
//           failbit (position kept) for a target outside [0, size].
//  tellg    -1 (and failbit) unless the stream is good.
struct Cursor {
    const std::string& t;
    long pos = 0;
    bool eof = false, bad = false;
    explicit Cursor(const std::string& text) : t(text) {}
    bool good() const { return !eof && !bad; }
    bool getline(std::string& line) {
        if (!good()) {
            bad = true;
            return false;
        }
        line.clear();
        const long n = (long)t.size();
        long got = 0;
        while (true) {
            if (pos >= n) {
                eof = true;
                break;
            }
            const char c = t[(size_t)pos++];
            got++;
            if (c == '\n') break;
            line.push_back(c);
        }
        if (got == 0) bad = true;
        return !bad;
    }
    void seek(long off) {
        eof = false;
        if (bad) return;
        const long np = pos + off;
        if (np < 0 || np > (long)t.size())
            bad = true;
        else
            pos = np;
    }
    long tell() {
        if (!good()) {
            bad = true;
            return -1;
        }
        return pos;
    }
    void clear() { eof = bad = false; }
};

struct Columns {
    std::vector<float> altitude, pressure, dewPoint, temperature, windSpeed, windDir;
};

int emit(const Columns& c, long rows, skt_sounding_table* out) {
    if (rows <= 0) return fail(SKT_ERR_INPUT, "sounding has no data rows");
    const size_t r = (size_t)rows;
    if (c.altitude.size() < r || c.pressure.size() < r || c.dewPoint.size() < r || c.temperature.size() < r ||
        c.windSpeed.size() < r || c.windDir.size() < r)
        return fail(SKT_ERR_INPUT, "sounding rows lack one or more columns");
    out->rows = (int32_t)rows;
    if (rows > out->capacity) return fail(SKT_ERR_ARG, "sounding table capacity too small (rows set to the need)");
    if (!out->altitude || !out->pressure || !out->dew_point || !out->temperature || !out->wind_speed ||
        !out->wind_dir)
        return fail(SKT_ERR_ARG, "null sounding table column");
    std::memcpy(out->altitude, c.altitude.data(), r * sizeof(float));
    std::memcpy(out->pressure, c.pressure.data(), r * sizeof(float));
    std::memcpy(out->dew_point, c.dewPoint.data(), r * sizeof(float));
    std::memcpy(out->temperature, c.temperature.data(), r * sizeof(float));
    std::memcpy(out->wind_speed, c.windSpeed.data(), r * sizeof(float));
    std::memcpy(out->wind_dir, c.windDir.data(), r * sizeof(float));
    return SKT_OK;
}

// ------------------------------------------------------------------ KNMI ---
// readTable::readKNMIFile, Game/readTable.cpp:48-129.
int parseKnmi(const std::string& text, skt_sounding_table* out) {
    Columns c;
    Cursor s(text);
    std::string line;
    std::vector<std::string> words;
    long row = 0;
    s.getline(line);  // header line
    // Value of a token 'Key="v"' with a key of 'skip' characters (the token
    // from position 'skip' on, without its last character).
    auto value = [&](const std::string& w, size_t skip, float& v) -> int {
        if (w.size() <= skip) return fail(SKT_ERR_INPUT, "KNMI token too short: " + w);
        if (!parseFloat(w.substr(skip, w.size() - skip - 1), v))
            return fail(SKT_ERR_INPUT, "KNMI token has no number: " + w);
        return SKT_OK;
    };
    while (s.getline(line)) {
        splitWords(line, words);
        for (size_t k = 0; k < words.size(); k++) {
            const std::string& w = words[k];
            float v;
            switch (k) {
                case 4:
                    if (int rc = value(w, 10, v)) return rc;
                    c.pressure.push_back(v);
                    break;
                case 5:
                    if (int rc = value(w, 13, v)) return rc;
                    c.temperature.push_back(v - 273.15f);
                    break;
                case 6: {
                    if (int rc = value(w, 10, v)) return rc;
                    if ((size_t)row >= c.temperature.size())
                        return fail(SKT_ERR_INPUT, "KNMI humidity before the temperature of its row");
                    const float T = c.temperature[(size_t)row];
                    c.dewPoint.push_back(T - ((100 - v) * 0.2f) * ((T + 273.15f) * 0.01f));
                    break;
                }
                case 7:
                    if (int rc = value(w, 9, v)) return rc;
                    c.windDir.push_back(v);
                    break;
                case 8:
                    if (int rc = value(w, 11, v)) return rc;
                    c.windSpeed.push_back(v);
                    break;
                default:
                    if (w.compare(0, 8, "Altitude") == 0) {
                        if (int rc = value(w, 10, v)) return rc;
                        c.altitude.push_back(v);
                    }
                    break;
            }
        }
        row++;
    }
    return emit(c, row - 1, out);
}

// ------------------------------------------------------------------- DWD ---
// readTable::readDWDFile, Game/readTable.cpp:134-354.
int dwdDates(const std::string& meta, std::vector<std::string>& dates) {
    Cursor s(meta);
    std::string line;
    std::vector<std::string> f;
    s.getline(line);  // header line
    dates.clear();
    while (s.getline(line)) {
        splitFields(line, ';', f);
        for (size_t k = 0; k < f.size(); k++) {
            if (k == 1) {
                if (f[k].size() < 2) return fail(SKT_ERR_INPUT, "DWD metadata date too short: " + f[k]);
                dates.push_back(f[k].substr(0, f[k].size() - 2));
            } else if (k == 8) {
                float gt;  // ground temperature at the launch, parsed (and required) but unused
                if (!parseFloat(f[k], gt)) return fail(SKT_ERR_INPUT, "DWD metadata ground temperature: " + f[k]);
                break;
            }
        }
    }
    if (dates.empty()) return fail(SKT_ERR_INPUT, "DWD metadata lists no dates");
    return SKT_OK;
}

int dwdProfile(const std::string& product, const std::string& target, skt_sounding_table* out) {
    Cursor s(product);
    std::string line;
    std::vector<std::string> f;
    const long kStride = 100000;
    bool reached = false;
    // Probe: skip to the next line start, read one line, compare its date.
    while (!reached) {
        s.getline(line);
        s.getline(line);
        splitFields(line, ';', f);
        if (f.size() > 1) {
            if (f[1] == target) {
                s.seek(-kStride);
                if (s.bad) s.clear();
                reached = true;
                s.getline(line);
            } else {
                s.seek(kStride);
            }
        }
        if (s.tell() == -1) return fail(SKT_ERR_INPUT, "DWD product file: date " + target + " not found");
    }
    Columns c;
    reached = false;
    bool quit = false;
    long row = 0;
    while (s.getline(line) && !quit) {
        splitFields(line, ';', f);
        int count = 0;
        for (const std::string& w : f) {
            if (!reached) {
                if (count == 1 && w == target) {
                    reached = true;
                    count--;
                }
                count++;
            }
            if (!reached) continue;
            float v;
            std::vector<float>* col = nullptr;
            switch (count) {
                case 1:
                    if (w != target) {
                        reached = false;
                        quit = true;
                    }
                    break;
                case 6: col = &c.altitude; break;
                case 7: col = &c.pressure; break;
                case 8: col = &c.temperature; break;
                case 10: col = &c.dewPoint; break;
                case 11: col = &c.windSpeed; break;
                case 12: col = &c.windDir; break;
                default: break;
            }
            if (col) {
                if (!parseFloat(w, v)) return fail(SKT_ERR_INPUT, "DWD product value has no number: " + w);
                col->push_back(v);
            }
            count++;
        }
        if (reached) row++;
        if (s.tell() == -1) return fail(SKT_ERR_INPUT, "DWD product file: date " + target + " not found");
    }
    return emit(c, row, out);
}

// Target date of launch 'dateIndex' and the product member it is read from.
int dwdTarget(const std::string& meta, int32_t dateIndex, const std::string& station, std::string& target,
              std::string& productName) {
    std::vector<std::string> dates;
    if (int rc = dwdDates(meta, dates)) return rc;
    if (dateIndex < 0 || (size_t)dateIndex >= dates.size())
        return fail(SKT_ERR_ARG, "DWD date index outside the metadata's " + std::to_string(dates.size()) + " dates");
    target = dates[(size_t)dateIndex] + "12";
    productName = "produkt_sec_aero_" + dates.front() + "_" + dates.back() + "_" + station + ".txt";
    return SKT_OK;
}

int tableOk(const skt_sounding_table* t) {
    if (!t) return fail(SKT_ERR_ARG, "null sounding table");
    if (t->rows < 1) return fail(SKT_ERR_ARG, "sounding table has no rows");
    if (!t->altitude || !t->pressure || !t->dew_point || !t->temperature || !t->wind_speed || !t->wind_dir)
        return fail(SKT_ERR_ARG, "null sounding table column");
    return SKT_OK;
}

// ------------------------------------------------- sounding -> levels ---
// Engine/source/math/meteoformulas.cpp:8-23 (es, ws) and :184-187 (potentialTemp),
// with the constants of Engine/include/math/constants.hpp:7-13,30.
const float kRsd = 287.0528f, kRsw = 461.5f, kCpd = 1003.5f, kPI = 3.14159265359f;
float esKpa(float T) { return 0.61078f * expf((17.27f * T) / (T + 237.3f)); }
float wsat(float T, float P) {
    const float Es = esKpa(T) * 10.0f;
    return ((kRsd / kRsw) * Es) / (P - Es);
}
float thetaC(float T, float Pk, float Pt) { return (T + 273.15f) * powf(Pt / Pk, kRsd / kCpd) - 273.15f; }

// lerpEnvValue, Game/readTable.cpp:356-362.
float lerpH(float H1, float H2, float HC, float V1, float V2) {
    if (HC == H1) return V1;
    if (HC == H2) return V2;
    const float t = (HC - H1) / (H2 - H1);
    return V1 + t * (V2 - V1);
}

}  // namespace

namespace skt {

// readTable::initEnvironment, Game/readTable.cpp:364-474, per vertical level.
int soundingLevels(const skt_sounding_table* t, int ny, float dx, SoundingLevels& L) {
    if (int rc = tableOk(t)) return rc;
    if (ny < 1 || !(dx > 0.0f)) return fail(SKT_ERR_ARG, "invalid grid for the sounding");
    const float* alt = t->altitude;
    const int rows = t->rows;
    const float H0 = alt[0];
    // getIndexAtHeight, Game/readTable.cpp:506-519: first row at or above h.
    auto rowAt = [&](float h) {
        for (int i = 0; i < rows; i++)
            if (h <= alt[i]) return i;
        return -1;
    };
    std::vector<int> idx((size_t)ny);
    L.theta.assign((size_t)ny, 0.0f);
    L.pSmall.assign((size_t)ny, 0.0f);
    // Levels at heights accumulated in float from 0 while below ny*dx, as the
    // reference's loop; a dx whose accumulation does not give exactly ny
    // levels is rejected (the reference would leave levels unset or write
    // past its level arrays).
    const float top = (float)ny * dx;
    float y = 0.0f;
    int j = 0;
    for (; y < top; y += dx, j++) {
        if (j >= ny) return fail(SKT_ERR_ARG, "dx accumulates to more than ny levels");
        const int i = rowAt(y + H0);
        if (i < 0) return fail(SKT_ERR_INPUT, "sounding ends below the top level of the grid");
        const int Pi = i == 0 ? 0 : i - 1;
        L.pSmall[(size_t)j] = lerpH(alt[Pi] - H0, alt[i] - H0, y, t->pressure[Pi], t->pressure[i]);
        L.theta[(size_t)j] = lerpH(alt[Pi] - H0, alt[i] - H0, y, t->temperature[Pi], t->temperature[i]);
        idx[(size_t)j] = i;
    }
    if (j != ny) return fail(SKT_ERR_ARG, "dx accumulates to fewer than ny levels");
    // Potential temperature relative to the first row's pressure, in kelvin.
    for (int k = 0; k < ny; k++) L.theta[(size_t)k] = thetaC(L.theta[(size_t)k], L.pSmall[(size_t)k], t->pressure[0]) + 273.15f;
    L.velX.assign((size_t)ny, 0.0f);
    L.velZ.assign((size_t)ny, 0.0f);
    L.qv.assign((size_t)ny, 0.0f);
    L.p.assign((size_t)ny, 0.0f);
    for (int k = 0; k < ny; k++) {
        const int i = idx[(size_t)k], Pi = i == 0 ? 0 : i - 1, Pk = k == 0 ? 0 : k - 1;
        const float a = (t->wind_dir[i] - 180.0f) * (kPI / 180.0f);
        const float u = sinf(a) * t->wind_speed[i], w = cosf(a) * t->wind_speed[i];
        // The level pressure interpolates the previous and current level
        // values with the sounding rows' heights as the interval
        // (Game/readTable.cpp:427), and the vapour uses that pressure.
        const float HC = (float)k * dx;
        const float ps = lerpH(alt[Pi] - H0, alt[i] - H0, HC, L.pSmall[(size_t)Pk], L.pSmall[(size_t)k]);
        const float dew = lerpH(alt[Pi] - H0, alt[i] - H0, HC, t->dew_point[Pi], t->dew_point[i]);
        L.qv[(size_t)k] = wsat(dew, ps);
        L.p[(size_t)k] = ps;
        // MAC faces average a cell with its +1 neighbour along the axis
        // (Game/readTable.cpp:443-460); the wind is uniform per level, so the
        // face value is (v + v) / 2 and the vertical component stays 0.
        L.velX[(size_t)k] = (u + u) / 2.0f;
        L.velZ[(size_t)k] = (w + w) / 2.0f;
    }
    L.groundT = t->temperature[0] + 273.15f;
    L.groundP = t->pressure[0];
    return SKT_OK;
}

}  // namespace skt

extern "C" {

const char* skt_input_last_error(void) { return g_err.c_str(); }

int skt_sounding_parse_knmi(const char* text, int64_t len, skt_sounding_table* table) {
    if (!table || (!text && len) || len < 0) return fail(SKT_ERR_ARG, "invalid KNMI parse arguments");
    return parseKnmi(std::string(text ? text : "", (size_t)len), table);
}

int skt_sounding_parse_dwd(const char* meta, int64_t meta_len, const char* product, int64_t product_len,
                           int32_t date_index, skt_sounding_table* table) {
    if (!table || (!meta && meta_len) || (!product && product_len) || meta_len < 0 || product_len < 0)
        return fail(SKT_ERR_ARG, "invalid DWD parse arguments");
    std::string target, name;
    if (int rc = dwdTarget(std::string(meta ? meta : "", (size_t)meta_len), date_index, "", target, name)) return rc;
    return dwdProfile(std::string(product ? product : "", (size_t)product_len), target, table);
}

int skt_sounding_read_knmi(const char* path, skt_sounding_table* table) {
    if (!path || !table) return fail(SKT_ERR_ARG, "null argument");
    std::string z, xml;
    if (int rc = readFile(path, z)) return rc;
    if (int rc = zipMember(z, "SynchronizedSoundingData.xml", xml)) return rc;
    return parseKnmi(xml, table);
}

int skt_sounding_read_dwd(const char* path, int32_t date_index, skt_sounding_table* table) {
    if (!path || !table) return fail(SKT_ERR_ARG, "null argument");
    // Station number: third '_'-separated token of the file name (the
    // reference splits the whole path, which agrees when the directories
    // hold no '_').
    std::string base(path);
    const size_t slash = base.find_last_of('/');
    if (slash != std::string::npos) base = base.substr(slash + 1);
    std::vector<std::string> tok;
    splitFields(base, '_', tok);
    if (tok.size() < 3) return fail(SKT_ERR_INPUT, "DWD archive name has no station number: " + base);
    const std::string station = tok[2];
    std::string z, meta, product, target, productName;
    if (int rc = readFile(path, z)) return rc;
    if (int rc = zipMember(z, "Metadaten_Sekunde_Aero_" + station + ".txt", meta)) return rc;
    if (int rc = dwdTarget(meta, date_index, station, target, productName)) return rc;
    if (int rc = zipMember(z, productName, product)) return rc;
    return dwdProfile(product, target, table);
}

int skt_sounding_init_fields(const skt_sounding_table* table, int32_t nx, int32_t ny, int32_t nz, float dx,
                             float* potTemps, float* velField, float* Qv, float* groundTemp, float* groundPres,
                             float* pressures, float* smallPressure) {
    if (nx < 1 || nz < 1 || !potTemps || !velField || !Qv || !groundTemp || !groundPres || !pressures ||
        !smallPressure)
        return fail(SKT_ERR_ARG, "invalid sounding init arguments");
    skt::SoundingLevels L;
    if (int rc = skt::soundingLevels(table, ny, dx, L)) return rc;
    const size_t X = (size_t)nx, Y = (size_t)ny, Z = (size_t)nz;
    for (size_t z = 0; z < Z; z++)
        for (size_t y = 0; y < Y; y++)
            for (size_t x = 0; x < X; x++) {
                const size_t i = x + X * (y + Y * z);
                potTemps[i] = L.theta[y];
                velField[3 * i + 0] = L.velX[y];
                velField[3 * i + 1] = 0.0f;
                velField[3 * i + 2] = L.velZ[y];
                Qv[i] = L.qv[y];
                pressures[i] = L.p[y];
            }
    for (size_t g = 0; g < X * Z; g++) {
        groundTemp[g] = L.groundT;
        groundPres[g] = L.groundP;
    }
    std::memcpy(smallPressure, L.pSmall.data(), Y * sizeof(float));
    return SKT_OK;
}

}  // extern "C"
