// Host-side data formats of the drivers: SU shot gathers (dataio.c:306-516),
// the JSON acquisition geometry (observation.c:422-532), the grid geometry of
// the records (observation.c:338-420) and the wavelet's peak frequency
// (common.c:363-413).  No device work: these run before and after the
// propagation entry points.
//
// SU trace format: a 240-byte trace header followed by ns native-endian
// float32 samples, one shot after another (what Seismic Unix's fputtr writes
// and fvgettr reads).  The header layout is SEG-Y rev 1's for bytes 0-179
// plus Seismic Unix's extension (d1, f1, d2, f2, ungpow, unscale, ntr, mark)
// from byte 180; SU's segy.h itself is not part of this build.
#include <math.h>
#include <stddef.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <string>
#include <vector>

#include "engine.hpp"

namespace fwiedu {

#pragma pack(push, 1)
struct SuTrace {
    int32_t tracl, tracr, fldr, tracf, ep, cdp, cdpt;
    int16_t trid, nvs, nhs, duse;
    int32_t offset, gelev, selev, sdepth, gdel, sdel, swdep, gwdep;
    int16_t scalel, scalco;
    int32_t sx, sy, gx, gy;
    int16_t counit, wevel, swevel, sut, gut, sstat, gstat, tstat, laga, lagb, delrt, muts, mute;
    uint16_t ns, dt;
    int16_t gain, igc, igi, corr, sfs, sfe, slen, styp, stas, stae, tatyp, afilf, afils, nofilf,
        nofils, lcf, hcf, lcs, hcs, year, day, hour, minute, sec, timbas, trwf, grnors, grnofr,
        grnlof, gaps, otrav;
    float d1, f1, d2, f2, ungpow, unscale;
    int32_t ntr;
    int16_t mark, shortpad;
    int16_t unass[14];
};
#pragma pack(pop)
static_assert(sizeof(SuTrace) == 240, "SU trace header is 240 bytes");
static_assert(offsetof(SuTrace, scalel) == 68 && offsetof(SuTrace, sx) == 72 &&
                  offsetof(SuTrace, gx) == 80 && offsetof(SuTrace, ns) == 114 &&
                  offsetof(SuTrace, dt) == 116 && offsetof(SuTrace, d1) == 180 &&
                  offsetof(SuTrace, d2) == 188 && offsetof(SuTrace, ntr) == 204,
              "SEG-Y / SU header byte positions");

static int sgn(int v) { return (v > 0) - (v < 0); }

// scale_header (dataio.c:374-398): keyword * 10^(sign(scale) * log10|scale|)
static float scale_header(int keyword, int scale)
{
    if (scale == 0) fail("<error> : scale must not be %d!", scale);
    const int temp = sgn(scale) * (int)log10((double)abs(scale));
    return (float)keyword * powf(10.0f, (float)temp);
}

// Seismic Unix allocation conventions: alloc1* = one block; alloc2float(n1,
// n2) = a table of n2 row pointers over one block of n2*n1 floats, released
// by free(p[0]); free(p).
static float *alloc1f(size_t n)
{
    float *p = (float *)calloc(n ? n : 1, sizeof(float));
    if (!p) fail("<error>: out of host memory");
    return p;
}

static float **alloc2f(int n1, int n2)
{
    float **p = (float **)malloc(sizeof(float *) * (size_t)(n2 ? n2 : 1));
    float *b = alloc1f((size_t)n1 * n2);
    if (!p) fail("<error>: out of host memory");
    for (int i = 0; i < n2; ++i) p[i] = b + (size_t)i * n1;
    return p;
}

// ---------------------------------------------------------------------------
// a small JSON reader: objects, arrays, numbers, strings, true/false/null

struct Json {
    enum Kind { NUL, BOOL, NUM, STR, ARR, OBJ } kind = NUL;
    double num = 0.0;
    std::string str;
    std::vector<Json> items;                 // array elements / object values
    std::vector<std::string> keys;           // object keys
    const Json *get(const char *k) const
    {
        for (size_t i = 0; i < keys.size(); ++i)
            if (keys[i] == k) return &items[i];
        return nullptr;
    }
};

struct JsonParser {
    const char *p, *end;
    std::string err;
    void ws()
    {
        while (p < end && (*p == ' ' || *p == '\n' || *p == '\r' || *p == '\t')) ++p;
    }
    bool lit(const char *s)
    {
        const size_t n = strlen(s);
        if ((size_t)(end - p) >= n && strncmp(p, s, n) == 0) {
            p += n;
            return true;
        }
        return false;
    }
    bool string(std::string &out)
    {
        if (p >= end || *p != '"') return false;
        ++p;
        while (p < end && *p != '"') {
            if (*p == '\\' && p + 1 < end) {
                ++p;
                const char c = *p;
                out.push_back(c == 'n' ? '\n' : c == 't' ? '\t' : c);
                ++p;
            } else {
                out.push_back(*p++);
            }
        }
        if (p >= end) return false;
        ++p;
        return true;
    }
    bool value(Json &v, int depth)
    {
        if (depth > 64) return false;
        ws();
        if (p >= end) return false;
        if (*p == '{') {
            v.kind = Json::OBJ;
            ++p;
            ws();
            if (p < end && *p == '}') {
                ++p;
                return true;
            }
            for (;;) {
                ws();
                std::string k;
                if (!string(k)) return false;
                ws();
                if (p >= end || *p != ':') return false;
                ++p;
                v.keys.push_back(k);
                v.items.emplace_back();
                if (!value(v.items.back(), depth + 1)) return false;
                ws();
                if (p < end && *p == ',') {
                    ++p;
                    continue;
                }
                if (p < end && *p == '}') {
                    ++p;
                    return true;
                }
                return false;
            }
        }
        if (*p == '[') {
            v.kind = Json::ARR;
            ++p;
            ws();
            if (p < end && *p == ']') {
                ++p;
                return true;
            }
            for (;;) {
                v.items.emplace_back();
                if (!value(v.items.back(), depth + 1)) return false;
                ws();
                if (p < end && *p == ',') {
                    ++p;
                    continue;
                }
                if (p < end && *p == ']') {
                    ++p;
                    return true;
                }
                return false;
            }
        }
        if (*p == '"') {
            v.kind = Json::STR;
            return string(v.str);
        }
        if (lit("true")) {
            v.kind = Json::BOOL;
            v.num = 1;
            return true;
        }
        if (lit("false")) {
            v.kind = Json::BOOL;
            return true;
        }
        if (lit("null")) return true;
        char *e = nullptr;
        v.num = strtod(p, &e);
        if (e == p) return false;
        v.kind = Json::NUM;
        p = e;
        return true;
    }
};

static Json parse_json_file(const char *path)
{
    FILE *f = fopen(path, "rb");
    if (!f) fail("<error>: failed to open file: %s!", path);
    std::string txt;
    char buf[65536];
    size_t n;
    while ((n = fread(buf, 1, sizeof(buf), f)) > 0) txt.append(buf, n);
    fclose(f);
    JsonParser ps{txt.data(), txt.data() + txt.size(), {}};
    Json root;
    if (!ps.value(root, 0)) fail("<error>: Parse JSON from file (%s) Error!", path);
    return root;
}

static double jnum(const Json *v, const char *what, int shot)
{
    if (!v || v->kind != Json::NUM) fail("<error>: JSON shot %d has no number \"%s\"", shot, what);
    return v->num;
}

}  // namespace fwiedu

using namespace fwiedu;

// calculate_peak_frequency, common.c:363-413: the first bin of maximum |rfft|
// (strict >), at i * df with df = 1 / (nt * dt) in float.  The spectrum is an
// exact-enough double DFT (the reference's FFTW single precision picks the same
// bin: neighbouring bin magnitudes of the wavelets differ by far more than
// either transform's rounding).
extern "C" float calculate_peak_frequency(const float *s, int nt, float dt)
{
    const float df = 1.0f / ((float)nt * dt);
    float fpeak = 0.0f, maxamp = -1.0f;
    const double w = 2.0 * 3.141592653589793 / nt;
    std::vector<double> cs(nt), sn(nt);
    for (int m = 0; m < nt; ++m) {
        cs[m] = cos(w * (double)m);
        sn[m] = sin(w * (double)m);
    }
    for (int i = 0; i <= nt / 2; ++i) {
        double re = 0.0, im = 0.0;
        long m = 0;  // (i * j) mod nt
        for (int j = 0; j < nt; ++j) {
            re += s[j] * cs[m];
            im -= s[j] * sn[m];
            m += i;
            if (m >= nt) m -= nt;
        }
        const float amp = (float)sqrt(re * re + im * im);
        if (amp > maxamp) {
            maxamp = amp;
            fpeak = (float)i * df;
        }
    }
    return fpeak;
}

// create_grid_observe2d, observation.c:338-420
extern "C" void create_grid_observe2d(Record2D *record2d, int Nshot, float dx, float dz, int nx)
{
    for (int s = 0; s < Nshot; ++s) {
        Record2D &r = record2d[s];
        if (!r.gx || !r.gz) fail("gx and gz must be allocated for shot %d\n", s);
        const int ntr = r.ntr;
        if (ntr <= 0) fail("ntr must be positive for shot %d\n", s);
        if (!r.igx) r.igx = (int *)calloc(ntr, sizeof(int));
        if (!r.igz) r.igz = (int *)calloc(ntr, sizeof(int));
        const int isx = (int)floorf(r.sx / dx), isz = (int)floorf(r.sz / dz);
        int gmin = nx - 1, gmax = 0;
        for (int i = 0; i < ntr; ++i) {
            r.igx[i] = (int)floorf(r.gx[i] / dx);
            r.igz[i] = (int)floorf(r.gz[i] / dz);
            gmin = std::min(gmin, r.igx[i]);
            gmax = std::max(gmax, r.igx[i]);
        }
        const int ix_min = std::max(gmin - 30, 0), ix_max = std::min(gmax + 30, nx - 1);
        r.isx = isx - ix_min;
        r.isz = isz;
        for (int i = 0; i < ntr; ++i) r.igx[i] -= ix_min;
        r.ix_min = ix_min;
        r.ix_max = ix_max;
    }
}

// write_acoustic2d_traces, dataio.c:306-372: every trace of every shot to
// "<prefix>_<Nshot>_<ns>_<total traces>_<dt ms %05.2f>ms.su"; coordinates in
// millimetres (scalel = scalco = -1000), dt in microseconds.
extern "C" void write_acoustic2d_traces(char *prefix, Record2D *record2d, int Nshot)
{
    int ntr = 0;
    for (int s = 0; s < Nshot; ++s) ntr += record2d[s].ntr;
    const int nt = record2d[0].ns;
    const float dtms = record2d[0].dt * 1000;
    char name[4096];
    snprintf(name, sizeof(name), "%s_%d_%d_%d_%05.2fms.su", prefix, Nshot, nt, ntr, dtms);
    FILE *fp = fopen(name, "wb");
    if (!fp) fail("<error>: cannot open %s for writing", name);
    SuTrace tr;
    memset(&tr, 0, sizeof(tr));
    tr.trid = 1;
    tr.scalel = -1000;
    tr.scalco = -1000;
    const float scalel = powf(10.0f, (float)(sgn(-tr.scalel) * (int)log10((double)abs(-tr.scalel))));
    const float scalco = powf(10.0f, (float)(sgn(-tr.scalco) * (int)log10((double)abs(-tr.scalco))));
    for (int s = 0; s < Nshot; ++s) {
        const Record2D &r = record2d[s];
        tr.fldr = r.shotID;
        tr.tracf = 0;
        tr.ntr = r.ntr;
        tr.ns = (uint16_t)r.ns;
        tr.dt = (uint16_t)(r.dt * 1000000);
        tr.d1 = r.dt;
        tr.d2 = r.dtr;
        tr.sdepth = (int32_t)(r.sz * scalel);
        tr.sx = (int32_t)(r.sx * scalco);
        for (int i = 0; i < r.ntr; ++i) {
            tr.tracl++;
            tr.tracr++;
            tr.tracf++;
            tr.gelev = (int32_t)(r.gz[i] * scalel);
            tr.gx = (int32_t)(r.gx[i] * scalco);
            tr.offset = tr.gx - tr.sx;
            tr.cdp = (tr.gx + tr.sx) / 2;
            if (fwrite(&tr, sizeof(tr), 1, fp) != 1 ||
                fwrite(r.traces_p[i], sizeof(float), (size_t)r.ns, fp) != (size_t)r.ns)
                fail("<error>: write error on %s", name);
        }
    }
    fclose(fp);
}

// Next trace of an SU file into tr and data (data grown as needed); false at
// the end of the file.
static bool next_trace(FILE *fp, SuTrace &tr, std::vector<float> &data)
{
    if (fread(&tr, sizeof(tr), 1, fp) != 1) return false;
    data.resize(tr.ns);
    if (fread(data.data(), sizeof(float), tr.ns, fp) != tr.ns)
        fail("<error>: truncated trace in SU file");
    return true;
}

// get_shots_number, dataio.c:400-443: first trace of each gather, then skip
// the gather's other ntr-1 traces.
extern "C" int get_shots_number(const char *filename)
{
    FILE *fp = fopen(filename, "rb");
    if (!fp) fail("<error>: failed to open file: %s!", filename);
    SuTrace tr;
    std::vector<float> data;
    int n = 0;
    while (next_trace(fp, tr, data)) {
        if (tr.ntr == 0) fail("<error>: the ntr of shot gather is 0!");
        ++n;
        fseek(fp, (long)(tr.ntr - 1) * (240 + (long)tr.ns * 4), SEEK_CUR);
    }
    fclose(fp);
    return n;
}

// read_gather2d_from_sufile, dataio.c:445-503: geometry, sampling and traces
// of Nshot gathers; sampleRate = (nt - 1) / (ns - 1), dtr = mean of d2.
// gx, gz (alloc1float) and traces_p (alloc2float) are allocated here.
extern "C" void read_gather2d_from_sufile(Record2D *record2d, const char *recordfile, int Nshot,
                                          int nt)
{
    if (!recordfile) fail("<error>: didn't specify recordfile");
    FILE *fp = fopen(recordfile, "rb");
    if (!fp) fail("<error>: fail to open recordfile: %s!", recordfile);
    SuTrace tr;
    std::vector<float> data;
    for (int s = 0; s < Nshot; ++s) {
        if (!next_trace(fp, tr, data)) fail("<error>: %s holds fewer than %d shots", recordfile, Nshot);
        if (tr.ntr <= 0) fail("<error>: the number of traces in shot%d is %d.", s + 1, tr.ntr);
        Record2D &r = record2d[s];
        r.shotID = s + 1;
        r.ntr = tr.ntr;
        r.sx = scale_header(tr.sx, tr.scalco);
        r.sz = scale_header(tr.sdepth, tr.scalel);
        r.gx = alloc1f(r.ntr);
        r.gz = alloc1f(r.ntr);
        r.ns = tr.ns;
        r.dt = (float)tr.dt / 1e6f;
        if (r.ns < 2) fail("<error>: shot %d has %d samples per trace", s + 1, r.ns);
        r.sampleRate = (nt - 1) / (r.ns - 1);
        r.traces_p = alloc2f(r.ns, r.ntr);
        double dtr = 0.0;
        for (int i = 0; i < r.ntr; ++i) {
            if (i > 0 && !next_trace(fp, tr, data))
                fail("<error>: shot %d of %s is truncated", s + 1, recordfile);
            r.gx[i] = scale_header(tr.gx, tr.scalco);
            r.gz[i] = scale_header(tr.gelev, tr.scalel);
            memcpy(r.traces_p[i], data.data(), (size_t)r.ns * sizeof(float));
            dtr += (double)tr.d2;
        }
        r.dtr = (float)(dtr / r.ntr);
    }
    fclose(fp);
}

// init_observe2d_from_JSON, observation.c:422-532, reading the "record" array
// of a JSON acquisition file (the reference receives it parsed by cJSON).
// x[nx], z[nz] are the model's coordinates, src the wavelet of nt samples.
extern "C" void fwiedu_observe2d_from_json(Record2D *record2d, const char *jsonfile, int Nshot,
                                           float *x, int nx, float *z, int nz, float dt, int nt,
                                           float *src, int sampleRate)
{
    const Json root = parse_json_file(jsonfile);
    const Json *recs = root.get("record");
    if (!recs || recs->kind != Json::ARR) fail("<error>: JSON file %s has no \"record\" array", jsonfile);
    if ((int)recs->items.size() != Nshot)
        fail("<error>: The number of shots in JSON (%d) is not equal to Nshot (%d)!\n",
             (int)recs->items.size(), Nshot);
    if (sampleRate < 1) fail("<error>: sampleRate must be >= 1");
    const float fpeak = calculate_peak_frequency(src, nt, dt);
    for (int s = 0; s < Nshot; ++s) {
        const Json &shot = recs->items[s];
        Record2D &r = record2d[s];
        const int shotID = (int)jnum(shot.get("shotID"), "shotID", s);
        const float sx = (float)jnum(shot.get("sx"), "sx", s);
        const float sz = (float)jnum(shot.get("sz"), "sz", s);
        const int ntr = (int)jnum(shot.get("ntr"), "ntr", s);
        if (ntr <= 0) fail("<error>: the number of trace in shot %d is not more than 0", s);
        if (sx < x[0] || sx > x[nx - 1]) fail("<error>: x-coordinate of shot %d is out of range!\n", s);
        if (sz < z[0] || sz > z[nz - 1]) fail("<error>: z-coordinate of shot %d is out of range!\n", s);
        r.shotID = shotID;
        r.sx = sx;
        r.sz = sz;
        r.ntr = ntr;
        r.gx = alloc1f(ntr);
        r.gz = alloc1f(ntr);
        const Json *gx = shot.get("gx"), *gz = shot.get("gz");
        if (!gx || gx->kind != Json::ARR || (int)gx->items.size() != ntr)
            fail("<error>: The number of gx (%d) is not equal to ntr (%d) in shot %d at JSON!\n",
                 gx ? (int)gx->items.size() : 0, ntr, s);
        for (int i = 0; i < ntr; ++i) {
            r.gx[i] = (float)jnum(&gx->items[i], "gx", s);
            if (r.gx[i] < x[0] || r.gx[i] > x[nx - 1])
                fail("<error>: x-coordinate of receiver %d in shot %d is out of range!\n", i, s);
        }
        if (!gz || gz->kind != Json::ARR || (int)gz->items.size() != ntr)
            fail("<error>: The number of gz (%d) is not equal to ntr (%d) in shot %d at JSON!\n",
                 gz ? (int)gz->items.size() : 0, ntr, s);
        for (int i = 0; i < ntr; ++i) {
            r.gz[i] = (float)jnum(&gz->items[i], "gz", s);
            if (r.gz[i] < z[0] || r.gz[i] > z[nz - 1])
                fail("<error>: z-coordinate of receiver %d in shot %d is out of range!\n", i, s);
        }
        r.sampleRate = sampleRate;
        r.ns = (nt - 1) / sampleRate + 1;
        r.dt = dt * sampleRate;
        // mean receiver spacing, summed from the last pair down (observation.c:512-516)
        r.dtr = 0;
        for (int i = ntr - 1; i > 0; --i) r.dtr += (r.gx[i] - r.gx[i - 1]);
        r.dtr /= ntr - 1;
        r.traces_p = alloc2f(r.ns, ntr);
        r.src = alloc1f(nt);
        memcpy(r.src, src, (size_t)nt * sizeof(float));
        r.fpeak = fpeak;
    }
}

// generate_shot_positions, observation.c:102-165: Nshot_in shots from fsx
// every ds at depth sdepth, cut to those up to x[nx-1].  sx, sz are
// allocated here (alloc1float); returns the shot count.
extern "C" int generate_shot_positions(const float *x, int nx, const float *z, int nz, float fsx,
                                       float ds, float sdepth, int Nshot_in, float **sx,
                                       float **sz)
{
    if (nx < 1 || nz < 1) fail("Model domain arrays x or z are empty!\n");
    if (fsx < x[0] || fsx > x[nx - 1]) fail("x-coordinate of first shot is out of range!\n");
    if (sdepth < z[0] || sdepth > z[nz - 1]) fail("z-coordinate of shots is out of range!\n");
    if (!(ds > 0.0f)) fail("<error>: the shot interval ds must be positive");
    const int maxShots = (int)((x[nx - 1] - fsx) / ds) + 1;
    const int n = Nshot_in < maxShots ? Nshot_in : maxShots;
    if (n < 1) fail("<error>: no shot: Nshot = %d", Nshot_in);
    *sx = alloc1f(n);
    *sz = alloc1f(n);
    for (int s = 0; s < n; ++s) {
        (*sx)[s] = fsx + s * ds;
        (*sz)[s] = sdepth;
    }
    return n;
}

// init_observe2d, observation.c:229-336: a fixed-offset spread per shot.
// The relative offsets are those of the reference's two float loops, from
// -offsmax up to -offsmin and from offsmin up to offsmax in steps of dtr, the
// last negative one dropped when offsmin is 0 (that slot is the zero offset's;
// :259-272); a shot keeps the receivers sx + offset inside [x[0], x[nx-1]]
// at depth gdepth.  The reference sizes its offset table from
// lroundf((offsmax - offsmin) / dtr) and reads every slot of it; where the
// loops write fewer slots (offsmax - offsmin not a multiple of dtr) it reads
// unwritten memory, and this version keeps only the offsets the loops give.
extern "C" void init_observe2d(Record2D *record2d, int Nshot, float *x, int nx, float *z, int nz,
                               float dt, int nt, float *sx, float *sz, float *src, float offsmin,
                               float offsmax, float dtr, float gdepth, int sampleRate)
{
    if (!(dtr > 0.0f)) fail("<error>: the receiver interval dtr must be positive");
    if (sampleRate < 1) fail("<error>: sampleRate must be >= 1");
    (void)z;
    (void)nz;
    std::vector<float> offs;
    for (float o = -offsmax; o <= -offsmin; o += dtr) offs.push_back(o);
    if (offsmin == 0.0f && !offs.empty()) offs.pop_back();
    for (float o = offsmin; o <= offsmax; o += dtr) offs.push_back(o);
    const float fpeak = calculate_peak_frequency(src, nt, dt);
    for (int s = 0; s < Nshot; ++s) {
        Record2D &r = record2d[s];
        r.shotID = s + 1;
        r.sx = sx[s];
        r.sz = sz[s];
        int ntr = 0;
        for (float o : offs) {
            const float g = r.sx + o;
            if (g >= x[0] && g <= x[nx - 1]) ++ntr;
        }
        if (ntr == 0) fail("the number of trace in shot %d is 0", s);
        r.ntr = ntr;
        r.gx = alloc1f(ntr);
        r.gz = alloc1f(ntr);
        int k = 0;
        for (float o : offs) {
            const float g = r.sx + o;
            if (g >= x[0] && g <= x[nx - 1]) {
                r.gx[k] = g;
                r.gz[k] = gdepth;
                ++k;
            }
        }
        r.sampleRate = sampleRate;
        r.ns = (nt - 1) / sampleRate + 1;
        r.dt = dt * sampleRate;
        // mean receiver spacing, summed from the last pair down (:318-323)
        r.dtr = 0;
        for (int i = ntr - 1; i > 0; --i) r.dtr += (r.gx[i] - r.gx[i - 1]);
        r.dtr /= ntr - 1;
        r.traces_p = alloc2f(r.ns, ntr);
        r.src = alloc1f(nt);
        memcpy(r.src, src, (size_t)nt * sizeof(float));
        r.fpeak = fpeak;
    }
}

// Shots in the "record" array of a JSON acquisition file.
extern "C" int fwiedu_json_shot_count(const char *jsonfile)
{
    const Json root = parse_json_file(jsonfile);
    const Json *recs = root.get("record");
    if (!recs || recs->kind != Json::ARR) fail("<error>: JSON file %s has no \"record\" array", jsonfile);
    return (int)recs->items.size();
}
