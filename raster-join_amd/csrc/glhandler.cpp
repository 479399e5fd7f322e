// glhandler.cpp -- the GLHandler-shaped C++ front end over the C ABI, and the
// polygon text reader of the reference's PolyHandler.
#include "glhandler.hpp"

#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <set>
#include <sstream>

namespace rjcpp {

GLHandler* GLHandler::instance = nullptr;

// ---------------------------------------------------------------------------
// polygons
// ---------------------------------------------------------------------------
namespace {

// Web Mercator at zoom 22 (common/Utils.h:9-20), centered on the US bbox.
void geo2world(double lat, double lon, double* x, double* y)
{
    double wy;
    if (lat == 90.0) wy = 256;
    else if (lat == -90.0) wy = 0;
    else wy = (M_PI - std::atanh(std::sin(lat * M_PI / 180))) / M_PI * 128;
    const double s = std::exp2(22.0);
    *x = (lon + 180) / 360.0 * 256 * s;
    *y = wy * s;
}

// transformPoint (common/UsefulFuncs.cpp:578-597): US bbox filter, then the
// projected point minus the projected bbox center.
bool transform_point(float lat, float lon, float* tx, float* ty)
{
    const double minlat = 24.396308, minlon = -124.848974, maxlat = 49.384358, maxlon = -66.885444;
    if (lat < minlat || lat > maxlat || lon < minlon || lon > maxlon) return false;
    double cx, cy, px, py;
    geo2world((minlat + maxlat) / 2, (minlon + maxlon) / 2, &cx, &cy);
    geo2world(lat, lon, &px, &py);
    *tx = (float)(px - cx);
    *ty = (float)(py - cy);
    return true;
}

// One record of the reference's binary datasets, the fields the join reads.
// TaxiRecordData (db/TaxiRecord.hpp:15-46), 88 bytes: db_idx @0, pick_x @8,
// pick_y @16, drop_x @24, drop_y @32 (doubles), pickup_time @40,
// dropoff_time @48 (u64), vendor @56, duration @60 (u32), miles @64 (float),
// fare @68 ... license_id @82 (u16), store_and_forward @84, payment_type @85,
// passengers @86, rate_code @87 (u8). Tweet (db/TwitterRecord.hpp:19-28),
// 32 bytes: user_id @0, time @8 (i64), loc_x @16, loc_y @20 (floats),
// favCt @24, rtCt @28 (i32).
constexpr size_t kTaxiRecord = 88, kTweetRecord = 32;

template <class T>
T field(const unsigned char* r, size_t at)
{
    T v;
    std::memcpy(&v, r + at, sizeof(T));
    return v;
}

// The 4-byte attribute `id` of a record as getAttributeAsBinary returns it
// (TaxiRecord.hpp:107-262, TwitterRecord.hpp:57-100); type receives RJ_ATTR_*.
// Returns false for an attribute that is not 4 bytes wide.
bool attribute(int dataset, int id, const unsigned char* r, uint32_t* v, int32_t* type)
{
    if (dataset == RJ_DATASET_TAXI) {
        switch (id) {
        case 3: *v = (uint32_t)field<uint64_t>(r, 40); *type = RJ_ATTR_UINT; return true;
        case 4: *v = (uint32_t)field<uint64_t>(r, 48); *type = RJ_ATTR_UINT; return true;
        case 6: *v = field<uint32_t>(r, 60); *type = RJ_ATTR_UINT; return true;
        case 7: *v = field<uint32_t>(r, 64); *type = RJ_ATTR_FLOAT; return true;
        case 8: *v = (uint32_t)field<uint16_t>(r, 68); *type = RJ_ATTR_UINT; return true;
        case 18: *v = (uint32_t)field<uint8_t>(r, 86); *type = RJ_ATTR_UINT; return true;
        default: return false;
        }
    }
    switch (id) {
    case 1: *v = (uint32_t)field<int64_t>(r, 8); *type = RJ_ATTR_UINT; return true;
    case 3: *v = field<uint32_t>(r, 24); *type = RJ_ATTR_INT; return true;
    case 4: *v = field<uint32_t>(r, 28); *type = RJ_ATTR_INT; return true;
    default: return false;
    }
}

}  // namespace

void PolyHandler::setPolygons(std::vector<float> xy, std::vector<int32_t> off)
{
    polys = std::move(xy);
    pindexes = std::move(off);
    float mn[2] = {polys[0], polys[1]}, mx[2] = {polys[0], polys[1]};
    for (size_t k = 0; k + 1 < polys.size(); k += 2)
        for (int d = 0; d < 2; d++) {
            mn[d] = std::min(mn[d], polys[k + d]);
            mx[d] = std::max(mx[d], polys[k + d]);
        }
    bounds[0] = mn[0]; bounds[1] = mn[1]; bounds[2] = mx[0]; bounds[3] = mx[1];
}

bool PolyHandler::readPolygons(const std::string& file, bool transform, std::string* err)
{
    std::ifstream in(file);
    if (!in) {
        *err = "could not open " + file;
        return false;
    }
    int tot = 0;
    in >> tot;
    std::vector<float> xy;
    std::vector<int32_t> off = {0};
    for (int i = 0; i < tot && in; i++) {
        int parts = 0;
        in >> parts;
        std::vector<float> poly;
        for (int j = 0; j < parts; j++) {
            int n = 0;
            in >> n;
            for (int k = 0; k < n; k++) {
                double x, y;
                in >> x >> y;
                float fx = (float)x, fy = (float)y;
                // consecutive duplicates are dropped (PolyHandler.cpp:156-166)
                if (!poly.empty() && k > 0 && poly[poly.size() - 2] == fx && poly.back() == fy) continue;
                poly.push_back(fx);
                poly.push_back(fy);
            }
        }
        if (poly.size() >= 4 && poly[0] == poly[poly.size() - 2] && poly[1] == poly.back()) {
            poly.pop_back();
            poly.pop_back();
        }
        if (transform) {
            std::vector<float> t;
            bool ok = true;
            for (size_t k = 0; k < poly.size() && ok; k += 2) {
                float tx, ty;
                ok = transform_point(poly[k], poly[k + 1], &tx, &ty);
                t.push_back(tx);
                t.push_back(ty);
            }
            if (!ok) continue;  // polygons leaving the US bbox are skipped (PolyHandler.cpp:167-175)
            poly.swap(t);
        }
        if (poly.size() < 6) continue;
        xy.insert(xy.end(), poly.begin(), poly.end());
        off.push_back((int32_t)(xy.size() / 2));
    }
    if (off.size() < 2) {
        *err = "no polygons in " + file;
        return false;
    }
    setPolygons(std::move(xy), std::move(off));
    return true;
}

bool PolyHandler::initFromFile(const std::string& list, const std::string& collection, bool transform,
                               std::string* err)
{
    std::ifstream in(list);
    if (!in) {
        *err = "could not open " + list;
        return false;
    }
    std::string dir = list.substr(0, list.find_last_of('/') == std::string::npos ? 0 : list.find_last_of('/') + 1);
    std::string line;
    while (std::getline(in, line)) {
        size_t c = line.find(',');
        if (c == std::string::npos) continue;
        std::string name = line.substr(0, c), file = line.substr(c + 1);
        while (!file.empty() && (file.back() == '\r' || file.back() == ' ')) file.pop_back();
        if (name == collection) return readPolygons(dir + file, transform, err);
    }
    *err = "collection " + collection + " not in " + list;
    return false;
}

// ---------------------------------------------------------------------------
// GLHandler
// ---------------------------------------------------------------------------

GLHandler::GLHandler(int64_t gpuMemMB, bool inMemory)
{
    if (rj_create(gpuMemMB, inMemory ? 1 : 0, &ctx_) != RJ_OK) ctx_ = nullptr;
}

GLHandler::~GLHandler()
{
    if (ctx_) rj_destroy(ctx_);
}

GLHandler* GLHandler::getInstance(int64_t gpuMemMB, bool inMemory)
{
    if (!instance) {
        instance = new GLHandler(gpuMemMB, inMemory);
        if (!instance->ctx_) {
            delete instance;
            instance = nullptr;
        }
    }
    return instance;
}

void GLHandler::destroyInstance()
{
    delete instance;
    instance = nullptr;
}

void GLHandler::setDataHandler(DataHandler* dataHandler)
{
    data_ = dataHandler;
    dirty_ = true;
}

void GLHandler::setAccuracyDistance(double meters, double groundRes) { rj_set_accuracy(ctx_, meters, groundRes); }

void GLHandler::setPolyIndexResolution(int resx, int resy) { rj_set_index_res(ctx_, resx, resy); }

// GLFunction::setupPoints (GLFunction.cpp:199-344): the location column plus
// the sorted set of filter and aggregate attributes, remapped to compact ids.
bool GLHandler::upload()
{
    PolyHandler* ph = data_->getPolyHandler();
    if (!ph || ph->getNoPolys() <= 0) {
        err_ = "no polygons";
        return false;
    }
    if (rj_set_polygons(ctx_, ph->polys.data(), ph->pindexes.data(), ph->getNoPolys(), ph->bounds) != RJ_OK) {
        err_ = rj_last_error(ctx_);
        return false;
    }
    std::set<int> used;
    for (auto& c : data_->constraints) used.insert(c.attribId);
    if (data_->aggr != Count) used.insert(data_->aggrAttrib);
    std::map<int, int> remap;
    std::vector<const void*> cols;
    std::vector<int32_t> types;
    for (int id : used) {
        auto it = data_->attributes.find(id);
        if (it == data_->attributes.end()) {
            err_ = "attribute " + std::to_string(id) + " not loaded";
            return false;
        }
        remap[id] = (int)cols.size();
        cols.push_back(it->second.data);
        types.push_back(it->second.type);
    }
    if (rj_set_points(ctx_, data_->n, data_->location, (int32_t)cols.size(), cols.data(), types.data()) != RJ_OK) {
        err_ = rj_last_error(ctx_);
        return false;
    }
    std::vector<rj_constraint> cons;
    for (auto& c : data_->constraints) {
        rj_constraint r;
        r.attrib_id = remap[c.attribId];
        r.type = c.type;
        r.val.uval = c.val.uval;
        r.dummy = 0;
        cons.push_back(r);
    }
    rj_set_constraints(ctx_, cons.data(), (int32_t)cons.size());
    rj_set_aggregation(ctx_, data_->aggr, data_->aggr == Count ? -1 : remap[data_->aggrAttrib]);
    dirty_ = false;
    return true;
}

std::vector<int> GLHandler::executeFunction(FunctionType fn)
{
    if (!data_) {
        err_ = "DataHandler not set.";
        return {};
    }
    if (dirty_ && !upload()) return {};
    const int np = data_->getPolyHandler()->getNoPolys();
    std::vector<int> legacy(3 * (size_t)np, 0);
    counts_.assign(np, 0);
    sums_.assign(np, 0.0);
    if (rj_execute(ctx_, fn, counts_.data(), sums_.data(), legacy.data()) != RJ_OK) {
        err_ = rj_last_error(ctx_);
        return {};
    }
    return legacy;
}

std::vector<int> GLHandler::getErrorBounds(FunctionType fn)
{
    if (fn != RasterJoinBoundFn || !data_) return {};
    const int np = data_->getPolyHandler()->getNoPolys();
    std::vector<int64_t> b(4 * (size_t)np);
    if (rj_error_bounds(ctx_, b.data()) != RJ_OK) return {};
    return std::vector<int>(b.begin(), b.end());
}

std::string GLHandler::printTimeStats(FunctionType fn)
{
    char buf[1024];
    if (rj_time_stats(ctx_, fn, buf, sizeof(buf)) != RJ_OK) return "";
    return buf;
}

}  // namespace rjcpp

// ---------------------------------------------------------------------------
// point records (Dataset::getNextRecord, db/Dataset.cpp:13-36)
// ---------------------------------------------------------------------------

extern "C" int rj_read_points(const char* path, int32_t dataset, int32_t loc_attrib, int32_t nattr,
                              const int32_t* attribs, int64_t max_records, int64_t* n_records, int64_t* n_kept,
                              float* xy, void* const* cols, int32_t* types)
{
    using namespace rjcpp;
    if (!path || !n_records || (dataset != RJ_DATASET_TAXI && dataset != RJ_DATASET_TWITTER) || nattr < 0 ||
        nattr > RJ_MAX_COLUMNS || (nattr > 0 && !attribs) || max_records < 0)
        return RJ_ERR_ARG;
    if (dataset == RJ_DATASET_TAXI ? (loc_attrib != 1 && loc_attrib != 2) : loc_attrib != 2) return RJ_ERR_ARG;
    const size_t rs = dataset == RJ_DATASET_TAXI ? kTaxiRecord : kTweetRecord;
    unsigned char probe[kTaxiRecord] = {0};
    for (int k = 0; k < nattr; k++) {
        uint32_t v;
        int32_t t;
        if (!attribute(dataset, attribs[k], probe, &v, &t)) return RJ_ERR_UNSUPPORTED;  // not a 4-byte attribute
    }
    FILE* f = std::fopen(path, "rb");
    if (!f) return RJ_ERR_ARG;
    std::fseek(f, 0, SEEK_END);
    const long bytes = std::ftell(f);
    std::fseek(f, 0, SEEK_SET);
    int64_t total = bytes > 0 ? (int64_t)bytes / (int64_t)rs : 0;
    if (max_records > 0) total = std::min(total, max_records);
    *n_records = total;
    if (!xy) {
        std::fclose(f);
        return RJ_OK;
    }
    if (!n_kept || (nattr > 0 && (!cols || !types))) {
        std::fclose(f);
        return RJ_ERR_ARG;
    }
    int64_t kept = 0;
    std::vector<unsigned char> buf(rs * 65536);
    for (int64_t done = 0; done < total;) {
        const int64_t n = std::min<int64_t>(65536, total - done);
        if (std::fread(buf.data(), rs, (size_t)n, f) != (size_t)n) {
            std::fclose(f);
            return RJ_ERR_ARG;
        }
        for (int64_t i = 0; i < n; i++) {
            const unsigned char* r = buf.data() + (size_t)i * rs;
            // index location: (lat, lon) as floats (TaxiRecord.hpp:277-285, TwitterRecord.hpp:103-110);
            // a kept record's location becomes the projected one (TaxiRecord.hpp:299-302)
            const float lat = dataset == RJ_DATASET_TAXI ? (float)field<double>(r, 16) : field<float>(r, 16);
            const float lon = dataset == RJ_DATASET_TAXI ? (float)field<double>(r, 8) : field<float>(r, 20);
            float tx, ty;
            if (!transform_point(lat, lon, &tx, &ty)) continue;  // outside the US bbox: skipped
            if (dataset == RJ_DATASET_TAXI && loc_attrib == 2) {
                // the dropoff location is not projected (TaxiRecord.hpp:134-145)
                tx = (float)field<double>(r, 32);
                ty = (float)field<double>(r, 24);
            }
            xy[2 * kept] = tx;
            xy[2 * kept + 1] = ty;
            for (int k = 0; k < nattr; k++) {
                uint32_t v;
                attribute(dataset, attribs[k], r, &v, &types[k]);
                static_cast<uint32_t*>(cols[k])[kept] = v;
            }
            kept++;
        }
        done += n;
    }
    std::fclose(f);
    *n_kept = kept;
    return RJ_OK;
}
