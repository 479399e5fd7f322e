// glhandler.cpp -- the GLHandler-shaped C++ front end over the C ABI, and the
// polygon text reader of the reference's PolyHandler.
#include "glhandler.hpp"

#include <cmath>
#include <cstdio>
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
