// cli_main.cpp -- `raster_join`, the --joinType driver (RasterJoin/main.cpp).
//
// Same flags, flow and outputs as the reference driver: parse --key value
// flags (main.cpp:254-351, 378-397), set up the data and the join handler
// (main.cpp:120-167), run nIter times (:169-176), then write the per-polygon
// CSV (:195-225), the error-bounds CSV (:178-193) and the timing line
// (:227-252). Points come from the seeded synthetic table (the disk backend
// index is not part of this build); polygons from the reference's polygon
// text format or from the Voronoi generator.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <map>
#include <string>
#include <vector>

#include "glhandler.hpp"
#include "rasterjoin_datagen.h"

using namespace rjcpp;

namespace {

// taxi attribute ids of the reference (TaxiRecord.hpp:60-105)
constexpr int kTimeAttr = 3, kFareAttr = 8, kPassengersAttr = 18;

std::map<std::string, std::string> parse(int argc, char** argv)
{
    std::map<std::string, std::string> args;
    for (int i = 1; i < argc; i++) {
        std::string a = argv[i];
        if (a.rfind("--", 0) != 0) continue;
        std::string v = " ";
        if (i + 1 < argc && std::string(argv[i + 1]).rfind("--", 0) != 0) v = argv[i + 1];
        args[a] = v;
    }
    return args;
}

double ground_resolution_us()
{
    // getGroundResolution() at the US bbox center, zoom 22 (UsefulFuncs.cpp:625-633, Utils.h:25-27)
    const double lat = (24.396308 + 49.384358) / 2;
    return std::cos(lat * M_PI / 180) * 6378137 * 2 * M_PI / std::exp2(8 + 22.0);
}

int usage()
{
    std::cout << "usage: raster_join --joinType raster|errorbounds|hybrid|index [--accuracy m] [--nIter k]\n"
                 "  [--indexRes r] [--nAttrib k] [--startTime t --endTime t] [--avg attrib] [--gpuMem MB]\n"
                 "  [--inmem] [--opAggregation dir] [--outputTime file] [--inputSize n] [--pointSeed s]\n"
                 "  (--polygonList file --polygonDataset name [--polyTransform] | --synthPolys sites)\n"
                 "  [--groundRes m_per_unit]\n";
    return 1;
}

}  // namespace

int main(int argc, char** argv)
{
    auto args = parse(argc, argv);
    auto has = [&](const char* k) { return args.count(k) > 0; };
    auto num = [&](const char* k, double d) { return has(k) ? std::atof(args[k].c_str()) : d; };
    if (has("--help") || !has("--joinType")) return usage();
    const std::string jt = args["--joinType"];
    GLHandler::FunctionType fn;
    if (jt == "raster") fn = GLHandler::RasterJoinFn;
    else if (jt == "hybrid") fn = GLHandler::HybridJoinFn;
    else if (jt == "index") fn = GLHandler::IndexJoinFn;
    else if (jt == "errorbounds") fn = GLHandler::RasterJoinBoundFn;
    else {
        std::cerr << "Wrong join type " << std::endl;
        return 1;
    }
    if (has("--backendIndexName")) {
        std::cerr << "the disk backend index is not part of this build: points come from the synthetic "
                     "table (--inputSize, --pointSeed)\n";
        return 1;
    }
    const int nIter = (int)num("--nIter", 1);
    int accuracy = (int)num("--accuracy", 100);
    const int gridRes = (int)num("--indexRes", 256);
    const int nAttrib = (int)num("--nAttrib", 0);
    const int64_t n = (int64_t)num("--inputSize", 1000000);
    const int aggrAttrib = has("--avg") ? (int)num("--avg", -1) : -1;
    const int64_t gpuMem = (int64_t)num("--gpuMem", 2048);

    // polygons
    PolyHandler poly;
    std::string err;
    double groundRes = 1.0;
    if (has("--polygonList")) {
        bool tr = has("--polyTransform");
        if (!poly.initFromFile(args["--polygonList"], args["--polygonDataset"], tr, &err)) {
            std::cerr << err << "\n";
            return 1;
        }
        if (tr) groundRes = ground_resolution_us();
    } else {
        rj_voronoi_params vp{(uint64_t)num("--polySeed", 1), (int32_t)num("--synthPolys", 260), 50000.0,
                             num("--polySpacing", 60.0), num("--polySpacing", 60.0), 30.0};
        rj_polys* h = nullptr;
        if (rj_gen_voronoi(&vp, &h) != 0) {
            std::cerr << "polygon generation failed\n";
            return 1;
        }
        int32_t np;
        int64_t nv;
        rj_polys_size(h, &np, &nv);
        std::vector<float> xy(2 * nv);
        std::vector<int32_t> off(np + 1);
        rj_polys_copy(h, xy.data(), off.data());
        rj_polys_free(h);
        poly.setPolygons(std::move(xy), std::move(off));
    }
    groundRes = num("--groundRes", groundRes);

    // points (synthetic table with the taxi attribute ids)
    rj_point_params pp{(uint64_t)num("--pointSeed", 7), 50000.0, 40, 0.7, 200.0, 2000.0, 1230768000u,
                       31536000u, 2.5, 0.6};
    std::vector<float> xy(2 * n), fare(n);
    std::vector<uint32_t> tm(n);
    std::vector<uint8_t> pas(n);
    rj_gen_points(&pp, 0, n, xy.data(), tm.data(), fare.data(), pas.data());
    DataHandler data;
    data.poly = &poly;
    data.setLocation(n, xy.data());
    data.addAttribute(kTimeAttr, Uint, tm.data());
    data.addAttribute(kFareAttr, Float, fare.data());
    data.addAttribute(kPassengersAttr, Uint8, pas.data());

    // constraints: exact time window, then the taxi filters of main.cpp:82-118
    std::vector<QueryConstraint> cons;
    uint32_t start_time = 0, end_time = 0;
    if (has("--startTime")) {
        start_time = (uint32_t)num("--startTime", 0);
        QueryConstraint c{kTimeAttr, GTE, {0}, 0};
        c.val.uval = start_time;
        cons.push_back(c);
    }
    if (has("--endTime")) {
        end_time = (uint32_t)num("--endTime", 0);
        QueryConstraint c{kTimeAttr, LTE, {0}, 0};
        c.val.uval = end_time;
        cons.push_back(c);
    }
    if (nAttrib >= 1) {
        QueryConstraint c{kPassengersAttr, EQ, {0}, 0};
        c.val.uval = 1;
        cons.push_back(c);
    }
    if (nAttrib >= 2) {  // fare > 1000 cents in the reference; dollars here
        QueryConstraint c{kFareAttr, GT, {0}, 0};
        c.val.fval = 10.0f;
        cons.push_back(c);
    }
    if (nAttrib >= 3) std::cerr << "--nAttrib > 2: the synthetic table has no miles/duration columns\n";
    data.setQueryConstraints(cons);
    if (aggrAttrib != -1) data.setAggregation(Avg, aggrAttrib);

    GLHandler* handler = GLHandler::getInstance(gpuMem, has("--inmem"));
    if (!handler) {
        std::cout << "Failed to obtain handler";
        return 1;
    }
    handler->setDataHandler(&data);
    handler->setAccuracyDistance(accuracy, groundRes);
    handler->setPolyIndexResolution(gridRes, gridRes);

    std::vector<int> agg, bounds;
    for (int k = 0; k < nIter; k++) {
        agg = handler->executeFunction(fn);
        if (agg.empty()) {
            std::cerr << "execute failed: " << handler->lastError() << "\n";
            return 1;
        }
        bounds = handler->getErrorBounds(fn);
    }

    // outputs (main.cpp:178-252)
    if (has("--opAggregation")) {
        const std::string dir = args["--opAggregation"];
        if (fn != GLHandler::RasterJoinFn && fn != GLHandler::RasterJoinBoundFn) accuracy = 0;
        const int np = poly.getNoPolys();
        const auto& counts = handler->counts();
        const auto& sums = handler->sums();
        if (fn == GLHandler::RasterJoinBoundFn) {
            std::ofstream op(dir + "/err_bound_" + std::to_string(end_time) + "_" + std::to_string(accuracy) + ".csv");
            for (int i = 0; i < np; i++) {
                int64_t fnv = bounds[i], fpv = bounds[i + np], fn1 = bounds[i + 2 * np], fp1 = bounds[i + 3 * np];
                op << i << "\t" << (counts[i] - fp1) << "\t" << (counts[i] + fn1) << "\t" << (counts[i] - fpv)
                   << "\t" << (counts[i] + fnv) << "\n";
            }
        } else {
            std::ofstream op(dir + "/raster_" + std::to_string(end_time) + "_" + std::to_string(accuracy) + ".csv");
            op.precision(10);
            for (int i = 0; i < np; i++) {
                if (aggrAttrib == -1) op << i << "\t" << counts[i] << "\n";
                else op << i << "\t" << (counts[i] == 0 ? 0.0 : sums[i] / (double)counts[i]) << "\n";
            }
        }
    }
    std::string timing = handler->printTimeStats(fn);
    std::cout << std::endl << timing << std::endl;
    if (has("--outputTime")) {
        std::ofstream fi(args["--outputTime"], std::ios::app);
        fi << timing << "\n";
    }
    GLHandler::destroyInstance();
    return 0;
}
