// cli_main.cpp -- `raster_join`, the --joinType driver (RasterJoin/main.cpp).
//
// Same flags, flow and outputs as the reference driver: parse --key value
// flags (main.cpp:254-351, 378-397), set up the data and the join handler
// (main.cpp:120-167), run nIter times (:169-176), then write the per-polygon
// CSV (:195-225), the error-bounds CSV (:178-193) and the timing line
// (:227-252). Points come from a file of the reference's binary point records
// (--pointFile, read as Dataset::getNextRecord does, db/Dataset.cpp:13-36,
// taxi or twitter records, --locAttrib as main.cpp:290-292) in place of the
// disk backend index, or from the seeded synthetic table; polygons from the
// reference's polygon text format or from the Voronoi generator.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <map>
#include <set>
#include <string>
#include <vector>

#include "glhandler.hpp"
#include "rasterjoin_datagen.h"

using namespace rjcpp;

namespace {

// taxi attribute ids of the reference (TaxiRecord.hpp:60-105); twitter's time (TwitterRecord.hpp:44)
constexpr int kTimeAttr = 3, kDurationAttr = 6, kMilesAttr = 7, kFareAttr = 8, kPassengersAttr = 18;
constexpr int kTweetTimeAttr = 1;

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
                 "  [--inmem] [--opAggregation dir] [--outputTime file] [--inputSize n]\n"
                 "  (--pointFile records [--pointDataset taxi|twitter] [--locAttrib id] | --pointSeed s)\n"
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
        std::cerr << "the disk backend index is not part of this build: points come from a record file "
                     "(--pointFile) or the synthetic table (--inputSize, --pointSeed)\n";
        return 1;
    }
    const bool file_points = has("--pointFile");
    const bool twitter = has("--pointDataset") && args["--pointDataset"] == "twitter";
    if (has("--pointDataset") && !twitter && args["--pointDataset"] != "taxi") {
        std::cerr << "--pointDataset is taxi or twitter\n";
        return 1;
    }
    const int nIter = (int)num("--nIter", 1);
    int accuracy = (int)num("--accuracy", 100);
    const int gridRes = (int)num("--indexRes", 256);
    const int nAttrib = (int)num("--nAttrib", 0);
    int64_t n = (int64_t)num("--inputSize", file_points ? 0 : 1000000);
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
    // record files hold Web-Mercator coordinates: the reference's ground
    // resolution at the US center (GLFunction::setMaxCellSize)
    if (has("--pointFile")) groundRes = ground_resolution_us();
    groundRes = num("--groundRes", groundRes);

    // constraints: exact time window, then the taxi filters of main.cpp:82-118
    // (the synthetic table holds fare in dollars as floats and no miles or
    // duration; record files hold the reference's columns)
    const int time_attr = twitter ? kTweetTimeAttr : kTimeAttr;
    std::vector<QueryConstraint> cons;
    uint32_t start_time = 0, end_time = 0;
    if (has("--startTime")) {
        start_time = (uint32_t)num("--startTime", 0);
        QueryConstraint c{time_attr, GTE, {0}, 0};
        c.val.uval = start_time;
        cons.push_back(c);
    }
    if (has("--endTime")) {
        end_time = (uint32_t)num("--endTime", 0);
        QueryConstraint c{time_attr, LTE, {0}, 0};
        c.val.uval = end_time;
        cons.push_back(c);
    }
    if (nAttrib > 0 && twitter) std::cerr << "--nAttrib: the hard-coded filters work only for taxi (main.cpp:83)\n";
    if (nAttrib >= 1 && !twitter) {
        QueryConstraint c{kPassengersAttr, EQ, {0}, 0};
        c.val.uval = 1;
        cons.push_back(c);
    }
    if (nAttrib >= 2 && !twitter) {  // fare > 1000 cents (records) or > 10 dollars (synthetic)
        QueryConstraint c{kFareAttr, GT, {0}, 0};
        if (file_points) c.val.uval = 1000;
        else c.val.fval = 10.0f;
        cons.push_back(c);
    }
    if (nAttrib >= 3 && !twitter) {
        if (!file_points) {
            std::cerr << "--nAttrib > 2: the synthetic table has no miles/duration columns\n";
        } else {
            QueryConstraint c{kMilesAttr, LT, {0}, 0};
            c.val.fval = 10.0f;
            cons.push_back(c);
        }
    }
    if (nAttrib >= 4 && !twitter && file_points) {
        QueryConstraint c{kDurationAttr, GT, {0}, 0};
        c.val.uval = 20;
        cons.push_back(c);
    }

    // points
    DataHandler data;
    data.poly = &poly;
    std::vector<float> xy;
    std::vector<std::vector<uint32_t>> fcols;
    std::vector<float> fare;
    std::vector<uint32_t> tm;
    std::vector<uint8_t> pas;
    if (file_points) {
        // the location, then every attribute a filter or the aggregate reads
        // (DataHandler::initData, DataHandler.cpp:24-59)
        std::set<int> ids;
        for (auto& c : cons) ids.insert(c.attribId);
        if (aggrAttrib != -1) ids.insert(aggrAttrib);
        std::vector<int32_t> attribs(ids.begin(), ids.end());
        const int32_t loc = (int32_t)num("--locAttrib", twitter ? 2 : 1);
        const int32_t ds = twitter ? RJ_DATASET_TWITTER : RJ_DATASET_TAXI;
        const std::string path = args["--pointFile"];
        int64_t nrec = 0, kept = 0;
        int rc = rj_read_points(path.c_str(), ds, loc, (int32_t)attribs.size(), attribs.data(), n, &nrec, nullptr,
                                nullptr, nullptr, nullptr);
        if (rc == RJ_OK) {
            xy.resize(2 * (size_t)nrec);
            fcols.assign(attribs.size(), std::vector<uint32_t>(nrec));
            std::vector<void*> ptrs;
            for (auto& c : fcols) ptrs.push_back(c.data());
            std::vector<int32_t> types(attribs.size());
            rc = rj_read_points(path.c_str(), ds, loc, (int32_t)attribs.size(), attribs.data(), n, &nrec, &kept,
                                xy.data(), ptrs.data(), types.data());
            for (size_t k = 0; k < attribs.size() && rc == RJ_OK; k++)
                data.addAttribute(attribs[k], (AttributeType)types[k], fcols[k].data());
        }
        if (rc != RJ_OK) {
            std::cerr << "could not read the point records of " << path << " (status " << rc << ")\n";
            return 1;
        }
        n = kept;
        std::cerr << "read " << nrec << " records, " << kept << " in the US bounding box\n";
    } else {
        // synthetic table with the taxi attribute ids
        rj_point_params pp{(uint64_t)num("--pointSeed", 7), 50000.0, 40, 0.7, 200.0, 2000.0, 1230768000u,
                           31536000u, 2.5, 0.6};
        xy.resize(2 * n);
        fare.resize(n);
        tm.resize(n);
        pas.resize(n);
        rj_gen_points(&pp, 0, n, xy.data(), tm.data(), fare.data(), pas.data());
        data.addAttribute(kTimeAttr, Uint, tm.data());
        data.addAttribute(kFareAttr, Float, fare.data());
        data.addAttribute(kPassengersAttr, Uint8, pas.data());
    }
    data.setLocation(n, xy.data());
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
