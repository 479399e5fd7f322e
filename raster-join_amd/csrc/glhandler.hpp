// glhandler.hpp -- C++ front end with the shape of the reference's join API,
// so a driver written against RasterJoin/join/GLHandler.hpp:20-78 keeps its
// flow: getInstance -> setDataHandler -> setAccuracyDistance ->
// setPolyIndexResolution -> executeFunction / getErrorBounds / printTimeStats.
// std::vector replaces QVector; everything runs through the C ABI.
#pragma once
#include <cstdint>
#include <map>
#include <string>
#include <vector>

#include "rasterjoin.h"

namespace rjcpp {

enum AttributeType { Location = 0, Uint = 1, Int = 2, Float = 3, Uint8 = 4 };  // Common.h:20
enum ConstraintType { EQ = 0, LT, LTE, GT, GTE };                            // Common.h:22
enum Aggregation { Count = 0, Avg, Sum };                                     // Common.h:26

// QueryConstraint (Common.h:34-39): attribId is a dataset attribute id.
struct QueryConstraint {
    int attribId;
    ConstraintType type;
    union { int32_t ival; float fval; uint32_t uval; } val;
    int dummy;
};

// A polygon collection (db/PolyHandler.hpp:29-57): open rings in CSR form.
struct PolyHandler {
    std::vector<float> polys;       // x,y per vertex
    std::vector<int32_t> pindexes;  // np+1 offsets
    double bounds[4] = {0, 0, 0, 0};
    int getNoPolys() const { return (int)pindexes.size() - 1; }
    // PolyHandler::readPolygons text format (PolyHandler.cpp:133-188): count,
    // then per polygon the number of parts and per part n then n "x y" pairs.
    // transform = true applies the reference's Web-Mercator projection with its
    // US bounding-box filter (common/UsefulFuncs.cpp:578-612).
    bool readPolygons(const std::string& file, bool transform, std::string* err);
    // `name,file` collection list (PolyHandler::initFromFile, PolyHandler.cpp:20-40).
    bool initFromFile(const std::string& list, const std::string& collection, bool transform, std::string* err);
    void setPolygons(std::vector<float> xy, std::vector<int32_t> off);
};

// Point columns, filters and aggregate (db/DataHandler.hpp:14-41).
struct DataHandler {
    struct Column {
        AttributeType type;
        const void* data;
    };
    int64_t n = 0;
    const float* location = nullptr;
    std::map<int, Column> attributes;   // dataset attribute id -> column
    std::vector<QueryConstraint> constraints;
    Aggregation aggr = Count;
    int aggrAttrib = -1;
    PolyHandler* poly = nullptr;

    void setLocation(int64_t npts, const float* xy) { n = npts; location = xy; }
    void addAttribute(int id, AttributeType t, const void* data) { attributes[id] = {t, data}; }
    void setQueryConstraints(std::vector<QueryConstraint> c) { constraints = std::move(c); }
    void setAggregation(Aggregation type = Count, int attribId = -1) { aggr = type; aggrAttrib = attribId; }
    PolyHandler* getPolyHandler() const { return poly; }
};

class GLHandler {
public:
    enum FunctionType { RasterJoinFn = 0, IndexJoinFn, HybridJoinFn, RasterJoinBoundFn, FunctionCount };

    static GLHandler* getInstance(int64_t gpuMemMB = 2048LL, bool inMemory = false);
    static void destroyInstance();

    void setDataHandler(DataHandler* dataHandler);
    void setAccuracyDistance(double meters, double groundRes = 1.0);
    void setPolyIndexResolution(int resx, int resy);
    // 3*np int32: [count | sum/100 | decimal*10], empty on error.
    std::vector<int> executeFunction(FunctionType fn);
    std::vector<int> getErrorBounds(FunctionType fn);  // 4*np or empty
    std::string printTimeStats(FunctionType fn);

    // exact results of the last executeFunction
    const std::vector<int64_t>& counts() const { return counts_; }
    const std::vector<double>& sums() const { return sums_; }
    const std::string& lastError() const { return err_; }
    rj_ctx* context() const { return ctx_; }

private:
    GLHandler(int64_t gpuMemMB, bool inMemory);
    ~GLHandler();
    bool upload();

    static GLHandler* instance;
    rj_ctx* ctx_ = nullptr;
    DataHandler* data_ = nullptr;
    bool dirty_ = true;
    std::vector<int64_t> counts_;
    std::vector<double> sums_;
    std::string err_;
};

}  // namespace rjcpp
