// ref_harness.cpp -- a C entry-point driver around the reference's own
// CPUJoin/GridIndexF.cpp, compiled from the reference tree into oracle/_ref/.
//
// TEST INFRASTRUCTURE ONLY: used by tests/ to pin the oracle's restatement of
// the grid index and of the per-point join against the reference's compiled
// code. Nothing in the product links it.
//
// The driver mirrors the call sequence of CPUJoin::setupPolyIndex
// (CPUJoin/CPUJoin.cpp:30-56: per-polygon PolygonF + BoundF, GridIndexF(bound,
// res, res), buildGrid) and the counting loop of CPUJoin::execute
// (CPUJoin.cpp:91-105: getRegion per point, one increment per returned id).
#include "GridIndexF.hpp"

#include <QString>
#include <QVector>
#include <algorithm>
#include <cstdint>
#include <vector>

struct RefGrid {
    GridIndexF grid;
    std::vector<int32_t> cell_off;
    std::vector<int32_t> cand;
};

extern "C" {

void* ref_grid_build(const float* xy, const int32_t* off, int32_t np, const float bbox[4],
                     int32_t res, const char* index_prefix)
{
    QVector<PolygonF> polys;
    QVector<BoundF> bounds;
    for (int i = 0; i < np; i++) {
        PolygonF poly;
        BoundF b;
        for (int j = off[i]; j < off[i + 1]; j++) {
            poly << PointF(xy[2 * j], xy[2 * j + 1]);
            b.updateBound(xy[2 * j], xy[2 * j + 1]);
        }
        polys << poly;
        bounds << b;
    }
    BoundF bf;
    bf.minx = bbox[0];
    bf.miny = bbox[1];
    bf.maxx = bbox[2];
    bf.maxy = bbox[3];
    RefGrid* g = new RefGrid();
    g->grid = GridIndexF(bf, res, res);
    g->grid.buildGrid(polys, bounds, QString(index_prefix));
    int xs = g->grid.xs, ys = g->grid.ys;
    g->cell_off.assign((size_t)xs * ys + 1, 0);
    for (int y = 0; y < ys; y++) {
        for (int x = 0; x < xs; x++) {
            std::vector<int32_t> c(g->grid.grid[x][y].begin(), g->grid.grid[x][y].end());
            std::sort(c.begin(), c.end());
            g->cand.insert(g->cand.end(), c.begin(), c.end());
            g->cell_off[(size_t)y * xs + x + 1] = (int32_t)g->cand.size();
        }
    }
    return g;
}

int64_t ref_grid_ncand(void* h) { return (int64_t)((RefGrid*)h)->cand.size(); }

void ref_grid_export(void* h, int32_t* cell_off, int32_t* cand)
{
    RefGrid* g = (RefGrid*)h;
    std::copy(g->cell_off.begin(), g->cell_off.end(), cell_off);
    std::copy(g->cand.begin(), g->cand.end(), cand);
}

// Counts per polygon over n points (x,y interleaved), CPUJoin::execute's loop.
int ref_join_count(void* h, int64_t n, const float* pts, int64_t* counts)
{
    RefGrid* g = (RefGrid*)h;
    int np = g->grid.noPolys;
    std::fill(counts, counts + np, 0);
    for (int64_t i = 0; i < n; i++) {
        QVector<int> res = g->grid.getRegion(pts[2 * i], pts[2 * i + 1]);
        for (int j = 0; j < res.size(); j++) counts[res[j]]++;
    }
    return 0;
}

void ref_grid_free(void* h) { delete (RefGrid*)h; }
}
