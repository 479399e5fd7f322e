/* rasterjoin_datagen.h -- seeded synthetic inputs for the benchmark
 * configurations (clustered points with time/fare/passengers columns, and
 * Voronoi polygon sets with shared jittered edges).
 *
 * These entry points have no counterpart in the reference, whose inputs come
 * from the NYC taxi / Twitter backend index (RasterJoin/db/HashGridIndex.cpp,
 * db/PolyHandler.cpp:133-188). They exist so the same inputs can be produced
 * on the host of any machine without a dataset download. Pure CPU code.
 */
#ifndef RASTERJOIN_DATAGEN_H
#define RASTERJOIN_DATAGEN_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct rj_point_params {
    uint64_t seed;
    double box;          /* square [0, box)^2 in projected meters */
    int32_t ncenters;    /* Gaussian clusters */
    double frac_cluster; /* share of clustered points (0.7) */
    double sigma_lo, sigma_hi; /* cluster sigma range (200, 2000) */
    uint32_t t0, tspan;  /* time = t0 + U[0, tspan) */
    double fare_mu, fare_sigma; /* lognormal parameters (2.5, 0.6) */
} rj_point_params;

/* Rows [first, first+n) of the seeded table. Any output pointer may be NULL.
 * xy: 2n floats interleaved. Returns 0 on success. */
int rj_gen_points(const rj_point_params* p, int64_t first, int64_t n, float* xy, uint32_t* time,
                  float* fare, uint8_t* passengers);

typedef struct rj_voronoi_params {
    uint64_t seed;
    int32_t nsites;
    double box;
    double spacing_lo, spacing_hi; /* subdivision spacing, log-uniform per edge */
    double jitter;                  /* max perpendicular jitter of shared edges */
} rj_voronoi_params;

typedef struct rj_polys rj_polys;
int rj_gen_voronoi(const rj_voronoi_params* p, rj_polys** out);
int rj_polys_size(const rj_polys* p, int32_t* np, int64_t* nverts);
/* xy: 2*nverts floats; off: np+1 offsets (CSR). */
int rj_polys_copy(const rj_polys* p, float* xy, int32_t* off);
void rj_polys_free(rj_polys* p);

#ifdef __cplusplus
}
#endif
#endif
