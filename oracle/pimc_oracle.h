/*
 * pimc_oracle.h -- test infrastructure only.
 *
 * The oracle library (liboracle_pimc.so) is a plain-C++ CPU restatement of
 * PIMC++'s action hot path.  It implements every entry point of
 * include/pimc_gpu.h in FP64 (the `precision` argument is accepted and
 * ignored) and adds the hooks below, which expose its building blocks to the
 * tests.  Only the tests, the smoke check and the benchmark's CPU-baseline leg
 * load it; the product library never does.
 */
#ifndef PIMC_ORACLE_H
#define PIMC_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Philox4x32-10 block: out[4] = philox(ctr[4], key[2]). */
void oracle_philox4x32_10(const uint32_t* ctr, const uint32_t* key, uint32_t* out);
/* Uniform double in [0,1) from two 32-bit words (w1 high). */
double oracle_uniform(uint32_t w0, uint32_t w1);
/* Gaussian pair `pair` of (seed, stream, move, level): writes g[2]. Returns attempts used. */
int oracle_gaussian_pair(uint64_t seed, uint32_t stream, uint32_t move, int32_t level,
                         uint32_t pair, double* g);

/* Grid::ReverseMap of a LINEAR (type 0) or LOG (type 1) grid built like the
 * .dm reader builds it; literal = 1 returns the reference's formula, 0 the
 * knot-corrected index both this oracle and the GPU library use. */
int oracle_reverse_map(int32_t type, double start, double end, int32_t npoints,
                       int32_t literal, int64_t n, const double* r, int32_t* out);
/* Grid points of the same grid. */
int oracle_grid_points(int32_t type, double start, double end, int32_t npoints, double* x);

/* MultiCubicSpline on a grid: y is [npoints][ncols]; start/end derivative per
 * column (> 0.99e30 means natural).  Writes d2y [npoints][ncols] and, for n
 * points x, the value of column `col` into out[n] (col < 0: all columns,
 * out is [n][ncols]). */
int oracle_multicubic(int32_t type, double start, double end, int32_t npoints, int32_t ncols,
                      const double* y, const double* start_deriv, const double* end_deriv,
                      double* d2y, int32_t col, int64_t n, const double* x, double* out);

/* Everything the oracle's bisection move did in NEWMODE sampling, for tests:
 * the bisection Sample of one stage on a replica (no action, no accept). */
int oracle_bisection_sample(void* ctx, int32_t replica, int32_t ptcl, int32_t slice1,
                            int32_t num_levels, int32_t level, uint64_t seed, uint32_t move,
                            double* log_sample_ratio);

/* PermuteTableClass::ConstructBosonCycleTable over an n x n float HTable (row-major), no
 * excluded particle: writes up to cap entries (len[e], rep[e][4] species-relative, P[e], C[e])
 * in the table's order and returns the entry count (which may exceed cap); *norm = the last C.
 * With xi >= 0, *found = FindEntry(xi) over the whole table. */
int oracle_cycle_table(const float* htable, int32_t n, const double* gamma, double epsilon, int32_t cap,
                       int32_t* len, int32_t* rep, float* P, float* C, float* norm, double xi, int32_t* found);

#ifdef __cplusplus
}
#endif

#endif
