/*
 * Synthetic directed weighted graphs for the clustered betweenness benchmarks.
 *
 * The reference reads graphs only as '<src> <dst> <weight>' edge lists
 * (libfastbc/include/DirectedWeightedGraph.h:67-96) and ships no generator, so
 * this file defines the workloads precisely.  Every random draw comes from one
 * PCG32 stream (XSH-RR 64/32, O'Neill 2014) seeded with
 * pcg32_srandom(initstate = seed, initseq = 54); bounded integers use the
 * unbiased rejection method of pcg32_boundedrand.
 *
 * Grid (R x C, id = r*C + c):
 *   1. Undirected links are enumerated in vertex order; for vertex v the link
 *      to v+1 (same row) comes before the link to v+C (next row).
 *   2. Road-like grids only: one PCG32 key per link in link order; links sorted
 *      by (key, index) feed Kruskal's algorithm, giving a spanning tree that
 *      keeps the graph strongly connected.  Then llround(0.10 * links) non-tree
 *      links are deleted, chosen by a partial Fisher-Yates shuffle of the
 *      non-tree links (in link order) with bounded draws.
 *   3. Each surviving link (a < b) gets w(a->b) = 1 + bounded(wmax), then
 *      w(b->a) = 1 + bounded(wmax).
 *   4. Road-like grids only: llround(0.02 * arcs) diagonal arcs.  Each attempt
 *      draws v = bounded(N) then a direction bounded(4) from
 *      {(-1,-1), (-1,+1), (+1,-1), (+1,+1)}; off-grid targets and arcs that
 *      already exist are skipped (the attempt's draws are spent).  An accepted
 *      arc draws u = 1 + bounded(wmax) and gets weight ceil(1.414 * u).
 *   5. Arcs are emitted sorted by (src, dst).
 *
 * R-MAT (2^scale vertices, ef * 2^scale draws, quadrant probabilities a, b, c,
 * d = 1-a-b-c): each draw descends 'scale' levels, most significant bit first;
 * each level draws x = pcg32() * 2^-32 and picks (0,0) if x < a, (0,1) if
 * x < a+b, (1,0) if x < a+b+c, else (1,1).  Self loops are dropped and repeated
 * arcs keep their first occurrence; labels are not permuted.  Arcs are sorted
 * by (src, dst) and then weighted 1 + bounded(wmax) in that order.  If vertex
 * N-1 has no out-arc, the arc (N-1 -> 0) is appended with one more weight
 * draw, so that the loader's vertex count (1 + largest source id) is N.
 */
#ifndef FBC_GEN_H
#define FBC_GEN_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct fbcgen_arcs {
    uint32_t n;      /* vertex count (1 + largest source id) */
    uint64_t m;      /* arc count */
    uint32_t *src;   /* sorted by (src, dst) */
    uint32_t *dst;
    uint32_t *w;     /* positive integer weights */
} fbcgen_arcs;

/* road != 0 adds the spanning-tree guarded 10% link deletion and the 2%
 * diagonal arcs.  Returns 0 on success, negative on bad arguments / OOM. */
int fbcgen_grid(uint32_t rows, uint32_t cols, uint32_t wmax, uint64_t seed,
                int road, fbcgen_arcs *out);

int fbcgen_rmat(uint32_t scale, uint32_t edge_factor, double a, double b,
                double c, uint32_t wmax, uint64_t seed, fbcgen_arcs *out);

/* The five benchmark workloads: cfg 0..4 (see fbc_gen.c). */
int fbcgen_config(int cfg, fbcgen_arcs *out);

/* Writes '<src> <dst> <w>\n' lines.  Returns 0 or a negative errno value. */
int fbcgen_write(const fbcgen_arcs *arcs, const char *path);

void fbcgen_free(fbcgen_arcs *arcs);

#ifdef __cplusplus
}
#endif

#endif
