/* Synthetic graph generators; the procedure is specified in fbc_gen.h. */
#include "fbc_gen.h"

#include <errno.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

/* ---- PCG32 (XSH-RR 64/32), O'Neill 2014 ---------------------------------- */
typedef struct { uint64_t state, inc; } pcg32;

static uint32_t pcg32_next(pcg32 *g)
{
    uint64_t old = g->state;
    g->state = old * 6364136223846793005ULL + g->inc;
    uint32_t xorshifted = (uint32_t)(((old >> 18u) ^ old) >> 27u);
    uint32_t rot = (uint32_t)(old >> 59u);
    return (xorshifted >> rot) | (xorshifted << ((-rot) & 31u));
}

static void pcg32_seed(pcg32 *g, uint64_t initstate, uint64_t initseq)
{
    g->state = 0u;
    g->inc = (initseq << 1u) | 1u;
    pcg32_next(g);
    g->state += initstate;
    pcg32_next(g);
}

/* Unbiased value in [0, bound). */
static uint32_t pcg32_bounded(pcg32 *g, uint32_t bound)
{
    uint32_t threshold = (uint32_t)(-bound) % bound;
    for (;;) {
        uint32_t r = pcg32_next(g);
        if (r >= threshold) return r % bound;
    }
}

/* ---- helpers ------------------------------------------------------------- */

/* LSD radix sort of (key, payload) pairs on the low 'bits' bits of key. */
static int radix_sort_pairs(uint64_t *key, uint32_t *pay, size_t n, int bits)
{
    uint64_t *tmpk = (uint64_t *)malloc(n * sizeof(uint64_t) + 8);
    uint32_t *tmpp = (uint32_t *)malloc(n * sizeof(uint32_t) + 4);
    size_t *cnt = (size_t *)malloc(65536 * sizeof(size_t));
    if (!tmpk || !tmpp || !cnt) { free(tmpk); free(tmpp); free(cnt); return -ENOMEM; }
    uint64_t *ks = key, *kd = tmpk;
    uint32_t *ps = pay, *pd = tmpp;
    for (int shift = 0; shift < bits; shift += 16) {
        memset(cnt, 0, 65536 * sizeof(size_t));
        for (size_t i = 0; i < n; ++i) cnt[(ks[i] >> shift) & 0xffff]++;
        size_t run = 0;
        for (int d = 0; d < 65536; ++d) { size_t c = cnt[d]; cnt[d] = run; run += c; }
        for (size_t i = 0; i < n; ++i) {
            size_t pos = cnt[(ks[i] >> shift) & 0xffff]++;
            kd[pos] = ks[i];
            pd[pos] = ps[i];
        }
        uint64_t *tk = ks; ks = kd; kd = tk;
        uint32_t *tp = ps; ps = pd; pd = tp;
    }
    if (ks != key) {
        memcpy(key, ks, n * sizeof(uint64_t));
        memcpy(pay, ps, n * sizeof(uint32_t));
    }
    free(tmpk); free(tmpp); free(cnt);
    return 0;
}

static int bits_for(uint64_t maxval)
{
    int b = 0;
    while (b < 64 && (maxval >> b)) ++b;
    return b < 16 ? 16 : b;
}

static uint32_t uf_find(uint32_t *parent, uint32_t x)
{
    while (parent[x] != x) {
        parent[x] = parent[parent[x]];
        x = parent[x];
    }
    return x;
}

static int finish_arcs(uint64_t *key, uint32_t *w, size_t m, uint32_t n,
                       fbcgen_arcs *out)
{
    int rc = radix_sort_pairs(key, w, m, bits_for(((uint64_t)n << 32) | n));
    if (rc) return rc;
    out->n = n;
    out->m = m;
    out->src = (uint32_t *)malloc(m * sizeof(uint32_t) + 1);
    out->dst = (uint32_t *)malloc(m * sizeof(uint32_t) + 1);
    out->w = (uint32_t *)malloc(m * sizeof(uint32_t) + 1);
    if (!out->src || !out->dst || !out->w) { fbcgen_free(out); return -ENOMEM; }
    for (size_t i = 0; i < m; ++i) {
        out->src[i] = (uint32_t)(key[i] >> 32);
        out->dst[i] = (uint32_t)key[i];
        out->w[i] = w[i];
    }
    return 0;
}

/* ---- grid ------------------------------------------------------------------ */

typedef struct { uint64_t key; uint32_t idx; } keyed_link;

static int cmp_keyed_link(const void *a, const void *b)
{
    const keyed_link *x = (const keyed_link *)a, *y = (const keyed_link *)b;
    if (x->key != y->key) return x->key < y->key ? -1 : 1;
    return x->idx < y->idx ? -1 : (x->idx > y->idx);
}

int fbcgen_grid(uint32_t rows, uint32_t cols, uint32_t wmax, uint64_t seed,
                int road, fbcgen_arcs *out)
{
    memset(out, 0, sizeof(*out));
    if (rows < 2 || cols < 2 || wmax < 1) return -EINVAL;
    uint64_t N = (uint64_t)rows * cols;
    if (N >= (1ull << 31)) return -EINVAL;
    size_t L = (size_t)rows * (cols - 1) + (size_t)cols * (rows - 1);

    uint32_t *la = (uint32_t *)malloc(L * sizeof(uint32_t));
    uint32_t *lb = (uint32_t *)malloc(L * sizeof(uint32_t));
    unsigned char *deleted = (unsigned char *)calloc(L, 1);
    if (!la || !lb || !deleted) { free(la); free(lb); free(deleted); return -ENOMEM; }
    size_t l = 0;
    for (uint32_t v = 0; v < N; ++v) {
        uint32_t r = v / cols, c = v % cols;
        if (c + 1 < cols) { la[l] = v; lb[l] = v + 1; ++l; }
        if (r + 1 < rows) { la[l] = v; lb[l] = v + cols; ++l; }
    }

    pcg32 g;
    pcg32_seed(&g, seed, 54u);

    if (road) {
        keyed_link *kl = (keyed_link *)malloc(L * sizeof(keyed_link));
        uint32_t *parent = (uint32_t *)malloc(N * sizeof(uint32_t));
        unsigned char *tree = (unsigned char *)calloc(L, 1);
        uint32_t *cand = (uint32_t *)malloc(L * sizeof(uint32_t));
        if (!kl || !parent || !tree || !cand) {
            free(kl); free(parent); free(tree); free(cand);
            free(la); free(lb); free(deleted);
            return -ENOMEM;
        }
        for (size_t i = 0; i < L; ++i) { kl[i].key = pcg32_next(&g); kl[i].idx = (uint32_t)i; }
        qsort(kl, L, sizeof(keyed_link), cmp_keyed_link);
        for (uint32_t v = 0; v < N; ++v) parent[v] = v;
        for (size_t i = 0; i < L; ++i) {
            uint32_t li = kl[i].idx;
            uint32_t ra = uf_find(parent, la[li]), rb = uf_find(parent, lb[li]);
            if (ra != rb) { parent[ra] = rb; tree[li] = 1; }
        }
        size_t ncand = 0;
        for (size_t i = 0; i < L; ++i)
            if (!tree[i]) cand[ncand++] = (uint32_t)i;
        size_t ndel = (size_t)llround(0.10 * (double)L);
        if (ndel > ncand) ndel = ncand;
        for (size_t i = 0; i < ndel; ++i) {
            size_t j = i + pcg32_bounded(&g, (uint32_t)(ncand - i));
            uint32_t t = cand[i]; cand[i] = cand[j]; cand[j] = t;
            deleted[cand[i]] = 1;
        }
        free(kl); free(parent); free(tree); free(cand);
    }

    size_t kept = 0;
    for (size_t i = 0; i < L; ++i) kept += !deleted[i];
    size_t narcs = 2 * kept;
    size_t ndiag = road ? (size_t)llround(0.02 * (double)narcs) : 0;
    size_t cap = narcs + ndiag;
    uint64_t *key = (uint64_t *)malloc(cap * sizeof(uint64_t) + 8);
    uint32_t *w = (uint32_t *)malloc(cap * sizeof(uint32_t) + 4);
    if (!key || !w) { free(key); free(w); free(la); free(lb); free(deleted); return -ENOMEM; }
    size_t m = 0;
    for (size_t i = 0; i < L; ++i) {
        if (deleted[i]) continue;
        key[m] = ((uint64_t)la[i] << 32) | lb[i];
        w[m++] = 1 + pcg32_bounded(&g, wmax);
        key[m] = ((uint64_t)lb[i] << 32) | la[i];
        w[m++] = 1 + pcg32_bounded(&g, wmax);
    }
    free(la); free(lb); free(deleted);

    if (ndiag) {
        static const int dr[4] = {-1, -1, 1, 1};
        static const int dc[4] = {-1, 1, -1, 1};
        unsigned char *used = (unsigned char *)calloc(N, 1);  /* 4 direction bits */
        if (!used) { free(key); free(w); return -ENOMEM; }
        size_t added = 0;
        while (added < ndiag) {
            uint32_t v = pcg32_bounded(&g, (uint32_t)N);
            uint32_t d = pcg32_bounded(&g, 4);
            int64_t r2 = (int64_t)(v / cols) + dr[d], c2 = (int64_t)(v % cols) + dc[d];
            if (r2 < 0 || r2 >= rows || c2 < 0 || c2 >= cols) continue;
            if (used[v] & (1u << d)) continue;
            used[v] |= (unsigned char)(1u << d);
            uint32_t u = 1 + pcg32_bounded(&g, wmax);
            key[m] = ((uint64_t)v << 32) | (uint32_t)(r2 * cols + c2);
            w[m++] = (uint32_t)ceil(1.414 * (double)u);
            ++added;
        }
        free(used);
    }
    int rc = finish_arcs(key, w, m, (uint32_t)N, out);
    free(key); free(w);
    return rc;
}

/* ---- R-MAT ----------------------------------------------------------------- */

int fbcgen_rmat(uint32_t scale, uint32_t edge_factor, double a, double b,
                double c, uint32_t wmax, uint64_t seed, fbcgen_arcs *out)
{
    memset(out, 0, sizeof(*out));
    if (scale < 1 || scale > 30 || edge_factor < 1 || wmax < 1) return -EINVAL;
    uint32_t N = 1u << scale;
    size_t draws = (size_t)edge_factor * N;
    uint64_t *key = (uint64_t *)malloc((draws + 1) * sizeof(uint64_t));
    uint32_t *w = (uint32_t *)calloc(draws + 1, sizeof(uint32_t));
    if (!key || !w) { free(key); free(w); return -ENOMEM; }
    pcg32 g;
    pcg32_seed(&g, seed, 54u);
    const double ab = a + b, abc = a + b + c;
    size_t m = 0;
    for (size_t i = 0; i < draws; ++i) {
        uint32_t u = 0, v = 0;
        for (uint32_t lvl = 0; lvl < scale; ++lvl) {
            double x = (double)pcg32_next(&g) * (1.0 / 4294967296.0);
            uint32_t bu, bv;
            if (x < a) { bu = 0; bv = 0; }
            else if (x < ab) { bu = 0; bv = 1; }
            else if (x < abc) { bu = 1; bv = 0; }
            else { bu = 1; bv = 1; }
            u = (u << 1) | bu;
            v = (v << 1) | bv;
        }
        if (u == v) continue;
        key[m++] = ((uint64_t)u << 32) | v;
    }
    int rc = radix_sort_pairs(key, w, m, bits_for(((uint64_t)N << 32) | N));
    if (rc) { free(key); free(w); return rc; }
    size_t uniq = 0;
    for (size_t i = 0; i < m; ++i)
        if (uniq == 0 || key[i] != key[uniq - 1]) key[uniq++] = key[i];
    m = uniq;
    for (size_t i = 0; i < m; ++i) w[i] = 1 + pcg32_bounded(&g, wmax);
    if (m == 0 || (uint32_t)(key[m - 1] >> 32) != N - 1) {
        key[m] = ((uint64_t)(N - 1) << 32) | 0u;
        w[m++] = 1 + pcg32_bounded(&g, wmax);
    }
    rc = finish_arcs(key, w, m, N, out);
    free(key); free(w);
    return rc;
}

/* ---- named workloads -------------------------------------------------------- */

int fbcgen_config(int cfg, fbcgen_arcs *out)
{
    switch (cfg) {
    case 0: return fbcgen_grid(64, 64, 100, 1, 0, out);
    case 1: return fbcgen_grid(256, 256, 1000, 2, 1, out);
    case 2: return fbcgen_grid(1024, 1024, 1000, 3, 1, out);
    case 3: return fbcgen_rmat(15, 16, 0.57, 0.19, 0.19, 255, 4, out);
    case 4: return fbcgen_grid(2048, 2048, 1000, 5, 1, out);
    default: memset(out, 0, sizeof(*out)); return -EINVAL;
    }
}

int fbcgen_write(const fbcgen_arcs *arcs, const char *path)
{
    FILE *f = fopen(path, "w");
    if (!f) return -errno;
    char buf[64];
    for (uint64_t i = 0; i < arcs->m; ++i) {
        int len = snprintf(buf, sizeof buf, "%u %u %u\n", arcs->src[i], arcs->dst[i], arcs->w[i]);
        if (fwrite(buf, 1, (size_t)len, f) != (size_t)len) { fclose(f); return -EIO; }
    }
    return fclose(f) ? -errno : 0;
}

void fbcgen_free(fbcgen_arcs *arcs)
{
    free(arcs->src); free(arcs->dst); free(arcs->w);
    arcs->src = arcs->dst = arcs->w = NULL;
    arcs->n = 0; arcs->m = 0;
}

#ifdef FBC_GEN_MAIN
int main(int argc, char **argv)
{
    if (argc != 3) {
        fprintf(stderr, "usage: fbc_gen <cfg 0..4> <out_edge_list>\n");
        return 2;
    }
    fbcgen_arcs a;
    int rc = fbcgen_config(atoi(argv[1]), &a);
    if (rc) { fprintf(stderr, "fbc_gen: error %d\n", rc); return 1; }
    rc = fbcgen_write(&a, argv[2]);
    fprintf(stderr, "fbc_gen: cfg%s n=%u m=%llu\n", argv[1], a.n, (unsigned long long)a.m);
    fbcgen_free(&a);
    return rc ? 1 : 0;
}
#endif
