/*
 * Topological classes, pivots and k-means super-classes on the host.
 *
 *   select_pivots   VertexInfoPivotSelector::selectPivots (VertexInfoPivotSelector.h:29-127)
 *   kmeans          PlusPlusKMeans::computeCentroids / _initPlusPlus / _centroidVariance
 *                   (kmeans/PlusPlusKMeans.h:70-229)
 *   kmeans_pivots   KMeansPivotSelector::selectPivots (KMeansPivotSelector.h:79-132)
 *
 * Class grouping uses a hash of the normalized row instead of the reference's
 * scan over all class representatives; the classes, their first-appearance
 * order and their members' order are the same.  The k-means keeps the
 * reference's FP64 order of operations in every distance (per border: the
 * length term, then the count term; pow(x, 2) as x * x) and its first-wins
 * tie rules; only independent distances are computed in parallel.
 */
#include "fbc_host.h"

#include <algorithm>
#include <climits>
#include <cstring>
#include <set>
#include <unordered_map>

namespace fbchost {

namespace {

uint64_t mix(uint64_t h, uint64_t x)
{
    h ^= x + 0x9e3779b97f4a7c15ull + (h << 6) + (h >> 2);
    return h * 0xff51afd7ed558ccdull;
}

/* VertexInfo<int,double>::squaredDistance between two rows (VertexInfo.h:293-307). */
inline double sq_points(const double *la, const int *ca, const double *lb, const int *cb, int nb)
{
    double sq = 0;
    for (int i = 0; i < nb; ++i) {
        const double dl = la[i] - lb[i];
        sq += dl * dl;
        const double dc = (double)(int)((uint32_t)ca[i] - (uint32_t)cb[i]);
        sq += dc * dc;
    }
    return sq;
}

/* VertexInfo<double,double> (a mean) against a row. */
inline double sq_mean(const double *ml, const double *mc, const double *lb, const int *cb, int nb)
{
    double sq = 0;
    for (int i = 0; i < nb; ++i) {
        const double dl = ml[i] - lb[i];
        sq += dl * dl;
        const double dc = mc[i] - (double)cb[i];
        sq += dc * dc;
    }
    return sq;
}

} // namespace

void select_pivots(int nv, int nb, double *len, const int *cnt, const double *bc_row,
                   const char *is_border, std::vector<int> &piv_rows, std::vector<int> &weights)
{
    std::vector<std::vector<int>> members;
    std::unordered_map<uint64_t, std::vector<int>> by_hash;
    by_hash.reserve((size_t)nv * 2);
    for (int i = 0; i < nv; ++i) {
        double *row = len + (size_t)i * nb;
        if (nb) {
            const double mn = *std::min_element(row, row + nb);
            for (int j = 0; j < nb; ++j) row[j] -= mn;
        }
        uint64_t h = 0x12345;
        for (int j = 0; j < nb; ++j) {
            const double x = row[j] + 0.0;   /* -0.0 and +0.0 compare equal */
            uint64_t bits;
            std::memcpy(&bits, &x, 8);
            h = mix(mix(h, (uint64_t)(uint32_t)cnt[(size_t)i * nb + j]), bits);
        }
        std::vector<int> &cands = by_hash[h];
        int found = -1;
        for (int c : cands) {
            const int r = members[c][0];
            bool eq = true;
            for (int j = 0; j < nb && eq; ++j)
                eq = cnt[(size_t)r * nb + j] == cnt[(size_t)i * nb + j] &&
                     len[(size_t)r * nb + j] == row[j];
            if (eq) {
                found = c;
                break;
            }
        }
        if (found >= 0) {
            members[found].push_back(i);
        } else {
            cands.push_back((int)members.size());
            members.push_back(std::vector<int>(1, i));
        }
    }
    piv_rows.clear();
    weights.clear();
    for (const auto &m : members) {
        size_t j = 0;
        while (j < m.size() && is_border[m[j]]) ++j;
        if (j == m.size()) {     /* FASTBC_BRANDES_ENABLE_PIVOT_BORDER (main.cpp:5) */
            piv_rows.push_back(m[0]);
            weights.push_back((int)m.size());
            continue;
        }
        int best = m[j];
        for (; j < m.size(); ++j)
            if (!is_border[m[j]] && bc_row[m[j]] < bc_row[best]) best = m[j];
        piv_rows.push_back(best);
        weights.push_back((int)m.size());
    }
}

void kmeans(int k, int P, int nb, const double *L, const int *C, const int *weights,
            double min_variance, size_t max_iteration, std::vector<int> &centroids,
            std::vector<int> &cweights)
{
    auto row_l = [&](int i) { return L + (size_t)i * nb; };
    auto row_c = [&](int i) { return C + (size_t)i * nb; };
    /* _initPlusPlus: c_0 = point 0, then the running-average farthest point */
    std::vector<int> next((size_t)k);
    next[0] = 0;
    {
        std::vector<double> cdist((size_t)P, 0.0);
        for (int i = 1; i < k; ++i) {
            const int last = next[(size_t)i - 1];
            const double p = 1.0 / i, q = 1.0 - p;
#pragma omp parallel for schedule(static)
            for (int v = 0; v < P; ++v)
                cdist[v] = cdist[v] * q + sq_points(row_l(last), row_c(last), row_l(v), row_c(v), nb) * p;
            int far = 0;
            for (int v = 0; v < P; ++v)
                if (cdist[v] > cdist[far]) far = v;
            next[i] = far;
        }
    }
    std::vector<int> cent((size_t)k), assign((size_t)P);
    std::vector<double> mlen((size_t)k * nb), mcnt((size_t)k * nb);
    std::vector<std::vector<int>> members((size_t)k);
    size_t iteration = 0;
    for (;;) {
        ++iteration;
        cent = next;
#pragma omp parallel for schedule(dynamic, 16)
        for (int v = 0; v < P; ++v) {
            int bc = 0;
            double bd = sq_points(row_l(cent[0]), row_c(cent[0]), row_l(v), row_c(v), nb);
            for (int c = 1; c < k; ++c) {
                const double d = sq_points(row_l(cent[c]), row_c(cent[c]), row_l(v), row_c(v), nb);
                if (d < bd) {
                    bc = c;
                    bd = d;
                }
            }
            assign[v] = bc;
        }
        std::fill(mlen.begin(), mlen.end(), 0.0);
        std::fill(mcnt.begin(), mcnt.end(), 0.0);
        for (auto &m : members) m.clear();
        for (int v = 0; v < P; ++v) {
            const int c = assign[v];
            double *ml = &mlen[(size_t)c * nb], *mc = &mcnt[(size_t)c * nb];
            const double *lv = row_l(v);
            const int *cv = row_c(v);
            for (int i = 0; i < nb; ++i) {
                ml[i] += lv[i];
                mc[i] += cv[i];
            }
            members[c].push_back(v);
        }
#pragma omp parallel for schedule(dynamic, 4)
        for (int c = 0; c < k; ++c) {
            if (members[c].empty()) continue;
            double *ml = &mlen[(size_t)c * nb], *mc = &mcnt[(size_t)c * nb];
            const double sz = (double)members[c].size();
            for (int i = 0; i < nb; ++i) {
                ml[i] /= sz;
                mc[i] /= sz;
            }
            int bv = members[c][0];
            double bd = sq_mean(ml, mc, row_l(bv), row_c(bv), nb);
            for (size_t j = 1; j < members[c].size(); ++j) {
                const int v = members[c][j];
                const double d = sq_mean(ml, mc, row_l(v), row_c(v), nb);
                if (d < bd) {
                    bv = v;
                    bd = d;
                }
            }
            next[c] = bv;
        }
        double var = 0;
        for (int c = 0; c < k; ++c) {
            const double d = sq_points(row_l(cent[c]), row_c(cent[c]), row_l(next[c]), row_c(next[c]), nb);
            if (d > var) var = d;
        }
        if (!(var > min_variance && iteration <= max_iteration)) break;
    }
    centroids = cent;
    cweights.assign((size_t)k, 0);
    for (int c = 0; c < k; ++c)
        for (int v : members[c]) cweights[c] += weights[v];
}

void kmeans_pivots(const std::vector<int> &rows, const std::vector<int> &w, int nb, const double *len,
                   const int *cnt, double kfrac, std::vector<int> &out_rows, std::vector<int> &out_w)
{
    const int P = (int)rows.size();
    const int k = std::max((int)(rows.size() * kfrac), 1);
    std::vector<double> L((size_t)P * nb);
    std::vector<int> C((size_t)P * nb);
    for (int p = 0; p < P; ++p) {
        std::memcpy(&L[(size_t)p * nb], len + (size_t)rows[p] * nb, sizeof(double) * nb);
        std::memcpy(&C[(size_t)p * nb], cnt + (size_t)rows[p] * nb, sizeof(int) * nb);
    }
    std::vector<int> cent, cw;
    kmeans(k, P, nb, L.data(), C.data(), w.data(), 0.0, 100, cent, cw);
    std::set<int> seen;
    out_rows.clear();
    out_w.clear();
    for (size_t i = 0; i < cent.size(); ++i) {
        const int r = rows[cent[i]];
        if (seen.insert(r).second) {
            out_rows.push_back(r);
            out_w.push_back(cw[i]);
        }
    }
}

} // namespace fbchost
