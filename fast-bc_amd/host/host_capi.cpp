/* C ABI of the host engine (include/fbc_host.h). */
#include "fbc_host.h"
#include "../../include/fbc_host.h"

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <exception>
#include <stdexcept>

using namespace fbchost;

namespace {

template <typename F>
int guarded(F f)
{
    try {
        return f();
    } catch (const std::invalid_argument &) {
        return FBC_E_ARG;
    } catch (const std::out_of_range &) {
        return FBC_E_ARG;
    } catch (const std::bad_alloc &) {
        return FBC_E_OOM;
    } catch (const std::exception &) {
        return FBC_E_HIP;
    }
}

} // namespace

extern "C" {

int fbc_louvain(uint64_t m, const int32_t *src, const int32_t *dst, const double *w,
                const uint64_t *seeds, int nseeds, double precision, int32_t *comm_of)
{
    if ((m && (!src || !dst || !w)) || !seeds || nseeds < 1 || !comm_of) return FBC_E_ARG;
    return guarded([&] {
        Graph g = from_arcs(m, src, dst, w);
        auto comms = louvain(g, std::vector<uint64_t>(seeds, seeds + nseeds), precision);
        for (size_t c = 0; c < comms.size(); ++c)
            for (int v : comms[c]) comm_of[v] = (int32_t)c;
        return (int)comms.size();
    });
}

int fbc_select_pivots(int nv, int nb, double *vi_len, const int32_t *vi_cnt, const double *bc_row,
                      const uint8_t *is_border, int32_t *piv_rows, int32_t *weights)
{
    if (nv < 0 || nb < 0 || (nv && (!bc_row || !is_border || !piv_rows || !weights)) ||
        (((size_t)nv * nb) != 0 && (!vi_len || !vi_cnt)))
        return FBC_E_ARG;
    return guarded([&] {
        std::vector<int> rows, w;
        select_pivots(nv, nb, vi_len, vi_cnt, bc_row, (const char *)is_border, rows, w);
        for (size_t i = 0; i < rows.size(); ++i) {
            piv_rows[i] = rows[i];
            weights[i] = w[i];
        }
        return (int)rows.size();
    });
}

int fbc_kmeans(int k, int P, int nb, const double *vi_len, const int32_t *vi_cnt, const int32_t *weights,
               double min_variance, int max_iteration, int32_t *centroids, int32_t *cweights)
{
    if (k < 1 || P < 1 || nb < 0 || !weights || !centroids || !cweights || max_iteration < 0 ||
        (((size_t)P * nb) != 0 && (!vi_len || !vi_cnt)))
        return FBC_E_ARG;
    return guarded([&] {
        std::vector<int> c, cw;
        kmeans(k, P, nb, vi_len, vi_cnt, weights, min_variance, (size_t)max_iteration, c, cw);
        for (int i = 0; i < k; ++i) {
            centroids[i] = c[i];
            cweights[i] = cw[i];
        }
        return k;
    });
}

int fbc_clustered_bc(int device, uint64_t m, const int32_t *src, const int32_t *dst, const double *w,
                     const uint64_t *seeds, int nseeds, double precision, double kfrac, int prec,
                     int form, double *bc, int32_t *comm_of, double *intra_bc, uint32_t *pivots,
                     uint32_t *weights, uint32_t *pivot_cluster, double *phase_seconds)
{
    if ((m && (!src || !dst || !w)) || !seeds || nseeds < 1 || !bc) return FBC_E_ARG;
    if (prec != FBC_PREC_F64 && prec != FBC_PREC_F32C) return FBC_E_ARG;
    if (form != FBC_FORM_BATCHED && form != FBC_FORM_PLAIN && form != FBC_FORM_SELECT_ONLY) return FBC_E_ARG;
    return guarded([&] {
        Graph g = from_arcs(m, src, dst, w);
        Device dev(device);
        ClusteredOptions opt;
        opt.seeds.assign(seeds, seeds + nseeds);
        opt.precision = precision;
        opt.kfrac = kfrac;
        opt.prec = prec;
        opt.form = form;
        if (std::getenv("FBC_PROGRESS"))   /* progress lines on stderr */
            opt.progress = [](int, const std::string &m) { std::fprintf(stderr, "[fbc] %s\n", m.c_str()); };
        ClusteredRun r = clustered_bc(dev, g, opt);
        std::memcpy(bc, r.bc.data(), r.bc.size() * sizeof(double));
        if (comm_of)
            for (size_t i = 0; i < r.comm.size(); ++i) comm_of[i] = r.comm[i];
        if (intra_bc) std::memcpy(intra_bc, r.intra.data(), r.intra.size() * sizeof(double));
        for (size_t i = 0; i < r.pivots.size(); ++i) {
            if (pivots) pivots[i] = r.pivots[i];
            if (weights) weights[i] = r.weights[i];
            if (pivot_cluster) pivot_cluster[i] = r.pivot_cluster[i];
        }
        if (phase_seconds) {
            phase_seconds[0] = r.t.louvain;
            phase_seconds[1] = r.t.cluster;
            phase_seconds[2] = r.t.pivots;
            phase_seconds[3] = r.t.pivot_phase;
            phase_seconds[4] = r.t.total;
            phase_seconds[5] = r.t.cluster_te;
            phase_seconds[6] = r.t.pivot_te;
            phase_seconds[7] = r.t.pivot_kernel_ms;
        }
        return (int)r.pivots.size();
    });
}

int fbc_exact_bc_arcs(int device, uint64_t m, const int32_t *src, const int32_t *dst, const double *w,
                      int prec, double *bc)
{
    if ((m && (!src || !dst || !w)) || !bc) return FBC_E_ARG;
    if (prec != FBC_PREC_F64 && prec != FBC_PREC_F32C) return FBC_E_ARG;
    return guarded([&] {
        Graph g = from_arcs(m, src, dst, w);
        Device dev(device);
        std::vector<double> r = exact_bc(dev, g, prec);
        std::memcpy(bc, r.data(), r.size() * sizeof(double));
        return 0;
    });
}

} // extern "C"
