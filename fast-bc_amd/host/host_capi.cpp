/* C ABI of the host engine (include/fbc_host.h). */
#include "fbc_host.h"
#include "../../include/fbc_host.h"

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <exception>
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>

using namespace fbchost;

namespace {

/* message of the calling thread's last failed host entry point */
thread_local std::string last_error;

template <typename F>
int guarded(F f)
{
    try {
        return f();
    } catch (const std::invalid_argument &e) {
        last_error = e.what();
        return FBC_E_ARG;
    } catch (const std::out_of_range &e) {
        last_error = e.what();
        return FBC_E_ARG;
    } catch (const std::bad_alloc &) {
        last_error = "host allocation failed";
        return FBC_E_OOM;
    } catch (const std::exception &e) {
        last_error = e.what();
        return FBC_E_HIP;
    }
}

/* What the engine keeps between calls: one device context per HIP device
 * (with the FBC_OPTIONS it was configured with) and the last graph it built,
 * with a copy of the arcs it was built from.  A call on the same arcs finds
 * the graph, its device copy, its relabelled copy and the per-source scratch
 * resident, as a caller holding a graph object would; other arcs replace it.
 * Contexts are released by fbc_host_release, not at exit (the HIP runtime
 * may be gone by then). */
struct Resident {
    std::mutex mu;
    std::map<int, std::pair<std::string, Device *>> devices;
    std::vector<int32_t> src, dst;
    std::vector<double> w;
    std::shared_ptr<Graph> graph;
    uint64_t tuning[8] = {};   /* of the last clustered run: cluster launch, then pivot launch */
};

Resident &resident()
{
    static Resident *r = new Resident;
    return *r;
}

/* The device context for 'device', created (or re-created when FBC_OPTIONS
 * changed) under the resident lock. */
Device &device_for(int device)
{
    Resident &r = resident();
    const char *env = std::getenv("FBC_OPTIONS");
    const std::string opts = env ? env : "";
    auto it = r.devices.find(device);
    if (it != r.devices.end() && it->second.first == opts) return *it->second.second;
    if (it != r.devices.end()) {
        delete it->second.second;
        r.devices.erase(it);
    }
    Device *d = new Device(device);
    r.devices[device] = {opts, d};
    return *d;
}

std::shared_ptr<Graph> graph_for(uint64_t m, const int32_t *src, const int32_t *dst, const double *w)
{
    Resident &r = resident();
    if (r.graph && r.src.size() == m && std::equal(src, src + m, r.src.begin()) &&
        std::equal(dst, dst + m, r.dst.begin()) &&
        std::memcmp(w, r.w.data(), m * sizeof(double)) == 0)
        return r.graph;
    r.graph.reset();
    auto g = std::make_shared<Graph>(from_arcs(m, src, dst, w));
    r.src.assign(src, src + m);
    r.dst.assign(dst, dst + m);
    r.w.assign(w, w + m);
    r.graph = g;
    return g;
}

} // namespace

namespace {

/* fbc_clustered_bc and fbc_clustered_bc_partition: Louvain over the seeds
 * when comm_in is null, the caller's partition otherwise. */
int clustered_entry(int device, uint64_t m, const int32_t *src, const int32_t *dst, const double *w,
                    const int32_t *comm_in, const uint64_t *seeds, int nseeds, double precision, double kfrac,
                    int prec, int form, double *bc, int32_t *comm_of, double *intra_bc, uint32_t *pivots,
                    uint32_t *weights, uint32_t *pivot_cluster, double *phase_seconds)
{
    if ((m && (!src || !dst || !w)) || !bc) return FBC_E_ARG;
    if (!comm_in && (!seeds || nseeds < 1)) return FBC_E_ARG;
    if (prec != FBC_PREC_F64 && prec != FBC_PREC_F32C) return FBC_E_ARG;
    if (form != FBC_FORM_BATCHED && form != FBC_FORM_PLAIN && form != FBC_FORM_SELECT_ONLY) return FBC_E_ARG;
    return guarded([&] {
        std::lock_guard<std::mutex> lk(resident().mu);
        const std::shared_ptr<Graph> gp = graph_for(m, src, dst, w);
        const Graph &g = *gp;
        Device &dev = device_for(device);
        ClusteredOptions opt;
        if (comm_in)
            opt.partition.assign(comm_in, comm_in + g.n);
        else
            opt.seeds.assign(seeds, seeds + nseeds);
        opt.precision = precision;
        opt.kfrac = kfrac;
        opt.prec = prec;
        opt.form = form;
        if (std::getenv("FBC_PROGRESS"))   /* progress lines on stderr */
            opt.progress = [](int, const std::string &m) { std::fprintf(stderr, "[fbc] %s\n", m.c_str()); };
        ClusteredRun r = clustered_bc(dev, g, opt);
        std::copy(r.t.cluster_tuning, r.t.cluster_tuning + 4, resident().tuning);
        std::copy(r.t.pivot_tuning, r.t.pivot_tuning + 4, resident().tuning + 4);
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

} // namespace

extern "C" {

void fbc_host_release(void)
{
    Resident &r = resident();
    std::lock_guard<std::mutex> lk(r.mu);
    for (auto &kv : r.devices) delete kv.second.second;
    r.devices.clear();
    r.graph.reset();
    r.src = {};
    r.dst = {};
    r.w = {};
}

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
    if (!seeds || nseeds < 1) return FBC_E_ARG;
    return clustered_entry(device, m, src, dst, w, nullptr, seeds, nseeds, precision, kfrac, prec, form, bc, comm_of,
                           intra_bc, pivots, weights, pivot_cluster, phase_seconds);
}

int fbc_clustered_bc_partition(int device, uint64_t m, const int32_t *src, const int32_t *dst, const double *w,
                               const int32_t *comm_in, double kfrac, int prec, int form, double *bc,
                               double *intra_bc, uint32_t *pivots, uint32_t *weights, uint32_t *pivot_cluster,
                               double *phase_seconds)
{
    if (!comm_in) return FBC_E_ARG;
    return clustered_entry(device, m, src, dst, w, comm_in, nullptr, 0, 0.0, kfrac, prec, form, bc, nullptr,
                           intra_bc, pivots, weights, pivot_cluster, phase_seconds);
}

int fbc_exact_bc_arcs(int device, uint64_t m, const int32_t *src, const int32_t *dst, const double *w,
                      int prec, double *bc)
{
    if ((m && (!src || !dst || !w)) || !bc) return FBC_E_ARG;
    if (prec != FBC_PREC_F64 && prec != FBC_PREC_F32C) return FBC_E_ARG;
    return guarded([&] {
        std::lock_guard<std::mutex> lk(resident().mu);
        const std::shared_ptr<Graph> gp = graph_for(m, src, dst, w);
        std::vector<double> r = exact_bc(device_for(device), *gp, prec);
        std::memcpy(bc, r.data(), r.size() * sizeof(double));
        return 0;
    });
}

const char *fbc_host_last_error(void) { return last_error.c_str(); }

int fbc_host_last_tuning(uint64_t *out)
{
    if (!out) return FBC_E_ARG;
    std::lock_guard<std::mutex> lk(resident().mu);
    std::copy(resident().tuning, resident().tuning + 3, out);
    std::copy(resident().tuning + 4, resident().tuning + 7, out + 3);
    return 0;
}

int fbc_host_last_tuning_ex(uint64_t *out, int count)
{
    if (!out || count < 0) return FBC_E_ARG;
    std::lock_guard<std::mutex> lk(resident().mu);
    const int k = count < 8 ? count : 8;
    std::copy(resident().tuning, resident().tuning + k, out);
    return k;
}

} // extern "C"
