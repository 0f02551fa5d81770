/*
 * Driver for the reference's own headers -- TEST INFRASTRUCTURE ONLY.
 *
 * Built by oracle/Makefile (target 'ref') against the headers under
 * /root/reference/libfastbc/include into oracle/_ref/ref_driver, only where
 * that tree is present; the built binaries travel with the repository tree
 * (they are git-ignored).  Only headers that compile with the image's own
 * toolchain are included; SubGraph.h, ClusteredBrandesBC.h, the two pivot
 * selectors and LouvainGraphPartition.h need spdlog, which is not installed,
 * so those parts are restated below (marked) or checked by other means.
 *
 * Modes (graph = '<src> <dst> <w>' edge list read by the reference loader):
 *   exact   <graph> <out>                       ExactBrandesBC::computeBC
 *   ssb     <graph> <pivots> <out>              DijkstraSSBrandesBC over "vertex weight" lines,
 *                                               OpenMP reduction as ClusteredBrandesBC.h:141-166
 *   cluster <graph> <vertices> <out>            DijkstraClusterEvaluator::evaluateCluster
 *   clusters <graph> <partition> <c>...         the same over listed communities, timed
 *   louvain <graph> <precision> <seed>...       Partition::one_level / partition2graph
 *   kmeans  <rows> <out>                        PlusPlusKMeans::computeCentroids
 * Outputs are text with %.17g values.  Timings go to stderr.
 */
#include <map>
#include <memory>
#include <random>
#include <set>
#include <vector>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <sstream>
#include <string>

#include <IGraphPartition.h>
#include <ISubGraph.h>
#include <DirectedWeightedGraph.h>
#include <brandes/ExactBrandesBC.h>
#include <brandes/DijkstraSSBrandesBC.h>
#include <brandes/DijkstraClusterEvaluator.h>
#include <brandes/VertexInfo.h>
#include <kmeans/PlusPlusKMeans.h>
#include <louvain/LouvainGraph.h>
#include <louvain/Partition.h>

#include <omp.h>

using G = fastbc::DirectedWeightedGraph<int, double>;

static std::shared_ptr<G> load(const char *path)
{
    std::ifstream in(path);
    if (!in.is_open()) {
        std::fprintf(stderr, "cannot open %s\n", path);
        std::exit(2);
    }
    return std::make_shared<G>(in);
}

static double now_s()
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

/* Driver-side ISubGraph (restated, not the reference's SubGraph.h): border =
 * arc leaving or entering the vertex set; forward/backward stars restricted
 * to the set. */
class DriverSubGraph : public fastbc::ISubGraph<int, double> {
public:
    DriverSubGraph(const std::vector<int> &verts, std::shared_ptr<const fastbc::IGraph<int, double>> g)
        : _g(g), _v(verts), _e(0)
    {
        std::set<int> in(verts.begin(), verts.end());
        for (int v : verts) {
            for (auto &e : g->forwardStar(v)) {
                if (in.count(e.first)) { _fs[v][e.first] = e.second; ++_e; }
                else _b.insert(v);
            }
            _fs[v];
            for (auto &e : g->backwardStar(v)) {
                if (in.count(e.first)) _bs[v][e.first] = e.second;
                else _b.insert(v);
            }
            _bs[v];
        }
    }
    double edge(int s, int d) const override
    {
        auto &f = forwardStar(s);
        auto it = f.find(d);
        return it == f.end() ? 0 : it->second;
    }
    const std::map<int, double> &forwardStar(int s) const override { return _fs.at(s); }
    const std::map<int, double> &backwardStar(int d) const override { return _bs.at(d); }
    const std::vector<int> &vertices() const override { return _v; }
    int edges() const override { return (int)_e; }
    const std::set<int> &borders() const override { return _b; }
    bool isBorder(int v) const override { return _b.count(v) > 0; }
    std::shared_ptr<const fastbc::IGraph<int, double>> referenceGraph() const override { return _g; }

private:
    std::shared_ptr<const fastbc::IGraph<int, double>> _g;
    std::vector<int> _v;
    long long _e;
    std::map<int, std::map<int, double>> _fs, _bs;
    std::set<int> _b;
};

static int mode_exact(int argc, char **argv)
{
    if (argc != 4) return 2;
    auto g = load(argv[2]);
    fastbc::brandes::ExactBrandesBC<int, double> ex;
    double t0 = now_s();
    std::vector<double> bc = ex.computeBC(g);
    std::fprintf(stderr, "exact: n=%zu m=%d threads=%d seconds=%.6f\n", bc.size(), g->edges(),
                 omp_get_max_threads(), now_s() - t0);
    FILE *f = std::fopen(argv[3], "w");
    for (double x : bc) std::fprintf(f, "%.17g\n", x);
    std::fclose(f);
    return 0;
}

/* The reference's pivot loop body (ClusteredBrandesBC.h:141-155) over an
 * explicit pivot list, without the intra-cluster subtraction. */
static int mode_ssb(int argc, char **argv)
{
    if (argc != 5) return 2;
    auto g = load(argv[2]);
    std::vector<int> piv, w;
    {
        std::ifstream in(argv[3]);
        int p, x;
        while (in >> p >> x) { piv.push_back(p); w.push_back(x); }
    }
    const size_t n = g->vertices().size();
    std::vector<double> bc(n, 0.0);
    double *B = bc.data();
    fastbc::brandes::DijkstraSSBrandesBC<int, double> ssb;
    double t0 = now_s();
#pragma omp parallel for reduction(+ : B[:n])
    for (size_t p = 0; p < piv.size(); ++p) {
        std::vector<double> d = ssb.singleSourceBrandes(piv[p], g);
        for (size_t v = 0; v < n; ++v) B[v] += d[v] * (double)w[p];
    }
    double dt = now_s() - t0;
    std::fprintf(stderr, "ssb: n=%zu m=%d pivots=%zu threads=%d seconds=%.6f\n", n, g->edges(),
                 piv.size(), omp_get_max_threads(), dt);
    FILE *f = std::fopen(argv[4], "w");
    std::fprintf(f, "# seconds %.6f threads %d pivots %zu\n", dt, omp_get_max_threads(), piv.size());
    for (double x : bc) std::fprintf(f, "%.17g\n", x);
    std::fclose(f);
    return 0;
}

static int mode_cluster(int argc, char **argv)
{
    if (argc != 5) return 2;
    auto g = load(argv[2]);
    std::vector<int> verts;
    {
        std::ifstream in(argv[3]);
        int v;
        while (in >> v) verts.push_back(v);
    }
    auto sub = std::make_shared<DriverSubGraph>(verts, g);
    fastbc::brandes::DijkstraClusterEvaluator<int, double> ce;
    std::vector<double> bc(g->vertices().size(), 0.0);
    std::vector<std::shared_ptr<fastbc::brandes::VertexInfo<int, double>>> vi(g->vertices().size());
    ce.evaluateCluster(bc, vi, sub);
    FILE *f = std::fopen(argv[4], "w");
    std::fprintf(f, "borders");
    for (int b : sub->borders()) std::fprintf(f, " %d", b);
    std::fprintf(f, "\n");
    for (int v : verts) {
        std::fprintf(f, "%d %.17g", v, bc[v]);
        for (int j = 0; j < vi[v]->borders(); ++j)
            std::fprintf(f, " %.17g %d", vi[v]->getBorderSPLength(j), vi[v]->getBorderSPCount(j));
        std::fprintf(f, "\n");
    }
    std::fclose(f);
    return 0;
}

/* DijkstraClusterEvaluator over several clusters of one partition, the graph
 * loaded once: partition = one community index per line (vertex order); the
 * listed communities are evaluated in turn (ClusteredBrandesBC.h:91-121 runs
 * them concurrently, each with its own evaluator call); prints per-cluster
 * seconds and the total. */
static int mode_clusters(int argc, char **argv)
{
    if (argc < 5) return 2;
    auto g = load(argv[2]);
    std::vector<int> comm;
    {
        std::ifstream in(argv[3]);
        int c;
        while (in >> c) comm.push_back(c);
    }
    if (comm.size() != g->vertices().size()) return 2;
    double total = 0;
    fastbc::brandes::DijkstraClusterEvaluator<int, double> ce;
    for (int a = 4; a < argc; ++a) {
        const int want = std::atoi(argv[a]);
        std::vector<int> verts;
        for (size_t v = 0; v < comm.size(); ++v)
            if (comm[v] == want) verts.push_back((int)v);
        double t0 = now_s();
        auto sub = std::make_shared<DriverSubGraph>(verts, g);
        std::vector<double> bc(g->vertices().size(), 0.0);
        std::vector<std::shared_ptr<fastbc::brandes::VertexInfo<int, double>>> vi(g->vertices().size());
        ce.evaluateCluster(bc, vi, sub);
        const double dt = now_s() - t0;
        total += dt;
        std::printf("cluster %d vertices %zu edges %d borders %zu seconds %.6f\n", want, verts.size(), sub->edges(),
                    sub->borders().size(), dt);
    }
    std::printf("total %.6f\n", total);
    return 0;
}

/* The multi-instance level loop of LouvainGraphPartition::partitionGraph is
 * restated here (that header needs spdlog); Partition and LouvainGraph are
 * the reference's. */
static int mode_louvain(int argc, char **argv)
{
    if (argc < 5) return 2;
    auto g = load(argv[2]);
    double precision = std::atof(argv[3]);
    std::set<std::mt19937::result_type> seeds;
    for (int i = 4; i < argc; ++i) seeds.insert(std::strtoul(argv[i], nullptr, 10));
    std::vector<std::mt19937> eng;
    for (auto s : seeds) eng.push_back(std::mt19937(s));
    const int P = (int)eng.size();
    using LG = fastbc::louvain::LouvainGraph<int, double>;
    using LP = fastbc::louvain::Partition<int, double>;
    LG lg(std::static_pointer_cast<const fastbc::IDegreeGraph<int, double>>(g));
    std::vector<LP> p(P, LP(lg, precision));
    std::vector<int> n2c(lg.nb_nodes);
    for (unsigned i = 0; i < lg.nb_nodes; ++i) n2c[i] = (int)i;
    std::vector<char> imp(P, 1);
    std::vector<double> mods(P);
    bool improvement;
    int level = 0;
    double t0 = now_s();
    do {
#pragma omp parallel for
        for (int i = 0; i < P; ++i) {
            imp[i] = p[i].one_level(eng[i]);
            mods[i] = p[i].modularity();
        }
        int best = 0;
        for (int i = 1; i < P; ++i)
            if (mods[i] > mods[best]) best = i;
        improvement = imp[best] != 0;
        std::fprintf(stderr, "level %d best %d modularity %.17g nodes %u\n", level, best,
                     p[best].modularity(), lg.nb_nodes);
        lg = p[best].partition2graph();
        {
            const std::vector<int> &lvl = p[best].n2c;
            std::vector<int> ren(lvl.size(), -1);
            for (size_t i = 0; i < lvl.size(); ++i) ren[lvl[i]]++;
            int fin = 0;
            for (size_t i = 0; i < lvl.size(); ++i)
                if (ren[i] != -1) ren[i] = fin++;
            for (size_t i = 0; i < n2c.size(); ++i) n2c[i] = ren[lvl[n2c[i]]];
        }
        for (int i = 0; i < P; ++i) p[i] = LP(lg, precision);
        ++level;
    } while (improvement);
    std::fprintf(stderr, "louvain: seconds=%.6f\n", now_s() - t0);
    for (int c : n2c) std::printf("%d\n", c);
    return 0;
}

/* rows file: "P nb k", then P lines "weight len_0 cnt_0 ... len_nb-1 cnt_nb-1" */
static int mode_kmeans(int argc, char **argv)
{
    if (argc != 4) return 2;
    std::ifstream in(argv[2]);
    int P, nb, k;
    in >> P >> nb >> k;
    std::vector<std::shared_ptr<fastbc::brandes::VertexInfo<int, double>>> vi(P);
    std::vector<int> pts(P), w(P);
    for (int p = 0; p < P; ++p) {
        vi[p] = std::make_shared<fastbc::brandes::VertexInfo<int, double>>(nb);
        in >> w[p];
        for (int j = 0; j < nb; ++j) {
            double l;
            int c;
            in >> l >> c;
            vi[p]->setBorderSPLength(j, l);
            vi[p]->setBorderSPCount(j, c);
        }
        pts[p] = p;
    }
    fastbc::kmeans::PlusPlusKMeans<int, double> km;
    auto r = km.computeCentroids(k, pts, w, vi, 0, 100);
    FILE *f = std::fopen(argv[3], "w");
    for (size_t i = 0; i < r.first.size(); ++i) std::fprintf(f, "%d %d\n", r.first[i], r.second[i]);
    std::fclose(f);
    return 0;
}

int main(int argc, char **argv)
{
    if (argc < 2) {
        std::fprintf(stderr, "usage: ref_driver exact|ssb|cluster|clusters|louvain|kmeans ...\n");
        return 2;
    }
    std::string m = argv[1];
    int rc = 2;
    if (m == "exact") rc = mode_exact(argc, argv);
    else if (m == "ssb") rc = mode_ssb(argc, argv);
    else if (m == "cluster") rc = mode_cluster(argc, argv);
    else if (m == "clusters") rc = mode_clusters(argc, argv);
    else if (m == "louvain") rc = mode_louvain(argc, argv);
    else if (m == "kmeans") rc = mode_kmeans(argc, argv);
    if (rc == 2) std::fprintf(stderr, "bad arguments for mode %s\n", m.c_str());
    return rc;
}
