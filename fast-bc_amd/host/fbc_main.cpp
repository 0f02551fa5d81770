/*
 * fbc -- the reference's command line (main.cpp:40-299) on the MI355X path.
 *
 *   fbc [options] <edge_list_path>
 *     -s, --louvain-seeds      comma-separated seeds (needs -e)
 *     -e, --louvain-instances  Louvain instances (default 4)
 *     -p, --louvain-precision  minimum modularity gain of a pass (default 0.01)
 *     -k, --kfrac              k-means class aggregation in (0, 1): 2-clustered mode
 *         --exact              exact Brandes (every vertex a source)
 *     -t, --threads            host threads (Louvain instances, pivot selection)
 *     -o, --output             output file (default bc.txt; must not exist)
 *     -d, --debug              log level (trace|debug|info|warning|error|critical|off)
 *   options beyond the reference's:
 *         --device             HIP device (default 0)
 *         --precision          f64 | f32c: dependency arithmetic (default f64)
 *         --form               batched | plain: all pivots in one launch (default),
 *                              or one cooperative launch per pivot (FP64)
 *         --full-output        also write full-precision, unclamped BC values
 *         --csr-cache          binary CSR cache of the input (read when it
 *                              matches the input file, else written)
 *
 * Option grammar (popl 1.2.0 as vendored by the reference, popl.hpp:919-1024):
 * '-x v', '-xv', '--long v', '--long=v'; unknown options are ignored; after
 * '--' the remaining arguments are also taken as non-option arguments.
 * Exit status: -1 for option and input errors, -2 when the output file exists
 * (main.cpp:93-97,100-104,123-130,136-181,200-205).  Output: one line per
 * vertex, default ostream formatting, negative values written as 0
 * (main.cpp:282-294).  The timer covers computeBC only (main.cpp:269-277).
 */
#include "fbc_host.h"

#include <chrono>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iostream>
#include <map>
#include <memory>
#include <set>
#include <sstream>
#include <string>
#include <vector>

#include <omp.h>

namespace {

enum Level { TRACE, DEBUG, INFO, WARN, ERR, CRIT, OFF };
Level g_level = INFO;

void log(Level l, const std::string &msg)
{
    static const char *names[] = {"trace", "debug", "info", "warning", "error", "critical", "off"};
    if (l < g_level || g_level == OFF) return;
    std::string name = names[l];
    const size_t pad = name.size() < 9 ? 9 - name.size() : 0;
    std::cout << "[" << std::string(pad / 2, ' ') << name << std::string(pad - pad / 2, ' ') << "] " << msg
              << std::endl;
}

Level level_from(const std::string &s)
{
    static const std::map<std::string, Level> m = {{"trace", TRACE}, {"debug", DEBUG}, {"info", INFO},
                                                   {"warning", WARN}, {"warn", WARN},   {"error", ERR},
                                                   {"err", ERR},      {"critical", CRIT}, {"off", OFF}};
    auto it = m.find(s);
    return it == m.end() ? OFF : it->second;
}

struct Opt {
    char shortname;
    std::string longname, desc;
    bool takes_value;
    bool set = false;
    std::string value;
};

struct InvalidOption : std::runtime_error {
    using std::runtime_error::runtime_error;
};

class Parser {
public:
    std::vector<Opt> opts;
    std::vector<std::string> args;

    Opt *find(char c)
    {
        for (auto &o : opts)
            if (c && o.shortname == c) return &o;
        return nullptr;
    }
    Opt *find(const std::string &l)
    {
        for (auto &o : opts)
            if (o.longname == l) return &o;
        return nullptr;
    }
    static std::string name(const Opt &o, bool is_long)
    {
        return is_long ? "--" + o.longname : std::string("-") + o.shortname;
    }
    void assign(Opt &o, const std::string &v, bool is_long)
    {
        if (o.takes_value && v.empty()) throw InvalidOption("missing argument for " + name(o, is_long));
        o.set = true;
        o.value = v;
    }
    void parse(int argc, char **argv)
    {
        for (int n = 1; n < argc; ++n) {
            const std::string arg(argv[n]);
            if (arg == "--") {
                for (int m = n + 1; m < argc; ++m) args.emplace_back(argv[m]);
            } else if (arg.rfind("--", 0) == 0) {
                std::string o = arg.substr(2), v;
                const size_t eq = o.find('=');
                if (eq != std::string::npos) {
                    v = o.substr(eq + 1);
                    o.resize(eq);
                }
                Opt *op = find(o);
                if (op && !op->takes_value && !v.empty()) op = nullptr;
                if (op && op->takes_value && v.empty() && n < argc - 1) v = argv[++n];
                if (op) assign(*op, v, true);
            } else if (arg.rfind("-", 0) == 0) {
                const std::string o = arg.substr(1);
                for (size_t m = 0; m < o.size(); ++m) {
                    Opt *op = find(o[m]);
                    if (!op) continue;
                    std::string v;
                    if (op->takes_value) {
                        v = o.substr(m + 1);
                        if (v.empty() && n < argc - 1) v = argv[++n];
                        m = o.size();
                    }
                    assign(*op, v, false);
                }
            } else {
                args.push_back(arg);
            }
        }
    }
    std::string help() const
    {
        std::ostringstream s;
        s << "Usage: fastbc [ options ] <edge_list_path>:\n";
        for (const auto &o : opts) {
            std::string head = o.shortname ? std::string("  -") + o.shortname + ", " : std::string("      ");
            head += "--" + o.longname + (o.takes_value ? " arg" : "");
            if (head.size() < 32) head.resize(32, ' ');
            s << head << " " << o.desc << "\n";
        }
        return s.str();
    }
};

/* popl's Value<T>::parse (popl.hpp:704-734): exactly one value of type T. */
template <typename T>
T parse_value(const Opt &o, bool is_long)
{
    std::istringstream is(o.value);
    T v{};
    int read = 0;
    while (is.good()) {
        if (is.peek() == EOF) break;
        is >> v;
        ++read;
    }
    if (is.fail()) throw InvalidOption("invalid argument for " + Parser::name(o, is_long) + ": '" + o.value + "'");
    if (read > 1) throw InvalidOption("too many arguments for " + Parser::name(o, is_long) + ": '" + o.value + "'");
    return v;
}

} // namespace

/* bc.txt as main.cpp:282-294 writes it: one `out << x << std::endl` per
 * vertex (default stream formatting, i.e. %g with 6 significant digits),
 * negative values as 0.  The lines are formatted in parallel and written in
 * order. */
static void write_bc(const std::string &path, const std::vector<double> &bc)
{
    const size_t n = bc.size();
    const int nchunk = n < 65536 ? 1 : 64;
    std::vector<std::string> parts((size_t)nchunk);
#pragma omp parallel for schedule(dynamic, 1)
    for (int k = 0; k < nchunk; ++k) {
        const size_t a = n * (size_t)k / nchunk, b = n * (size_t)(k + 1) / nchunk;
        std::string &s = parts[(size_t)k];
        s.reserve((b - a) * 12);
        char buf[64];
        for (size_t i = a; i < b; ++i) {
            const int len = bc[i] >= 0 ? std::snprintf(buf, sizeof buf, "%g\n", bc[i]) : std::snprintf(buf, sizeof buf, "0\n");
            s.append(buf, (size_t)len);
        }
    }
    std::ofstream out(path, std::ofstream::out);
    for (const std::string &s : parts) out.write(s.data(), (std::streamsize)s.size());
}

int main(int argc, char **argv)
{
    Parser op;
    op.opts = {
        {'s', "louvain-seeds", "Seeds to be used by each parallel louvain execution", true},
        {'e', "louvain-instances", "Number of parallel louvain instances", true},
        {'p', "louvain-precision", "Minimum precision value for louvain algorithm", true},
        {'k', "kfrac", "Topological classes aggregation factor (0-1). Enables 2-Clustered Brandes algorithm", true},
        {0, "exact", "Force exact betweenness computation (very long time)", false},
        {'t', "threads", "Maximum number of threads used in parallel computation", true},
        {'o', "output", "Output file path", true},
        {'d', "debug", "Logger level (trace|debug|info|warning|error|critical|off)", true},
        {0, "device", "HIP device index (not in the reference)", true},
        {0, "precision", "Dependency arithmetic: f64 | f32c (not in the reference)", true},
        {0, "form", "Pivot phase: batched | plain (not in the reference)", true},
        {0, "full-output", "Also write unclamped %.17g values to this file (not in the reference)", true},
        {0, "csr-cache", "Binary CSR cache of the input: read when it matches the input, else written (not in the reference)",
         true},
    };
    std::string seedstr, outpath = "bc.txt", level = "info", fullout, precstr = "f64", formstr = "batched", csrcache;
    int instances = 4, threads = 0, device = 0;
    double precision = 0.01, kfrac = 0;
    bool exact = false;
    try {
        op.parse(argc, argv);
        if (Opt *o = op.find('s'); o->set) seedstr = o->value;
        if (Opt *o = op.find('e'); o->set) instances = parse_value<int>(*o, false);
        if (Opt *o = op.find('p'); o->set) precision = parse_value<double>(*o, false);
        if (Opt *o = op.find('k'); o->set) kfrac = parse_value<double>(*o, false);
        if (Opt *o = op.find('t'); o->set) threads = parse_value<int>(*o, false);
        if (Opt *o = op.find('o'); o->set) outpath = o->value;
        if (Opt *o = op.find('d'); o->set) level = o->value;
        if (Opt *o = op.find(std::string("device")); o->set) device = parse_value<int>(*o, true);
        if (Opt *o = op.find(std::string("precision")); o->set) precstr = o->value;
        if (Opt *o = op.find(std::string("form")); o->set) formstr = o->value;
        if (Opt *o = op.find(std::string("full-output")); o->set) fullout = o->value;
        if (Opt *o = op.find(std::string("csr-cache")); o->set) csrcache = o->value;
        exact = op.find(std::string("exact"))->set;
        if (precstr != "f64" && precstr != "f32c")
            throw InvalidOption("invalid argument for --precision: '" + precstr + "'");
        if (formstr != "batched" && formstr != "plain")
            throw InvalidOption("invalid argument for --form: '" + formstr + "'");
        if (formstr == "plain" && precstr != "f64")
            throw InvalidOption("--form plain computes in f64");
    } catch (const InvalidOption &e) {
        std::cout << e.what() << "\n\n" << op.help();
        return -1;
    }
    if (op.args.size() != 1) {
        std::cout << "Missing input file path" << "\n\n" << op.help();
        return -1;
    }
    const std::string input = op.args.front();
    g_level = level_from(level);

    {
        std::ifstream test(outpath);
        if (test.good()) {
            log(CRIT, "File \"" + outpath + "\" already existing");
            return -2;
        }
    }

    std::set<unsigned long> seeds;
    if (op.find('s')->set) {
        if (!op.find('e')->set) {
            log(CRIT, "Louvain executors count must be set to allow executors seeds to be set.");
            return -1;
        }
        std::stringstream ss(seedstr);
        unsigned long s;
        while (ss >> s) {
            if (!seeds.insert(s).second) {
                log(CRIT, "Duplicate value in louvain seeds, each seed must be unique.");
                return -1;
            }
            if (ss.peek() == ',') ss.ignore();
        }
        if ((int)seeds.size() != instances) {
            log(CRIT, "Louvain seeds count is different from louvain executors count.");
            return -1;
        }
    } else {
        for (int i = 0; i < instances; ++i)
            seeds.insert((unsigned long)std::chrono::high_resolution_clock::now().time_since_epoch().count());
    }
    const bool kfrac_set = op.find('k')->set;
    if (kfrac_set && (kfrac <= 0.0 || kfrac >= 1.0)) {
        log(CRIT, "Kfrac value must be in range 0-1.");
        return -1;
    }
    if (op.find('t')->set) {
        log(INFO, "Maximum number of threads set to " + std::to_string(threads));
        omp_set_num_threads(threads);
    }

    std::ifstream in(input);
    if (!in.is_open()) {
        log(CRIT, "There was an error opening given edge list file path.");
        return -1;
    }
    /* like the reference, a loader exception (weight <= 0) is not caught */
    bool cache_hit = false;
    fbchost::Graph g = csrcache.empty() ? fbchost::load_edge_list(in)
                                        : fbchost::load_edge_list_cached(input, csrcache, &cache_hit);
    if (!csrcache.empty())
        log(DEBUG, std::string(cache_hit ? "Graph read from the CSR cache " : "CSR cache written to ") + csrcache);
    log(INFO, "Loaded graph contains " + std::to_string(g.n) + " vertices and " + std::to_string(g.edges) +
                  " edges");

    const int prec = precstr == "f32c" ? FBC_PREC_F32C : FBC_PREC_F64;
    std::unique_ptr<fbchost::Device> devp;
    try {
        devp.reset(new fbchost::Device(device));
    } catch (const std::exception &e) {
        log(CRIT, e.what());
        return -1;
    }
    fbchost::Device &dev = *devp;
    std::vector<double> bc;
    fbchost::ClusteredRun run;
    if (exact) log(INFO, "Algorithm: exact Brandes' betweenness centrality");
    else if (kfrac_set) log(INFO, "Algorithm: 2-clustered Brandes' betweenness centrality");
    else log(INFO, "Algorithm: clustered Brandes' betweenness centrality");

    const auto start = std::chrono::high_resolution_clock::now();
    if (exact) {
        bc = fbchost::exact_bc(dev, g, prec);
    } else {
        fbchost::ClusteredOptions opt;
        opt.seeds.assign(seeds.begin(), seeds.end());
        opt.precision = precision;
        opt.kfrac = kfrac_set ? kfrac : 0.0;
        opt.prec = prec;
        opt.form = formstr == "plain" ? FBC_FORM_PLAIN : FBC_FORM_BATCHED;
        opt.progress = [](int level, const std::string &m) { log(level == 0 ? INFO : DEBUG, m); };
        run = fbchost::clustered_bc(dev, g, opt);
        bc = run.bc;
    }
    const auto total = std::chrono::high_resolution_clock::now() - start;
    const long long ms = std::chrono::duration_cast<std::chrono::milliseconds>(total).count();
    const long long us = std::chrono::duration_cast<std::chrono::microseconds>(total).count() - ms * 1000;
    if (!exact) {
        char buf[256];
        std::snprintf(buf, sizeof buf,
                      "Phases: louvain %.3fs, clusters %.3fs, pivot selection %.3fs, pivot phase %.3fs",
                      run.t.louvain, run.t.cluster, run.t.pivots, run.t.pivot_phase);
        log(DEBUG, buf);
    }
    log(INFO, "Total computation time: " + std::to_string(ms) + "." + std::to_string(us) + "ms");

    write_bc(outpath, bc);
    if (!fullout.empty()) {
        FILE *f = std::fopen(fullout.c_str(), "w");
        if (f) {
            for (double x : bc) std::fprintf(f, "%.17g\n", x);
            std::fclose(f);
        }
    }
    log(INFO, "Results written to \"" + outpath + "\"");
    return 0;
}
