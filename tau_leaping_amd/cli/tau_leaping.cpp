// tau_leaping -- command-line driver over the C ABI (include/tl_api.h).
//
// Same surface as the reference driver (tau_leaping/tau_leaping.cpp:40-190,
// tau_leaping/cl_parser.cpp:19-130):
//   -f <StochKit XML>  -t <t_final>  -n <realizations>  -r <names...|all|-1>
//   -e <eps, 0.03>  -a <max iterations>  -d <device>  -b <log name>
//   -i <logging intervals, 1>
// and writes mean_<log>.txt / stddev_<log>.txt with lines "t v1 v2 ..."
// (StatToFile, statistics.cpp:89-111).  Extensions: --precision fp64|fp32,
// --seed N, --poisson normal|exact|fast|fast-exact (the firing-count
// sampler, see TL_POISSON_* in tl_api.h), --net-stoichiometry (a catalyst
// counts once, with its net coefficient, in the tau bound), --mode auto|persistent|plain (auto: persistent,
// plain where it runs a model too large for LDS faster), --block-size N, --sensitivities (writes
// dmean_<log>.txt: one line per tracked species, dE[X(t_final)]/dc_j),
// --json (one JSON record of the run on stdout).
//
// Differences from the reference driver, all fixes: realizations run
// concurrently on the GPU; the grid point t_final is written (the reference
// drops it); "-r -1" selects every species like "-r all"; -n below 10 is
// accepted; -d selects the device (the reference hard-codes device 1).
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <set>
#include <sstream>
#include <string>
#include <vector>

#include "tl_api.h"

namespace {

struct Args {
    std::string file;
    double t_end = 1.0;
    double eps = 0.03;
    long long n = 1;
    long long max_iter = 0;
    int device = 0;
    int intervals = 1;
    std::string log = "stat";
    bool log_set = false;
    std::vector<std::string> track;
    bool track_flag = false;
    int precision = TL_FP64;
    unsigned long long seed = 0;
    bool seed_set = false;
    int mode = TL_MODE_AUTO;
    int poisson = TL_POISSON_NORMAL;
    int block_size = 0;
    bool sens = false;
    bool net = false;
    bool json = false;
};

[[noreturn]] void usage(const char *msg) {
    if (msg) std::cerr << msg << "\n";
    std::cerr << "usage: tau_leaping -f model.xml -t t_end [-n realizations] [-r names...|all] [-e eps]\n"
                 "                   [-a max_iterations] [-d device] [-b log_name] [-i intervals]\n"
                 "                   [--precision fp64|fp32] [--seed N] [--mode auto|persistent|plain]\n"
                 "                   [--poisson normal|exact|fast|fast-exact] [--net-stoichiometry]\n"
                 "                   [--block-size N] [--sensitivities] [--json]\n";
    std::exit(EXIT_FAILURE);
}

const char *need(int argc, char **argv, int &a) {
    if (a + 1 >= argc) usage((std::string("missing value after ") + argv[a]).c_str());
    return argv[++a];
}

Args parse(int argc, char **argv) {
    Args A;
    for (int a = 1; a < argc; ++a) {
        const std::string k = argv[a];
        if (k == "-f") A.file = need(argc, argv, a);
        else if (k == "-t") A.t_end = atof(need(argc, argv, a));
        else if (k == "-n") A.n = atoll(need(argc, argv, a));
        else if (k == "-e") A.eps = atof(need(argc, argv, a));
        else if (k == "-a") A.max_iter = atoll(need(argc, argv, a));
        else if (k == "-d") A.device = atoi(need(argc, argv, a));
        else if (k == "-i") A.intervals = atoi(need(argc, argv, a));
        else if (k == "-b") {
            A.log = need(argc, argv, a);
            A.log_set = true;
        } else if (k == "-r") {
            // names up to the next option; "-1" and "all" select every species
            A.track_flag = true;
            while (a + 1 < argc && (argv[a + 1][0] != '-' || strcmp(argv[a + 1], "-1") == 0)) {
                std::string v = argv[++a];
                std::stringstream ss(v);
                std::string tok;
                while (std::getline(ss, tok, ','))
                    if (!tok.empty()) A.track.push_back(tok);
            }
        } else if (k == "--precision") {
            const std::string v = need(argc, argv, a);
            if (v == "fp64") A.precision = TL_FP64;
            else if (v == "fp32") A.precision = TL_FP32;
            else usage("precision must be fp64 or fp32");
        } else if (k == "--seed") {
            A.seed = strtoull(need(argc, argv, a), nullptr, 10);
            A.seed_set = true;
        } else if (k == "--mode") {
            const std::string v = need(argc, argv, a);
            if (v == "auto") A.mode = TL_MODE_AUTO;
            else if (v == "persistent") A.mode = TL_MODE_PERSISTENT;
            else if (v == "plain") A.mode = TL_MODE_PLAIN;
            else usage("mode must be auto, persistent or plain");
        } else if (k == "--poisson") {
            const std::string v = need(argc, argv, a);
            if (v == "normal") A.poisson = TL_POISSON_NORMAL;
            else if (v == "exact") A.poisson = TL_POISSON_EXACT;
            else if (v == "fast") A.poisson = TL_POISSON_FAST;
            else if (v == "fast-exact") A.poisson = TL_POISSON_FAST_EXACT;
            else usage("poisson must be normal, exact, fast or fast-exact");
        } else if (k == "--block-size") A.block_size = atoi(need(argc, argv, a));
        else if (k == "--sensitivities") A.sens = true;
        else if (k == "--net-stoichiometry") A.net = true;
        else if (k == "--json") A.json = true;
        else if (k == "-h" || k == "--help") usage(nullptr);
        else usage(("unknown option " + k).c_str());
    }
    if (A.file.empty()) usage("a model file is required (-f); random systems are not supported");
    if (!(A.t_end > 0)) usage("-t must be > 0");
    if (A.n < 1) usage("-n must be >= 1");
    if (A.intervals < 1) usage("-i must be >= 1");
    return A;
}

int fail(tl_status st) {
    std::cerr << "error " << st << ": " << tl_last_error() << std::endl;
    return EXIT_FAILURE;
}

}  // namespace

int main(int argc, char **argv) {
    Args A = parse(argc, argv);
    const std::string logfn = A.log + ".txt";

    tl_model *m = nullptr;
    tl_status st = tl_model_from_stochkit(A.file.c_str(), &m);
    if (st != TL_OK) return fail(st);
    if (A.net && (st = tl_model_set_net_stoichiometry(m, 1)) != TL_OK) {
        tl_model_destroy(m);
        return fail(st);
    }
    const int N = tl_model_n_species(m), M = tl_model_n_channels(m);

    std::vector<int32_t> track;
    const bool all = A.track.size() == 1 && (A.track[0] == "all" || A.track[0] == "-1");
    if (all) {
        for (int i = 0; i < N; ++i) track.push_back(i);
    } else {
        std::set<std::string> want(A.track.begin(), A.track.end());
        for (int i = 0; i < N; ++i)
            if (want.count(tl_model_species_name(m, i))) track.push_back(i);
        for (auto &w : want) {
            int32_t idx;
            if (tl_model_species_index(m, w.c_str(), &idx) != TL_OK) std::cerr << "unknown species " << w << "\n";
        }
    }
    if (!A.json) {
        if (!track.empty()) {
            std::cout << "Tracking reactants: " << std::endl;
            for (int32_t i : track) std::cout << tl_model_species_name(m, i) << std::endl;
        }
        std::cout << "Processing file " << A.file << std::endl;
        std::cout << "Finish time: " << A.t_end << std::endl;
        std::cout << "Running " << A.n << " system(s) in parallel" << std::endl;
        std::cout << "Logging intervals: " << A.intervals << std::endl;
        std::cout << "Iterations limit: " << (A.max_iter > 0 ? A.max_iter : 2147483647LL) << std::endl;
        std::cout << std::endl << "logging statistics into " << logfn << " file" << std::endl;
    }

    tl_sim_params p;
    tl_sim_params_default(&p);
    p.t_final = A.t_end;
    p.eps = A.eps;
    p.n_realizations = A.n;
    p.seed = A.seed_set ? A.seed
                        : (unsigned long long)std::chrono::steady_clock::now().time_since_epoch().count();
    p.max_iter = A.max_iter;
    p.precision = A.precision;
    // the engine needs at least one tracked species; untracked runs track S1
    std::vector<int32_t> sim_track = track.empty() ? std::vector<int32_t>{0} : track;
    p.n_track = (int32_t)sim_track.size();
    p.track = sim_track.data();
    p.n_intervals = A.intervals;
    p.sensitivities = A.sens ? 1 : 0;
    p.device = A.device;
    p.mode = A.mode;
    p.block_size = A.block_size;
    p.poisson = A.poisson;

    const int T = p.n_track;
    const int G = p.n_intervals > 0 ? p.n_intervals + 1 : 1;
    std::vector<double> mean((size_t)G * T), sd((size_t)G * T), times(G);
    std::vector<double> dmean(A.sens ? (size_t)T * M : 0), dse(A.sens ? (size_t)T * M : 0);
    tl_result r;
    memset(&r, 0, sizeof(r));
    r.mean = mean.data();
    r.stddev = sd.data();
    r.grid_times = times.data();
    if (A.sens) {
        r.dmean_dc = dmean.data();
        r.dmean_dc_se = dse.data();
    }
    const auto t0 = std::chrono::steady_clock::now();
    st = tl_simulate(m, &p, &r);
    const double wall = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (st != TL_OK) {
        tl_model_destroy(m);
        return fail(st);
    }
    const int rows = (int)times.size();
    const int Tout = (int)track.size();
    {
        std::ofstream fm("mean_" + logfn), fs("stddev_" + logfn);
        if (!fm || !fs) {
            std::cerr << "can not open the output files for " << logfn << std::endl;
            return EXIT_FAILURE;
        }
        for (int g = 0; g < rows; ++g) {
            fm << times[g] << " ";
            fs << times[g] << " ";
            for (int q = 0; q < Tout; ++q) {
                fm << mean[(size_t)g * T + q] << " ";
                fs << sd[(size_t)g * T + q] << " ";
            }
            fm << std::endl;
            fs << std::endl;
        }
    }
    if (A.sens) {
        std::ofstream fd("dmean_" + logfn);
        for (int q = 0; q < Tout; ++q) {
            for (int j = 0; j < M; ++j) fd << dmean[(size_t)q * M + j] << (j + 1 < M ? " " : "");
            fd << std::endl;
        }
    }
    const double leaps_per_s = r.kernel_ms > 0 ? (double)r.total_steps / (r.kernel_ms * 1e-3) : 0.0;
    if (A.json) {
        std::cout << "{\"file\": \"" << A.file << "\", \"species\": " << N << ", \"channels\": " << M
                  << ", \"realizations\": " << A.n << ", \"t_final\": " << A.t_end
                  << ", \"precision\": \"" << (A.precision == TL_FP64 ? "fp64" : "fp32") << "\", \"mode\": \""
                  << (A.mode == TL_MODE_PLAIN ? "plain" : (A.mode == TL_MODE_AUTO ? "auto" : "persistent"))
                  << "\", \"seed\": " << p.seed
                  << ", \"total_steps\": " << r.total_steps << ", \"ssa_events\": " << r.total_ssa_events
                  << ", \"rejections\": " << r.total_rejections << ", \"failed\": " << r.n_failed
                  << ", \"kernel_ms\": " << r.kernel_ms << ", \"wall_s\": " << wall
                  << ", \"leaps_per_s\": " << leaps_per_s << ", \"block_size\": " << r.block_size << "}"
                  << std::endl;
    } else {
        std::cout << "Simulation kernel time: " << r.kernel_ms / 1000.0 << " s (" << r.total_steps
                  << " realization-steps, " << leaps_per_s << " per s; " << r.total_ssa_events
                  << " direct events; " << r.total_rejections << " rejected leaps)" << std::endl;
        std::cout << "It took " << wall << " s to complete calculations" << std::endl;
        std::cout << "Computational time, s: " << (double)(A.n - r.n_failed) * A.t_end << std::endl;
        std::cout << "Computing statistics at time steps..." << std::endl << " done" << std::endl;
        if (r.n_failed)
            std::cout << r.n_failed << " realization(s) stopped early, first status " << r.first_failure
                      << std::endl;
    }
    tl_model_destroy(m);
    return EXIT_SUCCESS;
}
