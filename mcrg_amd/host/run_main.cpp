// `mcrg_run`: the reference's `run` driver (src/main.cpp:4-22) on the MI355X.
// With no arguments it runs the reference's call sequence: b = 2, N = 128,
// 1e4 equilibration and 1e6 measured samples, calc_critical_exponent at
// K = -0.44 and at K_c = -ln(1 + sqrt 2)/2.  Flags change the sizes and add the
// critical-point locator (src/mcrg.cpp:146-188).
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "mcrg.hpp"

static void usage()
{
    std::printf(
        "usage: mcrg_run [--N n] [--eq n] [--samples n] [--K k ...] [--replicas r] [--ops 2|3]\n"
        "                [--bins b] [--levels l] [--seed s] [--device d] [--outdir dir] [--quiet]\n"
        "                [--locate iterations --L l --K0 k]\n"
        "                [--sampler metropolis|sw] [--layout auto|int8|packed]\n"
        "                [--bisect rounds --n-k n --J-lo j --J-hi j --n-large n --L-large l\n"
        "                 --n-small n --L-small l --sampler-large sw|metropolis --eq-small n]\n"
        "                (--bisect prints one JSON line)\n"
        "K values are in the reference's sign convention (K < 0 ferromagnetic); 'Kc' means\n"
        "-ln(1+sqrt(2))/2.\n");
}

int main(int argc, char **argv)
{
    int b = 2, N = 128, n_eq = 10000, n = 1000000, replicas = 1, ops = 2, bins = 50, levels = -1, device = 0;
    int locate = 0, Lloc = 16;
    int bisect = 0, n_k = 16, n_large = 64, L_large = 1024, n_small = 1024, L_small = 128;
    double J_lo = 0.43, J_hi = 0.45;
    int layout = -1;
    int sampler = MCRG_SAMPLER_METROPOLIS, sampler_large = MCRG_SAMPLER_SWENDSEN_WANG, n_eq_small = -1;
    double K0 = -0.44;
    unsigned long long seed = 1;
    bool quiet = false;
    std::string outdir = ".";
    const double Kc = -std::log(1.0 + std::sqrt(2.0)) / 2.0;
    std::vector<double> Ks;
    for (int i = 1; i < argc; ++i) {
        const std::string a = argv[i];
        auto next = [&](void) -> const char * {
            if (i + 1 >= argc) {
                usage();
                std::exit(2);
            }
            return argv[++i];
        };
        if (a == "--N") N = std::atoi(next());
        else if (a == "--eq") n_eq = std::atoi(next());
        else if (a == "--samples") n = std::atoi(next());
        else if (a == "--K") {
            const std::string v = next();
            Ks.push_back(v == "Kc" ? Kc : std::atof(v.c_str()));
        } else if (a == "--replicas") replicas = std::atoi(next());
        else if (a == "--ops") ops = std::atoi(next());
        else if (a == "--bins") bins = std::atoi(next());
        else if (a == "--levels") levels = std::atoi(next());
        else if (a == "--seed") seed = std::strtoull(next(), nullptr, 10);
        else if (a == "--device") device = std::atoi(next());
        else if (a == "--outdir") outdir = next();
        else if (a == "--quiet") quiet = true;
        else if (a == "--locate") locate = std::atoi(next());
        else if (a == "--L") Lloc = std::atoi(next());
        else if (a == "--bisect") bisect = std::atoi(next());
        else if (a == "--n-k") n_k = std::atoi(next());
        else if (a == "--sampler") sampler = std::string(next()) == "sw" ? MCRG_SAMPLER_SWENDSEN_WANG : MCRG_SAMPLER_METROPOLIS;
        else if (a == "--sampler-large")
            sampler_large = std::string(next()) == "sw" ? MCRG_SAMPLER_SWENDSEN_WANG : MCRG_SAMPLER_METROPOLIS;
        else if (a == "--eq-small") n_eq_small = std::atoi(next());
        else if (a == "--layout") {
            const std::string v = next();
            if (v == "int8") layout = MCRG_LAYOUT_INT8;
            else if (v == "packed") layout = MCRG_LAYOUT_PACKED;
            else if (v == "auto") layout = -1;
            else {
                usage();
                return 2;
            }
        }
        else if (a == "--J-lo") J_lo = std::atof(next());
        else if (a == "--J-hi") J_hi = std::atof(next());
        else if (a == "--n-large") n_large = std::atoi(next());
        else if (a == "--L-large") L_large = std::atoi(next());
        else if (a == "--n-small") n_small = std::atoi(next());
        else if (a == "--L-small") L_small = std::atoi(next());
        else if (a == "--K0") {
            const std::string v = next();
            K0 = v == "Kc" ? Kc : std::atof(v.c_str());
        } else {
            usage();
            return a == "--help" || a == "-h" ? 0 : 2;
        }
    }
    if (Ks.empty()) Ks = {-0.44, Kc};
    try {
        mcrg_cpp::set_device(device);
        mcrg_cpp::set_seed(seed);
        MonteCarloRenormalizationGroup MCRG(b);
        MCRG.set_replicas(replicas);
        MCRG.set_operators(ops);
        MCRG.set_bins(bins);
        MCRG.set_levels(levels);
        MCRG.set_output_dir(outdir);
        MCRG.set_quiet(quiet);
        MCRG.set_sampler(sampler);
        MCRG.set_layout(layout);
        if (bisect > 0) {
            const auto r = MCRG.bisect_critical_point(bisect, n_k, J_lo, J_hi, n_large, L_large, n_small, L_small,
                                                      n_eq, n, sampler_large, n_eq_small);
            const double exact = std::log(1.0 + std::sqrt(2.0)) / 2.0;
            // systematic: change of the crossing between the two deepest matched levels
            const auto &lv0 = r.rounds.back().Jc_level;
            const double sys = lv0.size() >= 2 ? std::fabs(lv0[lv0.size() - 1] - lv0[lv0.size() - 2]) : 0.0;
            const double tot = std::sqrt(r.Jc_err * r.Jc_err + sys * sys);
            std::printf("{\"Jc\": %.10f, \"Jc_err\": %.10f, \"Jc_sys\": %.10f, \"exact\": %.10f, "
                        "\"deviation\": %.10f, \"deviation_sigma\": %.3f, "
                        "\"rounds\": %i, \"seconds\": %.3f, \"site_updates\": %.6e, \"updates_per_ns\": %.3f, "
                        "\"round_Jc\": [",
                        r.Jc, r.Jc_err, sys, exact, r.Jc - exact, tot > 0 ? (r.Jc - exact) / tot : 0.0,
                        (int)r.rounds.size(), r.seconds, r.site_updates, r.site_updates / (r.seconds * 1e9));
            for (size_t i = 0; i < r.rounds.size(); ++i)
                std::printf("%s[%.10f, %.10f, %.10f, %.10f]", i ? ", " : "", r.rounds[i].J_lo, r.rounds[i].J_hi,
                            r.rounds[i].Jc, r.rounds[i].Jc_err);
            std::printf("], \"last_round_Jc_by_level\": [");
            const auto &lv = r.rounds.back().Jc_level;
            for (size_t i = 0; i < lv.size(); ++i) std::printf("%s%.10f", i ? ", " : "", lv[i]);
            std::printf("]}\n");
        } else if (locate > 0) {
            MCRG.locate_critical_point(locate, n_eq, n, Lloc, K0);
        } else {
            for (double K : Ks) MCRG.calc_critical_exponent(n_eq, n, N, K);
        }
    } catch (const std::exception &e) {
        std::fprintf(stderr, "mcrg_run: %s\n", e.what());
        return 1;
    }
    return 0;
}
