// Exercises the reference-named classes and prints what the test compares
// with the oracle: the chain seed, spins after a known sequence of calls, and
// the operator values the classes report.
#include <cstdio>
#include <cstring>
#include <memory>

#include "../../mcrg_amd/host/mcrg.hpp"

int main(int argc, char **argv)
{
    const int N = argc > 1 ? std::atoi(argv[1]) : 16;
    const double K = argc > 2 ? std::atof(argv[2]) : -0.4406868;
    mcrg_cpp::set_seed(12345);
    auto lat = std::make_shared<Lattice>(N);
    IsingModel model(K);
    model.equilibrate(lat, 37, false);
    for (int i = 0; i < 5; ++i) model.sample_new_configuration(lat);
    lat->sync_spins();
    const vec2D S = lat->calc_interactions();
    std::printf("{\"seed\": %llu, \"N\": %d, \"NN_ref\": %.1f, \"NNN_ref\": %.1f, \"NN_only\": %.1f, ",
                (unsigned long long)lat->seed(), lat->N_, S(0), S(1), lat->calc_nearest_neighbor_interaction());
    std::printf("\"E\": %.17g, \"M\": %.17g, \"spins\": [", model.calc_energy(lat), model.calc_magnetization(lat));
    for (int j = 0; j < N; ++j)
        for (int i = 0; i < N; ++i) std::printf("%s%d", (i | j) ? ", " : "", lat->spins_(i, j));
    std::printf("], \"nn\": [");
    const imat nn = lat->nearest_neighbors(0, N - 1), nnn = lat->next_nearest_neighbors(0, N - 1);
    for (int k = 0; k < 4; ++k) std::printf("%s[%d, %d]", k ? ", " : "", nn(k, 0), nn(k, 1));
    std::printf("], \"nnn\": [");
    for (int k = 0; k < 4; ++k) std::printf("%s[%d, %d]", k ? ", " : "", nnn(k, 0), nnn(k, 1));
    std::printf("]");
    if (argc > 3 && std::strcmp(argv[3], "eq") == 0) {
        // equilibration log of a second chain (file written to the working directory)
        auto lat2 = std::make_shared<Lattice>(N);
        model.equilibrate(lat2, 100, true);
        std::printf(", \"seed2\": %llu", (unsigned long long)lat2->seed());
    }
    std::printf("}\n");
    return 0;
}
