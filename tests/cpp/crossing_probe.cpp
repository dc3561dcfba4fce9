// Reads "J D w" triples (one per line) on stdin and prints the crossing
// mcrg_cpp::zero_crossing finds, for tests/test_crossing_cpu.py.
#include <cstdio>
#include <vector>

#include "../../mcrg_amd/host/mcrg.hpp"

int main()
{
    std::vector<double> J, D, w;
    double a, b, c;
    while (std::scanf("%lf %lf %lf", &a, &b, &c) == 3) {
        J.push_back(a);
        D.push_back(b);
        w.push_back(c);
    }
    std::printf("%.17g\n", mcrg_cpp::zero_crossing(J, D, w));
    return 0;
}
