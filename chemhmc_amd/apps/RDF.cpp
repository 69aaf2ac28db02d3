// "RDF -i infile": the reference's offline radial distribution program (code/main/RDF.cpp,
// HMC_class::measure) over this engine: g(r) of every configuration of the run's xyz output.
#include <chrono>
#include <cstdio>
#include <exception>

#include "chmc/driver.hh"

int main(int argc, char** argv)
{
    const auto t0 = std::chrono::steady_clock::now();
    try {
        chmc_host::hmc_driver D;
        D.init(argc, argv, false);
        D.measure();
    } catch (const std::exception& e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
    const double s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    std::printf("total time = %f s\n", s);
    return 0;
}
