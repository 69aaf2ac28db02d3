// Parses an input file of the reference's schema and prints the values the drivers read (host-only
// check of the YAML front end; no device calls).
#include <cstdio>
#include <exception>

#include "chmc/read_infile.hh"

int main(int argc, char** argv)
{
    using namespace chmc_host;
    try {
        const yaml_node doc = read_params(argc, argv);
        const params_class p(doc, false, false);
        const yaml_node& pd = doc["particles"];
        std::printf("L %g %g %g\n", p.L[0], p.L[1], p.L[2]);
        std::printf("seed %d\nStartCondition %s\nstart_configuration_file %s\noutput_file %s\n", p.seed,
                    p.StartCondition.c_str(), p.start_configuration_file.c_str(), p.nameout.c_str());
        std::printf("append %d\nname_RDF %s\n", (int)p.append, p.name_RDF.c_str());
        std::printf("N %d\nmass %g\nbeta %g\ncutoff %g\neps %g\nsigma %g\nname_xyz %s\nalgorithm %s\n",
                    check_and_assign_value<int>(pd, "N"), check_and_assign_value<double>(pd, "mass"),
                    check_and_assign_value<double>(pd, "beta"), check_and_assign_value<double>(pd, "cutoff"),
                    check_and_assign_value<double>(pd, "eps"), check_and_assign_value<double>(pd, "sigma"),
                    check_and_assign_value<std::string>(pd, "name_xyz").c_str(),
                    check_and_assign_value<std::string>(pd, "algorithm").c_str());
        std::printf("RDF %g %g\n", check_and_assign_value<double>(pd["RDF"], "Lmax"), check_and_assign_value<double>(pd["RDF"], "size_bin"));
        const yaml_node& id = doc["integrator"];
        std::printf("integrator %s %g %d\n", check_and_assign_value<std::string>(id, "name").c_str(),
                    check_and_assign_value<double>(id, "dt"), check_and_assign_value<int>(id, "steps"));
        std::printf("Ntrajectories %d\nthermalization_steps %d\nsave_every %d\n", check_and_assign_value<int>(doc, "Ntrajectories"),
                    check_and_assign_value<int>(doc, "thermalization_steps"), check_and_assign_value<int>(doc, "save_every"));
        std::printf("randomize_trajectory %s\n", doc["randomize_trajectory"] ? doc["randomize_trajectory"].as<std::string>().c_str() : "absent");
    } catch (const std::exception& e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
    return 0;
}
