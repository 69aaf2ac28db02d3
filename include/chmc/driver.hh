// The reference's two programs over this engine: "main -i infile" (code/main/main.cpp +
// HMC_class::init / run, code/modules/HMC.cpp:6-136) and "RDF -i infile" (code/main/RDF.cpp +
// HMC_class::measure, HMC.cpp:139-166).  The input file has the reference's schema
// (code/builds/test.yaml); the trajectories run on the device in chunks that end where the reference
// saves (save_every), and every save writes the xyz configuration, the host mt19937_64 state (text,
// as HMC.cpp:52-59 writes it) and the device random state.
//
// The device random state is the heat bath's counter-based stream: the seed and the number of the last
// trajectory.  Restoring it (and istart from the xyz file) makes an appended run continue the same
// streams, which the reference cannot do (its device state save/load are commented out, Q15, and its
// istart is lost, Q14).  Extensions of the schema, all optional: particles.precision (fp64, the
// reference's arithmetic and the default, or fp32), integrator.lambda (OMF2, default 0.1938 as
// integrator.cpp:72).  StartCondition cold is the project's lattice without jitter and hot is uniform
// positions from the host mt19937_64(seed) (the reference's Kokkos pool cannot be reproduced).
#pragma once

#include <cstdio>
#include <fstream>
#include <memory>
#include <random>
#include <stdexcept>
#include <string>
#include <vector>

#include "hmc.hh"
#include "read_infile.hh"
#include "xyz.hh"

namespace chmc_host {

class hmc_driver {
public:
    yaml_node doc;
    params_class params;
    std::unique_ptr<identical_particles> particles;
    std::unique_ptr<integrator_type> integ;
    std::unique_ptr<HMC_class> hmc;
    int Ntrajectories = 0, thermalization_steps = 0, save_every = 1;
    bool randomize_traj = false;
    std::string name_xyz, filename_RDF;
    double LmaxRDF = 0, size_bRDF = 0;
    int acceptance = 0;

    ~hmc_driver()
    {
        if (params.fileout) std::fclose(params.fileout);
    }

    // HMC_class::init (HMC.cpp:6-46) with identical_particles' constructor and InitX
    // (identical_particles.cpp:17-87, 218-238)
    void init(int argc, char** argv, bool check_overwrite = true)
    {
        doc = read_params(argc, argv);
        params = params_class(doc, check_overwrite);
        const yaml_node& pd = doc["particles"];
        if (!pd) throw std::runtime_error("no particles in input file");
        const std::string pname = check_and_assign_value<std::string>(pd, "name");
        if (pname != "identical_particles") throw std::runtime_error("no valid name for particles: " + pname);
        particle_params pp;
        pp.N = check_and_assign_value<int>(pd, "N");
        for (int d = 0; d < 3; d++) pp.L[d] = params.L[d];
        pp.mass = check_and_assign_value<double>(pd, "mass");
        pp.beta = check_and_assign_value<double>(pd, "beta");
        pp.cutoff = check_and_assign_value<double>(pd, "cutoff");
        pp.eps = check_and_assign_value<double>(pd, "eps");
        pp.sigma = check_and_assign_value<double>(pd, "sigma");
        name_xyz = check_and_assign_value<std::string>(pd, "name_xyz");
        // every strategy of the reference computes the same forces; this engine has one
        const std::string algo = check_and_assign_value<std::string>(pd, "algorithm");
        if (algo != "all_neighbour" && algo != "all_neighbour_inner_parallel" && algo != "binning_serial" &&
            algo != "parallel_binning" && algo != "quick_sort")
            throw std::runtime_error("selected algorithm: " + algo + " is not a valid algorithm");
        const std::string prec = pd["precision"] ? pd["precision"].as<std::string>() : std::string("fp64");
        if (prec != "fp64" && prec != "fp32") throw std::runtime_error("particles.precision must be fp64 or fp32");
        pp.fp64_forces = prec == "fp64";
        if (pd["RDF"]) {
            size_bRDF = check_and_assign_value<double>(pd["RDF"], "size_bin");
            LmaxRDF = check_and_assign_value<double>(pd["RDF"], "Lmax");
            filename_RDF = check_and_assign_value<std::string>(pd["RDF"], "output_file");
            pp.rdf_size_bin = size_bRDF;
            pp.rdf_Lmax = LmaxRDF;
            pp.rdf_bins = (int)(LmaxRDF / size_bRDF);        // identical_particles.cpp:82 (truncates, Q17)
            std::printf("RDF: Lmax=%g  size_bin=%g  N=%d  \n", LmaxRDF, size_bRDF, pp.rdf_bins);
        }
        particles = std::make_unique<identical_particles>(pp);
        std::printf("particles: identical_particles N=%d mass=%g beta=%g cutoff=%g eps=%g sigma=%g (%s forces)\n", pp.N,
                    pp.mass, pp.beta, pp.cutoff, pp.eps, pp.sigma, prec.c_str());

        // InitX
        if (params.StartCondition == "cold") {
            particles->InitX((uint64_t)params.seed, 0.0);
        } else if (params.StartCondition == "hot") {
            std::mt19937_64 g((uint64_t)params.seed);
            std::uniform_real_distribution<double> u01(0.0, 1.0);
            config<coordinate> x(pp.N, 3);
            for (int i = 0; i < pp.N; i++) {
                double r[3];
                for (int d = 0; d < 3; d++) {
                    r[d] = u01(g) * params.L[d];
                    if (r[d] >= params.L[d]) r[d] = 0.0;
                }
                x[i] = coordinate(r[0], r[1], r[2]);
            }
            particles->set_x(x);
        } else if (params.StartCondition == "read") {
            std::vector<double> xyz;
            params.istart = read_xyz_last(params.start_configuration_file, pp.N, name_xyz, xyz);
            std::printf("reading last configuration (trajectory %d) from %s\n", params.istart, params.start_configuration_file.c_str());
            config<coordinate> x(pp.N, 3);
            for (int i = 0; i < pp.N; i++) x[i] = coordinate(xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]);
            particles->set_x(x);        // check_in_volume: chmc_set_positions rejects x outside [0, L)
        } else {
            throw std::runtime_error("StartCondition not implemented");
        }

        // integrator (integrator.cpp:7-30, HMC.cpp:10-27)
        const yaml_node& idoc = doc["integrator"];
        if (!idoc) throw std::runtime_error("no integrator in input file");
        const std::string iname = check_and_assign_value<std::string>(idoc, "name");
        const int steps = check_and_assign_value<int>(idoc, "steps");
        const double dt = check_and_assign_value<double>(idoc, "dt");
        if (iname == "LEAP") integ = std::make_unique<LEAP>(particles.get(), dt, steps);
        else if (iname == "OMF2") integ = std::make_unique<OMF2>(particles.get(), dt, steps, idoc["lambda"] ? idoc["lambda"].as<double>() : 0.1938);
        else if (iname == "OMF4") integ = std::make_unique<OMF4>(particles.get(), dt, steps);
        else throw std::runtime_error("no valid integrator name: " + iname);
        std::printf("integrator %s steps: %d dt: %g\n", iname.c_str(), steps, dt);

        Ntrajectories = check_and_assign_value<int>(doc, "Ntrajectories");
        thermalization_steps = check_and_assign_value<int>(doc, "thermalization_steps");
        save_every = check_and_assign_value<int>(doc, "save_every");
        if (save_every < 1) throw std::runtime_error("save_every must be >= 1");
        if (doc["randomize_trajectory"]) randomize_traj = check_and_assign_value<int>(doc, "randomize_trajectory") != 0;
        else std::printf("randomize_trajectory not given: fixed trajectory length\n");
        std::printf("Ntrajectories:%d\nthermalization_steps:%d\nsave_every:%d\n", Ntrajectories, thermalization_steps, save_every);

        hmc_params hp;
        hp.Ntrajectories = Ntrajectories;
        hp.thermalization_steps = thermalization_steps;
        hp.save_every = save_every;
        hp.istart = params.istart;
        hp.randomize_traj = randomize_traj;
        hp.seed = (uint64_t)params.seed;
        hmc = std::make_unique<HMC_class>(integ.get(), hp);
        hmc->gen64.seed((uint64_t)params.seed);
        if (params.append) {
            std::printf("Loading rng host...\n");
            std::ifstream fin(params.rng_host_state);
            fin >> hmc->gen64;
            if (!fin) throw std::runtime_error("invalid rng_host_state file " + params.rng_host_state);
            load_device_rng();
        }
    }

    // heat-bath stream state: seed and last trajectory number
    void save_device_rng() const
    {
        std::FILE* f = std::fopen(params.rng_device_state.c_str(), "w");
        if (!f) throw std::runtime_error("unable to open file " + params.rng_device_state);
        std::fprintf(f, "chmc_heatbath_philox4x64 seed %d trajectory %d\n", params.seed, params.istart);
        std::fclose(f);
    }
    void load_device_rng()
    {
        std::FILE* f = std::fopen(params.rng_device_state.c_str(), "r");
        if (!f) throw std::runtime_error("unable to open file " + params.rng_device_state);
        int seed = 0, traj = 0;
        const int n = std::fscanf(f, " chmc_heatbath_philox4x64 seed %d trajectory %d", &seed, &traj);
        std::fclose(f);
        if (n != 2) throw std::runtime_error("invalid rng_device_state file " + params.rng_device_state);
        if (seed != params.seed) throw std::runtime_error("rng_device_state was written with another seed");
        if (traj != params.istart)
            throw std::runtime_error("rng_device_state (trajectory " + std::to_string(traj) +
                                     ") does not match the last configuration of the xyz file (trajectory " +
                                     std::to_string(params.istart) + ")");
    }
    void save_host_rng_state() const
    {
        std::ofstream fout(params.rng_host_state);
        fout << hmc->gen64;
    }

    // HMC_class::run (HMC.cpp:61-136): chunks of trajectories up to each save point
    void run()
    {
        const int first = params.istart + 1, last = Ntrajectories + params.istart + 1;
        std::vector<double> xyz(3 * (size_t)particles->N);
        int i = first;
        while (i < last) {
            int j = i;                                          // last trajectory of the chunk
            while (j + 1 < last && j % save_every != 0) j++;
            hmc->params.istart = i - 1;
            hmc->params.Ntrajectories = j - i + 1;
            const std::vector<chmc_hmc_record> rec = hmc->run_device(true);
            for (int t = 0; t < j - i + 1; t++)
                std::printf("trajectory %d  dh= %.12g  exp_mdh= %.12g  %s  V= %.12g  K= %.12g\n", i + t, rec[t].dh,
                            std::exp(-rec[t].dh), rec[t].accepted ? "accepted" : "rejected", rec[t].u, rec[t].k);
            params.istart = j;
            if (j >= thermalization_steps && j % save_every == 0) {
                check(chmc_get_positions(particles->ctx(), 0, xyz.data()), particles->ctx(), "get_positions");
                print_xyz(params.fileout, name_xyz, j, rec.back().k, rec.back().u, xyz);
                save_host_rng_state();
                save_device_rng();
            }
            i = j + 1;
        }
        acceptance = hmc->acceptance;
        std::printf("Acceptance: %g\n", acceptance / ((double)(Ntrajectories - thermalization_steps)));
    }

    // HMC_class::measure (HMC.cpp:139-166) with write_header_RDF (identical_particles.cpp:784-796)
    void measure()
    {
        if (filename_RDF.empty() || particles->NbRDF <= 0) throw std::runtime_error("no RDF block in the input file");
        std::FILE* file = std::fopen(params.nameout.c_str(), "r");
        if (!file) throw std::runtime_error("error in opening file " + params.nameout);
        const int confs = how_many_confs_xyz(file, particles->N);
        std::printf("the input file contains %d configurations \n", confs);
        std::FILE* frdf = std::fopen(filename_RDF.c_str(), "w");
        if (!frdf) { std::fclose(file); throw std::runtime_error("error in opening file " + filename_RDF); }
        const identical_particles& p = *particles;
        std::fprintf(frdf, "%g\n%g\n%g\n", p.L[0], p.L[1], p.L[2]);
        std::fprintf(frdf, "%d\n", p.N);
        std::fprintf(frdf, "%g\n%g\n%g\n%g\n%g\n", p.mass, p.beta, p.cutoff, p.eps, p.sigma);
        std::fprintf(frdf, "%g\n%g\n%d\n", LmaxRDF, size_bRDF, confs);
        std::vector<double> xyz;
        config<coordinate> x(p.N, 3);
        for (int c = 0; c < confs; c++) {
            read_next_confs_xyz(file, p.N, name_xyz, xyz);
            for (int i = 0; i < p.N; i++) x[i] = coordinate(xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]);
            particles->set_x(x);
            particles->compute_RDF();
            particles->write_RDF(frdf, c);
        }
        std::fclose(frdf);
        std::fclose(file);
    }
};

}  // namespace chmc_host
