// C++ host driver over the C ABI, keeping the reference's class and method names so that a caller of
// the Kokkos engine finds the same surface:
//   identical_particles : the particles_type "plugin" interface (code/include/particles.hpp:73-98,
//                         code/modules/identical_particles.cpp) -- hb, compute_potential,
//                         evaluate_potential, compute_kinetic_E, compute_force, update_momenta,
//                         update_positions, create_binning, compute_RDF, h_x host mirror
//   LEAP / OMF2 / OMF4  : integrator_type::integrate (code/modules/integrator.cpp:52-131)
//   set_integrator      : code-carsten/integrator.hh:243-278 numbering (0 leapfrog, 2 omf4, 4 euler,
//                         5 ruth, 6 omf2) acting on a hamiltonian_field (integrator.hh:17-22)
//   HMC_class           : HMC_class::run (code/modules/HMC.cpp:61-136), host loop with the reference's
//                         mt19937_64 Metropolis; run_device() runs the same chain fully on the GPU
// Errors from the C ABI are raised as std::runtime_error (the reference aborts).
#pragma once

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <memory>
#include <random>
#include <stdexcept>
#include <string>
#include <vector>

#include "../chmc.h"
#include "config.hh"
#include "hamiltonian_field.hh"

namespace chmc_host {

inline void check(int rc, chmc_ctx* ctx, const char* what)
{
    if (rc != CHMC_OK) throw std::runtime_error(std::string(what) + ": " + chmc_last_error(ctx));
}

struct particle_params {
    int N = 0;
    double L[3] = { 0, 0, 0 };
    double mass = 1, beta = 1, cutoff = 2.5, eps = 1, sigma = 1;
    bool fp64_forces = false;
    // in-cell order by atom id instead of the reference's history-dependent stable order (Q8): two runs
    // from the same positions then see the same FP32 summation order whatever ran before
    bool order_atom_id = false;
    // RDF (identical_particles.cpp:78-86): bins of width size_bin up to Lmax
    double rdf_Lmax = 0, rdf_size_bin = 0;
    int rdf_bins = 0;
};

class identical_particles {
public:
    int N;
    double L[3];
    double mass, beta, cutoff, eps, sigma;
    config<coordinate> h_x;     // host mirror used to restore the position after a rejection
    int NbRDF;
    double size_bRDF, LmaxRDF;
    std::vector<double> RDF;

    explicit identical_particles(const particle_params& p, int device = -1)
        : N(p.N), L{ p.L[0], p.L[1], p.L[2] }, mass(p.mass), beta(p.beta), cutoff(p.cutoff), eps(p.eps),
          sigma(p.sigma), h_x(p.N, 3), NbRDF(p.rdf_bins), size_bRDF(p.rdf_size_bin), LmaxRDF(p.rdf_Lmax)
    {
        chmc_chain_desc d{};
        d.N = N;
        for (int i = 0; i < 3; i++) d.L[i] = L[i];
        d.cutoff = cutoff; d.sigma = sigma; d.eps = eps; d.mass = mass; d.beta = beta;
        const int flags = (p.fp64_forces ? CHMC_FORCE_FP64 : CHMC_FORCE_FP32) | (p.order_atom_id ? CHMC_ORDER_ATOM_ID : 0);
        check(chmc_create(&ctx_, 1, &d, device, flags), nullptr, "chmc_create");
        if (NbRDF == 0 && size_bRDF > 0) NbRDF = (int)(LmaxRDF / size_bRDF);
        RDF.assign(NbRDF > 0 ? NbRDF : 0, 0.0);
    }
    ~identical_particles() { if (ctx_) chmc_destroy(ctx_); }
    identical_particles(const identical_particles&) = delete;
    identical_particles& operator=(const identical_particles&) = delete;

    chmc_ctx* ctx() { return ctx_; }
    double get_beta() const { return beta; }

    // InitX: project lattice + jitter start (replaces cold/hot starts) or a given configuration
    void InitX(uint64_t seed, double jitter = 0.05) { check(chmc_init_lattice(ctx_, seed, jitter), ctx_, "InitX"); pull_x(); }
    void InitX(const config<coordinate>& x) { set_x(x); }

    void set_x(const config<coordinate>& x)
    {
        check(chmc_set_positions(ctx_, 0, reinterpret_cast<const double*>(x.data())), ctx_, "set_positions");
        pull_x();
    }
    void get_x(config<coordinate>& x) { check(chmc_get_positions(ctx_, 0, reinterpret_cast<double*>(x.data())), ctx_, "get_positions"); }
    void set_p(const config<coordinate>& p) { check(chmc_set_momenta(ctx_, 0, reinterpret_cast<const double*>(p.data())), ctx_, "set_momenta"); }
    void get_p(config<coordinate>& p) { check(chmc_get_momenta(ctx_, 0, reinterpret_cast<double*>(p.data())), ctx_, "get_momenta"); }
    void pull_x() { get_x(h_x); }                      // deep_copy(h_x, x)
    void push_x() { set_x(h_x); }                      // deep_copy(x, h_x)
    // the same copies kept on the device (HMC.cpp:72,103,107,111 without the host round trip); they
    // also keep the neighbour lists of the saved positions, so a restore needs no rebuild
    void save_x() { check(chmc_save_positions(ctx_), ctx_, "save_positions"); }
    void restore_x() { check(chmc_restore_positions(ctx_), ctx_, "restore_positions"); }

    void hb(uint64_t seed, uint64_t traj) { check(chmc_heatbath(ctx_, seed, traj), ctx_, "hb"); }
    double compute_potential() { double v; check(chmc_compute_potential(ctx_, &v), ctx_, "compute_potential"); return v; }
    double evaluate_potential() { double v; check(chmc_evaluate_potential(ctx_, &v), ctx_, "evaluate_potential"); return v; }
    double compute_kinetic_E() { double k; check(chmc_kinetic(ctx_, &k), ctx_, "compute_kinetic_E"); return k; }
    double compute_virial() { double w; check(chmc_virial(ctx_, &w), ctx_, "virial"); return w; }
    void compute_force(bool fp64 = false) { check(chmc_compute_force(ctx_, fp64 ? 1 : 0), ctx_, "compute_force"); }
    void get_force(config<coordinate>& f) { check(chmc_get_force(ctx_, 0, reinterpret_cast<double*>(f.data())), ctx_, "get_force"); }
    void update_momenta(double dt) { check(chmc_update_momenta(ctx_, dt), ctx_, "update_momenta"); }
    void update_positions(double dt) { check(chmc_update_positions(ctx_, dt), ctx_, "update_positions"); }
    void create_binning() { check(chmc_bin(ctx_), ctx_, "create_binning"); }
    void get_bins(std::vector<int>& bincount, std::vector<int>& binoffsets, std::vector<int>& permute_vector)
    {
        int nbin[3], bintot;
        double sb[3];
        check(chmc_get_geometry(ctx_, 0, nbin, sb, &bintot), ctx_, "get_geometry");
        bincount.resize(bintot); binoffsets.resize(bintot); permute_vector.resize(N);
        check(chmc_get_bins(ctx_, 0, bincount.data(), binoffsets.data(), permute_vector.data()), ctx_, "get_bins");
    }
    void compute_RDF()
    {
        if (NbRDF <= 0) throw std::runtime_error("compute_RDF: no RDF bins configured");
        check(chmc_rdf(ctx_, 0, NbRDF, size_bRDF, nullptr, RDF.data()), ctx_, "compute_RDF");
    }
    // write_RDF / print_RDF format of identical_particles.cpp:799-813
    void write_RDF(FILE* file, int iconf) const
    {
        std::fprintf(file, "%d\n", iconf);
        for (int i = 0; i < NbRDF; i++) std::fprintf(file, "%g  %g\n", i * size_bRDF, RDF[i]);
    }

private:
    chmc_ctx* ctx_ = nullptr;
};

// ---- integrators ------------------------------------------------------------------------------------
class integrator_type {
public:
    identical_particles* particles;
    int average_steps, steps;
    double dt;
    integrator_type(identical_particles* p, double dt_, int steps_) : particles(p), average_steps(steps_), steps(steps_), dt(dt_) {}
    virtual ~integrator_type() = default;
    // integrator.cpp:37-49
    void set_binomial_steps(std::mt19937_64& gen64)
    {
        const int K = 2, n = K * average_steps;
        const double pr = 1. / ((double)K);
        steps = 0;
        for (int i = 0; i < n; i++) {
            const double r = ((double)gen64() - gen64.min()) / (gen64.max() - gen64.min());
            if (r < pr) steps++;
        }
    }
    virtual int scheme() const = 0;
    virtual double lambda() const { return 0.0; }
    // the whole sequence on the device, each drift fused into the next binning
    virtual void integrate()
    {
        if (steps < 1) return;
        check(chmc_integrate(particles->ctx(), scheme(), dt, steps, lambda()), particles->ctx(), "integrate");
    }
};

class LEAP : public integrator_type {
public:
    using integrator_type::integrator_type;
    int scheme() const override { return CHMC_LEAPFROG; }
};

class OMF2 : public integrator_type {
public:
    double lambda_;
    OMF2(identical_particles* p, double dt_, int steps_, double lam = 0.1938) : integrator_type(p, dt_, steps_), lambda_(lam) {}
    int scheme() const override { return CHMC_OMF2; }
    double lambda() const override { return lambda_; }
};

class OMF4 : public integrator_type {
public:
    using integrator_type::integrator_type;
    int scheme() const override { return CHMC_OMF4; }
};

// code-carsten interface: dtau = tau / nsteps, integrate(h, params, restore) on a hamiltonian_field
struct md_params {
    double tau;
    size_t nsteps;
    double gettau() const { return tau; }
    size_t getnsteps() const { return nsteps; }
};

enum integrators { LEAPFROG = 0, LP_LEAPFROG = 1, OMF4_ = 2, LP_OMF4 = 3, EULER = 4, RUTH = 5, OMF2_ = 6 };

class integrator {
public:
    // exponent: n_prec of the lp_* schemes (their grid 10^-n_prec), ignored by the others
    explicit integrator(int scheme, double lam = 0.1938, int exponent = 16) : scheme_(scheme), lam_(lam), exponent_(exponent) {}
    // the system's force field takes the role of the monomial list; the field's configs are uploaded,
    // integrated on the GPU and downloaded again
    void integrate(identical_particles& sys, hamiltonian_field<coordinate>& h, const md_params& params, bool restore = true)
    {
        (void)restore;   // the wrap into the box is part of every drift (identical_particles.cpp:667)
        sys.set_x(*h.U);
        sys.set_p(*h.momenta);
        const double dtau = params.gettau() / double(params.getnsteps());
        if (scheme_ == CHMC_LP_LEAPFROG || scheme_ == CHMC_LP_OMF4)
            check(chmc_set_lp_exponent(sys.ctx(), exponent_), sys.ctx(), "set_lp_exponent");
        check(chmc_integrate(sys.ctx(), scheme_, dtau, (int)params.getnsteps(), lam_), sys.ctx(), "integrate");
        sys.get_x(*h.U);
        sys.get_p(*h.momenta);
    }
    int scheme() const { return scheme_; }
    int exponent() const { return exponent_; }

private:
    int scheme_;
    double lam_;
    int exponent_;
};

// integrator.hh:243-278: lp_leapfrog(exponent) and lp_omf4(exponent) round every update to the grid
// 10^-exponent; unknown values fall back to leapfrog, as there (:272-275)
inline integrator* set_integrator(size_t integs, size_t exponent = 16)
{
    switch (integs) {
    case LP_LEAPFROG: return new integrator(CHMC_LP_LEAPFROG, 0.1938, (int)exponent);
    case LP_OMF4: return new integrator(CHMC_LP_OMF4, 0.1938, (int)exponent);
    case OMF4_: return new integrator(CHMC_OMF4);
    case EULER: return new integrator(CHMC_EULER);
    case RUTH: return new integrator(CHMC_RUTH);
    case OMF2_: return new integrator(CHMC_OMF2);
    default: return new integrator(CHMC_LEAPFROG);
    }
}

// ---- HMC ----------------------------------------------------------------------------------------------
struct hmc_params {
    int Ntrajectories = 0, thermalization_steps = 0, save_every = 1, istart = 0;
    bool randomize_traj = false;
    uint64_t seed = 1;
};

struct traj_record {
    int traj;
    double Ki, Vi, Kf, Vf, dh;
    bool accepted;
};

class HMC_class {
public:
    integrator_type* integrator;
    hmc_params params;
    int acceptance = 0;
    std::mt19937_64 gen64;
    std::vector<traj_record> history;

    HMC_class(integrator_type* integ, const hmc_params& p) : integrator(integ), params(p), gen64(p.seed) {}

    double gen_random() { return (((double)gen64() - gen64.min()) / (gen64.max() - gen64.min())); }

    // HMC.cpp:61-136 with the host stepping every stage (one device round trip per quantity)
    void run(bool verbose = false)
    {
        identical_particles* P = integrator->particles;
        double Vi = P->compute_potential();
        const double beta = P->get_beta();
        P->save_x();
        P->pull_x();
        const int first = params.istart + 1, last = params.Ntrajectories + params.istart + 1;
        for (int i = first; i < last; i++) {
            P->hb(params.seed, (uint64_t)i);
            double Ki = P->compute_kinetic_E();
            if (params.randomize_traj) integrator->set_binomial_steps(gen64);
            integrator->integrate();
            const double Vf = P->compute_potential();
            const double Kf = P->compute_kinetic_E();
            const double dh = beta * (Kf + Vf - Ki - Vi);
            const double exp_mdh = std::exp(-dh);
            bool acc;
            if (i < params.thermalization_steps) {
                acc = true;
            } else {
                const double r = gen_random();
                acc = r < exp_mdh;
                if (acc) acceptance++;
            }
            if (acc) { Vi = Vf; Ki = Kf; P->save_x(); P->pull_x(); }
            else P->restore_x();
            history.push_back(traj_record{ i, Ki, Vi, Kf, Vf, dh, acc });
            if (verbose) std::printf("traj %d dh= %.12g accepted= %d V= %.12g\n", i, dh, (int)acc, Vi);
        }
    }

    // the same chain with every stage on the device (hipGraph-captured trajectory); the Metropolis
    // uniforms and binomial trajectory lengths come from this object's gen64 in the order run() draws
    // them, so run() and run_device() consume one stream identically
    std::vector<chmc_hmc_record> run_device(bool use_graph = true)
    {
        identical_particles* P = integrator->particles;
        std::vector<double> u(params.Ntrajectories, 2.0);
        std::vector<int> st;
        int q = 0;
        for (int t = 0; t < params.Ntrajectories; t++) {
            if (params.randomize_traj) {
                integrator->set_binomial_steps(gen64);
                st.push_back(integrator->steps);
            }
            if (params.istart + 1 + t >= params.thermalization_steps) u[q++] = gen_random();
        }
        chmc_hmc_params hp{};
        hp.uniforms = u.data();
        hp.traj_steps = params.randomize_traj ? st.data() : nullptr;
        hp.ntraj = params.Ntrajectories;
        hp.therm = params.thermalization_steps;
        hp.scheme = integrator->scheme();
        hp.steps = integrator->average_steps;
        hp.dt = integrator->dt;
        hp.lambda = integrator->lambda();
        hp.seed = params.seed;
        hp.use_graph = use_graph ? 1 : 0;
        hp.randomize = params.randomize_traj ? 1 : 0;
        std::vector<chmc_hmc_record> rec(params.Ntrajectories);
        int nacc = 0;
        check(chmc_set_trajectory(P->ctx(), params.istart), P->ctx(), "set_trajectory");
        check(chmc_hmc_run(P->ctx(), &hp, rec.data(), &nacc), P->ctx(), "hmc_run");
        acceptance += nacc;
        return rec;
    }
};

}  // namespace chmc_host
