// Exercises the C++ host layer (include/chmc/*.hh) on the GPU the way a caller of the reference would:
// code-carsten's config / potential_energy / set_integrator on a hamiltonian_field, and the Kokkos
// engine's identical_particles + OMF2 + HMC_class (host-stepped run() against the device run_device()).
// Prints one "CHECK name ok|FAIL detail" line per check and exits non-zero on any failure.
#include <cmath>
#include <cstdio>
#include <vector>

#include "chmc/config.hh"
#include "chmc/hamiltonian_field.hh"
#include "chmc/hmc.hh"
#include "chmc/lennardjones.hh"
#include "chmc/potential_energy.hh"

using namespace chmc_host;

static int failures = 0;
static void report(const char* name, bool ok, double detail)
{
    std::printf("CHECK %-28s %s (%.3e)\n", name, ok ? "ok" : "FAIL", detail);
    if (!ok) failures++;
}

int main()
{
    // ---- code-carsten: potential_energy over a config<coordinate> (open boundaries, all pairs) ----
    config<coordinate> U(64, 3);
    std::mt19937 eng(7);
    std::uniform_real_distribution<double> dist(0.0, 0.8);
    for (size_t i = 0; i < U.getN(); i++) U[i] = coordinate(dist(eng), dist(eng), dist(eng));
    double host_v = 0.0;
    for (size_t i = 0; i + 1 < U.getN(); i++)
        for (size_t j = i + 1; j < U.getN(); j++) host_v += LennardJones((U[i] - U[j]).norm());
    const double gpu = potential_energy(U);
    report("potential_energy", std::fabs(gpu - host_v) <= 1e-12 * std::fabs(host_v), std::fabs(gpu - host_v) / std::fabs(host_v));

    // ---- identical_particles + integrators ----
    const int N = 512;
    particle_params pp;
    pp.N = N;
    pp.L[0] = pp.L[1] = pp.L[2] = std::cbrt(N / 0.8);
    pp.beta = 1 / 1.5;
    pp.rdf_Lmax = pp.L[0] / 2;
    pp.rdf_size_bin = pp.L[0] / 400;
    pp.rdf_bins = 200;
    identical_particles P(pp);
    P.InitX(1, 0.05);
    const double V0 = P.compute_potential();
    report("compute_potential finite", std::isfinite(V0) && V0 < 0, V0);
    P.compute_force();
    config<coordinate> f(N, 3);
    P.get_force(f);
    double fsum[3] = { 0, 0, 0 };
    for (int i = 0; i < N; i++)
        for (int d = 0; d < 3; d++) fsum[d] += f[i][d];
    report("newton third law", std::fabs(fsum[0]) + std::fabs(fsum[1]) + std::fabs(fsum[2]) < 1e-2, std::fabs(fsum[0]));
    std::vector<int> cnt, off, perm;
    P.get_bins(cnt, off, perm);
    int tot = 0;
    for (int c : cnt) tot += c;
    report("bins cover all atoms", tot == N, tot);
    P.compute_RDF();
    report("rdf populated", P.RDF[150] > 0.5, P.RDF[150]);

    // set_integrator on a hamiltonian_field: forward then backward with flipped momenta
    config<coordinate> x(N, 3), p(N, 3), x0(N, 3);
    P.get_x(x);
    x0 = x;
    P.hb(3, 1);
    P.get_p(p);
    hamiltonian_field<coordinate> h(p, x);
    std::unique_ptr<integrator> integ(set_integrator(OMF2_));
    md_params md{ 0.1, 20 };
    P.compute_force(true);
    integ->integrate(P, h, md);
    for (int i = 0; i < N; i++)
        for (int d = 0; d < 3; d++) p[i][d] = -p[i][d];
    integ->integrate(P, h, md);
    double dmax = 0;
    for (int i = 0; i < N; i++)
        for (int d = 0; d < 3; d++) {
            double dd = x[i][d] - x0[i][d];
            dd -= pp.L[d] * std::round(dd / pp.L[d]);
            dmax = std::fmax(dmax, std::fabs(dd));
        }
    report("omf2 reversibility (fp32)", dmax < 1e-4, dmax);

    // set_integrator(LP_LEAPFROG / LP_OMF4, exponent): every update rounded to the grid 10^-exponent
    for (size_t integs : { (size_t)LP_LEAPFROG, (size_t)LP_OMF4 }) {
        std::unique_ptr<integrator> lp(set_integrator(integs, 6));
        lp->integrate(P, h, md);
        double off = 0;
        for (int i = 0; i < N; i++)
            for (int d = 0; d < 3; d++) {
                off = std::fmax(off, std::fabs(std::round(x[i][d] * 1e6) / 1e6 - x[i][d]));
                off = std::fmax(off, std::fabs(std::round(p[i][d] * 1e6) / 1e6 - p[i][d]));
            }
        report(integs == LP_LEAPFROG ? "lp_leapfrog(6) states on the 1e-6 grid" : "lp_omf4(6) states on the 1e-6 grid",
               off == 0.0 && lp->exponent() == 6 && lp->scheme() == (int)integs, off);
    }

    // ---- HMC_class: host-stepped loop vs the device-resident loop, same seeds ----
    // (in-cell order by atom id, so that both runs sum the FP32 forces in the same order although the
    // first one leaves a different binning history behind)
    particle_params pq = pp;
    pq.order_atom_id = true;
    identical_particles Q(pq);
    Q.InitX(5, 0.05);
    config<coordinate> xs(N, 3);
    Q.get_x(xs);
    LEAP leap(&Q, 0.005, 20);
    hmc_params hp;
    hp.Ntrajectories = 30;
    hp.thermalization_steps = 5;
    hp.seed = 11;
    HMC_class host(&leap, hp);
    host.run();
    config<coordinate> xh(N, 3);
    Q.get_x(xh);
    Q.set_x(xs);
    HMC_class dev(&leap, hp);
    std::vector<chmc_hmc_record> rec = dev.run_device(true);
    config<coordinate> xd(N, 3);
    Q.get_x(xd);
    int same_acc = 0;
    double max_ddh = 0;
    for (int t = 0; t < hp.Ntrajectories; t++) {
        same_acc += (int)host.history[t].accepted == rec[t].accepted;
        max_ddh = std::fmax(max_ddh, std::fabs(host.history[t].dh - rec[t].dh));
    }
    report("hmc host == device accepts", same_acc == hp.Ntrajectories, same_acc);
    report("hmc host == device dH", max_ddh < 1e-8, max_ddh);
    double xdiff = 0;
    for (int i = 0; i < N; i++)
        for (int d = 0; d < 3; d++) xdiff = std::fmax(xdiff, std::fabs(xh[i][d] - xd[i][d]));
    report("hmc host == device final x", xdiff < 1e-8, xdiff);
    report("acceptance counted", host.acceptance == dev.acceptance, host.acceptance - dev.acceptance);
    const double rate = host.acceptance / (double)(hp.Ntrajectories - hp.thermalization_steps);
    report("acceptance rate > 0.5", rate > 0.5, rate);
    std::printf("%s\n", failures ? "FAILED" : "ALL OK");
    return failures ? 1 : 0;
}
