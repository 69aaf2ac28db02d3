// C-ABI shim over the reference's own code-carsten headers (compiled from the reference tree by
// oracle/Makefile `ref`).  TEST INFRASTRUCTURE ONLY: it lets tests compare the oracle's LJ pair
// arithmetic with the reference's LennardJones() and potential_energy() on the same inputs.
#include "config.hh"
#include "coordinate.hh"
#include "lennardjones.hh"
#include "potential_energy.hh"

extern "C" double ref_lennard_jones(double r, int n, double a, double b)
{
  return LennardJones(r, n, a, b);
}

// potential_energy<coordinate>(config&) with the header's default LennardJones parameters.
extern "C" double ref_potential_energy(const double* xyz, int n_particles)
{
  config<coordinate> U(static_cast<size_t>(n_particles), 3);
  for (int i = 0; i < n_particles; i++) U[i] = coordinate(&xyz[3 * i]);
  return potential_energy(U);
}
