// potential_energy<T>(config<T>&): sum over i < j of LennardJones(|r_i - r_j|) with open boundaries and no
// cutoff, the function of code-carsten/potential_energy.hh:6-15, evaluated on the GPU by
// chmc_potential_energy_open (all pairs in FP64, per-block partials summed in a fixed order).
#pragma once
#include <stdexcept>
#include <string>

#include "../chmc.h"
#include "config.hh"
#include "lennardjones.hh"

template <class T> double potential_energy(config<T>& U, int n = 12, double a = 1.e-12, double b = 1.e-6)
{
    static_assert(sizeof(T) == 3 * sizeof(double), "potential_energy expects three-double coordinates");
    double res = 0.0;
    const int rc = chmc_potential_energy_open(reinterpret_cast<const double*>(U.data()), (int)U.getN(), n, a, b, &res);
    if (rc != CHMC_OK) throw std::runtime_error(std::string("potential_energy: ") + chmc_last_error(nullptr));
    return res;
}
