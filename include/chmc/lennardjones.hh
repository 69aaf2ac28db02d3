// LennardJones(r, n, a, b) = a / r^n - b / r^6 with the defaults of code-carsten/lennardjones.hh:5-9
// (n = 12, a = 1e-12, b = 1e-6: 4 eps [(sigma/r)^12 - (sigma/r)^6] with sigma = 0.1, eps = 0.25).
#pragma once
#include <cmath>

inline double LennardJones(const double r, const int n = 12, const double a = 1.e-12, const double b = 1.e-6)
{
    return a / std::pow(r, n) - b / std::pow(r, 6);
}
