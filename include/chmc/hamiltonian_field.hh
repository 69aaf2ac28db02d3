// hamiltonian_field<T>: non-owning pointers to the momenta and the positions of one system, the shape of
// code-carsten/hamiltonian_field.hh:6-11.
#pragma once
#include "config.hh"

template <class T> struct hamiltonian_field {
    config<T>* momenta;
    config<T>* U;
    hamiltonian_field(config<T>& momenta_, config<T>& U_) : momenta(&momenta_), U(&U_) {}
};
