// Host thermochemistry helpers for case set-up and synthetic inputs.
#pragma once
#include <cstdint>

#include "mechanism.hpp"

namespace hf {

// cp/R, h/RT, s/R per species (any pointer may be null).
void species_thermo(const Mechanism& m, double T, double* cp, double* h, double* s);
double mass_energy(const Mechanism& m, double T, const double* Y);   // J/kg
double mean_weight(const Mechanism& m, const double* Y);            // kg/kmol
void mole_to_mass(const Mechanism& m, const double* X, double* Y);
// Conserved state [rhoY, rho u, rho v, rho E] of a gas at (T, p, Y, u, v).
void state_from_TPY(const Mechanism& m, double T, double p, const double* Y, double u, double v, double* q);
// TP equilibrium mass fractions for the elements of mole fractions X. Throws on failure.
void equilibrate_tp(const Mechanism& m, double T, double p, const double* X, double* Y);

struct Rng {  // xoshiro256** seeded by splitmix64
    uint64_t s[4];
    explicit Rng(uint64_t seed);
    uint64_t next();
    double uniform();  // (x >> 11) * 2^-53
};

// Synthetic cells: U has ns+3 planes of n values; dt may be null.
void synth_cells(const Mechanism& m, int which, uint64_t seed, int64_t n, double* U, double* dt);

}  // namespace hf
