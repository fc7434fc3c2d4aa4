// Cantera-YAML mechanism loader (host side) and the flat coefficient tables the
// HIP kernels read.
//
// Replaces the mechanism load the reference does with
// `Cantera::newSolution(mechanism, "", "none")` (HyperFlow_V3/src/chemistry_batch.cpp:17-18,
// HyperFlow_V3/src/BlockAMR.cpp:29): species with NASA7 thermo (one or two
// ranges), elementary / three-body / Troe-falloff reactions (3- or 4-parameter
// Troe), duplicates and third-body efficiencies, in either unit system the two
// mechanism files use (SI-kmol default, or `units: {length: cm, quantity: mol}`).
// Transport data are ignored, as with the reference's "none" transport model.
#pragma once
#include <array>
#include <cstdint>
#include <string>
#include <vector>

namespace hf {

constexpr double kGasConstant = 8314.462618;   // J kmol^-1 K^-1 (N_A k_B, exact)
constexpr double kOneAtm = 101325.0;           // Pa, NASA7 reference pressure
constexpr double kCalToJ = 4.184;

struct Reaction {
    enum Kind { ELEMENTARY = 0, THREE_BODY = 1, FALLOFF = 2 };
    std::string equation;
    int kind = ELEMENTARY;
    bool reversible = true, duplicate = false;
    // Species index repeated per unit of stoichiometry ("OH + OH" -> {OH, OH}).
    std::vector<int> reactants, products;
    double A = 0, b = 0, EaR = 0;        // k (or k_inf for falloff); SI kmol units, Ea/R in K
    double A0 = 0, b0 = 0, Ea0R = 0;     // falloff low-pressure limit
    int troe = 0;                        // 0: Lindemann; 3 or 4: Troe parameter count
    double troeA = 0, troeT3 = 0, troeT1 = 0, troeT2 = 0;
    std::vector<std::pair<int, double>> efficiencies;  // non-default third-body efficiencies
    double default_efficiency = 1.0;
};

struct Mechanism {
    std::string path;
    std::vector<std::string> elements, species;
    std::vector<double> W;                           // kg/kmol
    std::vector<std::vector<double>> composition;    // [species][element]
    std::vector<double> Tmin, Tmid, Tmax;
    std::vector<std::array<double, 7>> low, high;    // NASA7 coefficient sets
    std::vector<Reaction> reactions;

    int ns() const { return (int)species.size(); }
    int nr() const { return (int)reactions.size(); }
    int index(const std::string& name) const;      // -1 when absent
    // Topology signature: species order, reaction structure and efficiency
    // species, without the numeric coefficients. Compiled kernels are
    // specialised on it.
    std::string topology() const;
};

// Throws std::runtime_error with a message naming the file and the problem.
Mechanism load_mechanism(const std::string& path);

// Standard atomic weight (kg/kmol) for the elements the mechanisms use.
double atomic_weight(const std::string& element);

}  // namespace hf
