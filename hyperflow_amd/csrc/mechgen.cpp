// Build-time tool: reads a Cantera-YAML mechanism with the product loader and
// writes the header that fixes its *topology* (species order, reactant and
// product lists, reaction kinds, efficiency species) as compile-time tables.
// The kernels are specialised on this topology so every species index in the
// rate and Jacobian code is a constant and per-cell arrays stay in registers.
// The numeric coefficients (NASA7, Arrhenius, Troe, efficiencies) are *not*
// compiled in: they are loaded from the YAML at run time by hf_create, which
// also checks that the file's topology signature equals the compiled one.
//
// usage: mechgen <mechanism.yaml> <struct_name> <out.inc>
#include <algorithm>
#include <cstdio>
#include <fstream>
#include <iostream>

#include "mechanism.hpp"

int main(int argc, char** argv) {
    if (argc != 4) {
        std::cerr << "usage: mechgen <mechanism.yaml> <struct_name> <out.inc>\n";
        return 2;
    }
    hf::Mechanism m;
    try {
        m = hf::load_mechanism(argv[1]);
    } catch (const std::exception& e) {
        std::cerr << "mechgen: " << e.what() << "\n";
        return 1;
    }
    const std::string name = argv[2];
    const int ns = m.ns(), nr = m.nr();
    int nrmax = 1, neffmax = 1;
    for (auto& r : m.reactions) {
        nrmax = std::max<int>(nrmax, std::max(r.reactants.size(), r.products.size()));
        neffmax = std::max<int>(neffmax, r.efficiencies.size());
    }
    std::ofstream o(argv[3]);
    o << "// Generated by mechgen from " << m.path << ". Topology only; coefficients load at run time.\n";
    o << "#pragma once\n";
    o << "namespace hf_gen {\n";
    o << "struct " << name << " {\n";
    o << "    static constexpr int NS = " << ns << ", NR = " << nr << ", NE = " << m.elements.size()
      << ", NRMAX = " << nrmax << ", NEFF = " << neffmax << ";\n";
    auto arr1 = [&](const char* type, const char* field, auto get) {
        o << "    static constexpr " << type << " " << field << "[NR] = {";
        for (int i = 0; i < nr; ++i) o << (i ? "," : "") << get(m.reactions[i]);
        o << "};\n";
    };
    auto arr2 = [&](const char* field, int width, auto get) {
        o << "    static constexpr int " << field << "[NR][" << width << "] = {";
        for (int i = 0; i < nr; ++i) {
            o << (i ? "," : "") << "{";
            auto v = get(m.reactions[i]);
            for (int j = 0; j < width; ++j) o << (j ? "," : "") << (j < (int)v.size() ? v[j] : -1);
            o << "}";
        }
        o << "};\n";
    };
    arr1("int", "kind", [](const hf::Reaction& r) { return r.kind; });
    arr1("int", "rev", [](const hf::Reaction& r) { return r.reversible ? 1 : 0; });
    arr1("int", "troe", [](const hf::Reaction& r) { return r.troe; });
    arr1("int", "nreac", [](const hf::Reaction& r) { return (int)r.reactants.size(); });
    arr1("int", "nprod", [](const hf::Reaction& r) { return (int)r.products.size(); });
    arr1("int", "neff", [](const hf::Reaction& r) { return (int)r.efficiencies.size(); });
    arr2("reac", nrmax, [](const hf::Reaction& r) { return r.reactants; });
    arr2("prod", nrmax, [](const hf::Reaction& r) { return r.products; });
    arr2("effsp", neffmax, [](const hf::Reaction& r) {
        std::vector<int> v;
        for (auto& e : r.efficiencies) v.push_back(e.first);
        return v;
    });
    o << "    static constexpr const char* topology = \"" << m.topology() << "\";\n";
    o << "};\n}  // namespace hf_gen\n";
    return o ? 0 : 1;
}
