// Host-side thermochemistry for case set-up: conserved state from (T, p, Y)
// (the reference's primitive() helper, HyperFlow_V3/src/BlockAMR.cpp:20-24,
// which uses Cantera's TPX/TPY setters), TP equilibrium compositions for the
// burnt driver gas and the synthetic-cell recipes, and the seeded generators.
#include "host_thermo.hpp"

#include <algorithm>
#include <cmath>
#include <stdexcept>
#include <thread>

namespace hf {

void species_thermo(const Mechanism& m, double T, double* cp, double* h, double* s) {
    const double lt = std::log(T);
    for (int k = 0; k < m.ns(); ++k) {
        const auto& a = T <= m.Tmid[k] ? m.low[k] : m.high[k];
        if (cp) cp[k] = a[0] + T * (a[1] + T * (a[2] + T * (a[3] + T * a[4])));
        if (h) h[k] = a[0] + T * (a[1] / 2 + T * (a[2] / 3 + T * (a[3] / 4 + T * a[4] / 5))) + a[5] / T;
        if (s) s[k] = a[0] * lt + T * (a[1] + T * (a[2] / 2 + T * (a[3] / 3 + T * a[4] / 4))) + a[6];
    }
}

double mass_energy(const Mechanism& m, double T, const double* Y) {
    std::vector<double> h(m.ns());
    species_thermo(m, T, nullptr, h.data(), nullptr);
    double e = 0;
    for (int k = 0; k < m.ns(); ++k) e += Y[k] / m.W[k] * (h[k] - 1.0) * kGasConstant * T;
    return e;
}

double mean_weight(const Mechanism& m, const double* Y) {
    double s = 0;
    for (int k = 0; k < m.ns(); ++k) s += Y[k] / m.W[k];
    return 1.0 / s;
}

void mole_to_mass(const Mechanism& m, const double* X, double* Y) {
    double w = 0;
    for (int k = 0; k < m.ns(); ++k) w += X[k] * m.W[k];
    for (int k = 0; k < m.ns(); ++k) Y[k] = X[k] * m.W[k] / w;
}

void state_from_TPY(const Mechanism& m, double T, double p, const double* Y, double u, double v, double* q) {
    const int ns = m.ns();
    const double rho = p * mean_weight(m, Y) / (kGasConstant * T);
    for (int k = 0; k < ns; ++k) q[k] = rho * Y[k];
    q[ns] = rho * u;
    q[ns + 1] = rho * v;
    q[ns + 2] = rho * (mass_energy(m, T, Y) + 0.5 * (u * u + v * v));
}

namespace {

// Solve the small dense system A x = b (n x n, row-major) by partial-pivot elimination.
bool solve_dense(int n, std::vector<double> A, std::vector<double>& b) {
    for (int k = 0; k < n; ++k) {
        int p = k;
        for (int i = k + 1; i < n; ++i)
            if (std::abs(A[i * n + k]) > std::abs(A[p * n + k])) p = i;
        if (A[p * n + k] == 0) return false;
        if (p != k) {
            for (int j = 0; j < n; ++j) std::swap(A[k * n + j], A[p * n + j]);
            std::swap(b[k], b[p]);
        }
        for (int i = k + 1; i < n; ++i) {
            const double l = A[i * n + k] / A[k * n + k];
            for (int j = k; j < n; ++j) A[i * n + j] -= l * A[k * n + j];
            b[i] -= l * b[k];
        }
    }
    for (int i = n - 1; i >= 0; --i) {
        double s = b[i];
        for (int j = i + 1; j < n; ++j) s -= A[i * n + j] * b[j];
        b[i] = s / A[i * n + i];
    }
    return true;
}

int species_with(const Mechanism& m, std::initializer_list<std::pair<const char*, double>> comp) {
    for (int k = 0; k < m.ns(); ++k) {
        bool match = true;
        double atoms = 0;
        for (auto& c : comp) {
            int e = -1;
            for (size_t j = 0; j < m.elements.size(); ++j)
                if (m.elements[j] == c.first) e = (int)j;
            if (e < 0 || m.composition[k][e] != c.second) { match = false; break; }
            atoms += c.second;
        }
        double total = 0;
        for (double a : m.composition[k]) total += a;
        if (match && total == atoms) return k;
    }
    return -1;
}

}  // namespace

// Element-potential Newton for the ideal-gas Gibbs minimum at fixed (T, p):
// n_k = N exp(sum_e a_ek lambda_e - g_k/RT - ln(p/p0)), with sum_k a_ek n_k = b_e
// and sum_k n_k / N = 1. Started from the complete-combustion products.
void equilibrate_tp(const Mechanism& m, double T, double p, const double* X, double* Y) {
    const int ns = m.ns(), ne = (int)m.elements.size();
    std::vector<double> b(ne, 0.0);
    for (int k = 0; k < ns; ++k)
        for (int e = 0; e < ne; ++e) b[e] += m.composition[k][e] * X[k];
    // elements present; species made only of present elements
    std::vector<int> el;
    for (int e = 0; e < ne; ++e)
        if (b[e] > 0) el.push_back(e);
    std::vector<int> sp;
    for (int k = 0; k < ns; ++k) {
        bool ok = true;
        for (int e = 0; e < ne; ++e)
            if (m.composition[k][e] > 0 && !(b[e] > 0)) ok = false;
        if (ok) sp.push_back(k);
    }
    const int E = (int)el.size(), S = (int)sp.size();
    std::vector<double> cp(ns), h(ns), s(ns), gt(ns);
    species_thermo(m, T, cp.data(), h.data(), s.data());
    for (int k = 0; k < ns; ++k) gt[k] = h[k] - s[k] + std::log(p / kOneAtm);

    // complete-combustion estimate
    std::vector<double> n0(ns, 0.0);
    auto idx = [&](const char* e) {
        for (int j = 0; j < ne; ++j)
            if (m.elements[j] == e) return j;
        return -1;
    };
    const int eC = idx("C"), eH = idx("H"), eO = idx("O"), eN = idx("N");
    const double bC = eC >= 0 ? b[eC] : 0, bH = eH >= 0 ? b[eH] : 0, bO = eO >= 0 ? b[eO] : 0, bN = eN >= 0 ? b[eN] : 0;
    const int kCO2 = species_with(m, {{"C", 1}, {"O", 2}}), kCO = species_with(m, {{"C", 1}, {"O", 1}});
    const int kH2O = species_with(m, {{"H", 2}, {"O", 1}}), kH2 = species_with(m, {{"H", 2}});
    const int kO2 = species_with(m, {{"O", 2}}), kN2 = species_with(m, {{"N", 2}});
    double co2 = bC, h2o = bH / 2, co = 0, h2 = 0, o2 = 0;
    double orem = bO - 2 * bC - bH / 2;
    if (orem >= 0) {
        o2 = orem / 2;
    } else {
        double def = -orem;
        const double k1 = std::min(def, co2);
        co2 -= k1; co += k1; def -= k1;
        h2o -= def; h2 += def;
    }
    if (kCO2 >= 0) n0[kCO2] = co2;
    if (kCO >= 0) n0[kCO] = co;
    if (kH2O >= 0) n0[kH2O] = h2o;
    if (kH2 >= 0) n0[kH2] = h2;
    if (kO2 >= 0) n0[kO2] = o2;
    if (kN2 >= 0) n0[kN2] = bN / 2;
    double N = 0;
    for (int k = 0; k < ns; ++k) N += n0[k];
    if (!(N > 0)) throw std::runtime_error("equilibrate: no products");
    // trace floor so that the element basis can always be completed
    for (int k : sp) n0[k] = std::max(n0[k], 1e-8 * N);
    // basis: the largest products that are linearly independent in element space
    std::vector<int> order(ns);
    for (int k = 0; k < ns; ++k) order[k] = k;
    std::sort(order.begin(), order.end(), [&](int a, int c) { return n0[a] > n0[c]; });
    std::vector<int> basis;
    std::vector<std::vector<double>> rows;
    for (int k : order) {
        if ((int)basis.size() == E || !(n0[k] > 0)) break;
        std::vector<double> r(E);
        for (int e = 0; e < E; ++e) r[e] = m.composition[k][el[e]];
        // Gram-Schmidt-like independence test
        std::vector<double> v = r;
        for (auto& q : rows) {
            double dq = 0, qq = 0;
            for (int e = 0; e < E; ++e) { dq += v[e] * q[e]; qq += q[e] * q[e]; }
            for (int e = 0; e < E; ++e) v[e] -= dq / qq * q[e];
        }
        double nv = 0;
        for (double x : v) nv += x * x;
        if (nv > 1e-12) { rows.push_back(v); basis.push_back(k); }
    }
    if ((int)basis.size() != E) throw std::runtime_error("equilibrate: no element basis");
    std::vector<double> A(E * E), lam(E);
    for (int i = 0; i < E; ++i) {
        const int k = basis[i];
        for (int e = 0; e < E; ++e) A[i * E + e] = m.composition[k][el[e]];
        lam[i] = std::log(n0[k] / N) + gt[k];
    }
    if (!solve_dense(E, A, lam)) throw std::runtime_error("equilibrate: singular basis");
    double lnN = std::log(N);
    std::vector<double> x(ns, 0.0);
    bool converged = false;
    for (int it = 0; it < 500 && !converged; ++it) {
        for (int k : sp) {
            double z = -gt[k];
            for (int e = 0; e < E; ++e) z += m.composition[k][el[e]] * lam[e];
            x[k] = std::exp(std::min(z, 600.0));
        }
        const double Nn = std::exp(lnN);
        std::vector<double> r(E + 1, 0.0), J((E + 1) * (E + 1), 0.0);
        double rmax = 0;
        for (int e = 0; e < E; ++e) {
            double s1 = 0;
            for (int k : sp) s1 += m.composition[k][el[e]] * x[k];
            r[e] = (Nn * s1 - b[el[e]]) / b[el[e]];
            rmax = std::max(rmax, std::abs(r[e]));
            for (int f = 0; f < E; ++f) {
                double s2 = 0;
                for (int k : sp) s2 += m.composition[k][el[e]] * m.composition[k][el[f]] * x[k];
                J[e * (E + 1) + f] = Nn * s2 / b[el[e]];
            }
            J[e * (E + 1) + E] = Nn * s1 / b[el[e]];
        }
        double sx = 0;
        for (int k : sp) sx += x[k];
        r[E] = sx - 1.0;
        rmax = std::max(rmax, std::abs(r[E]));
        for (int f = 0; f < E; ++f) {
            double s2 = 0;
            for (int k : sp) s2 += m.composition[k][el[f]] * x[k];
            J[E * (E + 1) + f] = s2;
        }
        J[E * (E + 1) + E] = 0;
        if (rmax < 1e-14) { converged = true; break; }
        for (double& v : r) v = -v;
        if (!solve_dense(E + 1, J, r)) throw std::runtime_error("equilibrate: singular Newton matrix");
        double scale = 1.0;
        for (int e = 0; e < E; ++e) scale = std::min(scale, 2.0 / std::max(std::abs(r[e]), 1e-300));
        scale = std::min(scale, 1.0 / std::max(std::abs(r[E]), 1e-300));
        scale = std::min(scale, 1.0);
        for (int e = 0; e < E; ++e) lam[e] += scale * r[e];
        lnN += scale * r[E];
        if (rmax < 1e-12 && scale == 1.0 && it > 100) converged = true;
    }
    if (!converged) throw std::runtime_error("equilibrate: Newton did not converge");
    std::vector<double> n(ns, 0.0);
    for (int k : sp) n[k] = x[k];
    double sx = 0;
    for (double v : n) sx += v;
    for (double& v : n) v /= sx;
    mole_to_mass(m, n.data(), Y);
}

// splitmix64 seeding of xoshiro256**
Rng::Rng(uint64_t seed) {
    uint64_t z = seed;
    for (int i = 0; i < 4; ++i) {
        z += 0x9e3779b97f4a7c15ull;
        uint64_t t = z;
        t = (t ^ (t >> 30)) * 0xbf58476d1ce4e5b9ull;
        t = (t ^ (t >> 27)) * 0x94d049bb133111ebull;
        s[i] = t ^ (t >> 31);
    }
}
uint64_t Rng::next() {
    auto rotl = [](uint64_t x, int k) { return (x << k) | (x >> (64 - k)); };
    const uint64_t result = rotl(s[1] * 5, 7) * 9;
    const uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl(s[3], 45);
    return result;
}
double Rng::uniform() { return (double)(next() >> 11) * 0x1.0p-53; }

void synth_cells(const Mechanism& m, int which, uint64_t seed, int64_t n, double* U, double* dt) {
    const int ns = m.ns();
    const int nv = ns + 3;
    int kF = -1;
    const int kO2 = m.index("O2"), kN2 = m.index("N2");
    double Tlo, Thi, plo, phi_p, flo, fhi;
    bool logp = true, progress = true;
    double fuel_per_phi, o2_amount, n2_amount;
    if (which == 0 || which == 2) {
        kF = m.index("H2");
        fuel_per_phi = 2.0; o2_amount = 1.0; n2_amount = 3.76;
        if (which == 0) { Tlo = 1100; Thi = 2400; plo = 0.5; phi_p = 50; flo = 0.5; fhi = 2.0; }
        else { Tlo = 1000; Thi = 1500; plo = 1; phi_p = 10; flo = 0.5; fhi = 2.0; logp = false; progress = false; }
    } else if (which == 1) {
        kF = m.index("CH3OCH3");
        fuel_per_phi = 1.0; o2_amount = 3.0; n2_amount = 11.28;
        Tlo = 700; Thi = 1800; plo = 1; phi_p = 40; flo = 0.6; fhi = 1.5;
    } else {
        throw std::runtime_error("unknown synthetic workload");
    }
    if (kF < 0 || kO2 < 0 || kN2 < 0) throw std::runtime_error("mechanism lacks the fuel/O2/N2 of this workload");
    // Draws are sequential per cell (one stream), so the cell set does not depend on threading.
    std::vector<double> draws((size_t)n * 5);
    Rng rng(seed);
    for (int64_t i = 0; i < n * 5; ++i) draws[i] = rng.uniform();
    auto work = [&](int64_t a, int64_t b) {
        std::vector<double> X(ns), Yf(ns), Ye(ns), Y(ns), q(nv);
        for (int64_t i = a; i < b; ++i) {
            const double* u = &draws[(size_t)i * 5];
            const double T = Tlo + (Thi - Tlo) * u[0];
            const double p = kOneAtm * (logp ? std::exp(std::log(plo) + (std::log(phi_p) - std::log(plo)) * u[1])
                                             : plo + (phi_p - plo) * u[1]);
            const double phi = flo + (fhi - flo) * u[2];
            const double f = progress ? u[3] : 0.0;
            std::fill(X.begin(), X.end(), 0.0);
            X[kF] = fuel_per_phi * phi;
            X[kO2] = o2_amount;
            X[kN2] = n2_amount;
            double sx = 0;
            for (double v : X) sx += v;
            for (double& v : X) v /= sx;
            mole_to_mass(m, X.data(), Yf.data());
            if (f > 0) {
                equilibrate_tp(m, T, p, X.data(), Ye.data());
                for (int k = 0; k < ns; ++k) Y[k] = (1 - f) * Yf[k] + f * Ye[k];
            } else {
                Y = Yf;
            }
            state_from_TPY(m, T, p, Y.data(), 0, 0, q.data());
            for (int c = 0; c < nv; ++c) U[(size_t)c * n + i] = q[c];
            if (dt) dt[i] = std::exp(std::log(1e-8) + (std::log(1e-6) - std::log(1e-8)) * u[4]);
        }
    };
    const int nt = (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    if (n < 256 || nt == 1) {
        work(0, n);
        return;
    }
    std::vector<std::thread> th;
    std::vector<std::exception_ptr> errs(nt);
    for (int t = 0; t < nt; ++t) {
        const int64_t a = n * t / nt, b = n * (t + 1) / nt;
        th.emplace_back([&, t, a, b] {
            try { work(a, b); } catch (...) { errs[t] = std::current_exception(); }
        });
    }
    for (auto& x : th) x.join();
    for (auto& e : errs)
        if (e) std::rethrow_exception(e);
}

}  // namespace hf
