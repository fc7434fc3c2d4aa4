// TEST INFRASTRUCTURE ONLY: the oracle and the product's host-side mechanism
// loader and set-up thermochemistry (hyperflow_amd/csrc/mechanism.cpp,
// host_thermo.cpp) built together under AddressSanitizer and
// UndefinedBehaviorSanitizer (oracle/Makefile, target `sanitize`). Loads each
// mechanism file given on the command line with the product loader, turns
// it into the oracle's tables, and runs the oracle's properties, HLLC, flux
// stage, chemistry batch and Robertson BDF paths on small inputs, plus the
// loader's error path. Any sanitizer report aborts the run; the exit status
// is 0 when every check ran.
#include <cmath>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <vector>

#include "host_thermo.hpp"
#include "mechanism.hpp"

extern "C" {
void* or_mech_create(int ns, int nr, const double* W, const double* Tmid, const double* low, const double* high,
                     const int* kind, const int* rev, const double* A, const double* b, const double* EaR,
                     const double* A0, const double* b0, const double* Ea0R, const int* troe, const double* tA,
                     const double* tT3, const double* tT1, const double* tT2, const int* nu_r, const int* nu_p,
                     const double* eff);
void or_mech_destroy(void* m);
void or_properties(void* mp, const double* U, long n, double* out, int* status);
void or_hllc(void* mp, const double* l, const double* r, int n, const double* pl, const double* pr, double* f);
int or_flux_stage(void* mp, int nx, int ny, const double* Uin, double* Uout, double dtdx, double dtdy,
                  int first_order, int walls, int wedge_i, int dims, long* fallbacks, long* first_failed,
                  double* fx_out, double* fy_out);
int or_chem_advance(void* mp, double* U, long ncell, long ld, const double* dt, int dt_per_cell, double rtol,
                    double atol, int maxsteps, int threads, long* stats, long* first_failed);
int or_bdf_robertson(double* y, double tout, double rtol, double atol, int maxsteps, long* stats);
const char* or_last_error();
}

namespace {

int failures = 0;

void check(bool ok, const std::string& what) {
    std::printf("%s %s\n", ok ? "ok" : "FAILED", what.c_str());
    if (!ok) ++failures;
}

// The oracle's table layout from the product loader's Mechanism.
void* oracle_tables(const hf::Mechanism& m) {
    const int ns = m.ns(), nr = m.nr();
    std::vector<double> low(7 * ns), high(7 * ns);
    for (int k = 0; k < ns; ++k)
        for (int i = 0; i < 7; ++i) {
            low[7 * k + i] = m.low[k][i];
            high[7 * k + i] = m.high[k][i];
        }
    std::vector<int> kind(nr), rev(nr), troe(nr), nu_r(nr * ns, 0), nu_p(nr * ns, 0);
    std::vector<double> A(nr), b(nr), EaR(nr), A0(nr), b0(nr), Ea0R(nr), tA(nr), tT3(nr), tT1(nr), tT2(nr),
        eff(nr * ns, 1.0);
    for (int i = 0; i < nr; ++i) {
        const hf::Reaction& r = m.reactions[i];
        kind[i] = r.kind;
        rev[i] = r.reversible ? 1 : 0;
        troe[i] = r.troe;
        A[i] = r.A; b[i] = r.b; EaR[i] = r.EaR;
        A0[i] = r.A0; b0[i] = r.b0; Ea0R[i] = r.Ea0R;
        tA[i] = r.troeA; tT3[i] = r.troeT3; tT1[i] = r.troeT1; tT2[i] = r.troeT2;
        for (int k : r.reactants) ++nu_r[i * ns + k];
        for (int k : r.products) ++nu_p[i * ns + k];
        if (r.kind != hf::Reaction::ELEMENTARY) {
            for (int k = 0; k < ns; ++k) eff[i * ns + k] = r.default_efficiency;
            for (const auto& e : r.efficiencies) eff[i * ns + e.first] = e.second;
        }
    }
    return or_mech_create(ns, nr, m.W.data(), m.Tmid.data(), low.data(), high.data(), kind.data(), rev.data(),
                          A.data(), b.data(), EaR.data(), A0.data(), b0.data(), Ea0R.data(), troe.data(), tA.data(),
                          tT3.data(), tT1.data(), tT2.data(), nu_r.data(), nu_p.data(), eff.data());
}

void run_mechanism(const std::string& path) {
    const hf::Mechanism m = hf::load_mechanism(path);
    const int ns = m.ns(), nv = ns + 3;
    check(ns > 1 && m.nr() > 0 && (int)m.W.size() == ns, "load " + path);
    void* mp = oracle_tables(m);
    check(mp != nullptr, "oracle tables from the product loader");
    if (!mp) return;

    // set-up thermochemistry: synthetic cells, a state from (T, p, Y), TP equilibrium
    const long n = 6;
    std::vector<double> U(nv * n), dt(n);
    hf::synth_cells(m, 0, 5, n, U.data(), dt.data());
    std::vector<double> X(ns, 0.0), Y(ns), q(nv);
    X[0] = 2.0;
    X[ns - 1] = 7.0;
    hf::mole_to_mass(m, X.data(), Y.data());
    hf::state_from_TPY(m, 1200.0, 2e5, Y.data(), 10.0, -5.0, q.data());
    std::vector<double> Yeq(ns);
    hf::equilibrate_tp(m, 2000.0, 1e5, X.data(), Yeq.data());
    double s = 0;
    for (double y : Yeq) s += y;
    check(std::fabs(s - 1.0) < 1e-10, "equilibrium mass fractions sum to one");

    // properties on AoS states, HLLC between two of them
    std::vector<double> aos(nv * n), pat(3 * n);
    std::vector<int> st(n);
    for (long c = 0; c < n; ++c)
        for (int v = 0; v < nv; ++v) aos[c * nv + v] = U[v * n + c];
    or_properties(mp, aos.data(), n, pat.data(), st.data());
    bool all = true;
    for (long c = 0; c < n; ++c) all = all && st[c] == 0 && pat[3 * c] > 0;
    check(all, "properties of synthetic cells");
    std::vector<double> f(nv);
    or_hllc(mp, aos.data(), aos.data() + nv, ns, pat.data(), pat.data() + 3, f.data());
    check(std::isfinite(f[0]), "hllc");

    // one flux stage on a small ghosted box from the synthetic states, with face fluxes
    const int nx = 5, ny = 4, gx = nx + 4, gy = ny + 4;
    std::vector<double> Uin(nv * gx * gy), Uout(nv * nx * ny), fx(nv * (nx + 1) * ny), fy(nv * nx * (ny + 1));
    for (int j = 0; j < gy; ++j)
        for (int i = 0; i < gx; ++i)
            for (int v = 0; v < nv; ++v) Uin[v * gx * gy + j * gx + i] = U[v * n + (i + j) % n];
    long fb = 0, ff = -1;
    const int rf = or_flux_stage(mp, nx, ny, Uin.data(), Uout.data(), 1e-4, 1e-4, 0, 15, 0, 2, &fb, &ff, fx.data(),
                                 fy.data());
    check(rf == 0, "flux stage");

    // a chemistry batch (one worker) and its all-or-nothing failure path
    long stats[7] = {0}, first = -1;
    const double h = 1e-7;
    std::vector<double> U1 = U;
    const int rc = or_chem_advance(mp, U1.data(), n, n, &h, 0, 1e-8, 1e-15, 5000, 1, stats, &first);
    check(rc == 0 && stats[0] > 0, "chemistry batch");
    std::vector<double> bad(nv * 2, 0.0);
    const int rb = or_chem_advance(mp, bad.data(), 2, 2, &h, 0, 1e-8, 1e-15, 5000, 1, stats, &first);
    check(rb != 0 && first == 0 && or_last_error()[0] != 0, "chemistry batch rejects an empty cell");
    or_mech_destroy(mp);
}

}  // namespace

int main(int argc, char** argv) {
    for (int a = 1; a < argc; ++a) run_mechanism(argv[a]);
    bool threw = false;
    try {
        hf::load_mechanism("/nonexistent/mechanism.yaml");
    } catch (const std::runtime_error&) {
        threw = true;
    }
    check(threw, "the loader reports a missing file");
    double y[3] = {1.0, 0.0, 0.0};
    long st[7] = {0};
    check(or_bdf_robertson(y, 40.0, 1e-6, 1e-10, 100000, st) == 0 && std::fabs(y[0] + y[1] + y[2] - 1.0) < 1e-6,
          "Robertson BDF");
    return failures ? 1 : 0;
}
