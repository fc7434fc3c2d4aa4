// Device-side thermodynamics and kinetics for gfx950, specialised on a
// mechanism topology T (generated by mechgen: species order, reactant/product
// lists, reaction kinds, efficiency species). Every species index is a
// compile-time constant, so per-cell arrays (concentrations, production rates,
// Jacobian rows) live in registers. Numeric coefficients come from MechData<T>
// in device memory and are read with uniform (scalar) loads.
//
// Physics restated from the reference's call sites:
//   properties   HyperFlow_V3/src/physics.cpp:4-15 (Y = rhoY/rho, not renormalised;
//                T from e by Newton; p = rho R T / Wbar; a = sqrt(cp/cv R T / Wbar))
//   minmod       HyperFlow_V3/src/physics.cpp:40-44
//   hllc         HyperFlow_V3/src/physics.cpp:17-38 (Davis wave speeds)
//   kinetics     Cantera GasKinetics net production rates, as called at
//                HyperFlow_V3/src/chemistry.cpp:45
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <type_traits>
#include <utility>

// Lambdas that carry per-reaction or per-face code must be inlined: a lambda
// left as a call would take the per-lane arrays it captures by reference into
// scratch memory.
#define HF_AI __attribute__((always_inline))

namespace hf {

constexpr double kR = 8314.462618;
constexpr double kP0 = 101325.0;
constexpr double kSmall = 1e-300;
constexpr double kLn10 = 2.302585092994045684;
// ln(p0 / R): ln(p0 / (R T)) = kLnP0R - ln T saves a logarithm per rate evaluation
constexpr double kLnP0R = 2.500336689979609446;

template <class T>
struct MechData {
    static constexpr int NS = T::NS, NR = T::NR, NEFF = T::NEFF;
    double W[NS], invW[NS], Tmid[NS];
    // NASA7 low/high coefficient sets a1..a7 (cp/R = a1 + a2 T + ... + a5 T^4)
    double lo[NS][7], hi[NS][7];
    // reaction coefficients, SI kmol units; Ea/R in K
    double lnA[NR], b[NR], EaR[NR];
    double lnA0[NR], b0[NR], Ea0R[NR];
    double troeA[NR], troeRT3[NR], troeRT1[NR], troeT2[NR];
    double defeff[NR];          // default third-body efficiency
    double effd[NR][NEFF];      // listed efficiency minus the default
    int common_tmid;            // every species switches range at the same T
    double tmid;
    // Runtime copy of the topology, for kernels that assign reactions and
    // species to lanes (indexed by a lane-dependent reaction or species):
    int8_t rk_kind[NR], rk_troe[NR], rk_rev[NR];
    // rate_slot[s]: the reaction evaluated in slot s of the one-cell-per-wave
    // rate loop (lane s % 64, round s / 64). Falloff reactions come first,
    // then three-body ones, so that the rounds after the first hold only
    // elementary reactions and skip those code paths on every lane.
    int16_t rate_slot[NR];
    int16_t rk_reac[NR][T::NRMAX], rk_prod[NR][T::NRMAX], rk_eff[NR][NEFF > 0 ? NEFF : 1];
    // species k appears in reactions inc_rxn[inc_start[k] .. inc_start[k+1]) with net
    // stoichiometric coefficient inc_nu (products minus reactants)
    int16_t inc_start[NS + 1];
    int16_t inc_rxn[2 * NR * T::NRMAX];
    int8_t inc_nu[2 * NR * T::NRMAX];
    // The analytic Jacobian of the one-cell-per-wave kernel by concentration
    // columns: column s takes the entries jc_start[s] .. jc_start[s+1), each
    // adding jc_coef x (derivative jc_m of reaction jc_rxn) to row jc_k of
    // d wdot / d C. Derivative m < 2 NRMAX is d q / d C through reaction slot
    // m (reactant slots, then product slots); m = 2 NRMAX is d q / d[M]
    // (for the listed third-body efficiencies; the default efficiency, common
    // to every column, is added per row). Rows are species < NS - 1.
    static constexpr int NJC = NR * 2 * T::NRMAX * (2 * T::NRMAX + (NEFF > 0 ? NEFF : 1));
    int32_t jc_start[NS + 1];
    int16_t jc_rxn[NJC];
    int8_t jc_m[NJC], jc_k[NJC];
    double jc_coef[NJC];
};

// ----------------------------------------------------------------- utilities
template <int B, int E, class F>
__device__ __forceinline__ void static_for(F&& f) {
    if constexpr (B < E) {
        f(std::integral_constant<int, B>{});
        static_for<B + 1, E>(f);
    }
}

__device__ __forceinline__ bool finite(double x) { return __builtin_isfinite(x); }

// ----------------------------------------------------------------- logarithms
// Double-precision ln x for the rate expressions (ln T, and log10 in the
// Troe falloff), written for the FP64 pipe: no table lookups, two thirds of
// the library routine's instructions, within 2 units in the last place.
// ln 2 split for exact e ln 2 products (Cody and Waite): the high part has
// 29 significant bits, so e * kLn2Hi is exact for |e| < 2^24.
constexpr double kLn2Hi = 0x1.62e42ffp-1;
constexpr double kLn2Lo = -4.2009150726810846e-11;   // ln 2 - kLn2Hi
constexpr double kInvLn10 = 0.4342944819032518;

// ln x: x = f 2^e with f in [sqrt(1/2), sqrt(2)); ln f = 2 atanh(s),
// s = (f - 1) / (f + 1), |s| <= 0.1716, by its series 2 s sum s^(2j) / (2j + 1)
// to j = 10 (remainder below 1e-18 relative); e ln 2 added in two parts.
// x = 0 gives -inf, +inf gives +inf, x < 0 or NaN gives NaN.
__device__ __forceinline__ double log_fast(double x) {
    double f = __builtin_amdgcn_frexp_mant(x);
    int e = __builtin_amdgcn_frexp_exp(x);
    const bool low = f < 0.70710678118654752440;
    f = low ? f + f : f;
    e = low ? e - 1 : e;
    const double s = (f - 1.0) / (f + 1.0);
    const double z = s * s;
    double p = 2.0 / 21.0;
    p = fma(p, z, 2.0 / 19.0);
    p = fma(p, z, 2.0 / 17.0);
    p = fma(p, z, 2.0 / 15.0);
    p = fma(p, z, 2.0 / 13.0);
    p = fma(p, z, 2.0 / 11.0);
    p = fma(p, z, 2.0 / 9.0);
    p = fma(p, z, 2.0 / 7.0);
    p = fma(p, z, 2.0 / 5.0);
    p = fma(p, z, 2.0 / 3.0);
    const double de = (double)e;
    const double r = fma(de, kLn2Hi, fma(de, kLn2Lo, fma(s * z, p, s + s)));
    if (x > 0 && x < __builtin_inf()) return r;
    return x == 0 ? -__builtin_inf() : (x == __builtin_inf() ? x : __builtin_nan(""));
}

__device__ __forceinline__ double log10_fast(double x) { return log_fast(x) * kInvLn10; }

// x^n for n = 0..7 (by the bits of n).
__device__ __forceinline__ double ipow(double x, int n) {
    const double x2 = x * x, x4 = x2 * x2;
    return ((n & 1) ? x : 1.0) * ((n & 2) ? x2 : 1.0) * ((n & 4) ? x4 : 1.0);
}

// x^(1/n) for n = 1..7, the root in CVODE's step-size ratios
// eta = 1 / ((c dsm)^(1/n) + 1e-6). x = m 2^(n k) with k near log2(x) / n
// (exact scalings), so m lies within a few powers of two of 1; a
// single-precision estimate of m^(1/n) from the hardware log2 and exp2
// (relative error below 1e-5) is refined by two double-precision Newton
// steps on r^n = m, r <- r (1 + (m / r^n - 1) / n), whose fixed point does
// not depend on the rounding of 1/n. The result is within 2 units in the
// last place of the exact root (pow(x, 1.0 / n) is up to about 100 from it
// at the ends of the double range, where the rounding of 1.0 / n shows), in
// a quarter of pow's instructions; every trip of a batched wave evaluates
// several of these. x = 0 gives 0, +inf gives +inf, a negative x or NaN
// gives NaN, as pow does.
__device__ __forceinline__ double root_n(double x, int n) {
    if (!(x > 0) || x == __builtin_inf()) return x == 0 || x == __builtin_inf() ? x : __builtin_nan("");
    const float rn = __builtin_amdgcn_rcpf((float)n);
    const int k = (int)((float)__builtin_amdgcn_frexp_exp(x) * rn);
    const double m = __builtin_ldexp(x, -n * k);
    double r = (double)__builtin_amdgcn_exp2f(__builtin_amdgcn_logf((float)m) * rn);
    const double c = (double)rn;
    #pragma unroll
    for (int it = 0; it < 2; ++it) r = fma(r * c, m / ipow(r, n) - 1.0, r);
    return __builtin_ldexp(r, k);
}

// NASA7 coefficients of species k for temperature t (low range for t <= Tmid).
// Both sets are read with uniform loads and the per-lane choice is a value
// select, so the coefficients never go through a per-lane pointer.
struct Nasa7 {
    double a[7];
};
template <class D>
__device__ __forceinline__ Nasa7 nasa_select(const D& d, int k, double t) {
    Nasa7 r;
    const bool low = t <= d.Tmid[k];
    #pragma unroll
    for (int i = 0; i < 7; ++i) r.a[i] = low ? d.lo[k][i] : d.hi[k][i];
    return r;
}

// ----------------------------------------------------------------- thermo
// Mixture energy polynomial for fixed mass fractions: with yw_k = Y_k / W_k,
// e/R = sum_k yw_k (h_k/RT - 1) T and cv/R = sum_k yw_k (cp_k/R - 1).
// When every species shares Tmid the sums collapse to one polynomial per range.
template <class T>
struct EnergyPoly {
    double lo[6], hi[6];   // e/R = c0 T + c1 T^2/2 + c2 T^3/3 + c3 T^4/4 + c4 T^5/5 + c5
    double tmid;
    double rw;             // sum_k Y_k / W_k (1/Wbar)

    __device__ __forceinline__ void build(const MechData<T>& d, const double* Y) {
        #pragma unroll
        for (int i = 0; i < 6; ++i) { lo[i] = 0; hi[i] = 0; }
        double s = 0;
        #pragma unroll
        for (int k = 0; k < T::NS; ++k) {
            const double yw = Y[k] * d.invW[k];
            s += yw;
            #pragma unroll
            for (int i = 0; i < 5; ++i) {
                lo[i] = fma(yw, d.lo[k][i], lo[i]);
                hi[i] = fma(yw, d.hi[k][i], hi[i]);
            }
            lo[5] = fma(yw, d.lo[k][5], lo[5]);
            hi[5] = fma(yw, d.hi[k][5], hi[5]);
        }
        lo[0] -= s;
        hi[0] -= s;
        tmid = d.tmid;
        rw = s;
    }
    // e (J/kg) and cv (J/kg/K) at temperature t
    __device__ __forceinline__ void eval(double t, double& e, double& cv) const {
        // select values, not arrays: a pointer select would force lo/hi to scratch
        const bool low = t <= tmid;
        double c[6];
        #pragma unroll
        for (int i = 0; i < 6; ++i) c[i] = low ? lo[i] : hi[i];
        cv = kR * fma(t, fma(t, fma(t, fma(t, c[4], c[3]), c[2]), c[1]), c[0]);
        const double p = fma(t, fma(t, fma(t, fma(t, c[4] * 0.2, c[3] * 0.25), c[2] * (1.0 / 3.0)), c[1] * 0.5), c[0]);
        e = kR * fma(t, p, c[5]);
    }
};

// Per-species evaluation (used when species switch range at different T).
template <class T>
__device__ __forceinline__ void energy_species(const MechData<T>& d, const double* Y, double t, double& e,
                                               double& cv) {
    e = 0;
    cv = 0;
    #pragma unroll
    for (int k = 0; k < T::NS; ++k) {
        const Nasa7 na = nasa_select(d, k, t);
        const double* a = na.a;
        const double yw = Y[k] * d.invW[k];
        const double cpr = fma(t, fma(t, fma(t, fma(t, a[4], a[3]), a[2]), a[1]), a[0]);
        const double hrt = fma(t, fma(t, fma(t, fma(t, a[4] * 0.2, a[3] * 0.25), a[2] * (1.0 / 3.0)), a[1] * 0.5), a[0]) + a[5] / t;
        cv = fma(yw, cpr - 1.0, cv);
        e = fma(yw, (hrt - 1.0) * t, e);
    }
    e *= kR;
    cv *= kR;
}

// Energy -> temperature Newton on the mixture polynomial (every species
// shares Tmid). Converged when the applied step is below 1e-9 T: Newton's
// quadratic convergence leaves the result at round-off level. Returns false
// when the iteration does not converge or leaves T <= 0.
template <class T>
__device__ __forceinline__ bool poly_newton(const MechData<T>& d, const double* Y, double e_target, double& t,
                                            EnergyPoly<T>& P, int max_iter) {
    P.build(d, Y);
    for (int it = 0; it < max_iter; ++it) {
        double e, cv;
        P.eval(t, e, cv);
        const double dT = (e - e_target) / cv;
        double tn = t - dT;
        tn = tn > 0 ? tn : 0.5 * t;
        const bool done = fabs(dT) <= 1e-9 * t;
        t = tn;
        if (done) return finite(t) && cv > 0;
    }
    return false;
}

// Which thermodynamic form a kernel carries: the mixture polynomial (every
// species shares Tmid), the per-species sums, or both behind a run-time test
// of the mechanism. A kernel specialised on one form holds none of the
// other's code, which keeps large kernels compact in the instruction cache.
enum ThermoForm { kThermoAny = 0, kThermoPoly = 1, kThermoSpecies = 2 };

// Energy -> temperature Newton (the per-species sums when the species switch
// range at different temperatures). Same convergence test as poly_newton.
template <class T, int TF = kThermoAny>
__device__ __forceinline__ bool temperature_from_e(const MechData<T>& d, const double* Y, double e_target,
                                                   double& t, int max_iter = 60) {
    if (!(t > 0) || !finite(t)) t = 1000.0;
    if (TF == kThermoPoly || (TF == kThermoAny && d.common_tmid)) {
        EnergyPoly<T> P;
        return poly_newton(d, Y, e_target, t, P, max_iter);
    }
    if constexpr (TF == kThermoPoly) return false;
    for (int it = 0; it < max_iter; ++it) {
        double e, cv;
        energy_species(d, Y, t, e, cv);
        const double dT = (e - e_target) / cv;
        double tn = t - dT;
        tn = tn > 0 ? tn : 0.5 * t;
        const bool done = fabs(dT) <= 1e-9 * t;
        t = tn;
        if (done) return finite(t) && cv > 0;
    }
    return false;
}

// Frozen cp, cv and 1/Wbar at (t, Y).
template <class T>
__device__ __forceinline__ void cp_cv_rw(const MechData<T>& d, const double* Y, double t, double& cp, double& cv,
                                          double& rw) {
    cp = 0;
    rw = 0;
    #pragma unroll
    for (int k = 0; k < T::NS; ++k) {
        const Nasa7 na = nasa_select(d, k, t);
        const double* a = na.a;
        const double yw = Y[k] * d.invW[k];
        rw += yw;
        cp = fma(yw, fma(t, fma(t, fma(t, fma(t, a[4], a[3]), a[2]), a[1]), a[0]), cp);
    }
    cv = kR * (cp - rw);
    cp *= kR;
}

struct Props {
    double p, a, T;
};

// properties(u) of physics.cpp:4-15 on nv = NS+3 conserved values.
// Returns false for an inadmissible state (the reference throws domain_error).
template <class T, int TF = kThermoAny>
__device__ __forceinline__ bool properties(const MechData<T>& d, const double* u, Props& out, double t_guess) {
    constexpr int NS = T::NS;
    double rho = 0;
    bool ok = true;
    #pragma unroll
    for (int k = 0; k < NS; ++k) {
        rho += u[k];
        ok = ok && finite(u[k]) && u[k] >= 0;
    }
    ok = ok && finite(u[NS]) && finite(u[NS + 1]) && finite(u[NS + 2]) && rho > 0;
    if (!ok) return false;
    double Y[NS];
    const double rr = 1.0 / rho;
    #pragma unroll
    for (int k = 0; k < NS; ++k) Y[k] = u[k] * rr;
    const double ux = u[NS] * rr, vy = u[NS + 1] * rr;
    const double e = u[NS + 2] * rr - 0.5 * (ux * ux + vy * vy);
    double t = t_guess;
    double cp, cv, rw;
    if (TF == kThermoPoly || (TF == kThermoAny && d.common_tmid)) {
        // cv at the converged T from the same polynomial; cp - cv = R / Wbar
        if (!(t > 0) || !finite(t)) t = 1000.0;
        EnergyPoly<T> P;
        if (!poly_newton(d, Y, e, t, P, 60)) return false;
        double ee;
        P.eval(t, ee, cv);
        rw = P.rw;
        cp = cv + kR * rw;
    } else if constexpr (TF != kThermoPoly) {
        if (!temperature_from_e<T, kThermoSpecies>(d, Y, e, t)) return false;
        cp_cv_rw(d, Y, t, cp, cv, rw);
    }
    out.T = t;
    out.p = rho * kR * t * rw;
    out.a = sqrt(cp / cv * kR * t * rw);
    return out.T > 0 && out.p > 0 && out.a > 0 && finite(out.a);
}

// Temperature of a conserved state (the inversion inside properties()), or
// `fallback` when the state is not admissible. Used to warm-start the
// inversions of nearby states, as Cantera's phase object keeps the last T.
template <class T, int TF = kThermoAny>
__device__ __forceinline__ double state_temperature(const MechData<T>& d, const double* u, double fallback) {
    constexpr int NS = T::NS;
    double rho = 0;
    bool ok = true;
    #pragma unroll
    for (int k = 0; k < NS; ++k) {
        rho += u[k];
        ok = ok && finite(u[k]) && u[k] >= 0;
    }
    ok = ok && finite(u[NS]) && finite(u[NS + 1]) && finite(u[NS + 2]) && rho > 0;
    if (!ok) return fallback;
    double Y[NS];
    const double rr = 1.0 / rho;
    #pragma unroll
    for (int k = 0; k < NS; ++k) Y[k] = u[k] * rr;
    const double ux = u[NS] * rr, vy = u[NS + 1] * rr;
    const double e = u[NS + 2] * rr - 0.5 * (ux * ux + vy * vy);
    double t = fallback;
    return temperature_from_e<T, TF>(d, Y, e, t) ? t : fallback;
}

__device__ __forceinline__ double minmod(double x, double y) {
    return x * y > 0 ? copysign(fmin(fabs(x), fabs(y)), x) : 0.0;
}

// hllc(l, r, n, pl, pr) of physics.cpp:17-38; N = normal momentum component.
template <int NS>
__device__ __forceinline__ void hllc(const double* l, const double* r, int N, const Props& pl, const Props& pr,
                                     double* f) {
    constexpr int NV = NS + 3, E = NS + 2;
    double rl = 0, rr = 0;
    #pragma unroll
    for (int k = 0; k < NS; ++k) { rl += l[k]; rr += r[k]; }
    const double ul = (N == NS ? l[NS] : l[NS + 1]) / rl;
    const double ur = (N == NS ? r[NS] : r[NS + 1]) / rr;
    const double sl = fmin(ul - pl.a, ur - pr.a), sr = fmax(ul + pl.a, ur + pr.a);
    if (sl >= 0 || sr <= 0) {
        const bool useL = sl >= 0;
        const double v = useL ? ul : ur, p = useL ? pl.p : pr.p;
        #pragma unroll
        for (int c = 0; c < NV; ++c) f[c] = v * (useL ? l[c] : r[c]);
        if (N == NS) f[NS] += p; else f[NS + 1] += p;
        f[E] += p * v;
        return;
    }
    const double dl = rl * (sl - ul), dr = rr * (sr - ur);
    const double sm = (pr.p - pl.p + dl * ul - dr * ur) / (dl - dr);
    const bool left = sm >= 0;
    const double v = left ? ul : ur, s = left ? sl : sr, p = left ? pl.p : pr.p, rq = left ? rl : rr;
    const double scale = (s - v) / (s - sm);
    double rs = 0;
    double q[NV];
    #pragma unroll
    for (int c = 0; c < NV; ++c) q[c] = left ? l[c] : r[c];
    #pragma unroll
    for (int k = 0; k < NS; ++k) rs += q[k] * scale;
    #pragma unroll
    for (int c = 0; c < NV; ++c) {
        double star = q[c] * scale;
        if (c == N) star = rs * sm;
        if (c == E) star += rs * (sm - v) * (sm + p / (rq * (s - v)));
        double fc = v * q[c];
        if (c == N) fc += p;
        if (c == E) fc += p * v;
        f[c] = fc + s * (star - q[c]);
    }
}

// ----------------------------------------------------------------- kinetics
// Per-species h/RT and g/RT = h/RT - s/R at temperature t. HIGH: the
// caller knows t is above every species' Tmid, so the high-range
// coefficients are used without a per-lane choice.
template <class T, bool HIGH = false>
__device__ __forceinline__ void species_h_g(const MechData<T>& d, double t, double logt, double* hrt, double* grt) {
    const double it = 1.0 / t;
    #pragma unroll
    for (int k = 0; k < T::NS; ++k) {
        Nasa7 na;
        if constexpr (HIGH) {
            #pragma unroll
            for (int i = 0; i < 7; ++i) na.a[i] = d.hi[k][i];
        } else {
            na = nasa_select(d, k, t);
        }
        const double* a = na.a;
        const double h = fma(t, fma(t, fma(t, fma(t, a[4] * 0.2, a[3] * 0.25), a[2] * (1.0 / 3.0)), a[1] * 0.5), a[0]) + a[5] * it;
        const double s = fma(a[0], logt, fma(t, fma(t, fma(t, fma(t, a[4] * 0.25, a[3] * (1.0 / 3.0)), a[2] * 0.5), a[1]), a[6]));
        hrt[k] = h;
        grt[k] = h - s;
    }
}

// The equilibrium-constant factors of every species, exp(g_k/RT) and its
// inverse, and the standard concentration c0 = p0 / (R T) and its inverse:
// 1/Kc of a reaction is the product of its products' factors, its
// reactants' inverse factors and c0^(-dn) (dn = products - reactants), the
// exponential of sum nu g/RT - dn ln c0 that it equals. One exponential and
// one division per species replace an exponential per reversible reaction.
// Above 200 K |g/RT| stays below 170 for H2_Li's species and below 325 for
// dme_sk39's, so a product-reactant pair contributes at most e^650 and
// every partial product is a finite double (the NASA fits hold from 200 or
// 300 K).
template <class T>
struct GibbsFactors {
    double eg[T::NS], ieg[T::NS], c0, ic0;
    __device__ __forceinline__ GibbsFactors(const double* grt, double t, double it) {
        #pragma unroll
        for (int k = 0; k < T::NS; ++k) {
            eg[k] = exp(grt[k]);
            ieg[k] = 1.0 / eg[k];
        }
        c0 = (kP0 / kR) * it;
        ic0 = (kR / kP0) * t;
    }
};

// Rate constants of reaction I: forward kf (including [M] for three-body and
// the falloff blend), and 1/Kc. Optional derivatives for the Jacobian:
// dlnkf_dT (at fixed [M]), dkf_dM, dln_rKc_dT.
template <class T, int I, bool DERIV>
struct RateI {
    double kf, rkc, dlnkf_dT, dkf_dM, dlnrkc_dT;

    __device__ __forceinline__ RateI(const MechData<T>& d, double t, double logt, double it, double logc0,
                                     const double* hrt, const double* grt, const GibbsFactors<T>& gf, double Mc,
                                     double lnA_shift) {
        constexpr int kind = T::kind[I];
        double kinf = exp(d.lnA[I] + lnA_shift + d.b[I] * logt - d.EaR[I] * it);
        dlnkf_dT = 0;
        dkf_dM = 0;
        if constexpr (DERIV) dlnkf_dT = (d.b[I] + d.EaR[I] * it) * it;
        if constexpr (kind == 0) {
            kf = kinf;
        } else if constexpr (kind == 1) {
            kf = kinf * Mc;
            if constexpr (DERIV) dkf_dM = kinf;
        } else {
            const double k0 = exp(d.lnA0[I] + lnA_shift + d.b0[I] * logt - d.Ea0R[I] * it);
            const double pr = Mc * k0 / (kinf + kSmall);
            double F = 1.0, g_lpr = 0.0, g_lc = 0.0, dlc_dT = 0.0;
            if constexpr (T::troe[I] > 0) {
                const double A = d.troeA[I];
                const double e3 = exp(-t * d.troeRT3[I]), e1 = exp(-t * d.troeRT1[I]);
                double fc = (1.0 - A) * e3 + A * e1;
                // a zero T3 or T1 is an infinite rate whose term vanishes (0 x inf would be NaN)
                double dfc = -(1.0 - A) * (e3 > 0 ? d.troeRT3[I] * e3 : 0.0) - A * (e1 > 0 ? d.troeRT1[I] * e1 : 0.0);
                if constexpr (T::troe[I] == 4) {
                    const double e2 = exp(-d.troeT2[I] * it);
                    fc += e2;
                    dfc += d.troeT2[I] * it * it * e2;
                }
                const double fcc = fmax(fc, kSmall);
                const double lc = log10_fast(fcc);
                const double lpr = log10_fast(fmax(pr, kSmall));
                const double c = -0.4 - 0.67 * lc, n = 0.75 - 1.27 * lc;
                const double u = lpr + c;
                const double den = n - 0.14 * u;
                const double f1 = u / den;
                const double q = 1.0 / (1.0 + f1 * f1);
                const double lgf = lc * q;
                F = exp(kLn10 * lgf);
                if constexpr (DERIV) {
                    const double df1_du = n / (den * den), df1_dn = -u / (den * den);
                    const double dlgf_df1 = -2.0 * lc * f1 * q * q;
                    g_lpr = pr > kSmall ? dlgf_df1 * df1_du : 0.0;
                    g_lc = q + dlgf_df1 * (df1_du * (-0.67) + df1_dn * (-1.27));
                    dlc_dT = fc > kSmall ? dfc / (fcc * kLn10) : 0.0;
                }
            }
            const double opr = 1.0 / (1.0 + pr);
            kf = kinf * pr * opr * F;
            if constexpr (DERIV) {
                const double dlnk0 = (d.b0[I] + d.Ea0R[I] * it) * it;
                const double dlnpr_dT = dlnk0 - dlnkf_dT;
                // d ln F / dT = ln10 (dlog10F/dlPr dlPr/dT + dlog10F/dLc dLc/dT)
                const double dlnF_dT = g_lpr * dlnpr_dT + kLn10 * g_lc * dlc_dT;
                dlnkf_dT = dlnkf_dT + opr * dlnpr_dT + dlnF_dT;
                dkf_dM = k0 * F * opr * (opr + g_lpr);
            }
        }
        // 1/Kc = exp(sum nu g/RT - dn ln(p0/RT)), as the product of the
        // species' factors (products and reactants alternate, so the partial
        // products stay near the result's magnitude) and c0^(-dn)
        if constexpr (T::rev[I]) {
            double dh = 0, r_kc = 1.0;
            int dn = 0;
            static_for<0, T::NRMAX>([&](auto j) HF_AI {
                constexpr int r = T::reac[I][j], p = T::prod[I][j];
                if constexpr (p >= 0) { r_kc *= gf.eg[p]; dn += 1; if constexpr (DERIV) dh += hrt[p]; }
                if constexpr (r >= 0) { r_kc *= gf.ieg[r]; dn -= 1; if constexpr (DERIV) dh -= hrt[r]; }
            });
            for (int m = 0; m < dn; ++m) r_kc *= gf.ic0;
            for (int m = 0; m < -dn; ++m) r_kc *= gf.c0;
            rkc = r_kc;
            if constexpr (DERIV) dlnrkc_dT = (dn - dh) * it;
        } else {
            rkc = 0;
            dlnrkc_dT = 0;
        }
    }
};

template <class T, int I>
__device__ __forceinline__ double third_body(const MechData<T>& d, const double* C, double Ctot) {
    double M = d.defeff[I] * Ctot;
    static_for<0, T::NEFF>([&](auto j) HF_AI {
        constexpr int k = T::effsp[I][j];
        if constexpr (k >= 0) M = fma(d.effd[I][j], C[k], M);
    });
    return M;
}

// Net production rates wdot (kmol/m^3/s) at (t, C). lnA_shift (optional, per
// reaction) multiplies both k and k0 of reaction i by exp(lnA_shift[i]).
template <class T, bool HIGH = false>
__device__ __forceinline__ void production_rates(const MechData<T>& d, double t, const double* C, double* wdot,
                                                 const double* lnA_shift = nullptr, double* qout = nullptr) {
    constexpr int NS = T::NS;
    const double logt = log_fast(t), it = 1.0 / t;
    const double logc0 = kLnP0R - logt;
    double hrt[NS], grt[NS];
    species_h_g<T, HIGH>(d, t, logt, hrt, grt);
    const GibbsFactors<T> gf(grt, t, it);
    double Ctot = 0;
    #pragma unroll
    for (int k = 0; k < NS; ++k) { Ctot += C[k]; wdot[k] = 0; }
    static_for<0, T::NR>([&](auto ic) HF_AI {
        constexpr int I = decltype(ic)::value;
        double Mc = 0;
        if constexpr (T::kind[I] != 0) Mc = third_body<T, I>(d, C, Ctot);
        const RateI<T, I, false> k(d, t, logt, it, logc0, hrt, grt, gf, Mc, lnA_shift ? lnA_shift[I] : 0.0);
        double fwd = k.kf, bwd = k.kf * k.rkc;
        static_for<0, T::NRMAX>([&](auto j) HF_AI {
            constexpr int r = T::reac[I][j], p = T::prod[I][j];
            if constexpr (r >= 0) fwd *= C[r];
            if constexpr (p >= 0) bwd *= C[p];
        });
        const double q = fwd - bwd;
        if (qout) qout[I] = q;
        static_for<0, T::NRMAX>([&](auto j) HF_AI {
            constexpr int r = T::reac[I][j], p = T::prod[I][j];
            if constexpr (r >= 0) wdot[r] -= q;
            if constexpr (p >= 0) wdot[p] += q;
        });
    });
}

// Analytic Jacobian of the V3 chemistry right-hand side
//   f_k(y) = wdot_k(T(y), C(y)) W_k / rho,   k < NS-1,
// with y = Y[0..NS-2], Y[NS-1] = 1 - sum(y), C_k = rho Y_k / W_k and T(y)
// defined by e(T, Y) = e at fixed e (HyperFlow_V3/src/chemistry.cpp:36-53):
//   df_k/dy_j = W_k [dwdot_k/dC_j / W_j - dwdot_k/dC_last / W_last]
//             + W_k/rho * dwdot_k/dT * dT/dy_j,  dT/dy_j = -(e_j - e_last)/cv.
// Writes J (row-major N x N, N = NS-1) and f. Rate constants, their
// derivatives and f are FP64; J is accumulated in MT (double, or float for
// the mixed-precision Newton matrix). Optional outputs for sensitivities:
// fT[k] = W_k/rho dwdot_k/dT at fixed concentrations, qout[i] = net rate of
// progress of reaction i.
template <class T, class MT>
__device__ __forceinline__ void chem_jacobian(const MechData<T>& d, double t, double rho, const double* Y,
                                              MT* J, double* f, const double* lnA_shift = nullptr,
                                              double* fT = nullptr, double* qout = nullptr) {
    constexpr int NS = T::NS, N = NS - 1, L = NS - 1;
    const double logt = log_fast(t), it = 1.0 / t;
    const double logc0 = kLnP0R - logt;
    double hrt[NS], grt[NS], C[NS], wd[NS], dwdT[NS], dwdCL[N];
    species_h_g(d, t, logt, hrt, grt);
    const GibbsFactors<T> gf(grt, t, it);
    double Ctot = 0;
    #pragma unroll
    for (int k = 0; k < NS; ++k) {
        C[k] = rho * Y[k] * d.invW[k];
        Ctot += C[k];
        wd[k] = 0;
        dwdT[k] = 0;
    }
    #pragma unroll
    for (int k = 0; k < N; ++k) {
        dwdCL[k] = 0;
        #pragma unroll
        for (int j = 0; j < N; ++j) J[k * N + j] = MT(0);
    }
    // J accumulates dwdot_k/dC_j (k, j < N) in this pass.
    static_for<0, T::NR>([&](auto ic) HF_AI {
        constexpr int I = decltype(ic)::value;
        double Mc = 0;
        if constexpr (T::kind[I] != 0) Mc = third_body<T, I>(d, C, Ctot);
        const RateI<T, I, true> k(d, t, logt, it, logc0, hrt, grt, gf, Mc, lnA_shift ? lnA_shift[I] : 0.0);
        const double kr = k.kf * k.rkc;
        double pf = 1.0, pb = 1.0;
        static_for<0, T::NRMAX>([&](auto j) HF_AI {
            constexpr int r = T::reac[I][j], p = T::prod[I][j];
            if constexpr (r >= 0) pf *= C[r];
            if constexpr (p >= 0) pb *= C[p];
        });
        const double q = k.kf * pf - kr * pb;
        if (qout) qout[I] = q;
        const double dqdT = k.kf * k.dlnkf_dT * pf - kr * (k.dlnkf_dT + k.dlnrkc_dT) * pb;
        // d q / d C_s for the species in the reaction's concentration products
        // (the kf/kr factors that carry [M] are handled separately below).
        auto add = [&](auto sc, double dq) HF_AI {
            // species touched by reaction I: accumulate nu_k * dq into column s
            constexpr int s = decltype(sc)::value;
            static_for<0, T::NRMAX>([&](auto jj) HF_AI {
                constexpr int r = T::reac[I][jj], p = T::prod[I][jj];
                if constexpr (r >= 0 && r < N) {
                    if constexpr (s < N) J[r * N + s] -= (MT)dq; else dwdCL[r] -= dq;
                }
                if constexpr (p >= 0 && p < N) {
                    if constexpr (s < N) J[p * N + s] += (MT)dq; else dwdCL[p] += dq;
                }
            });
        };
        // reactant concentrations: d(prod C^nu')/dC_s = product with one C_s removed
        static_for<0, T::NRMAX>([&](auto j) HF_AI {
            constexpr int s = T::reac[I][j];
            if constexpr (s >= 0) {
                double pr_ = k.kf;
                static_for<0, T::NRMAX>([&](auto m) HF_AI {
                    constexpr int r2 = T::reac[I][m];
                    if constexpr (r2 >= 0 && m != j) pr_ *= C[r2];
                });
                add(std::integral_constant<int, s>{}, pr_);
            }
            constexpr int p = T::prod[I][j];
            if constexpr (p >= 0) {
                double pp_ = kr;
                static_for<0, T::NRMAX>([&](auto m) HF_AI {
                    constexpr int p2 = T::prod[I][m];
                    if constexpr (p2 >= 0 && m != j) pp_ *= C[p2];
                });
                add(std::integral_constant<int, p>{}, -pp_);
            }
        });
        // third-body concentration dependence: dq/dC_s += eff_s * dkf/d[M] * (pf - rkc pb)
        if constexpr (T::kind[I] != 0) {
            const double dqdM = k.dkf_dM * (pf - k.rkc * pb);
            // default efficiency on every species, plus the listed corrections
            static_for<0, NS>([&](auto s) HF_AI {
                double eff = d.defeff[I];
                static_for<0, T::NEFF>([&](auto j) HF_AI {
                    if constexpr (T::effsp[I][j] == decltype(s)::value) eff += d.effd[I][j];
                });
                add(s, eff * dqdM);
            });
        }
        static_for<0, T::NRMAX>([&](auto j) HF_AI {
            constexpr int r = T::reac[I][j], p = T::prod[I][j];
            if constexpr (r >= 0) { wd[r] -= q; dwdT[r] -= dqdT; }
            if constexpr (p >= 0) { wd[p] += q; dwdT[p] += dqdT; }
        });
    });
    // dT/dy_j = -(e_j - e_last)/cv with e_k = (h_k/RT - 1) R T / W_k
    double cv = 0, ek[NS];
    #pragma unroll
    for (int k = 0; k < NS; ++k) {
        const Nasa7 na = nasa_select(d, k, t);
        const double* a = na.a;
        const double cpr = fma(t, fma(t, fma(t, fma(t, a[4], a[3]), a[2]), a[1]), a[0]);
        cv = fma(Y[k] * d.invW[k], cpr - 1.0, cv);
        ek[k] = (hrt[k] - 1.0) * kR * t * d.invW[k];
    }
    cv *= kR;
    const double rcv = 1.0 / cv, rrho = 1.0 / rho;
    #pragma unroll
    for (int k = 0; k < N; ++k) {
        const double Wk = d.W[k];
        f[k] = wd[k] * Wk * rrho;
        const MT cl = (MT)(dwdCL[k] * d.invW[L]);
        const MT tk = (MT)(Wk * rrho * dwdT[k]);
        if (fT) fT[k] = Wk * rrho * dwdT[k];
        const MT wk = (MT)Wk;
        #pragma unroll
        for (int j = 0; j < N; ++j) {
            const MT dTdy = (MT)(-(ek[j] - ek[L]) * rcv);
            J[k * N + j] = wk * (J[k * N + j] * (MT)d.invW[j] - cl) + tk * dTdy;
        }
    }
}

}  // namespace hf
