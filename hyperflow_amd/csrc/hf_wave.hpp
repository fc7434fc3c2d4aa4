// Wave-per-cell chemistry for large mechanisms (DME: 39 species, 175
// reactions): one wave integrates one cell. Lane k < N owns unknown k (its
// value, Nordsieck history, error weight and correction live in that lane's
// registers; norms are wave reductions), lane N holds the closing species
// for the thermodynamics, and the reactions are spread over the lanes (lane
// l evaluates rate slots l, l + 64, ..., falloff and three-body reactions in
// the first round; production rates are then gathered per species in a fixed
// order). Every step-control decision is uniform
// across the wave, so the wave never diverges on control flow.
//
// The integrator is the same CVODE-equivalent BDF as hf_bdf.hpp (the
// reference's CVODE call, HyperFlow_V3/src/chemistry.cpp:10-26,64-68):
// Nordsieck form, orders 1-5, WRMS error test, Newton with CVODE's
// convergence test and Jacobian reuse. The Jacobian is analytic (the
// reference's CVODE forms it by difference quotients, cvLsDenseDQJac, at N
// rate evaluations each; see WaveGas::jacobian). It is kept in
// a per-wave global slot (row = lane, one entry per column); I - gamma J is
// formed and factored in LDS: in FP32 by a blocked LU whose trailing updates
// run on the matrix cores, in FP64 row by row against the broadcast pivot
// row.
#pragma once
#include "hf_device.hpp"

namespace hf {

// The value of x in the first active lane, in scalar registers. Every
// step-control decision of the integrator (Newton exits, error tests, the
// temperature iteration's stop) must see bitwise the same value on all 64
// lanes: if two lanes disagree in the last bit on a test that sits on its
// threshold, the wave splits, and the reductions of the lanes still running
// read stale values from the lanes that left (seen as a temperature solve
// jumping by 1000 K). A butterfly sum is not bitwise uniform by itself once
// the compiler fuses its first add with the product that feeds it, so each
// reduction ends with this broadcast.
__device__ __forceinline__ double uniform(double x) {
    const int lo = __builtin_amdgcn_readfirstlane(__double2loint(x));
    const int hi = __builtin_amdgcn_readfirstlane(__double2hiint(x));
    return __hiloint2double(hi, lo);
}
// Wave reductions by data-parallel primitives (DPP) on the vector pipe, not
// by LDS permutes: pairs and quads by quad permutations, then 8 and 16 lanes
// by half-row and row mirrors, then rows 0 + 1 into row 1 and rows 2 + 3
// into row 3 (row broadcast of lane 15), and rows 0-1 into rows 2-3 (row
// broadcast of lane 31): lane 63 holds the total, which every lane reads.
// A disabled row of a broadcast step keeps `old`. All 64 lanes take part
// (every call site is in wave-uniform control flow).
template <int CTRL, int ROWS = 0xF>
__device__ __forceinline__ double dpp(double x, double old) {
    const int lo = __builtin_amdgcn_update_dpp(__double2loint(old), __double2loint(x), CTRL, ROWS, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(__double2hiint(old), __double2hiint(x), CTRL, ROWS, 0xF, false);
    return __hiloint2double(hi, lo);
}
__device__ __forceinline__ double wave_sum(double x) {
    x += dpp<0xB1>(x, 0.0);           // quad_perm [1, 0, 3, 2]
    x += dpp<0x4E>(x, 0.0);           // quad_perm [2, 3, 0, 1]
    x += dpp<0x141>(x, 0.0);          // row_half_mirror
    x += dpp<0x140>(x, 0.0);          // row_mirror
    x += dpp<0x142, 0xA>(x, 0.0);     // row_bcast:15 into rows 1 and 3
    x += dpp<0x143, 0xC>(x, 0.0);     // row_bcast:31 into rows 2 and 3
    const int lo = __builtin_amdgcn_readlane(__double2loint(x), 63);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(x), 63);
    return __hiloint2double(hi, lo);
}
__device__ __forceinline__ double wave_max(double x) {
    x = fmax(x, dpp<0xB1>(x, x));
    x = fmax(x, dpp<0x4E>(x, x));
    x = fmax(x, dpp<0x141>(x, x));
    x = fmax(x, dpp<0x140>(x, x));
    x = fmax(x, dpp<0x142, 0xA>(x, x));
    x = fmax(x, dpp<0x143, 0xC>(x, x));
    const int lo = __builtin_amdgcn_readlane(__double2loint(x), 63);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(x), 63);
    return __hiloint2double(hi, lo);
}
// value of x in lane l (l wave-uniform)
__device__ __forceinline__ double lane_of(double x, int l) {
    const int lo = __builtin_amdgcn_readlane(__double2loint(x), l);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(x), l);
    return __hiloint2double(hi, lo);
}
__device__ __forceinline__ float lane_of(float x, int l) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), l));
}

// The cell's thermochemistry with one species per lane.
template <class T>
struct WaveGas {
    static constexpr int NS = T::NS, N = NS - 1, NR = T::NR;
    const MechData<T>& d;
    double* lds;                 // per wave: C[NS], grt[NS], q[NR], h/RT[NS]
    const int* inc;              // per wave (LDS): species incidences, reaction + nu * 65536
    int inc_b, inc_e;            // this lane's species: inc[inc_b .. inc_e)
    int lane;
    bool sp;                     // this lane holds a species (lane < NS)
    double W, invW, tmid, lo[7], hi[7];
    double rho, e, t;            // cell constants and the warm-started temperature

    // inc_lds: LDS room for the incidence lists (2 NR NRMAX ints), filled
    // here once per wave: the production-rate gather walks them on every
    // evaluation, and from global memory each step of that walk was a
    // dependent load.
    __device__ __forceinline__ WaveGas(const MechData<T>& dd, double* scratch, int* inc_lds)
        : d(dd), lds(scratch), inc(inc_lds) {
        lane = threadIdx.x & 63;
        sp = lane < NS;
        const int k = sp ? lane : 0;
        const int ntot = d.inc_start[NS];
        for (int p = lane; p < ntot; p += 64) inc_lds[p] = (int)d.inc_rxn[p] + (int)d.inc_nu[p] * 65536;
        inc_b = lane < N ? (int)d.inc_start[lane] : 0;
        inc_e = lane < N ? (int)d.inc_start[lane + 1] : 0;
        __syncthreads();
        W = d.W[k];
        invW = d.invW[k];
        tmid = d.Tmid[k];
        #pragma unroll
        for (int i = 0; i < 7; ++i) {
            lo[i] = d.lo[k][i];
            hi[i] = d.hi[k][i];
        }
        rho = 1;
        e = 0;
        t = 1000;
    }

    // Mass fraction of this lane's species from the unknowns (lane N closes the mass).
    __device__ __forceinline__ double mass_fraction(double y) const {
        const double s = wave_sum(lane < N ? y : 0.0);
        return lane < N ? y : (lane == N ? 1.0 - s : 0.0);
    }
    // T from e at fixed Y (Newton, warm start), as temperature_from_e.
    __device__ __forceinline__ bool temperature(double Yk, double e_target, double& tt) const {
        if (!(tt > 0) || !finite(tt)) tt = 1000.0;
        const double yw = sp ? Yk * invW : 0.0;
        for (int it = 0; it < 60; ++it) {
            const bool low = tt <= tmid;
            double a[6];
            #pragma unroll
            for (int i = 0; i < 6; ++i) a[i] = low ? lo[i] : hi[i];
            const double cpr = fma(tt, fma(tt, fma(tt, fma(tt, a[4], a[3]), a[2]), a[1]), a[0]);
            const double er = fma(tt, fma(tt, fma(tt, fma(tt, fma(tt, a[4] * 0.2, a[3] * 0.25), a[2] * (1.0 / 3.0)),
                                                   a[1] * 0.5), a[0]) - 1.0, a[5]);
            const double ev = kR * wave_sum(yw * er), cv = kR * wave_sum(yw * (cpr - 1.0));
            const double dT = (ev - e_target) / cv;
            double tn = tt - dT;
            tn = tn > 0 ? tn : 0.5 * tt;
            const bool done = fabs(dT) <= 1e-9 * tt;
            tt = tn;
            if (done) return finite(tt) && cv > 0;
        }
        return false;
    }

    // Net rate of progress of reaction i from C and g/RT in LDS (the runtime
    // form of RateI and production_rates in hf_device.hpp).
    __device__ __forceinline__ double rate(int i, double logt, double it, double logc0, double ctot) const {
        const double* C = lds;
        const double* grt = lds + NS;
        const int kind = d.rk_kind[i];
        const double kinf = exp(d.lnA[i] + d.b[i] * logt - d.EaR[i] * it);
        double Mc = 0;
        if (kind != 0) {
            Mc = d.defeff[i] * ctot;
            #pragma unroll
            for (int j = 0; j < (T::NEFF > 0 ? T::NEFF : 1); ++j) {
                const int s = d.rk_eff[i][j];
                if (s >= 0) Mc = fma(d.effd[i][j], C[s], Mc);
            }
        }
        double kf = kinf;
        if (kind == 1) {
            kf = kinf * Mc;
        } else if (kind == 2) {
            const double k0 = exp(d.lnA0[i] + d.b0[i] * logt - d.Ea0R[i] * it);
            const double pr = Mc * k0 / (kinf + kSmall);
            double F = 1.0;
            const int troe = d.rk_troe[i];
            if (troe > 0) {
                const double A = d.troeA[i];
                double fc = (1.0 - A) * exp(-t * d.troeRT3[i]) + A * exp(-t * d.troeRT1[i]);
                if (troe == 4) fc += exp(-d.troeT2[i] * it);
                const double lc = log10_fast(fmax(fc, kSmall));
                const double lpr = log10_fast(fmax(pr, kSmall));
                const double c = -0.4 - 0.67 * lc, n = 0.75 - 1.27 * lc;
                const double u = lpr + c;
                const double f1 = u / (n - 0.14 * u);
                F = exp(kLn10 * (lc / (1.0 + f1 * f1)));
            }
            kf = kinf * (pr / (1.0 + pr)) * F;
        }
        double pf = 1.0, pb = 1.0, dg = 0.0;
        int dn = 0;
        #pragma unroll
        for (int j = 0; j < T::NRMAX; ++j) {
            const int r = d.rk_reac[i][j], p = d.rk_prod[i][j];
            if (r >= 0) {
                pf *= C[r];
                dg -= grt[r];
                --dn;
            }
            if (p >= 0) {
                pb *= C[p];
                dg += grt[p];
                ++dn;
            }
        }
        const double kr = d.rk_rev[i] ? kf * exp(dg - dn * logc0) : 0.0;
        return kf * pf - kr * pb;
    }

    // f_k = wdot_k W_k / rho on lane k < N (chemistry.cpp:36-53); y on lanes
    // k < N. Returns the wave-uniform success flag.
    __device__ __forceinline__ bool rhs(double y, double& f) {
        const double Yk = mass_fraction(y);
        double tt = t;
        if (!temperature(Yk, e, tt)) return false;
        t = tt;
        const double logt = log_fast(t), it = 1.0 / t;
        const double logc0 = kLnP0R - logt;
        double C = 0;
        if (sp) {
            const bool low = t <= tmid;
            double a[7];
            #pragma unroll
            for (int i = 0; i < 7; ++i) a[i] = low ? lo[i] : hi[i];
            const double hrt = fma(t, fma(t, fma(t, fma(t, a[4] * 0.2, a[3] * 0.25), a[2] * (1.0 / 3.0)), a[1] * 0.5),
                                   a[0]) + a[5] * it;
            const double sr = fma(a[0], logt, fma(t, fma(t, fma(t, fma(t, a[4] * 0.25, a[3] * (1.0 / 3.0)), a[2] * 0.5),
                                                     a[1]), a[6]));
            C = rho * Yk * invW;
            lds[lane] = C;
            lds[NS + lane] = hrt - sr;
        }
        const double ctot = wave_sum(C);
        __syncthreads();
        for (int sl = lane; sl < NR; sl += 64) {
            const int i = d.rate_slot[sl];
            lds[2 * NS + i] = rate(i, logt, it, logc0, ctot);
        }
        __syncthreads();
        double w = 0;
        if (lane < N) {
            // four incidences per trip, their loads issued together; the sum
            // runs in list order
            const double* q = lds + 2 * NS;
            int p = inc_b;
            for (; p + 4 <= inc_e; p += 4) {
                const int v0 = inc[p], v1 = inc[p + 1], v2 = inc[p + 2], v3 = inc[p + 3];
                const double q0 = q[v0 & 0xffff], q1 = q[v1 & 0xffff], q2 = q[v2 & 0xffff], q3 = q[v3 & 0xffff];
                w = fma((double)(v0 >> 16), q0, w);
                w = fma((double)(v1 >> 16), q1, w);
                w = fma((double)(v2 >> 16), q2, w);
                w = fma((double)(v3 >> 16), q3, w);
            }
            for (; p < inc_e; ++p) {
                const int v = inc[p];
                w = fma((double)(v >> 16), q[v & 0xffff], w);
            }
            f = w * W / rho;
        } else {
            f = 0;
        }
        __syncthreads();   // the scratch is rewritten by the next evaluation
        return !__any(!finite(f));
    }

    // ---------------------------------------------------------------- analytic Jacobian
    // Derivatives of reaction i's net rate of progress (the runtime form of
    // RateI<., ., true> in hf_device.hpp) into deriv[m * NR + i]: m < NRMAX
    // d q / d C through reactant slot m (k_f times the product of the other
    // reactant concentrations), NRMAX <= m < 2 NRMAX through product slot
    // m - NRMAX (-k_r times the others), 2 NRMAX d q / d[M], 2 NRMAX + 1
    // d q / d T at fixed concentrations. C, g/RT and h/RT come from LDS.
    __device__ __forceinline__ void rate_derivs(int i, double logt, double it, double logc0, double ctot,
                                                double* __restrict__ deriv) const {
        constexpr int R = T::NRMAX;
        const double* C = lds;
        const double* grt = lds + NS;
        const double* hrt = lds + 2 * NS + NR;
        const int kind = d.rk_kind[i];
        const double kinf = exp(d.lnA[i] + d.b[i] * logt - d.EaR[i] * it);
        const double dlnkinf = (d.b[i] + d.EaR[i] * it) * it;
        double Mc = 0;
        if (kind != 0) {
            Mc = d.defeff[i] * ctot;
            #pragma unroll
            for (int j = 0; j < (T::NEFF > 0 ? T::NEFF : 1); ++j) {
                const int sp_ = d.rk_eff[i][j];
                if (sp_ >= 0) Mc = fma(d.effd[i][j], C[sp_], Mc);
            }
        }
        double kf = kinf, dlnkf = dlnkinf, dkf_dM = 0;
        if (kind == 1) {
            kf = kinf * Mc;
            dkf_dM = kinf;
        } else if (kind == 2) {
            const double k0 = exp(d.lnA0[i] + d.b0[i] * logt - d.Ea0R[i] * it);
            const double pr = Mc * k0 / (kinf + kSmall);
            double F = 1.0, g_lpr = 0.0, g_lc = 0.0, dlc_dT = 0.0;
            const int troe = d.rk_troe[i];
            if (troe > 0) {
                const double A = d.troeA[i];
                const double e3 = exp(-t * d.troeRT3[i]), e1 = exp(-t * d.troeRT1[i]);
                double fc = (1.0 - A) * e3 + A * e1;
                // a zero T3 or T1 is an infinite rate whose term vanishes (0 x inf would be NaN)
                double dfc = -(1.0 - A) * (e3 > 0 ? d.troeRT3[i] * e3 : 0.0) - A * (e1 > 0 ? d.troeRT1[i] * e1 : 0.0);
                if (troe == 4) {
                    const double e2 = exp(-d.troeT2[i] * it);
                    fc += e2;
                    dfc += d.troeT2[i] * it * it * e2;
                }
                const double fcc = fmax(fc, kSmall);
                const double lc = log10_fast(fcc);
                const double lpr = log10_fast(fmax(pr, kSmall));
                const double c = -0.4 - 0.67 * lc, n = 0.75 - 1.27 * lc;
                const double u = lpr + c;
                const double den = n - 0.14 * u;
                const double f1 = u / den;
                const double qq = 1.0 / (1.0 + f1 * f1);
                F = exp(kLn10 * (lc * qq));
                const double df1_du = n / (den * den), df1_dn = -u / (den * den);
                const double dlgf_df1 = -2.0 * lc * f1 * qq * qq;
                g_lpr = pr > kSmall ? dlgf_df1 * df1_du : 0.0;
                g_lc = qq + dlgf_df1 * (df1_du * (-0.67) + df1_dn * (-1.27));
                dlc_dT = fc > kSmall ? dfc / (fcc * kLn10) : 0.0;
            }
            const double opr = 1.0 / (1.0 + pr);
            kf = kinf * (pr / (1.0 + pr)) * F;
            const double dlnpr = (d.b0[i] + d.Ea0R[i] * it) * it - dlnkinf;
            dlnkf = dlnkinf + opr * dlnpr + g_lpr * dlnpr + kLn10 * g_lc * dlc_dT;
            dkf_dM = k0 * F * opr * (opr + g_lpr);
        }
        double cr[R], cp[R];
        double pf = 1.0, pb = 1.0, dg = 0.0, dh = 0.0;
        int dn = 0;
        #pragma unroll
        for (int j = 0; j < R; ++j) {
            const int r = d.rk_reac[i][j], p = d.rk_prod[i][j];
            cr[j] = r >= 0 ? C[r] : 1.0;
            cp[j] = p >= 0 ? C[p] : 1.0;
            pf *= cr[j];
            pb *= cp[j];
            if (r >= 0) {
                dg -= grt[r];
                dh -= hrt[r];
                --dn;
            }
            if (p >= 0) {
                dg += grt[p];
                dh += hrt[p];
                ++dn;
            }
        }
        const bool rev = d.rk_rev[i] != 0;
        const double rkc = rev ? exp(dg - dn * logc0) : 0.0;
        const double kr = kf * rkc;
        const double dlnrkc = rev ? (dn - dh) * it : 0.0;
        #pragma unroll
        for (int j = 0; j < R; ++j) {
            double a = kf, b = -kr;
            #pragma unroll
            for (int m = 0; m < R; ++m) {
                if (m != j) {
                    a *= cr[m];
                    b *= cp[m];
                }
            }
            deriv[j * NR + i] = d.rk_reac[i][j] >= 0 ? a : 0.0;
            deriv[(R + j) * NR + i] = d.rk_prod[i][j] >= 0 ? b : 0.0;
        }
        deriv[2 * R * NR + i] = dkf_dM * (pf - rkc * pb);
        deriv[(2 * R + 1) * NR + i] = kf * dlnkf * pf - kr * (dlnkf + dlnrkc) * pb;
    }

    // J = d f / d y at y (f_k = wdot_k W_k / rho, k < N; y the unknowns of
    // lanes < N, the last species closing the mass; T from e at fixed e):
    //   J_ks = W_k / W_s D_ks - W_k / W_L D_kL + W_k / rho dwdot_k/dT dT/dy_s,
    //   dT/dy_s = -(e_s - e_L) / cv,
    // with D = d wdot / d C. D is assembled by concentration columns, lane s
    // walking its column's list (MechData::jc_*) into LDS, from the
    // reactions' derivatives, which the reaction lanes write to this wave's
    // global scratch `deriv` ((2 NRMAX + 2) x NR doubles); the part of d[M]/d
    // C_s common to every column (the default efficiency) and dwdot/dT are
    // gathered per row over the species' incidences. Writes J into the
    // wave's slot jg (entry (row lane, column s) at jg[s * 64 + lane]); Dl:
    // LDS room for NS x 64 entries (the Newton matrix's, re-formed after this).
    // Returns the wave-uniform success flag.
    template <class MT>
    __device__ __forceinline__ bool jacobian(double y, MT* __restrict__ jg, MT* Dl, double* __restrict__ deriv) {
        constexpr int R = T::NRMAX;
        const double Yk = mass_fraction(y);
        double tt = t;
        if (!temperature(Yk, e, tt)) return false;
        t = tt;
        const double logt = log_fast(t), it = 1.0 / t;
        const double logc0 = kLnP0R - logt;
        double C = 0, ek = 0, cvk = 0;
        if (sp) {
            const bool low = t <= tmid;
            double a[7];
            #pragma unroll
            for (int i = 0; i < 7; ++i) a[i] = low ? lo[i] : hi[i];
            const double hrt = fma(t, fma(t, fma(t, fma(t, a[4] * 0.2, a[3] * 0.25), a[2] * (1.0 / 3.0)), a[1] * 0.5),
                                   a[0]) + a[5] * it;
            const double sr = fma(a[0], logt, fma(t, fma(t, fma(t, fma(t, a[4] * 0.25, a[3] * (1.0 / 3.0)), a[2] * 0.5),
                                                     a[1]), a[6]));
            const double cpr = fma(t, fma(t, fma(t, fma(t, a[4], a[3]), a[2]), a[1]), a[0]);
            C = rho * Yk * invW;
            lds[lane] = C;
            lds[NS + lane] = hrt - sr;
            lds[2 * NS + NR + lane] = hrt;
            ek = (hrt - 1.0) * kR * t * invW;
            cvk = Yk * invW * (cpr - 1.0);
        }
        const double ctot = wave_sum(C);
        const double cv = kR * wave_sum(cvk);
        #pragma unroll 1
        for (int s = 0; s < NS; ++s) Dl[s * 64 + lane] = MT(0);
        __syncthreads();
        for (int sl = lane; sl < NR; sl += 64) rate_derivs(d.rate_slot[sl], logt, it, logc0, ctot, deriv);
        __syncthreads();
        // column s of D on lane s
        if (sp) {
            const int b = d.jc_start[lane], en = d.jc_start[lane + 1];
            MT* col = Dl + lane * 64;
            for (int p = b; p < en; ++p) {
                const double v = d.jc_coef[p] * deriv[d.jc_m[p] * NR + d.jc_rxn[p]];
                col[d.jc_k[p]] += (MT)v;
            }
        }
        // row factors: dwdot_k/dT and the default-efficiency part of d wdot_k / d C
        double ak = 0, uk = 0;
        if (lane < N) {
            for (int p = inc_b; p < inc_e; ++p) {
                const int v = inc[p];
                const int r = v & 0xffff;
                const double nu = (double)(v >> 16);
                ak = fma(nu, deriv[(2 * R + 1) * NR + r], ak);
                if (d.rk_kind[r] != 0) uk = fma(nu * d.defeff[r], deriv[2 * R * NR + r], uk);
            }
        }
        __syncthreads();
        const double rc = 1.0 / cv, eL = lane_of(ek, N);
        const double dL = lane < N ? (double)Dl[N * 64 + lane] + uk : 0.0;
        const double wl = lane < N ? W * d.invW[N] : 0.0, tk = lane < N ? W / rho * ak : 0.0;
        bool ok = true;
        #pragma unroll 2
        for (int s = 0; s < N; ++s) {
            const double dTdy = -(lane_of(ek, s) - eL) * rc;
            double J = 0.0;
            if (lane < N) {
                J = W * d.invW[s] * ((double)Dl[s * 64 + lane] + uk) - wl * dL + tk * dTdy;
                ok = ok && finite(J);
            }
            jg[s * 64 + lane] = (MT)J;
        }
        return !__any(!ok);
    }

    // properties() on the cell state held one species per lane: rhoY_k on
    // lane k, momenta and energy (uniform). Returns admissibility; T warm-started from t.
    __device__ __forceinline__ bool properties(double rhoY, double mx, double my, double E, double& tt) const {
        bool ok = !sp || (finite(rhoY) && rhoY >= 0);
        ok = !__any(!ok) && finite(mx) && finite(my) && finite(E);
        const double r = wave_sum(sp ? rhoY : 0.0);
        if (!ok || !(r > 0)) return false;
        const double Yk = sp ? rhoY / r : 0.0;
        const double ux = mx / r, vy = my / r;
        const double ee = E / r - 0.5 * (ux * ux + vy * vy);
        if (!temperature(Yk, ee, tt)) return false;
        const bool low = tt <= tmid;
        double a[5];
        #pragma unroll
        for (int i = 0; i < 5; ++i) a[i] = low ? lo[i] : hi[i];
        const double yw = sp ? Yk * invW : 0.0;
        const double cp = wave_sum(yw * fma(tt, fma(tt, fma(tt, fma(tt, a[4], a[3]), a[2]), a[1]), a[0]));
        const double rw = wave_sum(yw);
        const double p = r * kR * tt * rw;
        const double cv = cp - rw;
        const double aa = sqrt(cp / cv * kR * tt * rw);
        return tt > 0 && p > 0 && aa > 0 && finite(aa);
    }
};

struct WaveCounters {
    int steps = 0, rhs = 0, jac = 0, lu = 0, newton = 0, ncf = 0, netf = 0;
};

// CVODE-equivalent BDF over one cell, the unknown of lane k in lane k.
// MT: precision of the Jacobian and the Newton matrix factors.
template <class T, class MT>
struct WaveBdf {
    static constexpr int NS = T::NS, N = NS - 1, QMAX = 5;
    static constexpr double UROUND = 2.220446049250313e-16;
    WaveGas<T> g;                // held by value: the whole integrator stays in registers
    MT* jg;                      // global, this wave's slot: Jacobian entry (lane, j) at jg[j * 64 + lane]
    MT* lu;                      // LDS: factored Newton matrix, entry (lane, j) at lu[j * 64 + lane] (NS x 64
                                 // entries: the Jacobian's assembly uses the extra column)
    double* jd;                  // global, this wave's reaction-derivative scratch ((2 NRMAX + 2) x NR)
    double* coef;                // LDS: tau[QMAX + 2], l[QMAX + 1]
    int lane;
    bool own;                    // lane < N
    double z[QMAX + 1];          // Nordsieck history of this lane's unknown
    double ewt, acor;
    double tq1, tq2, tq3, tq4, tq5;
    double tn, h, hprime, hscale, eta, etamax, gamma, gammap, gamrat, rl1, crate, acnrm, saved_tq5;
    double rtol, atol;
    int q, qprime, L, qwait, nst, nstlp, nstlj;
    bool force_setup, jfresh;
    WaveCounters c;

    __device__ __forceinline__ WaveBdf(const MechData<T>& d, double* scratch, int* inc_lds, MT* jslot, MT* lds_lu,
                                       double* deriv, double* cf, double rt, double at)
        : g(d, scratch, inc_lds), jg(jslot), lu(lds_lu), jd(deriv), coef(cf), rtol(rt), atol(at) {
        lane = threadIdx.x & 63;
        own = lane < N;
    }
    __device__ __forceinline__ double& TAU(int i) const { return coef[i]; }
    __device__ __forceinline__ double& LC(int i) const { return coef[QMAX + 2 + i]; }
    __device__ __forceinline__ double wrms(double v) const {
        const double x = own ? v * ewt : 0.0;
        return sqrt(wave_sum(x * x) * (1.0 / N));
    }
    __device__ __forceinline__ void set_ewt() { ewt = own ? 1.0 / fma(rtol, fabs(z[0]), atol) : 0.0; }
    // z[q] with a wave-uniform q, as a sum of values selected against zero
    // (exact): a select between array elements would be folded into one
    // load through a selected address, which keeps the history in scratch.
    __device__ __forceinline__ double zq(int qq) const {
        double v = 0;
        #pragma unroll
        for (int j = 0; j <= QMAX; ++j) v += j == qq ? z[j] : 0.0;
        return v;
    }

    __device__ __forceinline__ void rescale() {
        double factor = eta;
        #pragma unroll
        for (int j = 1; j <= QMAX; ++j) {
            if (j <= q) z[j] *= factor;
            factor *= eta;
        }
        h = hscale * eta;
        hprime = h;
        hscale = h;
    }
    __device__ __forceinline__ void predict() {
        tn += h;
        #pragma unroll
        for (int k = 1; k <= QMAX; ++k)
            #pragma unroll
            for (int j = QMAX; j >= k; --j)
                if (k <= q && j <= q) z[j - 1] += z[j];
    }
    __device__ __forceinline__ void restore(double saved_t) {
        tn = saved_t;
        #pragma unroll
        for (int k = 1; k <= QMAX; ++k)
            #pragma unroll
            for (int j = QMAX; j >= k; --j)
                if (k <= q && j <= q) z[j - 1] -= z[j];
    }
    __device__ __forceinline__ void decrease_order() {
        if (q <= 2) return;
        double lc[QMAX + 1];
        #pragma unroll
        for (int j = 0; j <= QMAX; ++j) lc[j] = 0;
        lc[2] = 1;
        double hsum = 0;
        #pragma unroll
        for (int j = 1; j <= QMAX - 2; ++j) {
            if (j <= q - 2) {
                hsum += TAU(j);
                const double xi = hsum / hscale;
                #pragma unroll
                for (int i = j + 2; i >= 2; --i) lc[i] = lc[i] * xi + lc[i - 1];
            }
        }
        const double top = zq(q);
        #pragma unroll
        for (int j = 2; j < QMAX; ++j)
            if (j < q) z[j] -= top * lc[j];
    }
    __device__ __forceinline__ void increase_order() {
        double lc[QMAX + 1];
        #pragma unroll
        for (int j = 0; j <= QMAX; ++j) lc[j] = 0;
        lc[2] = 1;
        double alpha1 = 1, prod = 1, xiold = 1, alpha0 = -1, hsum = hscale;
        #pragma unroll
        for (int j = 1; j < QMAX; ++j) {
            if (j < q) {
                hsum += TAU(j + 1);
                const double xi = hsum / hscale;
                prod *= xi;
                alpha0 -= 1.0 / (j + 1);
                alpha1 += 1.0 / xi;
                #pragma unroll
                for (int i = j + 2; i >= 2; --i) lc[i] = lc[i] * xiold + lc[i - 1];
                xiold = xi;
            }
        }
        const double A1 = (-alpha0 - alpha1) / prod;
        const double nz = A1 * z[QMAX];
        #pragma unroll
        for (int j = 2; j <= QMAX; ++j) {
            if (j <= q) z[j] += lc[j] * nz;
            if (j == q + 1) z[j] = nz;
        }
    }
    __device__ __forceinline__ void adjust_order(int dq) {
        if (q == 2 && dq != 1) return;
        if (dq == 1) increase_order();
        else if (dq == -1) decrease_order();
    }
    __device__ __forceinline__ void set_bdf() {
        // coefficients are uniform: lane 0 computes them into LDS
        if (lane == 0) {
            #pragma unroll
            for (int i = 0; i <= QMAX; ++i) LC(i) = 0;
            LC(0) = LC(1) = 1;
        }
        __syncthreads();
        double xi_inv = 1, xistar_inv = 1, alpha0 = -1, alpha0_hat = -1, hsum = h;
        if (q > 1) {
            for (int j = 2; j < q; ++j) {
                hsum += TAU(j - 1);
                xi_inv = h / hsum;
                alpha0 -= 1.0 / j;
                __syncthreads();
                if (lane == 0)
                    for (int i = j; i >= 1; --i) LC(i) += LC(i - 1) * xi_inv;
                __syncthreads();
            }
            alpha0 -= 1.0 / q;
            xistar_inv = -LC(1) - alpha0;
            hsum += TAU(q - 1);
            xi_inv = h / hsum;
            alpha0_hat = -LC(1) - xi_inv;
            __syncthreads();
            if (lane == 0)
                for (int i = q; i >= 1; --i) LC(i) += LC(i - 1) * xistar_inv;
            __syncthreads();
        }
        const double lq = LC(q);
        const double A1 = 1.0 - alpha0_hat + alpha0;
        const double A2 = 1.0 + q * A1;
        tq2 = fabs(A1 / (alpha0 * A2));
        tq5 = fabs(A2 * xistar_inv / (lq * xi_inv));
        if (qwait == 1) {
            if (q > 1) {
                const double C = xistar_inv / lq;
                const double A3 = alpha0 + 1.0 / q;
                const double A4 = alpha0_hat + xi_inv;
                tq1 = fabs(C * (1.0 - A4 + A3) / A3);
            } else {
                tq1 = 1.0;
            }
            hsum += TAU(q);
            xi_inv = h / hsum;
            const double A5 = alpha0 - 1.0 / (q + 1);
            const double A6 = alpha0_hat - xi_inv;
            const double Cppinv = (1.0 - A6 + A5) / A2;
            tq3 = fabs(Cppinv / (xi_inv * (q + 2) * A5));
        }
        tq4 = 0.1 / tq2;
    }

    // ---------------------------------------------------------------- linear algebra
    // lu = I - gamma J from the wave's global Jacobian slot; rows N..63 are
    // zero. (Loading 16 columns at a time measured slower than this rolled
    // loop: the wider batch spills.)
    __device__ __forceinline__ void form_newton() {
        const MT gm = (MT)gamma;
        #pragma unroll 2
        for (int j = 0; j < N; ++j)
            lu[j * 64 + lane] = lane < N ? (lane == j ? MT(1) : MT(0)) - gm * jg[j * 64 + lane] : MT(0);
    }
    __device__ __forceinline__ bool factor() {
        if constexpr (std::is_same<MT, float>::value) return factor_blocked();
        else return factor_rows();
    }
    // I - gamma J factored in LDS without pivoting (reciprocal pivots on the
    // diagonal), one row per lane. Step k reads pivot row k as broadcasts
    // (one address for the whole wave) and each lane below it updates its own
    // row; rolled loops keep the register footprint small, which the
    // unrolled register form (38 values per lane and ~700 lane broadcasts
    // into scalar registers) did not. Used for the FP64 Newton matrix.
    __device__ __forceinline__ bool factor_rows() {
        form_newton();
        __syncthreads();
        bool ok = true;
        #pragma unroll 1
        for (int k = 0; k < N; ++k) {
            const MT p = lu[k * 64 + k];
            ok = ok && fabs(p) > __FLT_MIN__ && __builtin_isfinite(p);
            const MT inv = MT(1) / p;
            const bool below = lane > k;
            MT l = 0;
            if (below) {
                l = lu[k * 64 + lane] * inv;
                lu[k * 64 + lane] = l;
            }
            #pragma unroll 4
            for (int j = k + 1; j < N; ++j) {
                const MT pj = lu[j * 64 + k];
                if (below) lu[j * 64 + lane] = fma(-l, pj, lu[j * 64 + lane]);
            }
            if (lane == k) lu[k * 64 + k] = inv;
            __syncthreads();
        }
        return ok;
    }
    // The FP32 Newton matrix: blocked right-looking LU without pivoting, in
    // panels of 16 columns. For the panel starting at column k0:
    //  (a) rows k0.. of its 16 columns are factored in registers, one row per
    //      lane, each pivot row broadcast with readlane (no barrier per pivot);
    //  (b) the row block right of it, U12 = L11^-1 A12, is solved one column
    //      per lane, L11 read from the panel registers;
    //  (c) the trailing matrix A22 -= L21 U12 is updated on the matrix cores,
    //      16 x 16 tiles by v_mfma_f32_16x16x4f32 (lane l holds A[l%16][l/16],
    //      B[l/16][l%16] and D[4(l/16)+v][l%16]; the four D rows of a lane are
    //      contiguous in the column-major LDS matrix).
    // Rows N..63 of every column are zero, so the rows of a tile past the
    // matrix contribute nothing; its columns past the matrix are masked per
    // lane and never stored. The factors are those of factor_rows up to FP32
    // summation order: unit L below the diagonal, U above, reciprocal pivots
    // on it. lu must be 16-byte aligned.
    __device__ __forceinline__ bool factor_blocked() {
        typedef float f4 __attribute__((ext_vector_type(4)));
        form_newton();
        __syncthreads();
        bool ok = true;
        static_for<0, (N + 15) / 16>([&](auto kbc) HF_AI {
            constexpr int k0 = 16 * decltype(kbc)::value, R0 = k0 + 16;
            // (a) the panel
            float p[16];
            #pragma unroll
            for (int c = 0; c < 16; ++c) p[c] = k0 + c < N ? (float)lu[(k0 + c) * 64 + lane] : 0.0f;
            #pragma unroll
            for (int c = 0; c < 16; ++c) {
                if (k0 + c < N) {
                    const int k = k0 + c;
                    const float piv = lane_of(p[c], k);
                    ok = ok && fabsf(piv) > __FLT_MIN__ && __builtin_isfinite(piv);
                    const float inv = 1.0f / piv;
                    const float l = p[c] * inv;
                    const bool below = lane > k;
                    #pragma unroll
                    for (int j = c + 1; j < 16; ++j) {
                        if (k0 + j < N) {
                            const float pkj = lane_of(p[j], k);
                            if (below) p[j] = fmaf(-l, pkj, p[j]);
                        }
                    }
                    p[c] = below ? l : (lane == k ? inv : p[c]);
                }
            }
            if (lane >= k0 && lane < N) {
                #pragma unroll
                for (int c = 0; c < 16; ++c)
                    if (k0 + c < N) lu[(k0 + c) * 64 + lane] = p[c];
            }
            if constexpr (R0 < N) {
                // (b) U12: lane j holds column R0 + j (lanes past the matrix
                // work on a copy of column R0 and store nothing)
                constexpr int NC = N - R0;
                const int col = lane < NC ? R0 + lane : R0;
                float x[16];
                #pragma unroll
                for (int r = 0; r < 16; ++r) x[r] = lu[col * 64 + k0 + r];
                #pragma unroll
                for (int c = 0; c < 16; ++c)
                    #pragma unroll
                    for (int r = c + 1; r < 16; ++r) x[r] = fmaf(-lane_of(p[c], k0 + r), x[c], x[r]);
                if (lane < NC) {
                    #pragma unroll
                    for (int r = 0; r < 16; ++r) lu[col * 64 + k0 + r] = x[r];
                }
                __syncthreads();
                // (c) trailing tiles; every MFMA runs on the whole wave
                constexpr int NT = (NC + 15) / 16;
                const int i = lane & 15, q = lane >> 4;
                #pragma unroll
                for (int rb = 0; rb < NT; ++rb) {
                    #pragma unroll
                    for (int cb = 0; cb < NT; ++cb) {
                        const int r0 = R0 + 16 * rb, cc = R0 + 16 * cb + i;
                        const bool cv = cc < N;
                        f4* dst = reinterpret_cast<f4*>(&lu[(cv ? cc : R0) * 64 + r0 + 4 * q]);
                        f4 acc = cv ? *dst : f4{0.0f, 0.0f, 0.0f, 0.0f};
                        #pragma unroll
                        for (int ks = 0; ks < 4; ++ks) {
                            const int kk = k0 + 4 * ks + q;
                            const float a = -(float)lu[kk * 64 + r0 + i];
                            const float b = cv ? (float)lu[cc * 64 + kk] : 0.0f;
                            acc = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, acc, 0, 0, 0);
                        }
                        if (cv) *dst = acc;
                    }
                }
                __syncthreads();
            }
        });
        return ok;
    }
    // Solve (I - gamma J) x = b with b, x one component per lane.
    // Each substitution step depends on the previous one through lane k's b,
    // so the lane's row entries are loaded from LDS 16 at a time ahead of the
    // steps that use them (loaded per step, every step waited out an LDS
    // round trip). The arithmetic is the plain row-by-row substitution.
    __device__ __forceinline__ double solve(double bd) const {
        MT b = (MT)bd;
        #pragma unroll
        for (int k0 = 0; k0 < N; k0 += 16) {
            MT r[16];
            #pragma unroll
            for (int c = 0; c < 16; ++c)
                if (k0 + c < N) r[c] = lu[(k0 + c) * 64 + lane];
            #pragma unroll
            for (int c = 0; c < 16; ++c) {
                if (k0 + c < N) {
                    const MT bk = lane_of(b, k0 + c);
                    if (lane > k0 + c) b = fma(-r[c], bk, b);
                }
            }
        }
        #pragma unroll
        for (int k1 = ((N - 1) / 16) * 16; k1 >= 0; k1 -= 16) {
            MT r[16];
            #pragma unroll
            for (int c = 0; c < 16; ++c)
                if (k1 + c < N) r[c] = lu[(k1 + c) * 64 + lane];
            #pragma unroll
            for (int c = 15; c >= 0; --c) {
                const int k = k1 + c;
                if (k < N) {
                    if (lane == k) b *= r[c];
                    const MT bk = lane_of(b, k);
                    if (lane < k) b = fma(-r[c], bk, b);
                }
            }
        }
        return own ? (double)b : 0.0;
    }
    __device__ __forceinline__ bool lsetup(double y, double f, bool jbad) {
        ++c.lu;
        if (jbad) {
            ++c.jac;
            if (!g.jacobian(y, jg, lu, jd)) return false;
            nstlj = nst;
            jfresh = true;
        } else {
            jfresh = false;
        }
        __syncthreads();
        gammap = gamma;
        gamrat = 1;
        crate = 1;
        nstlp = nst;
        return factor();
    }

    // ---------------------------------------------------------------- Newton and step
    // 0 converged, 1 retry with a fresh matrix, 2 reduce the step, 3 rhs/setup failure
    __device__ __forceinline__ int newton(int& convfail) {
        double y = own ? z[0] : 0.0, f;
        ++c.rhs;
        if (!g.rhs(y, f)) return 3;
        const bool callsetup = nst == 0 || nst >= nstlp + 20 || fabs(gamrat - 1.0) > 0.3 || convfail != 0 || force_setup;
        const bool jbad = callsetup && (nst == 0 || nst >= nstlj + 51 || convfail == 2 ||
                                        (convfail == 1 && fabs(gamma / gammap - 1.0) < 0.2));
        force_setup = false;
        jfresh = false;
        if (callsetup && !lsetup(y, f, jbad)) return 3;
        acor = 0;
        double delp = 0;
        for (int m = 0; m < 3; ++m) {
            ++c.newton;
            double delta = own ? -(fma(rl1, z[1], acor) - gamma * f) : 0.0;
            delta = solve(delta);
            if (gamrat != 1.0) delta *= 2.0 / (1.0 + gamrat);
            acor += delta;
            y = own ? z[0] + acor : 0.0;
            const double del = wrms(delta);
            if (m > 0) crate = fmax(0.3 * crate, del / delp);
            const double dcon = del * fmin(1.0, crate) / tq4;
            if (dcon <= 1.0) {
                acnrm = (m == 0) ? del : wrms(acor);
                return 0;
            }
            if (m == 2 || (m >= 1 && del > 2.0 * delp)) break;
            delp = del;
            ++c.rhs;
            if (!g.rhs(y, f)) return 3;
        }
        if (!jfresh) {
            convfail = 1;
            return 1;
        }
        return 2;
    }

    __device__ __forceinline__ bool step(double tout) {
        const double saved_t = tn;
        int ncf = 0, nef = 0, convfail = 0;
        double dsm = 0;
        for (;;) {
            predict();
            if ((tn - tout) * h > 0) tn = tout;
            set_bdf();
            rl1 = 1.0 / LC(1);
            gamma = h * rl1;
            if (nst == 0) gammap = gamma;
            gamrat = (nst > 0) ? gamma / gammap : 1.0;
            int nf = newton(convfail);
            while (nf == 1) nf = newton(convfail);
            if (nf != 0) {
                restore(saved_t);
                ++ncf;
                ++c.ncf;
                etamax = 1;
                if (ncf >= 10) return false;
                eta = 0.25;
                convfail = 2;
                rescale();
                continue;
            }
            convfail = 0;
            dsm = acnrm * tq2;
            if (dsm <= 1.0) break;
            ++nef;
            ++c.netf;
            restore(saved_t);
            force_setup = true;
            if (nef >= 7) return false;
            etamax = 1;
            if (nef <= 3) {
                eta = 1.0 / (root_n(6.0 * dsm, L) + 1e-6);
                eta = fmax(0.1, eta);
                if (nef >= 2) eta = fmin(eta, 0.2);
                rescale();
                continue;
            }
            if (q > 1) {
                eta = 0.1;
                adjust_order(-1);
                L = q;
                --q;
                qwait = L;
                rescale();
                continue;
            }
            eta = 0.1;
            h *= eta;
            hprime = h;
            hscale = h;
            qwait = 10;
            double f;
            ++c.rhs;
            if (!g.rhs(own ? z[0] : 0.0, f)) return false;
            z[1] = h * f;
        }
        // accept
        ++nst;
        ++c.steps;
        __syncthreads();
        if (lane == 0) {
            for (int i = q; i >= 2; --i) TAU(i) = TAU(i - 1);
            if (q == 1 && nst > 1) TAU(2) = TAU(1);
            TAU(1) = h;
        }
        __syncthreads();
        #pragma unroll
        for (int j = 0; j <= QMAX; ++j)
            if (j <= q) z[j] = fma(LC(j), acor, z[j]);
        --qwait;
        if (qwait == 1 && q != QMAX) {
            z[QMAX] = acor;
            saved_tq5 = tq5;
        }
        prepare_next(dsm);
        return true;
    }

    __device__ __forceinline__ void prepare_next(double dsm) {
        if (etamax == 1.0) {
            qwait = max(qwait, 2);
            hprime = h;
            eta = 1.0;
            etamax = 10.0;
            return;
        }
        const double etaq = 1.0 / (root_n(6.0 * dsm, L) + 1e-6);
        double e = etaq;
        if (qwait == 0) {
            qwait = 2;
            double etaqm1 = 0, etaqp1 = 0;
            if (q > 1) {
                const double ddn = wrms(zq(q)) * tq1;
                etaqm1 = 1.0 / (root_n(6.0 * ddn, q) + 1e-6);
            }
            if (q != QMAX && saved_tq5 != 0) {
                const double cquot = (tq5 / saved_tq5) * ipow(h / TAU(2), L);
                const double dup = wrms(acor - cquot * z[QMAX]) * tq3;
                etaqp1 = 1.0 / (root_n(10.0 * dup, L + 1) + 1e-6);
            }
            const double etam = fmax(etaqm1, fmax(etaq, etaqp1));
            if (etam < 1.5) {
                e = 1.0;
            } else if (etam == etaq) {
                e = etaq;
            } else if (etam == etaqm1) {
                e = etaqm1;
                qprime = q - 1;
            } else {
                e = etaqp1;
                qprime = q + 1;
                z[QMAX] = acor;
            }
        }
        eta = e < 1.5 ? 1.0 : fmin(e, etamax);
        hprime = e < 1.5 ? h : h * eta;
        etamax = 10.0;
    }

    __device__ __forceinline__ double initial_step(double f0, double tout) {
        const double tdist = fabs(tout);
        const double tround = UROUND * tdist;
        if (tdist < 2.0 * tround) return -1;
        const double hlb = 100.0 * tround;
        const double hub_inv = wave_max(own ? fabs(f0) / fma(0.1, fabs(z[0]), atol) : 0.0);
        double hub = 0.1 * tdist;
        if (hub * hub_inv > 1.0) hub = 1.0 / hub_inv;
        double hg = sqrt(hlb * hub);
        if (hub < hlb) return hg;
        double hnew = hg;
        for (int count = 1;; ++count) {
            bool ok = false;
            double yddnrm = 0;
            for (int tries = 0; tries < 4 && !ok; ++tries) {
                double f1;
                ++c.rhs;
                if (g.rhs(own ? fma(hg, f0, z[0]) : 0.0, f1)) {
                    yddnrm = wrms((f1 - f0) / hg);
                    ok = __builtin_isfinite(yddnrm);
                }
                if (!ok) hg *= 0.2;
            }
            if (!ok) return -1;
            hnew = (yddnrm * hub * hub > 2.0) ? sqrt(2.0 / yddnrm) : sqrt(hg * hub);
            if (count >= 4) break;
            const double hrat = hnew / hg;
            if (hrat > 0.5 && hrat < 2.0) break;
            if (count > 1 && hrat > 2.0) {
                hnew = hg;
                break;
            }
            hg = hnew;
        }
        return fmin(fmax(0.5 * hnew, hlb), hub);
    }

    // Integrate this lane's unknown y from 0 to tout; false on failure.
    __device__ __forceinline__ bool integrate(double& y, double tout, int maxsteps) {
        c = WaveCounters{};
        z[0] = own ? y : 0.0;
        #pragma unroll
        for (int j = 1; j <= QMAX; ++j) z[j] = 0;
        tn = 0;
        double f0;
        ++c.rhs;
        if (!g.rhs(z[0], f0)) return false;
        set_ewt();
        const double h0 = initial_step(f0, tout);
        if (!(h0 > 0)) return false;
        h = h0;
        z[1] = h * f0;
        q = qprime = 1;
        L = 2;
        qwait = 2;
        hscale = hprime = h;
        eta = 1;
        etamax = 10000.0;
        __syncthreads();
        if (lane == 0)
            for (int i = 0; i < QMAX + 2; ++i) TAU(i) = 0;
        __syncthreads();
        tq1 = tq2 = tq3 = tq4 = tq5 = 0;
        nst = nstlp = nstlj = 0;
        gammap = 0;
        gamrat = 1;
        crate = 1;
        saved_tq5 = 0;
        force_setup = false;
        jfresh = false;
        bool first = true;
        for (;;) {
            if (nst >= maxsteps) return false;
            if (!first) set_ewt();
            if ((tn + hprime - tout) * hprime > 0) {
                hprime = (tout - tn) * (1.0 - 4.0 * UROUND);
                eta = hprime / h;
            }
            if (hprime != h) {
                if (!first && qprime != q) {
                    adjust_order(qprime - q);
                    q = qprime;
                    L = q + 1;
                    qwait = L;
                }
                rescale();
            }
            first = false;
            if (!step(tout)) return false;
            const double troundoff = 100.0 * UROUND * (fabs(tn) + fabs(h));
            if (fabs(tn - tout) <= troundoff || (tn - tout) * h >= 0) {
                const double s = (tout - tn) / h;
                double v = zq(q);
                #pragma unroll
                for (int j = QMAX - 1; j >= 0; --j)
                    if (j < q) v = fma(s, v, z[j]);
                y = v;
                return true;
            }
        }
    }
};

}  // namespace hf
