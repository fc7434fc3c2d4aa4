// Chemistry system of the V3 reaction step and the per-cell advance that
// replaces Chemistry::advance (HyperFlow_V3/src/chemistry.cpp:55-76).
#pragma once
#include "hf_bdf.hpp"
#include "hf_device.hpp"
#include "hyperflow.h"

namespace hf {

// Constant-volume adiabatic cell: unknowns y = Y[0..NS-2] at fixed rho and
// specific internal energy e; the last species closes the mass,
// Y[NS-1] = 1 - sum(y) (chemistry.cpp:36-53). T is recovered from e at every
// evaluation, warm-started from the previous evaluation's T.
template <class T>
struct ChemSys {
    static constexpr int NS = T::NS, N = NS - 1;
    // analytic Jacobian for small mechanisms; difference quotients otherwise
    static constexpr bool kAnalyticJac = NS <= 16;
    const MechData<T>& d;
    double rho, e, t;
    const double* shift;   // optional per-reaction ln A shift (sensitivities)

    __device__ __forceinline__ const MechData<T>& tables() const { return d; }
    __device__ __forceinline__ bool temp(const MechData<T>& dd, const double* y, double* Y) {
        double s = 0;
        #pragma unroll
        for (int k = 0; k < N; ++k) { Y[k] = y[k]; s += y[k]; }
        Y[N] = 1.0 - s;
        double tt = t;
        if (!temperature_from_e(dd, Y, e, tt)) return false;
        t = tt;
        return true;
    }
    __device__ __forceinline__ bool rhs(const double* y, double* f) {
        const MechData<T>& dd = tables();
        double Y[NS], C[NS], w[NS];
        if (!temp(dd, y, Y)) return false;
        #pragma unroll
        for (int k = 0; k < NS; ++k) C[k] = rho * Y[k] * dd.invW[k];
        // when every active lane of the wave is above every species' Tmid (as
        // in ignition and burnt gas), the rates take the high-range NASA
        // coefficients without a per-lane choice (a wave-uniform branch)
        if (__all(t > dd.tmid) && dd.common_tmid) production_rates<T, true>(dd, t, C, w, shift);
        else production_rates<T, false>(dd, t, C, w, shift);
        bool ok = true;
        const double rr = 1.0 / rho;
        #pragma unroll
        for (int k = 0; k < N; ++k) {
            f[k] = w[k] * dd.W[k] * rr;
            ok = ok && finite(f[k]);
        }
        return ok;
    }
    template <class MT>
    __device__ __forceinline__ bool jac(const double* y, MT* J, double* f) {
        const MechData<T>& dd = tables();
        double Y[NS];
        if (!temp(dd, y, Y)) return false;
        chem_jacobian(dd, t, rho, Y, J, f, shift);
        bool ok = true;
        #pragma unroll
        for (int i = 0; i < N * N; ++i) ok = ok && finite(J[i]);
        return ok;
    }
};

struct ChemParams {
    double rtol, atol;
    int maxsteps;
};

// One cell: U holds nv = NS+3 component planes U[n*ld + i]; the advanced
// species planes go to out[k*ldo + i]. Momenta and energy are not changed.
// Returns HF_OK, HF_EINVAL (bad dt), HF_EDOMAIN (inadmissible state before or
// after, or accepted species below -1e-13) or HF_ECONV (integrator failure).
// A lane's integration of one cell, split so that a persistent kernel can
// interleave cells: start() validates and begins, step() makes one trip of the
// integrator (one rate evaluation and the bookkeeping it enables), finish()
// applies the reference's post-step checks and writes the species.
// zs: this lane's slot of the wave's Nordsieck slab ([order][unknown][64
// lanes]), in LDS (ZAS 3) or in a global per-wave tile (ZAS 1).
template <class T, class MT = double, int ZAS = 3>
struct CellRun {
    static constexpr int NS = T::NS, NV = NS + 3, N = NS - 1;
    ChemSys<T> sys;
    Bdf<N, ChemSys<T>, MT, 64, ZAS> bdf;
    double y[N];

    // jsave/lu: this lane's saved-Jacobian and LU-factor slots in a per-wave
    // tile (element k at Bdf::SI(k) from tile + lane * Bdf::SV).
    __device__ __forceinline__ CellRun(const MechData<T>& d, double* zs, const ChemParams& p, MT* jsave, MT* lu)
        : sys{d, 1.0, 0.0, 1000.0, nullptr} {
        bdf.zs = zs;
        bdf.rtol = p.rtol;
        bdf.atol = p.atol;
        bdf.jsave = jsave;
        bdf.lu = lu;
    }

    // HF_OK when the integration is running, else the cell's error code.
    // q: component 0 of the cell, component n at q[n * plane].
    __device__ __forceinline__ int start(const double* __restrict__ q, long plane, double dt, int maxsteps) {
        bdf.c = BdfCounters{};
        if (!(dt > 0) || !finite(dt)) return HF_EINVAL;
        double u[NV];
        #pragma unroll
        for (int n = 0; n < NV; ++n) u[n] = q[n * plane];
        Props pr;
        if (!properties(sys.d, u, pr, 1000.0)) return HF_EDOMAIN;
        double rho = 0;
        #pragma unroll
        for (int k = 0; k < NS; ++k) rho += u[k];
        const double ux = u[NS] / rho, vy = u[NS + 1] / rho;
        sys.rho = rho;
        sys.e = u[NS + 2] / rho - 0.5 * (ux * ux + vy * vy);
        sys.t = pr.T;
        #pragma unroll
        for (int k = 0; k < N; ++k) y[k] = u[k] / rho;
        return bdf.begin(sys, y, dt, maxsteps) == BDF_RUNNING ? HF_OK : HF_ECONV;
    }

    __device__ __forceinline__ int step() { return bdf.trip(sys, y); }

    // Post-step checks of chemistry.cpp:68-75 and the write of the species planes.
    __device__ __forceinline__ int finish(const double* __restrict__ q, long plane, double* __restrict__ out, long i,
                                          long ldo) {
        double u[NV];
        double sum = 0;
        bool ok = true;
        #pragma unroll
        for (int k = 0; k < N; ++k) {
            ok = ok && finite(y[k]) && y[k] >= -1e-13;
            u[k] = sys.rho * fmax(0.0, y[k]);
            sum += u[k];
        }
        if (!ok) return HF_EDOMAIN;
        u[N] = sys.rho - sum;
        u[NS] = q[NS * plane];
        u[NS + 1] = q[(NS + 1) * plane];
        u[NS + 2] = q[(NS + 2) * plane];
        Props pr;
        if (!properties(sys.d, u, pr, sys.t)) return HF_EDOMAIN;
        #pragma unroll
        for (int k = 0; k < NS; ++k) out[k * ldo + i] = u[k];
        return HF_OK;
    }
};

// counters: [steps, rhs, jac, lu, newton, conv fails, err fails, max steps of a cell]
__device__ __forceinline__ void reduce_counters(const BdfCounters& c, int max_cell_steps,
                                                unsigned long long* counters) {
    unsigned long long v[7] = {(unsigned long long)c.steps, (unsigned long long)c.rhs, (unsigned long long)c.jac,
                               (unsigned long long)c.lu, (unsigned long long)c.newton, (unsigned long long)c.ncf,
                               (unsigned long long)c.netf};
    int mx = max_cell_steps;
    #pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        #pragma unroll
        for (int k = 0; k < 7; ++k) v[k] += __shfl_xor(v[k], off, 64);
        mx = max(mx, __shfl_xor(mx, off, 64));
    }
    if ((threadIdx.x & 63) == 0) {
        #pragma unroll
        for (int k = 0; k < 7; ++k)
            if (v[k]) atomicAdd(&counters[k], v[k]);
        atomicMax(&counters[7], (unsigned long long)mx);
    }
}

}  // namespace hf
