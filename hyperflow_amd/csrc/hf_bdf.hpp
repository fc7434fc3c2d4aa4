// Per-lane variable-order (1-5) BDF integrator in fixed-leading-coefficient
// Nordsieck form, written for gfx950: the Nordsieck history lives in LDS
// (one [order][unknown][lane] slab per wave, conflict-free), everything else
// has a compile-time size (N unknowns, order <= 5) with loops over the current
// order unrolled and predicated, so the Newton matrix, weights and correction
// stay in VGPRs.
//
// This is the integrator that stands in for the reference's CVODE call
// (HyperFlow_V3/src/chemistry.cpp:10-26,64-68: CV_BDF, dense Newton, stop time
// = dt, 5000 max steps). Step/order control, the WRMS error test with
// w_i = 1/(rtol |y_i| + atol), the Newton convergence test and the exact stop
// at tout follow CVODE's published algorithm (Hindmarsh et al., ACM TOMS
// 31(3), 2005), including CVODE's Jacobian reuse: J is re-evaluated only on
// the first step, after 51 steps, or after a Newton failure, and is otherwise
// reloaded to re-form I - gamma J when gamma moves. GPU-specific choices: J is
// analytic (not a difference quotient), it and the factored Newton matrix
// are kept lane-contiguously in global memory (so they occupy registers only
// while they are formed or applied, not across rate evaluations), and
// I - gamma J is factored without pivoting (a zero or non-finite pivot is a
// recoverable failure that reduces the step, where I - gamma J tends to I).
// Every decision of a lane depends on its own cell only: the batched form may
// hold a lane back for some trips so that Jacobian evaluations of several
// lanes fall into one trip (pre() / wants_jac()), which changes when a trip
// runs, never what it computes, so results do not depend on the grouping.
#pragma once
#include <hip/hip_runtime.h>

#include "hf_device.hpp"

namespace hf {

struct BdfCounters {
    int steps = 0, rhs = 0, jac = 0, lu = 0, newton = 0, ncf = 0, netf = 0;
};

enum { BDF_OK = 0, BDF_RUNNING = 1, BDF_FAIL = 3 };

template <int N, class Sys, class MT = double, int ZSTRIDE = 64, int ZAS = 3>
struct Bdf {
    static constexpr int QMAX = 5;
    static constexpr double UROUND = 2.220446049250313e-16;
    // doubles per lane in the LDS slab: Nordsieck history, tau, l, and the
    // saved correction for an order increase
    static constexpr int SLAB = (QMAX + 1) * N + (QMAX + 2) + (QMAX + 1) + N;

    // Nordsieck history: element (j, i) at zs[(j*N + i) * ZSTRIDE]; the caller
    // points zs at this lane's slot of an array laid out [j][i][lane], so
    // consecutive lanes touch consecutive doubles. ZAS is its address space:
    // 3 (LDS) for small mechanisms, 1 (a per-wave tile in global memory) when
    // the history outgrows LDS. Columns above the current order q are kept
    // at zero, so the loops over the history (predict, restore, rescale, the
    // corrector update and the interpolation at tout) run over every column
    // without a per-lane test of q: the extra terms add or scale exact
    // zeros. CVODE keeps the correction saved for an order increase in its
    // column qmax; here it has a column of its own (ZS), so column qmax
    // obeys the same rule.
    double* zs;
    double ewt[N], acor[N];
    double tq1, tq2, tq3, tq4, tq5;
    double tn, h, hprime, hscale, eta, etamax, gamma, gammap, gamrat, rl1, crate, acnrm, saved_tq5;
    double rtol, atol;
    int q, qprime, L, qwait, nst, nstlp, nstlj;
    bool force_setup, jfresh;
    // Saved Jacobian and factored Newton matrix, in a per-wave tile of
    // global memory: SV consecutive elements of a lane form one 16-byte
    // vector, element k at [SI(k)] from the lane's base (the callers point
    // jsave and lu at tile + lane * SV), so the offsets are compile-time
    // constants and a row is read with 16-byte loads, each coalesced across
    // the wave.
    static constexpr int SV = 16 / sizeof(MT);
    static __device__ __forceinline__ constexpr int SI(int k) { return (k / SV) * (ZSTRIDE * SV) + k % SV; }
    static_assert((N * N) % SV == 0, "a lane's N x N slot must fill whole vectors (the tile holds N N ZSTRIDE)");
    MT* jsave = nullptr;
    MT* lu = nullptr;
    BdfCounters c;

    // ------------------------------------------------------------ helpers
    // Address zs in its own address space so the accesses are ds_read/ds_write
    // (or global loads) rather than flat loads through a generic pointer.
    __device__ __forceinline__ double& Z(int j, int i) const {
        typedef __attribute__((address_space(ZAS))) double zs_double;
        return *(double*)((zs_double*)zs + (j * N + i) * ZSTRIDE);
    }
    __device__ __forceinline__ double zrow(int j, int i) const { return Z(j, i); }
    // The Nordsieck column of unknown i staged in registers: the updates below
    // read and write it with one batch of slab loads and one of stores per
    // unknown, instead of a dependent slab round trip per update.
    __device__ __forceinline__ void zload(int i, double* zz) const {
        #pragma unroll
        for (int j = 0; j <= QMAX; ++j) zz[j] = Z(j, i);
    }
    __device__ __forceinline__ void zstore(int i, const double* zz) const {
        #pragma unroll
        for (int j = 0; j <= QMAX; ++j) Z(j, i) = zz[j];
    }
    // Step-size history tau and BDF coefficients l share the lane's LDS slab,
    // after the Nordsieck array: they are indexed by the current order.
    __device__ __forceinline__ double& slab(int k) const {
        typedef __attribute__((address_space(ZAS))) double zs_double;
        return *(double*)((zs_double*)zs + k * ZSTRIDE);
    }
    __device__ __forceinline__ double& TAU(int i) const { return slab((QMAX + 1) * N + i); }
    __device__ __forceinline__ double& LC(int i) const { return slab((QMAX + 1) * N + QMAX + 2 + i); }
    __device__ __forceinline__ double& ZS(int i) const { return slab((QMAX + 1) * N + 2 * QMAX + 3 + i); }
    __device__ __forceinline__ double wrms(const double* v) const {
        double s = 0;
        #pragma unroll
        for (int i = 0; i < N; ++i) {
            const double x = v[i] * ewt[i];
            s = fma(x, x, s);
        }
        return sqrt(s * (1.0 / N));
    }
    __device__ __forceinline__ void set_ewt() {
        #pragma unroll
        for (int i = 0; i < N; ++i) ewt[i] = 1.0 / fma(rtol, fabs(Z(0, i)), atol);
    }

    __device__ __forceinline__ void rescale() {
        double fac[QMAX + 1];
        fac[0] = 1.0;
        fac[1] = eta;
        #pragma unroll
        for (int j = 2; j <= QMAX; ++j) fac[j] = fac[j - 1] * eta;
        #pragma unroll
        for (int i = 0; i < N; ++i) {
            double zz[QMAX + 1];
            zload(i, zz);
            #pragma unroll
            for (int j = 1; j <= QMAX; ++j) zz[j] *= fac[j];
            zstore(i, zz);
        }
        h = hscale * eta;
        hprime = h;
        hscale = h;
    }
    __device__ __forceinline__ void predict() {
        tn += h;
        #pragma unroll
        for (int i = 0; i < N; ++i) {
            double zz[QMAX + 1];
            zload(i, zz);
            #pragma unroll
            for (int k = 1; k <= QMAX; ++k)
                #pragma unroll
                for (int j = QMAX; j >= k; --j) zz[j - 1] += zz[j];
            zstore(i, zz);
        }
    }
    __device__ __forceinline__ void restore(double saved_t) {
        tn = saved_t;
        #pragma unroll
        for (int i = 0; i < N; ++i) {
            double zz[QMAX + 1];
            zload(i, zz);
            #pragma unroll
            for (int k = 1; k <= QMAX; ++k)
                #pragma unroll
                for (int j = QMAX; j >= k; --j) zz[j - 1] -= zz[j];
            zstore(i, zz);
        }
    }

    // Order q -> q-1: subtract the top column's contribution (for q > 2, as
    // CVODE does) and clear it.
    __device__ __forceinline__ void decrease_order() {
        if (q <= 2) {
            #pragma unroll
            for (int i = 0; i < N; ++i) Z(2, i) = 0.0;
            return;
        }
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
        #pragma unroll
        for (int i = 0; i < N; ++i) {
            double zz[QMAX + 1];
            zload(i, zz);
            double top = 0;
            #pragma unroll
            for (int j = 0; j <= QMAX; ++j) top += j == q ? zz[j] : 0.0;
            #pragma unroll
            for (int j = 2; j <= QMAX; ++j) {
                if (j < q) zz[j] -= top * lc[j];
                if (j == q) zz[j] = 0.0;
            }
            zstore(i, zz);
        }
    }
    // Order q -> q+1: new top column from the saved correction.
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
        #pragma unroll
        for (int i = 0; i < N; ++i) {
            double zz[QMAX + 1];
            zload(i, zz);
            const double nz = A1 * ZS(i);
            #pragma unroll
            for (int j = 2; j <= QMAX; ++j) {
                if (j <= q) zz[j] += lc[j] * nz;
                if (j == q + 1) zz[j] = nz;
            }
            zstore(i, zz);
        }
    }
    __device__ __forceinline__ void adjust_order(int dq) {
        if (dq == 1) increase_order();
        else if (dq == -1) decrease_order();
    }

    __device__ __forceinline__ void set_bdf() {
        #pragma unroll
        for (int i = 0; i <= QMAX; ++i) LC(i) = 0;
        LC(0) = LC(1) = 1;
        double xi_inv = 1, xistar_inv = 1, alpha0 = -1, alpha0_hat = -1, hsum = h;
        if (q > 1) {
            #pragma unroll
            for (int j = 2; j < QMAX; ++j) {
                if (j < q) {
                    hsum += TAU(j - 1);
                    xi_inv = h / hsum;
                    alpha0 -= 1.0 / j;
                    #pragma unroll
                    for (int i = j; i >= 1; --i) LC(i) += LC(i - 1) * xi_inv;
                }
            }
            alpha0 -= 1.0 / q;
            xistar_inv = -LC(1) - alpha0;
            hsum += TAU(q - 1);
            xi_inv = h / hsum;
            alpha0_hat = -LC(1) - xi_inv;
            #pragma unroll
            for (int i = QMAX; i >= 1; --i)
                if (i <= q) LC(i) += LC(i - 1) * xistar_inv;
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

    // ------------------------------------------------------------ linear algebra
    // The saved Jacobian and the LU factors live in global memory; `opaque`
    // hides a pointer's provenance from the optimiser so that every use
    // reloads from memory instead of keeping 64 matrix values (128 VGPRs in
    // FP64) live in registers across the rate evaluations around it.
    // The slots are in global memory: addressing them in the global address
    // space makes the accesses global (not flat) loads and stores.
    typedef __attribute__((address_space(1))) MT GMT;
    template <class P>
    static __device__ __forceinline__ P* opaque(P* p) {
        asm volatile("" : "+v"(p));
        return static_cast<P*>(__builtin_assume_aligned(p, 16));
    }
    static __device__ __forceinline__ const GMT* gslot(const MT* p) { return (const GMT*)opaque(p); }
    static __device__ __forceinline__ GMT* gslot(MT* p) { return (GMT*)opaque(p); }

    // Small systems are unrolled into registers; large ones (N > 16) run
    // rolled loops over per-lane arrays.
    static constexpr bool SMALL = N <= 16;

    // Forward and back substitution with the unit-lower / upper factors in
    // `lu` (reciprocal pivots on the diagonal). MT is the Newton-matrix
    // precision (double, or float for the mixed form, where the residuals and
    // the corrections stay double).
    __device__ __forceinline__ void lu_solve(double* bd) const {
        const GMT* lp = gslot(lu);
        MT b[N];
        #pragma unroll
        for (int i = 0; i < N; ++i) b[i] = (MT)bd[i];
        if constexpr (SMALL) {
            #pragma unroll
            for (int k = 0; k < N; ++k)
                #pragma unroll
                for (int i = k + 1; i < N; ++i) b[i] = fma(-lp[SI(i * N + k)], b[k], b[i]);
            #pragma unroll
            for (int i = N - 1; i >= 0; --i) {
                MT s = b[i];
                #pragma unroll
                for (int j = i + 1; j < N; ++j) s = fma(-lp[SI(i * N + j)], b[j], s);
                b[i] = s * lp[SI(i * N + i)];
            }
        } else {
            #pragma unroll 1
            for (int i = 1; i < N; ++i) {
                MT s = b[i];
                #pragma unroll 1
                for (int k = 0; k < i; ++k) s = fma(-lp[SI(i * N + k)], b[k], s);
                b[i] = s;
            }
            #pragma unroll 1
            for (int i = N - 1; i >= 0; --i) {
                MT s = b[i];
                #pragma unroll 1
                for (int j = i + 1; j < N; ++j) s = fma(-lp[SI(i * N + j)], b[j], s);
                b[i] = s * lp[SI(i * N + i)];
            }
        }
        #pragma unroll
        for (int i = 0; i < N; ++i) bd[i] = (double)b[i];
    }

    // Row i of I - gamma J eliminated against the finished rows 0..i-1 of
    // `lu` and stored; returns false on a zero or non-finite pivot.
    __device__ __forceinline__ bool lu_row(int i, MT g, const GMT* js) {
        MT row[N];
        #pragma unroll
        for (int j = 0; j < N; ++j) row[j] = (i == j ? MT(1) : MT(0)) - g * js[SI(i * N + j)];
        const GMT* up = gslot(lu);
        #pragma unroll 1
        for (int k = 0; k < i; ++k) {
            const MT l = row[k] * up[SI(k * N + k)];
            row[k] = l;
            #pragma unroll 1
            for (int j = k + 1; j < N; ++j) row[j] = fma(-l, up[SI(k * N + j)], row[j]);
        }
        const MT p = row[i];
        const bool ok = fabs(p) > __FLT_MIN__ && __builtin_isfinite(p);
        row[i] = MT(1) / p;
        GMT* wp = gslot(lu);
        #pragma unroll
        for (int j = 0; j < N; ++j) wp[SI(i * N + j)] = row[j];
        return ok;
    }

    // Newton-matrix setup. A fresh Jacobian (jbad) is evaluated and saved in
    // the lane's global slot (jsave, lane-contiguous); otherwise the saved one
    // is reused, as CVODE reuses J across steps and only re-forms I - gamma J
    // when gamma changes. Small mechanisms use the analytic Jacobian; large
    // ones CVODE's own difference quotients (the reference's configuration,
    // HyperFlow_V3/src/chemistry.cpp:20-24: dense matrix, no Jacobian
    // function). f = f(y), already evaluated. I - gamma J is factored row by
    // row (Doolittle, no pivoting) straight into the lane's `lu` slot: row i is
    // formed from the saved J, eliminated against the finished rows 0..i-1
    // read back from `lu`, and stored, so one row at a time is live.
    __device__ __forceinline__ bool lsetup(Sys& sys, const double* y, const double* f, bool jbad) {
        ++c.lu;
        if (jbad) {
            ++c.jac;
            if constexpr (Sys::kAnalyticJac) {
                MT J[N * N];
                double fj[N];
                if (!sys.jac(y, J, fj)) return false;
                GMT* jw = gslot(jsave);
                #pragma unroll
                for (int i = 0; i < N * N; ++i) jw[SI(i)] = J[i];
            } else {
                // cvLsDenseDQJac: column j from f(y + inc_j e_j) with
                // inc_j = max(sqrt(uround) |y_j|, minInc / ewt_j),
                // minInc = 1000 |h| uround N ||f||_wrms (1 when f = 0).
                const double fnorm = wrms(f);
                const double min_inc = fnorm != 0 ? 1000.0 * fabs(h) * UROUND * N * fnorm : 1.0;
                const double srur = sqrt(UROUND);
                double yj[N], fj[N];
                #pragma unroll
                for (int i = 0; i < N; ++i) yj[i] = y[i];
                #pragma unroll 1
                for (int j = 0; j < N; ++j) {
                    const double ysave = yj[j];
                    const double inc = fmax(srur * fabs(ysave), min_inc / ewt[j]);
                    yj[j] = ysave + inc;
                    ++c.rhs;
                    if (!sys.rhs(yj, fj)) return false;
                    yj[j] = ysave;
                    const double rinc = 1.0 / inc;
                    GMT* js = gslot(jsave);
                    #pragma unroll
                    for (int i = 0; i < N; ++i) js[SI(i * N + j)] = (MT)((fj[i] - f[i]) * rinc);
                }
            }
            nstlj = nst;
            jfresh = true;
        } else {
            jfresh = false;
        }
        gammap = gamma;
        gamrat = 1;
        crate = 1;
        nstlp = nst;
        const MT g = (MT)gamma;
        const GMT* js = gslot(jsave);
        bool ok = true;
        if constexpr (SMALL) {
            // The finished rows stay in registers for the rows below them
            // (N*N values, local to this block); reading them back from the
            // slot would put a dependent memory round trip on every row.
            MT fin[N * N];
            #pragma unroll
            for (int i = 0; i < N; ++i) {
                MT row[N];
                #pragma unroll
                for (int j = 0; j < N; ++j) row[j] = (i == j ? MT(1) : MT(0)) - g * js[SI(i * N + j)];
                #pragma unroll
                for (int k = 0; k < i; ++k) {
                    const MT l = row[k] * fin[k * N + k];
                    row[k] = l;
                    #pragma unroll
                    for (int j = k + 1; j < N; ++j) row[j] = fma(-l, fin[k * N + j], row[j]);
                }
                const MT p = row[i];
                ok = ok && fabs(p) > __FLT_MIN__ && __builtin_isfinite(p);
                row[i] = MT(1) / p;
                GMT* wp = gslot(lu);
                #pragma unroll
                for (int j = 0; j < N; ++j) {
                    fin[i * N + j] = row[j];
                    wp[SI(i * N + j)] = row[j];
                }
            }
        } else {
            #pragma unroll 1
            for (int i = 0; i < N; ++i) ok = lu_row(i, g, js) && ok;
        }
        return ok;
    }

    // ------------------------------------------------------------ trip-driven integration
    // The integration is a state machine advanced by trip(): every call makes
    // exactly one right-hand-side evaluation, at a point chosen by the lane's
    // phase, and then the bookkeeping that evaluation enables. A wave of lanes
    // at different points of their steps (first Newton iteration, a later one,
    // a retry after a failed error test, a new cell's initial-step search)
    // therefore shares one evaluation per trip instead of serialising each
    // lane's whole step. The algorithm per lane is CVODE's: the order of
    // evaluations and decisions is unchanged, only their grouping into calls.
    enum { PH_F0 = 0, PH_HIN = 1, PH_START = 2, PH_NEWTON = 3, PH_ERR_RHS = 4, PH_DEAD = 5 };
    int phase, m, convfail, ncf, nef, hin_count, hin_tries;
    bool retry, callsetup, jbad;
    double saved_t, delp, hg, hub, hlb;
    double tout;
    int maxsteps;
    bool first;

    // Start of a Newton solve (iteration m = 0), after the predictor or for
    // a retry with a fresh setup: CVODE's setup decision (MSBP 20, DGMAX 0.3)
    // and Jacobian decision (MSBJ 51, DGMAXLSETUP 0.2).
    __device__ __forceinline__ void begin_newton() {
        m = 0;
        delp = 0;
        #pragma unroll
        for (int i = 0; i < N; ++i) acor[i] = 0;
        callsetup = nst == 0 || nst >= nstlp + 20 || fabs(gamrat - 1.0) > 0.3 || convfail != 0 || force_setup;
        jbad = callsetup && (nst == 0 || nst >= nstlj + 51 || convfail == 2 ||
                             (convfail == 1 && fabs(gamma / gammap - 1.0) < 0.2));
        force_setup = false;
        jfresh = false;
        phase = PH_NEWTON;
    }

    // A step attempt: between accepted steps, the tstop clamp and pending
    // order/step changes first; then the predictor and the BDF coefficients.
    __device__ __forceinline__ bool start_attempt() {
        if (!retry) {
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
            ncf = nef = 0;
            convfail = 0;
        }
        saved_t = tn;
        predict();
        if ((tn - tout) * h > 0) tn = tout;
        set_bdf();
        rl1 = 1.0 / LC(1);
        gamma = h * rl1;
        if (nst == 0) gammap = gamma;
        gamrat = (nst > 0) ? gamma / gammap : 1.0;
        begin_newton();
        return true;
    }

    // The step after its last Newton failure (CVODE: eta 0.25, convfail OTHER).
    __device__ __forceinline__ int newton_failed() {
        restore(saved_t);
        ++ncf;
        ++c.ncf;
        etamax = 1;
        if (ncf >= 10) return BDF_FAIL;
        eta = 0.25;
        convfail = 2;
        rescale();
        retry = true;
        phase = PH_START;
        return BDF_RUNNING;
    }

    // Newton converged: the local error test, then acceptance or a retry.
    __device__ __forceinline__ int error_test(double* yout) {
        convfail = 0;
        const double dsm = acnrm * tq2;
        if (dsm > 1.0) {
            ++nef;
            ++c.netf;
            restore(saved_t);
            force_setup = true;
            if (nef >= 7) return BDF_FAIL;
            etamax = 1;
            retry = true;
            phase = PH_START;
            if (nef <= 3) {
                eta = 1.0 / (root_n(6.0 * dsm, L) + 1e-6);
                eta = fmax(0.1, eta);
                if (nef >= 2) eta = fmin(eta, 0.2);
                rescale();
            } else if (q > 1) {
                eta = 0.1;
                adjust_order(-1);
                L = q;
                --q;
                qwait = L;
                rescale();
            } else {
                // order 1 after repeated failures: restart the history from a
                // fresh derivative at the restored state (next trip)
                eta = 0.1;
                h *= eta;
                hprime = h;
                hscale = h;
                qwait = 10;
                phase = PH_ERR_RHS;
            }
            return BDF_RUNNING;
        }
        // accept
        ++nst;
        ++c.steps;
        #pragma unroll
        for (int i = QMAX; i >= 2; --i)
            if (i <= q) TAU(i) = TAU(i - 1);
        if (q == 1 && nst > 1) TAU(2) = TAU(1);
        TAU(1) = h;
        double lcv[QMAX + 1];
        #pragma unroll
        for (int j = 0; j <= QMAX; ++j) lcv[j] = LC(j);
        --qwait;
        const bool save_acor = qwait == 1 && q != QMAX;
        #pragma unroll
        for (int i = 0; i < N; ++i) {
            double zz[QMAX + 1];
            zload(i, zz);
            // l_j = 0 above q (set_bdf), so those columns stay zero
            #pragma unroll
            for (int j = 0; j <= QMAX; ++j) zz[j] = fma(lcv[j], acor[i], zz[j]);
            zstore(i, zz);
            if (save_acor) ZS(i) = acor[i];
        }
        if (save_acor) saved_tq5 = tq5;
        prepare_next(dsm);
        retry = false;
        phase = PH_START;
        const double troundoff = 100.0 * UROUND * (fabs(tn) + fabs(h));
        if (fabs(tn - tout) <= troundoff || (tn - tout) * h >= 0) {
            const double s = (tout - tn) / h;
            #pragma unroll
            for (int i = 0; i < N; ++i) {
                // Horner from column qmax: the zero columns above q leave
                // v = Z(q) exactly where the sum of the order-q history starts
                double v = Z(QMAX, i);
                #pragma unroll
                for (int j = QMAX - 1; j >= 0; --j) v = fma(s, v, Z(j, i));
                yout[i] = v;
            }
            return BDF_OK;
        }
        return BDF_RUNNING;
    }

    // Every branch computes into locals and the state is written once at the
    // end, so no path-dependent store addresses appear (they would keep the
    // lane's integrator state out of registers).
    __device__ __forceinline__ void prepare_next(double dsm) {
        double n_eta, n_hprime;
        int n_qprime = q, n_qwait = qwait;
        if (etamax == 1.0) {
            n_qwait = max(qwait, 2);
            n_hprime = h;
            n_eta = 1.0;
        } else {
            const double etaq = 1.0 / (root_n(6.0 * dsm, L) + 1e-6);
            double e = etaq;
            if (qwait == 0) {
                n_qwait = 2;
                double etaqm1 = 0, etaqp1 = 0;
                if (q > 1) {
                    double tmp[N];
                    #pragma unroll
                    for (int i = 0; i < N; ++i) tmp[i] = zrow(q, i);
                    const double ddn = wrms(tmp) * tq1;
                    etaqm1 = 1.0 / (root_n(6.0 * ddn, q) + 1e-6);
                }
                if (q != QMAX && saved_tq5 != 0) {
                    const double cquot = (tq5 / saved_tq5) * ipow(h / TAU(2), L);
                    double tmp[N];
                    #pragma unroll
                    for (int i = 0; i < N; ++i) tmp[i] = acor[i] - cquot * ZS(i);
                    const double dup = wrms(tmp) * tq3;
                    etaqp1 = 1.0 / (root_n(10.0 * dup, L + 1) + 1e-6);
                }
                const double etam = fmax(etaqm1, fmax(etaq, etaqp1));
                if (etam < 1.5) {
                    e = 1.0;
                } else if (etam == etaq) {
                    e = etaq;
                } else if (etam == etaqm1) {
                    e = etaqm1;
                    n_qprime = q - 1;
                } else {
                    e = etaqp1;
                    n_qprime = q + 1;
                    #pragma unroll
                    for (int i = 0; i < N; ++i) ZS(i) = acor[i];
                }
            }
            n_eta = e < 1.5 ? 1.0 : fmin(e, etamax);
            n_hprime = e < 1.5 ? h : h * n_eta;
        }
        eta = n_eta;
        hprime = n_hprime;
        qprime = n_qprime;
        qwait = n_qwait;
        etamax = 10.0;
    }

    // ------------------------------------------------------------ driver
    // begin() starts an integration of y from t = 0 to t_out; every trip()
    // makes one rate evaluation and returns BDF_RUNNING, BDF_OK (t_out
    // reached, y written) or BDF_FAIL.
    __device__ __forceinline__ int begin(Sys& sys, const double* y, double t_out, int max_steps) {
        tout = t_out;
        maxsteps = max_steps;
        c = BdfCounters{};
        #pragma unroll
        for (int i = 0; i < N; ++i) {
            Z(0, i) = y[i];
            #pragma unroll
            for (int j = 2; j <= QMAX; ++j) Z(j, i) = 0.0;
        }
        tn = 0;
        phase = PH_F0;
        return BDF_RUNNING;
    }

    // Initial step accepted (CVODE's cvHin result): the history starts at order 1.
    __device__ __forceinline__ void finish_begin(double h0) {
        h = h0;
        #pragma unroll
        for (int i = 0; i < N; ++i) Z(1, i) *= h;   // Z(1) held f(y0) during the search
        q = qprime = 1;
        L = 2;
        qwait = 2;
        hscale = hprime = h;
        eta = 1;
        etamax = 10000.0;
        #pragma unroll
        for (int i = 0; i < QMAX + 2; ++i) TAU(i) = 0;
        tq1 = tq2 = tq3 = tq4 = tq5 = 0;
        nst = nstlp = nstlj = 0;
        gammap = 0;
        gamrat = 1;
        crate = 1;
        saved_tq5 = 0;
        force_setup = false;
        jfresh = false;
        first = true;
        retry = false;
        phase = PH_START;
    }

    // The part of a trip before its rate evaluation that needs none: the
    // start of a step attempt (predictor, coefficients, the setup and
    // Jacobian decisions). After it, wants_jac() tells whether the coming
    // trip evaluates a Jacobian. A lane may then sit out trips: its state
    // does not change until it runs one.
    __device__ __forceinline__ void pre() {
        if (phase == PH_START && !start_attempt()) phase = PH_DEAD;
    }
    __device__ __forceinline__ bool wants_setup() const { return phase == PH_NEWTON && m == 0 && callsetup; }
    __device__ __forceinline__ bool wants_jac() const { return wants_setup() && jbad; }

    __device__ __forceinline__ int trip(Sys& sys, double* yout) {
        pre();
        if (phase == PH_DEAD) return BDF_FAIL;
        double yv[N], f[N];
        #pragma unroll
        for (int i = 0; i < N; ++i) {
            const double z1 = Z(1, i);
            const double dz = phase == PH_NEWTON ? acor[i] : phase == PH_HIN ? hg * z1 : 0.0;
            yv[i] = Z(0, i) + dz;
        }
        ++c.rhs;
        const bool ok = sys.rhs(yv, f);

        if (phase == PH_NEWTON) {
            if (ok && (m != 0 || !callsetup || lsetup(sys, yv, f, jbad))) {
                ++c.newton;
                double delta[N];
                #pragma unroll
                for (int i = 0; i < N; ++i) delta[i] = -(fma(rl1, Z(1, i), acor[i]) - gamma * f[i]);
                lu_solve(delta);
                if (gamrat != 1.0) {
                    const double s = 2.0 / (1.0 + gamrat);
                    #pragma unroll
                    for (int i = 0; i < N; ++i) delta[i] *= s;
                }
                #pragma unroll
                for (int i = 0; i < N; ++i) acor[i] += delta[i];
                const double del = wrms(delta);
                if (m > 0) crate = fmax(0.3 * crate, del / delp);
                const double dcon = del * fmin(1.0, crate) / tq4;
                if (dcon <= 1.0) {
                    acnrm = (m == 0) ? del : wrms(acor);
                    return error_test(yout);
                }
                if (!(m == 2 || (m >= 1 && del > 2.0 * delp))) {
                    delp = del;
                    ++m;
                    return BDF_RUNNING;
                }
                if (!jfresh) {
                    // stale Jacobian: retry the solve with a fresh setup (CVODE's TRY_AGAIN)
                    convfail = 1;
                    begin_newton();
                    return BDF_RUNNING;
                }
            }
            return newton_failed();
        }
        if (phase == PH_F0) {
            if (!ok) return BDF_FAIL;
            #pragma unroll
            for (int i = 0; i < N; ++i) Z(1, i) = f[i];
            set_ewt();
            // CVODE's initial step (cvHin): bounds, then up to four refinements
            const double tdist = fabs(tout), tround = UROUND * tdist;
            if (tdist < 2.0 * tround) return BDF_FAIL;
            hlb = 100.0 * tround;
            double hub_inv = 0;
            #pragma unroll
            for (int i = 0; i < N; ++i) hub_inv = fmax(hub_inv, fabs(f[i]) / fma(0.1, fabs(yv[i]), atol));
            hub = 0.1 * tdist;
            if (hub * hub_inv > 1.0) hub = 1.0 / hub_inv;
            hg = sqrt(hlb * hub);
            if (hub < hlb) {
                finish_begin(hg);
            } else {
                hin_count = 1;
                hin_tries = 0;
                phase = PH_HIN;
            }
            return BDF_RUNNING;
        }
        if (phase == PH_HIN) {
            double yddnrm = 0;
            bool good = false;
            if (ok) {
                double d2[N];
                #pragma unroll
                for (int i = 0; i < N; ++i) d2[i] = (f[i] - Z(1, i)) / hg;
                yddnrm = wrms(d2);
                good = __builtin_isfinite(yddnrm);
            }
            if (!good) {
                hg *= 0.2;
                return ++hin_tries >= 4 ? BDF_FAIL : BDF_RUNNING;
            }
            hin_tries = 0;
            double hnew = (yddnrm * hub * hub > 2.0) ? sqrt(2.0 / yddnrm) : sqrt(hg * hub);
            bool done = hin_count >= 4;
            if (!done) {
                const double hrat = hnew / hg;
                if (hrat > 0.5 && hrat < 2.0) {
                    done = true;
                } else if (hin_count > 1 && hrat > 2.0) {
                    hnew = hg;
                    done = true;
                }
            }
            if (done) {
                finish_begin(fmin(fmax(0.5 * hnew, hlb), hub));
            } else {
                hg = hnew;
                ++hin_count;
            }
            return BDF_RUNNING;
        }
        // PH_ERR_RHS: new first derivative at the restored state
        if (!ok) return BDF_FAIL;
        #pragma unroll
        for (int i = 0; i < N; ++i) Z(1, i) = h * f[i];
        retry = true;
        phase = PH_START;
        return BDF_RUNNING;
    }
};


}  // namespace hf
