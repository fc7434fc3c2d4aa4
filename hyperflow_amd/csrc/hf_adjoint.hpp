// Discrete adjoint and forward tangent of fixed-step backward-Euler chemistry
// (no reference counterpart: rate-parameter sensitivities of the V3 reaction
// step, HyperFlow_V3/src/chemistry.cpp:36-53 for the right-hand side).
//
// Integration, per cell at fixed rho and e = e(T0, Y0):
//   y_{n+1} = y_n + h f(y_{n+1}; theta, e),  y_0 = Y0[0..N-1],  n = 0..S-1
// solved by full Newton with the analytic Jacobian to the roundoff floor.
// theta_i = ln A_i multiplies reaction i's rate (forward and reverse; k0 and
// k_inf of falloff reactions alike), so dq_i/dtheta_i = q_i.
// Objective: J = T(y_S) + sum_{n=1..S} h qdot(y_n),
//   qdot = -sum_k u_k wdot_k = -rho sum_{k<N} (e_k - e_N) f_k
// (u_k molar, e_k mass internal energy; the second form uses mass
// conservation, sum_k W_k wdot_k = 0, so the last species drops out).
//
// Adjoint recursion (reverse over n = S..1, lambda = dJ/dy_n from the future):
//   lbar = lambda + h dqdot/dy(y_n);  (I - h J_n)^T mu = lbar;  lambda <- mu
//   g_theta += h [mu . df/dtheta + dqdot/dtheta](y_n)
//   g_e     += h [mu . df/de     + dqdot/de    ](y_n)
// started from lambda = dT/dy(y_S), g_e = dT/de(y_S) = 1/cv, and closed by
// dJ/dT0 = g_e cv(T0, Y0), dJ/dY0_k = lambda_k [k < N] + g_e e_k(T0).
// The tangent runs the same linearisation forward along one direction.
#pragma once
#include "hf_device.hpp"

namespace hf {

// Everything the linearisation needs at one converged state.
template <class T>
struct BeState {
    static constexpr int NS = T::NS, N = NS - 1, NR = T::NR;
    double t, cv;
    double f[N], J[N * N], fT[N], q[NR], de[N], dcv[N];   // de_k = e_k - e_N, dcv_k = cv_k - cv_N
};

// f, J, fT, q and the species energies at y (T warm-started from t).
template <class T>
__device__ __forceinline__ bool be_eval(const MechData<T>& d, double rho, double e, const double* shift,
                                        const double* y, double& t, BeState<T>& s) {
    constexpr int NS = T::NS, N = NS - 1;
    double Y[NS], sum = 0;
    #pragma unroll
    for (int k = 0; k < N; ++k) {
        Y[k] = y[k];
        sum += y[k];
    }
    Y[N] = 1.0 - sum;
    if (!temperature_from_e(d, Y, e, t)) return false;
    chem_jacobian<T, double>(d, t, rho, Y, s.J, s.f, shift, s.fT, s.q);
    double ek[NS], cvk[NS], cv = 0;
    #pragma unroll
    for (int k = 0; k < NS; ++k) {
        const Nasa7 na = nasa_select(d, k, t);
        const double* a = na.a;
        const double cpr = fma(t, fma(t, fma(t, fma(t, a[4], a[3]), a[2]), a[1]), a[0]);
        const double hrt = fma(t, fma(t, fma(t, fma(t, a[4] * 0.2, a[3] * 0.25), a[2] * (1.0 / 3.0)), a[1] * 0.5),
                               a[0]) + a[5] / t;
        ek[k] = (hrt - 1.0) * kR * t * d.invW[k];
        cvk[k] = (cpr - 1.0) * kR * d.invW[k];
        cv = fma(Y[k], cvk[k], cv);
    }
    s.t = t;
    s.cv = cv;
    bool ok = cv > 0 && finite(t);
    #pragma unroll
    for (int k = 0; k < N; ++k) {
        s.de[k] = ek[k] - ek[N];
        s.dcv[k] = cvk[k] - cvk[N];
        ok = ok && finite(s.f[k]) && finite(s.fT[k]);
    }
    #pragma unroll
    for (int i = 0; i < N * N; ++i) ok = ok && finite(s.J[i]);
    return ok;
}

// M = I - h J factored in place without pivoting (h J is small next to I for
// the step sizes this integration is used with; a tiny pivot is a failure).
template <int N>
__device__ __forceinline__ bool be_factor(double* M) {
    #pragma unroll
    for (int k = 0; k < N; ++k) {
        const double p = M[k * N + k];
        if (!(fabs(p) > 1e-300) || !finite(p)) return false;
        const double inv = 1.0 / p;
        M[k * N + k] = inv;
        #pragma unroll
        for (int i = k + 1; i < N; ++i) {
            const double l = M[i * N + k] * inv;
            M[i * N + k] = l;
            #pragma unroll
            for (int j = k + 1; j < N; ++j) M[i * N + j] = fma(-l, M[k * N + j], M[i * N + j]);
        }
    }
    return true;
}
template <int N>
__device__ __forceinline__ void be_solve(const double* M, double* b) {
    #pragma unroll
    for (int k = 0; k < N; ++k)
        #pragma unroll
        for (int i = k + 1; i < N; ++i) b[i] = fma(-M[i * N + k], b[k], b[i]);
    #pragma unroll
    for (int i = N - 1; i >= 0; --i) {
        double s = b[i];
        #pragma unroll
        for (int j = i + 1; j < N; ++j) s = fma(-M[i * N + j], b[j], s);
        b[i] = s * M[i * N + i];
    }
}
// Solve M^T x = b with M = L U: U^T z = b (forward), then L^T x = z (backward).
template <int N>
__device__ __forceinline__ void be_solve_t(const double* M, double* b) {
    #pragma unroll
    for (int i = 0; i < N; ++i) {
        double s = b[i];
        #pragma unroll
        for (int k = 0; k < i; ++k) s = fma(-M[k * N + i], b[k], s);
        b[i] = s * M[i * N + i];
    }
    #pragma unroll
    for (int i = N - 1; i >= 0; --i) {
        double s = b[i];
        #pragma unroll
        for (int k = i + 1; k < N; ++k) s = fma(-M[k * N + i], b[k], s);
        b[i] = s;
    }
}

template <class T>
__device__ __forceinline__ void be_matrix(const BeState<T>& s, double h, double* M) {
    constexpr int N = T::NS - 1;
    #pragma unroll
    for (int i = 0; i < N * N; ++i) M[i] = -h * s.J[i];
    #pragma unroll
    for (int i = 0; i < N; ++i) M[i * N + i] += 1.0;
}

// One backward-Euler step from yold to y (in place), leaving the state
// evaluated at the converged y in s. Newton runs to the roundoff floor: it
// stops when every correction is below 1e-14 of its component (floor 1e-30),
// or when the largest relative correction stops halving once below 1e-8.
template <class T>
__device__ __forceinline__ bool be_step(const MechData<T>& d, double rho, double e, const double* shift, double h,
                                        const double* yold, double* y, double& t, BeState<T>& s) {
    constexpr int N = T::NS - 1;
    double prev = HUGE_VAL;
    for (int it = 0; it < 30; ++it) {
        if (!be_eval(d, rho, e, shift, y, t, s)) return false;
        double M[N * N], r[N];
        be_matrix(s, h, M);
        if (!be_factor<N>(M)) return false;
        #pragma unroll
        for (int k = 0; k < N; ++k) r[k] = -(y[k] - yold[k] - h * s.f[k]);
        be_solve<N>(M, r);
        double wn = 0;
        #pragma unroll
        for (int k = 0; k < N; ++k) {
            y[k] += r[k];
            wn = fmax(wn, fabs(r[k]) / (fabs(y[k]) + 1e-30));
        }
        if (!finite(wn)) return false;
        if (wn <= 1e-14 || (it >= 2 && wn < 1e-8 && wn >= 0.5 * prev)) return be_eval(d, rho, e, shift, y, t, s);
        prev = wn;
    }
    return false;
}

// qdot and its partial derivatives at a state.
template <class T>
__device__ __forceinline__ double be_qdot(const BeState<T>& s, double rho) {
    constexpr int N = T::NS - 1;
    double q = 0;
    #pragma unroll
    for (int k = 0; k < N; ++k) q = fma(s.de[k], s.f[k], q);
    return -rho * q;
}
// dqdot/dy_j (N values) and dqdot/de.
template <class T>
__device__ __forceinline__ void be_dqdot(const BeState<T>& s, double rho, double* dqdy, double& dqde) {
    constexpr int N = T::NS - 1;
    double a = 0, b = 0;   // sum (cv_k - cv_N) f_k ; sum (e_k - e_N) fT_k
    #pragma unroll
    for (int k = 0; k < N; ++k) {
        a = fma(s.dcv[k], s.f[k], a);
        b = fma(s.de[k], s.fT[k], b);
    }
    const double rcv = 1.0 / s.cv;
    #pragma unroll
    for (int j = 0; j < N; ++j) {
        double c = 0;
        #pragma unroll
        for (int k = 0; k < N; ++k) c = fma(s.de[k], s.J[k * N + j], c);
        // dT/dy_j = -(e_j - e_N)/cv
        dqdy[j] = -rho * (c - a * s.de[j] * rcv);
    }
    dqde = -rho * (b + a) * rcv;
}

// sum_k nu_ki W_k w_k over the species k < N of reaction I's reactant and
// product lists (net stoichiometry, repeated entries counted).
template <class T, int I>
__device__ __forceinline__ double nu_w_dot(const MechData<T>& d, const double* w) {
    constexpr int N = T::NS - 1;
    double acc = 0;
    static_for<0, T::NRMAX>([&](auto jc) HF_AI {
        constexpr int r = T::reac[I][decltype(jc)::value], p = T::prod[I][decltype(jc)::value];
        if constexpr (r >= 0 && r < N) acc -= d.W[r] * w[r];
        if constexpr (p >= 0 && p < N) acc += d.W[p] * w[p];
    });
    return acc;
}

}  // namespace hf
