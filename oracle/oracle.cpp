// ORACLE — TEST INFRASTRUCTURE ONLY. Never linked into, loaded by or called
// from the product library; only tests/, __graft_entry__.smoke() and the
// cpu_baseline leg of bench.py use it, as the checker and the timed CPU path.
//
// A plain C++/OpenMP restatement of the reference's hot path:
//   * properties / minmod / admissible / hllc      HyperFlow_V3/src/physics.cpp:4-48
//   * density / internalEnergy                      HyperFlow_V3/src/solver.hpp:21-25
//   * NND face loop + conservative update           HyperFlow_V3/src/BlockAMR.cpp:210-243
//   * Chemistry::advance / Chemistry::rhs           HyperFlow_V3/src/chemistry.cpp:36-76
//   * ChemistryBatch::advance (dynamic,8; trial copy; lowest failing index)
//                                                   HyperFlow_V3/src/chemistry_batch.cpp:21-50
// The third-party arithmetic behind those call sites is restated from its
// published algorithms (the libraries are not vendored and not installed):
//   * Cantera 3.2.0 thermo (NASA7, setState_UV) and GasKinetics net production
//     rates (Arrhenius, three-body, Lindemann/Troe falloff, Kc reverse rates);
//   * SUNDIALS 7.5.0 CVODE: variable-order (1-5) fixed-leading-coefficient
//     Nordsieck BDF, Newton with the dense difference-quotient Jacobian and
//     dense partial-pivot LU, WRMS error test, exact stop time
//     (Hindmarsh et al., ACM TOMS 31(3), 2005; CVODE user guide).
//
// Mechanism tables come from oracle/mech.py (PyYAML), not from the product's
// loader.
#include <omp.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <limits>
#include <string>
#include <vector>

namespace {

constexpr double R = 8314.462618;
constexpr double P0 = 101325.0;
constexpr double SMALL = 1e-300;
constexpr double UROUND = std::numeric_limits<double>::epsilon();

enum { OK = 0, EINVAL_ = 1, EDOMAIN_ = 2, ECONV_ = 3 };

struct Mech {
    int ns, nr;
    std::vector<double> W, Tmid, low, high;    // low/high: ns*7
    std::vector<int> kind, rev, troe, nu_r, nu_p; // nu: nr*ns
    std::vector<double> A, b, EaR, A0, b0, Ea0R, tA, tT3, tT1, tT2, eff;  // eff: nr*ns
};

struct Thermo {
    // Species cp/R, h/RT, s/R at T.
    static void nasa(const Mech& m, double T, double* cp, double* h, double* s) {
        const double lt = std::log(T);
        for (int k = 0; k < m.ns; ++k) {
            const double* a = (T <= m.Tmid[k] ? &m.low[7 * k] : &m.high[7 * k]);
            if (cp) cp[k] = a[0] + T * (a[1] + T * (a[2] + T * (a[3] + T * a[4])));
            if (h) h[k] = a[0] + T * (a[1] / 2 + T * (a[2] / 3 + T * (a[3] / 4 + T * a[4] / 5))) + a[5] / T;
            if (s) s[k] = a[0] * lt + T * (a[1] + T * (a[2] / 2 + T * (a[3] / 3 + T * a[4] / 4))) + a[6];
        }
    }
};

// e and cv per unit mass for (possibly unnormalised) Y.
void e_cv(const Mech& m, double T, const double* Y, double& e, double& cv) {
    double cp[64], h[64];
    Thermo::nasa(m, T, cp, h, nullptr);
    e = 0;
    cv = 0;
    for (int k = 0; k < m.ns; ++k) {
        const double yw = Y[k] / m.W[k];
        e += yw * (h[k] - 1.0) * R * T;
        cv += yw * (cp[k] - 1.0) * R;
    }
}

// Energy -> temperature Newton (restates Cantera setState_UV's result:
// T with e(T, Y) = e to the iteration's convergence limit).
bool temperature(const Mech& m, double e, const double* Y, double& T) {
    if (!(T > 0) || !std::isfinite(T)) T = 1000.0;
    for (int it = 0; it < 100; ++it) {
        double ee, cv;
        e_cv(m, T, Y, ee, cv);
        if (!(cv > 0) || !std::isfinite(ee)) return false;
        double dT = (ee - e) / cv;
        double Tn = T - dT;
        if (!(Tn > 0)) Tn = 0.5 * T;
        const bool done = std::abs(dT) <= 1e-13 * T;
        T = Tn;
        if (done) return std::isfinite(T) && T > 0;
    }
    return false;
}

struct Props { double p, a, T; };

// physics.cpp:4-15
int properties(const Mech& m, const double* u, Props& out, double Tguess = 1000.0) {
    const int ns = m.ns;
    double rho = 0;
    for (int k = 0; k < ns; ++k) rho += u[k];
    bool finite = true, neg = false;
    for (int n = 0; n < ns + 3; ++n) finite = finite && std::isfinite(u[n]);
    for (int k = 0; k < ns; ++k) neg = neg || u[k] < 0;
    if (!finite || !(rho > 0) || neg) return EDOMAIN_;
    double Y[64];
    for (int k = 0; k < ns; ++k) Y[k] = u[k] / rho;
    const double ux = u[ns] / rho, vy = u[ns + 1] / rho;
    const double e = u[ns + 2] / rho - 0.5 * (std::pow(ux, 2) + std::pow(vy, 2));
    double T = Tguess;
    if (!temperature(m, e, Y, T)) return EDOMAIN_;
    double cp[64];
    Thermo::nasa(m, T, cp, nullptr, nullptr);
    double sw = 0, cpm = 0, cvm = 0;
    for (int k = 0; k < ns; ++k) {
        const double yw = Y[k] / m.W[k];
        sw += yw;
        cpm += yw * cp[k] * R;
        cvm += yw * (cp[k] - 1.0) * R;
    }
    const double wbar = 1.0 / sw;
    out.T = T;
    out.p = rho * R * T / wbar;
    out.a = std::sqrt(cpm / cvm * R * T / wbar);
    if (!(out.T > 0 && out.p > 0 && out.a > 0 && std::isfinite(out.a))) return EDOMAIN_;
    return OK;
}

// physics.cpp:40-44
inline double minmod(double x, double y) {
    return x * y > 0 ? std::copysign(std::min(std::abs(x), std::abs(y)), x) : 0.;
}

// physics.cpp:17-38 (Davis wave speeds; one star state on the sm >= 0 side).
void hllc(int ns, const double* l, const double* r, int n, const Props& pl, const Props& pr, double* f) {
    const int nv = ns + 3, E = ns + 2;
    double rl = 0, rr = 0;
    for (int k = 0; k < ns; ++k) { rl += l[k]; rr += r[k]; }
    const double ul = l[n] / rl, ur = r[n] / rr;
    const double sl = std::min(ul - pl.a, ur - pr.a), sr = std::max(ul + pl.a, ur + pr.a);
    auto flux = [&](const double* q, double v, double p, double* out) {
        for (int c = 0; c < nv; ++c) out[c] = v * q[c];
        out[n] += p;
        out[E] += p * v;
    };
    if (sl >= 0) { flux(l, ul, pl.p, f); return; }
    if (sr <= 0) { flux(r, ur, pr.p, f); return; }
    const double dl = rl * (sl - ul), dr = rr * (sr - ur);
    const double sm = (pr.p - pl.p + dl * ul - dr * ur) / (dl - dr);
    const bool left = sm >= 0;
    const double* q = left ? l : r;
    const double v = left ? ul : ur, s = left ? sl : sr, p = left ? pl.p : pr.p;
    double star[64];
    const double scale = (s - v) / (s - sm);
    double rs = 0, rq = 0;
    for (int c = 0; c < nv; ++c) star[c] = q[c] * scale;
    for (int k = 0; k < ns; ++k) { rs += star[k]; rq += q[k]; }
    star[n] = rs * sm;
    star[E] += rs * (sm - v) * (sm + p / (rq * (s - v)));
    flux(q, v, p, f);
    for (int c = 0; c < nv; ++c) f[c] = f[c] + s * (star[c] - q[c]);
}

// ---------------------------------------------------------------- kinetics
void rates(const Mech& m, double T, const double* C, double* wdot, const double* lnA_shift = nullptr,
           double* qout = nullptr) {
    const int ns = m.ns;
    double h[64], s[64];
    Thermo::nasa(m, T, nullptr, h, s);
    const double logT = std::log(T), invT = 1.0 / T;
    const double logc0 = std::log(P0 / (R * T));
    for (int k = 0; k < ns; ++k) wdot[k] = 0;
    for (int i = 0; i < m.nr; ++i) {
        const double sh = lnA_shift ? lnA_shift[i] : 0.0;
        double kf = m.A[i] * std::exp(m.b[i] * logT - m.EaR[i] * invT + sh);
        double Mc = 0;
        if (m.kind[i] != 0)
            for (int k = 0; k < ns; ++k) Mc += m.eff[i * ns + k] * C[k];
        if (m.kind[i] == 1) kf *= Mc;
        if (m.kind[i] == 2) {
            const double k0 = m.A0[i] * std::exp(m.b0[i] * logT - m.Ea0R[i] * invT + sh);
            const double pr = Mc * k0 / (kf + SMALL);
            double F = 1.0;
            if (m.troe[i]) {
                const double rt3 = std::abs(m.tT3[i]) < SMALL ? std::numeric_limits<double>::infinity() : 1.0 / m.tT3[i];
                const double rt1 = std::abs(m.tT1[i]) < SMALL ? std::numeric_limits<double>::infinity() : 1.0 / m.tT1[i];
                double fc = (1 - m.tA[i]) * std::exp(-T * rt3) + m.tA[i] * std::exp(-T * rt1);
                if (m.tT2[i] != 0) fc += std::exp(-m.tT2[i] / T);
                const double lfc = std::log10(std::max(fc, SMALL));
                const double lpr = std::log10(std::max(pr, SMALL));
                const double c = -0.4 - 0.67 * lfc, nn = 0.75 - 1.27 * lfc;
                const double f1 = (lpr + c) / (nn - 0.14 * (lpr + c));
                F = std::pow(10.0, lfc / (1.0 + f1 * f1));
            }
            kf = kf * (pr / (1.0 + pr)) * F;
        }
        double dG = 0, dn = 0;
        for (int k = 0; k < ns; ++k) {
            const int nu = m.nu_p[i * ns + k] - m.nu_r[i * ns + k];
            if (nu) { dG += nu * (h[k] - s[k]); dn += nu; }
        }
        const double kr = m.rev[i] ? kf * std::exp(dG - dn * logc0) : 0.0;
        double fwd = kf, bwd = kr;
        for (int k = 0; k < ns; ++k) {
            for (int c = 0; c < m.nu_r[i * ns + k]; ++c) fwd *= C[k];
            for (int c = 0; c < m.nu_p[i * ns + k]; ++c) bwd *= C[k];
        }
        const double q = fwd - bwd;
        if (qout) qout[i] = q;
        for (int k = 0; k < ns; ++k) wdot[k] += (m.nu_p[i * ns + k] - m.nu_r[i * ns + k]) * q;
    }
}

// ---------------------------------------------------------------- chemistry RHS (chemistry.cpp:36-53)
struct Cell {
    const Mech* m;
    double rho, e, T;      // T: last inversion result (Cantera keeps the phase temperature)
    const double* shift = nullptr;
    long nrhs = 0;
    // returns 0 ok, 1 recoverable failure
    int rhs(const double* y, double* dy) {
        const int ns = m->ns;
        double Y[64], C[64], w[64];
        double sum = 0;
        for (int k = 0; k < ns - 1; ++k) { Y[k] = y[k]; sum += y[k]; }
        Y[ns - 1] = 1 - sum;
        ++nrhs;
        double Tn = T;
        if (!temperature(*m, e, Y, Tn)) return 1;
        T = Tn;
        for (int k = 0; k < ns; ++k) C[k] = rho * Y[k] / m->W[k];
        rates(*m, T, C, w, shift);
        for (int k = 0; k < ns - 1; ++k) {
            const double r = w[k] * m->W[k] / rho;
            if (!std::isfinite(r)) return 1;
            dy[k] = r;
        }
        return 0;
    }
};

// ---------------------------------------------------------------- dense LU (partial pivoting)
bool lu_factor(int n, double* a, int* piv) {
    for (int k = 0; k < n; ++k) {
        int p = k;
        double best = std::abs(a[k * n + k]);
        for (int i = k + 1; i < n; ++i)
            if (std::abs(a[i * n + k]) > best) { best = std::abs(a[i * n + k]); p = i; }
        piv[k] = p;
        if (best == 0) return false;
        if (p != k)
            for (int j = 0; j < n; ++j) std::swap(a[k * n + j], a[p * n + j]);
        const double inv = 1.0 / a[k * n + k];
        for (int i = k + 1; i < n; ++i) {
            const double l = a[i * n + k] * inv;
            a[i * n + k] = l;
            for (int j = k + 1; j < n; ++j) a[i * n + j] -= l * a[k * n + j];
        }
    }
    return true;
}
void lu_solve(int n, const double* a, const int* piv, double* b) {
    for (int k = 0; k < n; ++k) {
        if (piv[k] != k) std::swap(b[k], b[piv[k]]);
        for (int i = k + 1; i < n; ++i) b[i] -= a[i * n + k] * b[k];
    }
    for (int i = n - 1; i >= 0; --i) {
        double s = b[i];
        for (int j = i + 1; j < n; ++j) s -= a[i * n + j] * b[j];
        b[i] = s / a[i * n + i];
    }
}

// ---------------------------------------------------------------- CVODE-equivalent BDF
struct BdfStats { long steps = 0, rhs = 0, jac = 0, lu = 0, newton = 0, ncf = 0, netf = 0; };

template <class Sys>
class Bdf {
public:
    Bdf(int n, Sys& sys, double rtol, double atol, int maxsteps)
        : n_(n), sys_(sys), rtol_(rtol), atol_(atol), maxsteps_(maxsteps) {
        for (auto& z : zn_) z.assign(n, 0.0);
        ewt_.assign(n, 0); acor_.assign(n, 0); y_.assign(n, 0); ftemp_.assign(n, 0);
        tempv_.assign(n, 0); delta_.assign(n, 0);
        J_.assign(n * n, 0); M_.assign(n * n, 0); piv_.assign(n, 0);
    }

    // Integrate y from t=0 to tout; y in/out. Returns OK or ECONV_.
    int integrate(double* y0, double tout, BdfStats& st) {
        const int n = n_;
        for (int i = 0; i < n; ++i) zn_[0][i] = y0[i];
        tn_ = 0;
        if (sys_.rhs(zn_[0].data(), ftemp_.data()) != 0) return ECONV_;
        ++st.rhs;
        set_ewt(zn_[0].data());
        // initial step size (CVODE's cvHin estimate)
        double h0 = initial_step(tout, st);
        if (!(h0 > 0)) return ECONV_;
        h_ = h0;
        for (int i = 0; i < n; ++i) zn_[1][i] = h_ * ftemp_[i];
        q_ = 1; qprime_ = 1; L_ = 2; qwait_ = L_;
        hscale_ = h_; hprime_ = h_; eta_ = 1; etamax_ = 10000.0;
        for (double& t : tau_) t = 0;
        for (double& t : tq_) t = 0;
        nst_ = 0; nstlp_ = 0; nstlj_ = 0; gammap_ = 0; gamrat_ = 1; crate_ = 1;
        saved_tq5_ = 0; jcur_ = false; indx_acor_ = qmax_;
        bool first = true;
        for (;;) {
            if (nst_ >= maxsteps_) return ECONV_;
            if (!first) {
                set_ewt(zn_[0].data());
            }
            // stop-time control: never step past tout
            if ((tn_ + hprime_ - tout) * hprime_ > 0) {
                hprime_ = (tout - tn_) * (1.0 - 4.0 * UROUND);
                eta_ = hprime_ / h_;
            }
            if (!first && hprime_ != h_) adjust_params();
            else if (first && hprime_ != h_) { rescale(); }
            first = false;
            int flag = step(tout, st);
            if (flag != OK) return flag;
            // reached tout (within roundoff): interpolate there
            const double troundoff = 100.0 * UROUND * (std::abs(tn_) + std::abs(h_));
            if (std::abs(tn_ - tout) <= troundoff || (tn_ - tout) * h_ >= 0) {
                interpolate(tout, y0);
                return OK;
            }
        }
    }

private:
    static constexpr int qmax_ = 5;
    int n_;
    Sys& sys_;
    double rtol_, atol_;
    int maxsteps_;
    std::vector<double> zn_[qmax_ + 1];
    std::vector<double> ewt_, acor_, y_, ftemp_, tempv_, delta_, J_, M_;
    std::vector<int> piv_;
    double tn_ = 0, h_ = 0, hprime_ = 0, hscale_ = 0, eta_ = 1, etamax_ = 1e4;
    double tau_[qmax_ + 2], tq_[6], l_[qmax_ + 1];
    int q_ = 1, qprime_ = 1, L_ = 2, qwait_ = 2;
    long nst_ = 0, nstlp_ = 0, nstlj_ = 0;
    double gamma_ = 0, gammap_ = 0, gamrat_ = 1, rl1_ = 1, crate_ = 1, acnrm_ = 0, saved_tq5_ = 0;
    bool jcur_ = false, force_setup_ = false;
    int indx_acor_ = qmax_;
    double hu_ = 0;

    void set_ewt(const double* y) {
        for (int i = 0; i < n_; ++i) ewt_[i] = 1.0 / (rtol_ * std::abs(y[i]) + atol_);
    }
    double wrms(const double* v) const {
        double s = 0;
        for (int i = 0; i < n_; ++i) { const double x = v[i] * ewt_[i]; s += x * x; }
        return std::sqrt(s / n_);
    }

    double initial_step(double tout, BdfStats& st) {
        const int n = n_;
        const double tdist = std::abs(tout);
        const double tround = UROUND * std::abs(tout);
        if (tdist < 2.0 * tround) return -1;
        const double hlb = 100.0 * tround;
        // upper bound: 0.1*tdist limited so that the first step changes no component
        // by more than ~0.1 |y| + atol
        double hub_inv = 0;
        for (int i = 0; i < n; ++i) {
            const double denom = 0.1 * std::abs(zn_[0][i]) + atol_;
            hub_inv = std::max(hub_inv, std::abs(ftemp_[i]) / denom);
        }
        double hub = 0.1 * tdist;
        if (hub * hub_inv > 1.0) hub = 1.0 / hub_inv;
        double hg = std::sqrt(hlb * hub);
        if (hub < hlb) return hg;
        double hnew = hg;
        for (int count = 1;; ++count) {
            // second-derivative estimate ||(f(y + hg f) - f(y)) / hg||
            bool ok = false;
            double yddnrm = 0;
            for (int tries = 0; tries < 4 && !ok; ++tries) {
                for (int i = 0; i < n; ++i) y_[i] = zn_[0][i] + hg * ftemp_[i];
                ++st.rhs;
                if (sys_.rhs(y_.data(), tempv_.data()) == 0) {
                    for (int i = 0; i < n; ++i) tempv_[i] = (tempv_[i] - ftemp_[i]) / hg;
                    yddnrm = wrms(tempv_.data());
                    ok = std::isfinite(yddnrm);
                }
                if (!ok) hg *= 0.2;
            }
            if (!ok) return -1;
            hnew = (yddnrm * hub * hub > 2.0) ? std::sqrt(2.0 / yddnrm) : std::sqrt(hg * hub);
            if (count >= 4) break;
            const double hrat = hnew / hg;
            if (hrat > 0.5 && hrat < 2.0) break;
            if (count > 1 && hrat > 2.0) { hnew = hg; break; }
            hg = hnew;
        }
        double h0 = 0.5 * hnew;
        h0 = std::min(std::max(h0, hlb), hub);
        return h0;
    }

    void rescale() {
        double factor = eta_;
        for (int j = 1; j <= q_; ++j) {
            for (int i = 0; i < n_; ++i) zn_[j][i] *= factor;
            factor *= eta_;
        }
        h_ = hscale_ * eta_;
        hprime_ = h_;
        hscale_ = h_;
    }

    // Order decrease for BDF: remove the top Nordsieck component's contribution.
    void decrease_order() {
        if (q_ == 2) return;  // zn[2] is simply dropped
        // coefficients of x^2 (x + xi_1) ... (x + xi_{q-2})
        double l[qmax_ + 1];
        for (int j = 0; j <= qmax_; ++j) l[j] = 0;
        l[2] = 1;
        double hsum = 0;
        for (int j = 1; j <= q_ - 2; ++j) {
            hsum += tau_[j];
            const double xi = hsum / hscale_;
            for (int i = j + 2; i >= 2; --i) l[i] = l[i] * xi + l[i - 1];
        }
        for (int j = 2; j < q_; ++j)
            for (int i = 0; i < n_; ++i) zn_[j][i] -= zn_[q_][i] * l[j];
    }
    void increase_order() {
        // new column zn[q+1] = acor * A where A fits the order-(q+1) polynomial
        double l[qmax_ + 1];
        for (int j = 0; j <= qmax_; ++j) l[j] = 0;
        l[2] = 1;
        double alpha1 = 1, prod = 1, xiold = 1, alpha0 = -1, hsum = hscale_;
        if (q_ > 1) {
            for (int j = 1; j < q_; ++j) {
                hsum += tau_[j + 1];
                const double xi = hsum / hscale_;
                prod *= xi;
                alpha0 -= 1.0 / (j + 1);
                alpha1 += 1.0 / xi;
                for (int i = j + 2; i >= 2; --i) l[i] = l[i] * xiold + l[i - 1];
                xiold = xi;
            }
        }
        const double A1 = (-alpha0 - alpha1) / prod;
        for (int i = 0; i < n_; ++i) zn_[L_][i] = A1 * zn_[indx_acor_][i];
        for (int j = 2; j <= q_; ++j)
            for (int i = 0; i < n_; ++i) zn_[j][i] += l[j] * zn_[L_][i];
    }
    void adjust_order(int deltaq) {
        if (q_ == 2 && deltaq != 1) return;  // order 2 -> 1: zn[2] is simply dropped
        if (deltaq == 1) increase_order();
        else if (deltaq == -1) decrease_order();
    }
    void adjust_params() {
        if (qprime_ != q_) {
            adjust_order(qprime_ - q_);
            q_ = qprime_;
            L_ = q_ + 1;
            qwait_ = L_;
        }
        rescale();
    }

    void predict() {
        tn_ += h_;
        for (int k = 1; k <= q_; ++k)
            for (int j = q_; j >= k; --j)
                for (int i = 0; i < n_; ++i) zn_[j - 1][i] += zn_[j][i];
    }
    void restore(double saved_t) {
        tn_ = saved_t;
        for (int k = 1; k <= q_; ++k)
            for (int j = q_; j >= k; --j)
                for (int i = 0; i < n_; ++i) zn_[j - 1][i] -= zn_[j][i];
    }

    // Coefficients l and test constants tq for the current order and step history.
    void set_bdf() {
        for (int i = 0; i <= qmax_; ++i) l_[i] = 0;
        l_[0] = l_[1] = 1;
        double xi_inv = 1, xistar_inv = 1, alpha0 = -1, alpha0_hat = -1;
        double hsum = h_;
        if (q_ > 1) {
            for (int j = 2; j < q_; ++j) {
                hsum += tau_[j - 1];
                xi_inv = h_ / hsum;
                alpha0 -= 1.0 / j;
                for (int i = j; i >= 1; --i) l_[i] += l_[i - 1] * xi_inv;
            }
            alpha0 -= 1.0 / q_;
            xistar_inv = -l_[1] - alpha0;
            hsum += tau_[q_ - 1];
            xi_inv = h_ / hsum;
            alpha0_hat = -l_[1] - xi_inv;
            for (int i = q_; i >= 1; --i) l_[i] += l_[i - 1] * xistar_inv;
        }
        const double A1 = 1.0 - alpha0_hat + alpha0;
        const double A2 = 1.0 + q_ * A1;
        tq_[2] = std::abs(A1 / (alpha0 * A2));
        tq_[5] = std::abs(A2 * xistar_inv / (l_[q_] * xi_inv));
        if (qwait_ == 1) {
            if (q_ > 1) {
                const double C = xistar_inv / l_[q_];
                const double A3 = alpha0 + 1.0 / q_;
                const double A4 = alpha0_hat + xi_inv;
                const double Cpinv = (1.0 - A4 + A3) / A3;
                tq_[1] = std::abs(C * Cpinv);
            } else {
                tq_[1] = 1.0;
            }
            hsum += tau_[q_];
            xi_inv = h_ / hsum;
            const double A5 = alpha0 - 1.0 / (q_ + 1);
            const double A6 = alpha0_hat - xi_inv;
            const double Cppinv = (1.0 - A6 + A5) / A2;
            tq_[3] = std::abs(Cppinv / (xi_inv * (q_ + 2) * A5));
        }
        tq_[4] = 0.1 / tq_[2];  // nonlinear-solver coefficient 0.1
    }

    void jac_dq(const double* y, const double* fy, BdfStats& st) {
        // CVODE dense difference-quotient Jacobian, column by column.
        const int n = n_;
        const double fnorm = wrms(fy);
        const double minInc = fnorm != 0 ? 1000.0 * std::abs(h_) * UROUND * n * fnorm : 1.0;
        const double srur = std::sqrt(UROUND);
        std::vector<double> yy(y, y + n), ft(n);
        for (int j = 0; j < n; ++j) {
            const double ys = yy[j];
            const double inc = std::max(srur * std::abs(ys), minInc / ewt_[j]);
            yy[j] = ys + inc;
            ++st.rhs;
            if (sys_.rhs(yy.data(), ft.data()) != 0) { yy[j] = ys; throw 1; }
            for (int i = 0; i < n; ++i) J_[i * n + j] = (ft[i] - fy[i]) / inc;
            yy[j] = ys;
        }
        ++st.jac;
    }

    // Build M = I - gamma J (recomputing J when stale) and factor it.
    bool lsetup(int convfail, const double* ypred, const double* fpred, BdfStats& st) {
        const int n = n_;
        const double dgamma = std::abs(gamma_ / gammap_ - 1.0);
        const bool jbad = nst_ == 0 || nst_ >= nstlj_ + 51 || (convfail == 1 && dgamma < 0.2) || convfail == 2;
        if (jbad) {
            try { jac_dq(ypred, fpred, st); } catch (int) { return false; }
            nstlj_ = nst_;
            jcur_ = true;
        } else {
            jcur_ = false;
        }
        for (int i = 0; i < n * n; ++i) M_[i] = -gamma_ * J_[i];
        for (int i = 0; i < n; ++i) M_[i * n + i] += 1.0;
        ++st.lu;
        gammap_ = gamma_;
        gamrat_ = 1;
        crate_ = 1;
        nstlp_ = nst_;
        return lu_factor(n, M_.data(), piv_.data());
    }

    // Newton iteration for the corrector. Returns 0 converged, 1 retry with new J, 2 fail (reduce h), 3 rhs fail.
    int newton(int& convfail, BdfStats& st) {
        const int n = n_;
        // predicted value and f there
        for (int i = 0; i < n; ++i) y_[i] = zn_[0][i];
        ++st.rhs;
        if (sys_.rhs(y_.data(), ftemp_.data()) != 0) return 3;
        const bool callsetup = nst_ == 0 || nst_ >= nstlp_ + 20 || std::abs(gamrat_ - 1.0) > 0.3 || convfail != 0 ||
                               force_setup_;
        force_setup_ = false;
        if (callsetup) {
            if (!lsetup(convfail, y_.data(), ftemp_.data(), st)) return 3;
        }
        for (int i = 0; i < n; ++i) acor_[i] = 0;
        double delp = 0;
        for (int m = 0; m < 3; ++m) {
            ++st.newton;
            for (int i = 0; i < n; ++i) delta_[i] = -(rl1_ * zn_[1][i] + acor_[i] - gamma_ * ftemp_[i]);
            lu_solve(n, M_.data(), piv_.data(), delta_.data());
            if (gamrat_ != 1.0)
                for (int i = 0; i < n; ++i) delta_[i] *= 2.0 / (1.0 + gamrat_);
            for (int i = 0; i < n; ++i) { acor_[i] += delta_[i]; y_[i] = zn_[0][i] + acor_[i]; }
            const double del = wrms(delta_.data());
            if (m > 0) crate_ = std::max(0.3 * crate_, del / delp);
            const double dcon = del * std::min(1.0, crate_) / tq_[4];
            if (dcon <= 1.0) {
                acnrm_ = (m == 0) ? del : wrms(acor_.data());
                jcur_ = false;
                return 0;
            }
            if (m == 2 || (m >= 1 && del > 2.0 * delp)) break;
            delp = del;
            ++st.rhs;
            if (sys_.rhs(y_.data(), ftemp_.data()) != 0) return 3;
        }
        if (!jcur_) { convfail = 1; return 1; }
        return 2;
    }

    int step(double tout, BdfStats& st) {
        const double saved_t = tn_;
        int ncf = 0, nef = 0;
        int convfail = 0;
        double dsm = 0;
        for (;;) {
            predict();
            if ((tn_ - tout) * h_ > 0) tn_ = tout;
            set_bdf();
            rl1_ = 1.0 / l_[1];
            gamma_ = h_ * rl1_;
            if (nst_ == 0) gammap_ = gamma_;
            gamrat_ = (nst_ > 0) ? gamma_ / gammap_ : 1.0;
            int nf = newton(convfail, st);
            while (nf == 1) nf = newton(convfail, st);  // stale Jacobian: same prediction, fresh J
            if (nf != 0) {
                restore(saved_t);
                ++ncf;
                ++st.ncf;
                etamax_ = 1;
                if (ncf >= 10) return ECONV_;
                eta_ = 0.25;
                convfail = 2;
                rescale();
                continue;
            }
            convfail = 0;
            dsm = acnrm_ * tq_[2];
            if (dsm <= 1.0) break;
            // error test failed
            ++nef;
            ++st.netf;
            restore(saved_t);
            force_setup_ = true;  // refactor with the new gamma, keep J
            if (nef >= 7) return ECONV_;
            etamax_ = 1;
            if (nef <= 3) {
                eta_ = 1.0 / (std::pow(6.0 * dsm, 1.0 / L_) + 1e-6);
                eta_ = std::max(0.1, eta_);
                if (nef >= 2) eta_ = std::min(eta_, 0.2);
                rescale();
                continue;
            }
            if (q_ > 1) {
                eta_ = 0.1;
                adjust_order(-1);
                L_ = q_;
                --q_;
                qwait_ = L_;
                rescale();
                continue;
            }
            eta_ = 0.1;
            h_ *= eta_;
            hprime_ = h_;
            hscale_ = h_;
            qwait_ = 10;
            ++st.rhs;
            if (sys_.rhs(zn_[0].data(), tempv_.data()) != 0) return ECONV_;
            for (int i = 0; i < n_; ++i) zn_[1][i] = h_ * tempv_[i];
        }
        // complete the step
        ++nst_;
        ++st.steps;
        hu_ = h_;
        for (int i = q_; i >= 2; --i) tau_[i] = tau_[i - 1];
        if (q_ == 1 && nst_ > 1) tau_[2] = tau_[1];
        tau_[1] = h_;
        for (int j = 0; j <= q_; ++j)
            for (int i = 0; i < n_; ++i) zn_[j][i] += l_[j] * acor_[i];
        --qwait_;
        if (qwait_ == 1 && q_ != qmax_) {
            for (int i = 0; i < n_; ++i) zn_[qmax_][i] = acor_[i];
            saved_tq5_ = tq_[5];
            indx_acor_ = qmax_;
        }
        prepare_next(dsm);
        return OK;
    }

    void prepare_next(double dsm) {
        if (etamax_ == 1.0) {
            qwait_ = std::max(qwait_, 2);
            qprime_ = q_;
            hprime_ = h_;
            eta_ = 1.0;
        } else {
            const double etaq = 1.0 / (std::pow(6.0 * dsm, 1.0 / L_) + 1e-6);
            if (qwait_ != 0) {
                eta_ = etaq;
                qprime_ = q_;
            } else {
                qwait_ = 2;
                double etaqm1 = 0, etaqp1 = 0;
                if (q_ > 1) {
                    const double ddn = wrms(zn_[q_].data()) * tq_[1];
                    etaqm1 = 1.0 / (std::pow(6.0 * ddn, 1.0 / q_) + 1e-6);
                }
                if (q_ != qmax_ && saved_tq5_ != 0) {
                    const double cquot = (tq_[5] / saved_tq5_) * std::pow(h_ / tau_[2], L_);
                    for (int i = 0; i < n_; ++i) tempv_[i] = acor_[i] - cquot * zn_[qmax_][i];
                    const double dup = wrms(tempv_.data()) * tq_[3];
                    etaqp1 = 1.0 / (std::pow(10.0 * dup, 1.0 / (L_ + 1)) + 1e-6);
                }
                const double etam = std::max(etaqm1, std::max(etaq, etaqp1));
                if (etam < 1.5) {
                    eta_ = 1.0;
                    qprime_ = q_;
                } else if (etam == etaq) {
                    eta_ = etaq;
                    qprime_ = q_;
                } else if (etam == etaqm1) {
                    eta_ = etaqm1;
                    qprime_ = q_ - 1;
                } else {
                    eta_ = etaqp1;
                    qprime_ = q_ + 1;
                    // keep acor for the order increase
                    for (int i = 0; i < n_; ++i) zn_[qmax_][i] = acor_[i];
                }
            }
            // step-size bounds
            if (eta_ < 1.5) {
                eta_ = 1.0;
                hprime_ = h_;
            } else {
                eta_ = std::min(eta_, etamax_);
                hprime_ = h_ * eta_;
            }
        }
        etamax_ = 10.0;
    }

    void interpolate(double t, double* y) const {
        const double s = (t - tn_) / h_;
        for (int i = 0; i < n_; ++i) {
            double v = zn_[q_][i];
            for (int j = q_ - 1; j >= 0; --j) v = zn_[j][i] + s * v;
            y[i] = v;
        }
    }
};

// Chemistry::advance (chemistry.cpp:55-76) on one conserved state u (nv values).
int chem_advance_cell(const Mech& m, double* u, double dt, double rtol, double atol, int maxsteps,
                      BdfStats& st, const char** why) {
    const int ns = m.ns;
    if (!(dt > 0) || !std::isfinite(dt)) { *why = "invalid chemistry interval"; return EINVAL_; }
    Props pr;
    if (properties(m, u, pr) != OK) { *why = "inadmissible conserved state"; return EDOMAIN_; }
    double rho = 0;
    for (int k = 0; k < ns; ++k) rho += u[k];
    const double ux = u[ns] / rho, vy = u[ns + 1] / rho;
    const double e = u[ns + 2] / rho - 0.5 * (std::pow(ux, 2) + std::pow(vy, 2));
    Cell cell{&m, rho, e, pr.T};
    std::vector<double> y(ns - 1);
    for (int k = 0; k < ns - 1; ++k) y[k] = u[k] / rho;
    Bdf<Cell> bdf(ns - 1, cell, rtol, atol, maxsteps);
    const int flag = bdf.integrate(y.data(), dt, st);
    if (flag != OK) { *why = "CVode failed"; return ECONV_; }
    for (int k = 0; k < ns - 1; ++k) {
        const double v = y[k];
        if (!std::isfinite(v) || v < -1e-13) { *why = "invalid accepted species"; return EDOMAIN_; }
        u[k] = rho * std::max(0., v);
    }
    double sum = 0;
    for (int k = 0; k < ns - 1; ++k) sum += u[k];
    u[ns - 1] = rho - sum;
    if (properties(m, u, pr) != OK) { *why = "inadmissible conserved state"; return EDOMAIN_; }
    return OK;
}

thread_local std::string g_msg;

}  // namespace

extern "C" {

void* or_mech_create(int ns, int nr, const double* W, const double* Tmid, const double* low, const double* high,
                     const int* kind, const int* rev, const double* A, const double* b, const double* EaR,
                     const double* A0, const double* b0, const double* Ea0R, const int* troe, const double* tA,
                     const double* tT3, const double* tT1, const double* tT2, const int* nu_r, const int* nu_p,
                     const double* eff) {
    if (ns > 64 || ns < 2) return nullptr;
    auto* m = new Mech;
    m->ns = ns;
    m->nr = nr;
    m->W.assign(W, W + ns);
    m->Tmid.assign(Tmid, Tmid + ns);
    m->low.assign(low, low + 7 * ns);
    m->high.assign(high, high + 7 * ns);
    m->kind.assign(kind, kind + nr);
    m->rev.assign(rev, rev + nr);
    m->troe.assign(troe, troe + nr);
    m->A.assign(A, A + nr); m->b.assign(b, b + nr); m->EaR.assign(EaR, EaR + nr);
    m->A0.assign(A0, A0 + nr); m->b0.assign(b0, b0 + nr); m->Ea0R.assign(Ea0R, Ea0R + nr);
    m->tA.assign(tA, tA + nr); m->tT3.assign(tT3, tT3 + nr); m->tT1.assign(tT1, tT1 + nr); m->tT2.assign(tT2, tT2 + nr);
    m->nu_r.assign(nu_r, nu_r + nr * ns);
    m->nu_p.assign(nu_p, nu_p + nr * ns);
    m->eff.assign(eff, eff + nr * ns);
    return m;
}

void or_mech_destroy(void* m) { delete static_cast<Mech*>(m); }

// e, cv at (T, Y).
void or_e_cv(void* mp, double T, const double* Y, double* e, double* cv) { e_cv(*static_cast<Mech*>(mp), T, Y, *e, *cv); }

// T from (e, Y); returns 0 or EDOMAIN.
int or_temperature(void* mp, double e, const double* Y, double* T) {
    double t = *T;
    if (!temperature(*static_cast<Mech*>(mp), e, Y, t)) return EDOMAIN_;
    *T = t;
    return OK;
}

// properties(u) on n states stored AoS (nv each); out = n x {p, a, T}; status per state.
void or_properties(void* mp, const double* U, long n, double* out, int* status) {
    const Mech& m = *static_cast<Mech*>(mp);
    const int nv = m.ns + 3;
    for (long c = 0; c < n; ++c) {
        Props p{0, 0, 0};
        status[c] = properties(m, U + c * nv, p);
        out[3 * c] = p.p; out[3 * c + 1] = p.a; out[3 * c + 2] = p.T;
    }
}

void or_hllc(void* mp, const double* l, const double* r, int n, const double* pl, const double* pr, double* f) {
    const Mech& m = *static_cast<Mech*>(mp);
    hllc(m.ns, l, r, n, Props{pl[0], pl[1], pl[2]}, Props{pr[0], pr[1], pr[2]}, f);
}

void or_wdot(void* mp, double T, const double* C, double* w, const double* lnA_shift) {
    rates(*static_cast<Mech*>(mp), T, C, w, lnA_shift);
}

// One Euler stage on one box (BlockAMR.cpp:210-243).
//  Uin : ghosted SoA, component plane n at Uin + n*(nx+4)*(ny+4), i fastest, 2 ghosts.
//  Uout: nx*ny per component plane, i fastest.
//  walls: bit0 x-lo, bit1 x-hi, bit2 y-lo, bit3 y-hi faces of this box are walls.
//  wedge_i: the y-lo wall holds only for faces of columns i >= wedge_i (the
//           wedge bottom, BlockAMR.cpp:225-226: x = lo + (i + 0.5) dx >= wedgeStart).
//  dims: 1 -> x sweep only (1-D), 2 -> both.
//  fx_out, fy_out (optional): the face fluxes after wall zeroing, the arrays
//           BlockAMR::euler keeps for its flux registers (BlockAMR.cpp:230):
//           fx[((j*(nx+1) + i)*nv + c)], fy[((j*nx + i)*nv + c)] (y: j <= ny).
//  fallbacks: number of face sides that fell back to first order.
// Returns 0, or EDOMAIN with *first_failed = lowest cell index whose updated state is inadmissible.
int or_flux_stage(void* mp, int nx, int ny, const double* Uin, double* Uout, double dtdx, double dtdy,
                  int first_order, int walls, int wedge_i, int dims, long* fallbacks, long* first_failed,
                  double* fx_out, double* fy_out) {
    const Mech& m = *static_cast<Mech*>(mp);
    const int ns = m.ns, nv = ns + 3;
    const int gx = nx + 4, gy = ny + 4;
    const long plane = (long)gx * gy;
    auto at = [&](int i, int j, double* q) {  // valid-index coordinates, ghosts at -2..-1 and n..n+1
        for (int c = 0; c < nv; ++c) q[c] = Uin[c * plane + (long)(j + 2) * gx + (i + 2)];
    };
    std::vector<double> fx((long)(nx + 1) * ny * nv), fy((long)nx * (ny + 1) * nv);
    long fb = 0;
    for (int d = 0; d < dims; ++d) {
        const int di = d == 0, dj = d == 1;
        const int fnx = nx + di, fny = ny + dj;
        const int nrm = d == 0 ? ns : ns + 1, tan = d == 0 ? ns + 1 : ns;
        for (int j = 0; j < fny; ++j)
            for (int i = 0; i < fnx; ++i) {
                double qa[64], qb[64], qc[64], qd[64], L[64], Rr[64], f[64];
                at(i - 2 * di, j - 2 * dj, qa);
                at(i - di, j - dj, qb);
                at(i, j, qc);
                at(i + di, j + dj, qd);
                for (int c = 0; c < nv; ++c) {
                    L[c] = qb[c] + 0.5 * minmod(qb[c] - qa[c], qc[c] - qb[c]);
                    Rr[c] = qc[c] - 0.5 * minmod(qc[c] - qb[c], qd[c] - qc[c]);
                }
                if (first_order) { std::memcpy(L, qb, nv * 8); std::memcpy(Rr, qc, nv * 8); }
                Props pl, pr;
                if (properties(m, L, pl) != OK) {
                    std::memcpy(L, qb, nv * 8);
                    ++fb;
                    if (properties(m, L, pl) != OK) return EDOMAIN_;
                }
                if (properties(m, Rr, pr) != OK) {
                    std::memcpy(Rr, qc, nv * 8);
                    ++fb;
                    if (properties(m, Rr, pr) != OK) return EDOMAIN_;
                }
                hllc(ns, L, Rr, nrm, pl, pr, f);
                const int coord = d == 0 ? i : j, len = d == 0 ? nx : ny;
                const bool lo = (walls >> (2 * d)) & 1, hi = (walls >> (2 * d + 1)) & 1;
                if ((coord == 0 && lo && (d == 0 || i >= wedge_i)) || (coord == len && hi)) {
                    for (int k = 0; k < ns; ++k) f[k] = 0;
                    f[tan] = 0;
                    f[ns + 2] = 0;
                }
                double* dst = d == 0 ? &fx[((long)j * (nx + 1) + i) * nv] : &fy[((long)j * nx + i) * nv];
                std::memcpy(dst, f, nv * 8);
            }
    }
    long bad = -1;
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) {
            double q[64];
            at(i, j, q);
            const double* f0 = &fx[((long)j * (nx + 1) + i) * nv];
            const double* f1 = &fx[((long)j * (nx + 1) + i + 1) * nv];
            for (int c = 0; c < nv; ++c) q[c] = q[c] - dtdx * (f1[c] - f0[c]);
            if (dims > 1) {
                const double* g0 = &fy[((long)j * nx + i) * nv];
                const double* g1 = &fy[((long)(j + 1) * nx + i) * nv];
                for (int c = 0; c < nv; ++c) q[c] = q[c] - dtdy * (g1[c] - g0[c]);
            }
            for (int c = 0; c < nv; ++c) Uout[c * (long)nx * ny + (long)j * nx + i] = q[c];
            Props p;
            if (bad < 0 && properties(m, q, p) != OK) bad = (long)j * nx + i;
        }
    if (fx_out) std::memcpy(fx_out, fx.data(), sizeof(double) * fx.size());
    if (fy_out && dims > 1) std::memcpy(fy_out, fy.data(), sizeof(double) * fy.size());
    if (fallbacks) *fallbacks = fb;
    if (bad >= 0) {
        if (first_failed) *first_failed = bad;
        return EDOMAIN_;
    }
    return OK;
}

// ChemistryBatch::advance over SoA planes U[n*ld + i] (n < ns+3, i < ncell).
// dt: one value, or one per cell when dt_per_cell != 0.
// stats (optional, 7 longs): steps, rhs, jac, lu, newton, conv failures, error-test failures.
// On failure the field is unchanged, the lowest failing cell is reported, and
// the message is available from or_last_error().
int or_chem_advance(void* mp, double* U, long ncell, long ld, const double* dt, int dt_per_cell, double rtol,
                    double atol, int maxsteps, int threads, long* stats, long* first_failed) {
    const Mech& m = *static_cast<Mech*>(mp);
    const int ns = m.ns, nv = ns + 3;
    if (ncell < 0 || ld < ncell || !dt) { g_msg = "invalid chemistry batch extent or interval"; return EINVAL_; }
    if (!dt_per_cell && (!(dt[0] > 0) || !std::isfinite(dt[0]))) { g_msg = "invalid chemistry batch extent or interval"; return EINVAL_; }
    if (ncell == 0) return OK;
    std::vector<double> trial((size_t)ncell * nv);
    for (long c = 0; c < ncell; ++c)
        for (int n = 0; n < nv; ++n) trial[(size_t)c * nv + n] = U[(size_t)n * ld + c];
    std::vector<int> err(ncell, 0);
    std::vector<const char*> why(ncell, nullptr);
    long agg[7] = {0, 0, 0, 0, 0, 0, 0};
    const int team = (int)std::min<long>(threads > 0 ? threads : omp_get_max_threads(), ncell);
#pragma omp parallel num_threads(team)
    {
        long loc[7] = {0, 0, 0, 0, 0, 0, 0};
#pragma omp for schedule(dynamic, 8)
        for (long c = 0; c < ncell; ++c) {
            BdfStats st;
            const double h = dt_per_cell ? dt[c] : dt[0];
            err[c] = chem_advance_cell(m, &trial[(size_t)c * nv], h, rtol, atol, maxsteps, st, &why[c]);
            loc[0] += st.steps; loc[1] += st.rhs; loc[2] += st.jac; loc[3] += st.lu;
            loc[4] += st.newton; loc[5] += st.ncf; loc[6] += st.netf;
        }
#pragma omp critical
        for (int i = 0; i < 7; ++i) agg[i] += loc[i];
    }
    if (stats) for (int i = 0; i < 7; ++i) stats[i] = agg[i];
    for (long c = 0; c < ncell; ++c)
        if (err[c]) {
            g_msg = "chemistry cell " + std::to_string(c) + ": " + (why[c] ? why[c] : "unknown");
            if (first_failed) *first_failed = c;
            return err[c];
        }
    for (long c = 0; c < ncell; ++c)
        for (int n = 0; n < nv; ++n) U[(size_t)n * ld + c] = trial[(size_t)c * nv + n];
    return OK;
}

const char* or_last_error() { return g_msg.c_str(); }

// Fixed-step backward-Euler chemistry (the integration the adjoint differentiates):
// y_{n+1} = y_n + h f(y_{n+1}; lnA), Newton solved to machine precision.
// Inputs per cell: rho, T0, Y0 (ns). e = e(T0, Y0) is held fixed.
// Objective J = T(t_end) + sum_n h * qdot(y_{n+1}), qdot = -sum_k e_k(T) W_k wdot_k  [W/m^3]
// (e_k the species internal energy per unit mass). Returns J per cell in Jout.
int or_chem_fixed_be(void* mp, long ncell, const double* rho, const double* T0, const double* Y0, double h,
                     int nsteps, const double* lnA_shift, double* Jout, double* Yend, double* Tend) {
    const Mech& m = *static_cast<Mech*>(mp);
    const int ns = m.ns, N = ns - 1;
    int fail = 0;
#pragma omp parallel for schedule(dynamic, 8) reduction(| : fail)
    for (long c = 0; c < ncell; ++c) {
        double e, cv;
        e_cv(m, T0[c], &Y0[c * ns], e, cv);
        Cell cell{&m, rho[c], e, T0[c], lnA_shift};
        std::vector<double> y(N), yn(N), f(N), f2(N), Jm(N * N), r(N);
        std::vector<int> piv(N);
        for (int k = 0; k < N; ++k) y[k] = Y0[c * ns + k];
        double J = 0;
        bool ok = true;
        for (int s = 0; s < nsteps && ok; ++s) {
            yn = y;
            // Newton with a finite-difference Jacobian, iterated to the
            // roundoff floor: stop when every correction is below 1e-14 of its
            // component (floor 1e-30), or when the largest relative correction
            // stops halving once below 1e-8 (roundoff reached).
            bool conv = false;
            double prev = HUGE_VAL;
            for (int it = 0; it < 60; ++it) {
                if (cell.rhs(yn.data(), f.data())) { ok = false; break; }
                for (int k = 0; k < N; ++k) r[k] = -(yn[k] - y[k] - h * f[k]);
                for (int j = 0; j < N; ++j) {
                    const double inc = 1e-8 * std::max(std::abs(yn[j]), 1e-6);
                    const double ys = yn[j];
                    yn[j] = ys + inc;
                    if (cell.rhs(yn.data(), f2.data())) { ok = false; break; }
                    yn[j] = ys;
                    for (int i = 0; i < N; ++i) Jm[i * N + j] = (i == j ? 1.0 : 0.0) - h * (f2[i] - f[i]) / inc;
                }
                if (!ok) break;
                if (!lu_factor(N, Jm.data(), piv.data())) { ok = false; break; }
                lu_solve(N, Jm.data(), piv.data(), r.data());
                double wn = 0;
                for (int k = 0; k < N; ++k) {
                    yn[k] += r[k];
                    wn = std::max(wn, std::abs(r[k]) / (std::abs(yn[k]) + 1e-30));
                }
                if (!std::isfinite(wn)) { ok = false; break; }
                if (wn <= 1e-14 || (it >= 2 && wn < 1e-8 && wn >= 0.5 * prev)) { conv = true; break; }
                prev = wn;
            }
            if (ok && !conv) ok = false;
            if (!ok) break;
            y = yn;
            // heat release at the new state
            double Y[64], C[64], w[64], cp[64], hh[64];
            double sum = 0;
            for (int k = 0; k < N; ++k) { Y[k] = y[k]; sum += y[k]; }
            Y[N] = 1 - sum;
            double T = cell.T;
            if (!temperature(m, e, Y, T)) { ok = false; break; }
            cell.T = T;
            for (int k = 0; k < ns; ++k) C[k] = rho[c] * Y[k] / m.W[k];
            rates(m, T, C, w, lnA_shift);
            Thermo::nasa(m, T, cp, hh, nullptr);
            double qdot = 0;
            for (int k = 0; k < ns; ++k) qdot -= (hh[k] - 1.0) * R * T * w[k];
            J += h * qdot;
        }
        if (!ok) { fail |= 1; Jout[c] = std::numeric_limits<double>::quiet_NaN(); continue; }
        double Y[64], sum = 0;
        for (int k = 0; k < N; ++k) { Y[k] = y[k]; sum += y[k]; }
        Y[N] = 1 - sum;
        double T = cell.T;
        temperature(m, e, Y, T);
        Jout[c] = T + J;
        if (Tend) Tend[c] = T;
        if (Yend) for (int k = 0; k < ns; ++k) Yend[c * ns + k] = Y[k];
    }
    return fail ? ECONV_ : OK;
}

// Plain scalar ODE driver for integrator checks: y' = A y (dense n x n, row-major).
int or_bdf_linear(int n, const double* A, double* y, double tout, double rtol, double atol, int maxsteps, long* stats) {
    struct Lin {
        int n; const double* A;
        int rhs(const double* y, double* f) {
            for (int i = 0; i < n; ++i) { double s = 0; for (int j = 0; j < n; ++j) s += A[i * n + j] * y[j]; f[i] = s; }
            return 0;
        }
    } sys{n, A};
    BdfStats st;
    Bdf<Lin> bdf(n, sys, rtol, atol, maxsteps);
    const int r = bdf.integrate(y, tout, st);
    if (stats) { stats[0] = st.steps; stats[1] = st.rhs; stats[2] = st.jac; stats[3] = st.lu; stats[4] = st.newton; stats[5] = st.ncf; stats[6] = st.netf; }
    return r;
}

// Robertson problem (Hairer & Wanner) for integrator checks.
int or_bdf_robertson(double* y, double tout, double rtol, double atol, int maxsteps, long* stats) {
    struct Rob {
        int rhs(const double* y, double* f) {
            f[0] = -0.04 * y[0] + 1e4 * y[1] * y[2];
            f[2] = 3e7 * y[1] * y[1];
            f[1] = -f[0] - f[2];
            return 0;
        }
    } sys;
    BdfStats st;
    Bdf<Rob> bdf(3, sys, rtol, atol, maxsteps);
    const int r = bdf.integrate(y, tout, st);
    if (stats) { stats[0] = st.steps; stats[1] = st.rhs; stats[2] = st.jac; stats[3] = st.lu; stats[4] = st.newton; stats[5] = st.ncf; stats[6] = st.netf; }
    return r;
}

}  // extern "C"
