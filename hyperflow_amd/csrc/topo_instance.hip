// Kernels and launchers for one mechanism topology. Built once per mechanism
// with -DHF_TOPO=<struct name> -DHF_TOPO_INC="<generated header>"; the
// generated header fixes the topology (see mechgen.cpp), the coefficients are
// uploaded at run time.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>

#include "hf_adjoint.hpp"
#include "hf_chem.hpp"
#include "hf_wave.hpp"
#include "hf_flux.hpp"
#include "topo_ops.hpp"

#define HF_STR2(x) #x
#define HF_STR(x) HF_STR2(x)
#include HF_STR(HF_TOPO_INC)

namespace hf {
namespace {

using Topo = hf_gen::HF_TOPO;
using MD = MechData<Topo>;

void fill(const Mechanism& m, void* host) {
    MD& d = *static_cast<MD*>(host);
    std::memset(&d, 0, sizeof(MD));
    const int ns = Topo::NS, nr = Topo::NR;
    d.common_tmid = 1;
    d.tmid = m.Tmid[0];
    for (int k = 0; k < ns; ++k) {
        d.W[k] = m.W[k];
        d.invW[k] = 1.0 / m.W[k];
        d.Tmid[k] = m.Tmid[k];
        for (int i = 0; i < 7; ++i) {
            d.lo[k][i] = m.low[k][i];
            d.hi[k][i] = m.high[k][i];
        }
        if (m.Tmid[k] != d.tmid) d.common_tmid = 0;
    }
    for (int i = 0; i < nr; ++i) {
        const Reaction& r = m.reactions[i];
        d.lnA[i] = std::log(r.A);
        d.b[i] = r.b;
        d.EaR[i] = r.EaR;
        if (r.kind == Reaction::FALLOFF) {
            d.lnA0[i] = std::log(r.A0);
            d.b0[i] = r.b0;
            d.Ea0R[i] = r.Ea0R;
        }
        d.troeA[i] = r.troeA;
        d.troeRT3[i] = std::abs(r.troeT3) < 1e-300 ? INFINITY : 1.0 / r.troeT3;
        d.troeRT1[i] = std::abs(r.troeT1) < 1e-300 ? INFINITY : 1.0 / r.troeT1;
        d.troeT2[i] = r.troeT2;
        d.defeff[i] = r.default_efficiency;
        for (size_t j = 0; j < r.efficiencies.size(); ++j)
            d.effd[i][j] = r.efficiencies[j].second - r.default_efficiency;
    }
    // runtime topology (the same tables the kernels unroll over at compile time)
    constexpr int NEFF = Topo::NEFF > 0 ? Topo::NEFF : 1;
    for (int i = 0; i < nr; ++i) {
        d.rk_kind[i] = (int8_t)Topo::kind[i];
        d.rk_troe[i] = (int8_t)Topo::troe[i];
        d.rk_rev[i] = (int8_t)Topo::rev[i];
        for (int j = 0; j < Topo::NRMAX; ++j) {
            d.rk_reac[i][j] = (int16_t)Topo::reac[i][j];
            d.rk_prod[i][j] = (int16_t)Topo::prod[i][j];
        }
        for (int j = 0; j < NEFF; ++j) d.rk_eff[i][j] = Topo::NEFF > 0 ? (int16_t)Topo::effsp[i][j] : (int16_t)-1;
    }
    {
        int s = 0;
        for (int kind = 2; kind >= 0; --kind)
            for (int i = 0; i < nr; ++i)
                if (Topo::kind[i] == kind) d.rate_slot[s++] = (int16_t)i;
    }
    int n = 0;
    for (int k = 0; k < ns; ++k) {
        d.inc_start[k] = (int16_t)n;
        for (int i = 0; i < nr; ++i) {
            int nu = 0;
            for (int j = 0; j < Topo::NRMAX; ++j) {
                if (Topo::reac[i][j] == k) --nu;
                if (Topo::prod[i][j] == k) ++nu;
            }
            if (nu != 0) {
                d.inc_rxn[n] = (int16_t)i;
                d.inc_nu[n] = (int8_t)nu;
                ++n;
            }
        }
    }
    d.inc_start[ns] = (int16_t)n;
    // Jacobian columns: for each slot m of reaction i holding species s,
    // d q_i / d C_s gets that slot's derivative, and every slot m2 of the
    // reaction passes it to its species with the slot's sign (reactants -1,
    // products +1; a species in two slots gets two entries). A listed
    // third-body efficiency of s adds its correction times d q_i / d[M].
    {
        constexpr int R = Topo::NRMAX;
        auto slot = [](int i, int m) { return m < R ? Topo::reac[i][m] : Topo::prod[i][m - R]; };
        int q = 0;
        for (int sp = 0; sp < ns; ++sp) {
            d.jc_start[sp] = q;
            for (int i = 0; i < nr; ++i) {
                for (int m = 0; m < 2 * R; ++m) {
                    if (slot(i, m) != sp) continue;
                    for (int m2 = 0; m2 < 2 * R; ++m2) {
                        const int k = slot(i, m2);
                        if (k < 0 || k >= ns - 1) continue;
                        d.jc_rxn[q] = (int16_t)i;
                        d.jc_m[q] = (int8_t)m;
                        d.jc_k[q] = (int8_t)k;
                        d.jc_coef[q] = m2 < R ? -1.0 : 1.0;
                        ++q;
                    }
                }
                if (Topo::kind[i] != 0) {
                    for (int j = 0; j < NEFF; ++j) {
                        if (Topo::NEFF == 0 || Topo::effsp[i][j] != sp) continue;
                        for (int m2 = 0; m2 < 2 * R; ++m2) {
                            const int k = slot(i, m2);
                            if (k < 0 || k >= ns - 1) continue;
                            d.jc_rxn[q] = (int16_t)i;
                            d.jc_m[q] = (int8_t)(2 * R);
                            d.jc_k[q] = (int8_t)k;
                            d.jc_coef[q] = (m2 < R ? -1.0 : 1.0) * d.effd[i][j];
                            ++q;
                        }
                    }
                }
            }
        }
        // q <= NJC: a reaction's slots hold at most 2 NRMAX species
        // occurrences, each passing to at most 2 NRMAX rows, and each listed
        // efficiency to at most 2 NRMAX rows
        d.jc_start[ns] = q;
    }
}

// ------------------------------------------------------------------ chemistry
// Fill-kernel reset of the failure index and counters (graph-capturable).
__global__ void chem_zero_kernel(unsigned long long* first_failed, unsigned long long* counters) {
    const int t = threadIdx.x;
    if (t == 0) *first_failed = ~0ull;
    if (counters && t < 8) counters[t] = 0;
}

// ------------------------------------------------------------------ chemistry (batched)
// Persistent form: cells are binned by an a-priori work estimate and queued
// in decreasing order of work (longest first); a fixed grid of waves drains
// the queue, and each lane claims a new cell as soon as its current one
// finishes, so a wave never idles behind its slowest lane.
constexpr int kMaxBins = 256;

struct Workspace {
    uint8_t* bins;
    int32_t* order;
    unsigned* hist;
    unsigned* offs;
    unsigned* cursor;
    unsigned* head;
    void* jsave;          // saved Jacobians: one N x N x 64 tile per wave, [element / SV][lane][SV]
                          // (SV elements of a lane in 16 bytes; Bdf::SI)
    void* lu;             // factored Newton matrices, same layout
    void* zs;             // Nordsieck slabs, [element][64 lanes] per wave, when they do not fit in LDS
};

// The per-lane BDF slab (Nordsieck history, step-size history, BDF
// coefficients, saved correction) of a wave: (7N + 13) doubles x 64 lanes.
// It lives in LDS while that allows four waves per CU (35 KB per wave for
// H2), otherwise in a per-wave tile of the workspace (140 KB per wave for
// DME, N = 38).
constexpr int kSlab = Bdf<Topo::NS - 1, ChemSys<Topo>>::SLAB;
constexpr bool kSlabLds = kSlab * 64 * 8 <= 40 * 1024;

inline size_t align256(size_t x) { return (x + 255) & ~size_t(255); }

// Lanes that hold a Jacobian slot: every cell's lane (rounded up to whole
// waves) for the plain form, every resident lane of the persistent grid for
// the batched one; the larger of the two, so one workspace serves both.
int64_t slot_lanes(int64_t ncell, int num_cus, int waves_per_cu) {
    return std::max<int64_t>((ncell + 63) / 64 * 64, (int64_t)num_cus * (waves_per_cu > 0 ? waves_per_cu : 4) * 64);
}

Workspace carve(void* base, int64_t ncell, int64_t lanes) {
    constexpr int N = Topo::NS - 1;
    char* p = static_cast<char*>(base);
    Workspace w;
    w.bins = reinterpret_cast<uint8_t*>(p);
    p += align256(ncell);
    w.order = reinterpret_cast<int32_t*>(p);
    p += align256(sizeof(int32_t) * ncell);
    w.hist = reinterpret_cast<unsigned*>(p);
    p += align256(sizeof(unsigned) * kMaxBins);
    w.offs = reinterpret_cast<unsigned*>(p);
    p += align256(sizeof(unsigned) * kMaxBins);
    w.cursor = reinterpret_cast<unsigned*>(p);
    p += align256(sizeof(unsigned) * kMaxBins);
    w.head = reinterpret_cast<unsigned*>(p);
    p += 256;
    w.jsave = p;
    p += align256(sizeof(double) * N * N * lanes);
    w.lu = p;
    p += align256(sizeof(double) * N * N * lanes);
    w.zs = kSlabLds ? nullptr : p;
    return w;
}

// Queue arrays plus a saved Jacobian and an LU-factor slot (FP64-sized) per
// lane, and the BDF slabs when they live in global memory.
size_t chem_workspace(int64_t ncell, int, int num_cus, int waves_per_cu) {
    constexpr int N = Topo::NS - 1;
    const int64_t lanes = slot_lanes(ncell, num_cus, waves_per_cu);
    return align256(ncell) + align256(sizeof(int32_t) * ncell) + 3 * align256(sizeof(unsigned) * kMaxBins) + 256 +
           2 * align256(sizeof(double) * N * N * lanes) + (kSlabLds ? 0 : align256(sizeof(double) * kSlab * lanes));
}

__global__ void chem_init_kernel(Workspace w, unsigned long long* first_failed, unsigned long long* counters) {
    const int t = threadIdx.x;
    if (t < kMaxBins) {
        w.hist[t] = 0;
        w.cursor[t] = 0;
    }
    if (t == 0) {
        *w.head = 0;
        *first_failed = ~0ull;
    }
    if (t < 8) counters[t] = 0;
}

// Work estimate: dt times the weighted RMS of the initial rate of change,
// binned on a log10 scale over [1e-2, 1e10].
template <class TT>
__global__ void __launch_bounds__(256)
chem_estimate_kernel(const MechData<TT>* __restrict__ dm, CellMap cm, long ncell, const double* __restrict__ dt,
                     int per_cell, ChemParams p, int nbins, Workspace w) {
    constexpr int NS = TT::NS, NV = NS + 3, N = NS - 1;
    __shared__ unsigned lh[kMaxBins];
    for (int b = threadIdx.x; b < nbins; b += blockDim.x) lh[b] = 0;
    __syncthreads();
    const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    const double* q = nullptr;
    int64_t plane = 0;
    if (i < ncell && !cm.locate(i, q, plane)) {
        w.bins[i] = 0;                             // not in the batch: queued last, skipped at once
        atomicAdd(&lh[0], 1u);
    } else if (i < ncell) {
        const MechData<TT>& d = *dm;
        double u[NV];
        #pragma unroll
        for (int n = 0; n < NV; ++n) u[n] = q[n * plane];
        const double h = per_cell ? dt[i] : dt[0];
        int bin = nbins - 1;   // unusable cells fail fast: queue them first
        Props pr;
        if (properties(d, u, pr, 1000.0) && h > 0) {
            double rho = 0;
            #pragma unroll
            for (int k = 0; k < NS; ++k) rho += u[k];
            const double ux = u[NS] / rho, vy = u[NS + 1] / rho;
            ChemSys<TT> sys{d, rho, u[NS + 2] / rho - 0.5 * (ux * ux + vy * vy), pr.T, nullptr};
            double y[N], f[N];
            #pragma unroll
            for (int k = 0; k < N; ++k) y[k] = u[k] / rho;
            if (sys.rhs(y, f)) {
                double s = 0;
                #pragma unroll
                for (int k = 0; k < N; ++k) {
                    const double x = f[k] / fma(p.rtol, fabs(y[k]), p.atol);
                    s = fma(x, x, s);
                }
                const double key = h * sqrt(s / N);
                const double lg = log10(fmax(key, 1e-30));
                bin = (int)floor((lg + 2.0) / 12.0 * nbins);
                bin = min(max(bin, 0), nbins - 1);
            }
        }
        w.bins[i] = (uint8_t)bin;
        atomicAdd(&lh[bin], 1u);
    }
    __syncthreads();
    for (int b = threadIdx.x; b < nbins; b += blockDim.x)
        if (lh[b]) atomicAdd(&w.hist[b], lh[b]);
}

// Queue order: highest bin (most work) first.
__global__ void chem_scan_kernel(Workspace w, int nbins) {
    if (threadIdx.x == 0) {
        unsigned acc = 0;
        for (int b = nbins - 1; b >= 0; --b) {
            w.offs[b] = acc;
            acc += w.hist[b];
        }
    }
}

__global__ void __launch_bounds__(256) chem_scatter_kernel(long ncell, int nbins, Workspace w) {
    __shared__ unsigned lh[kMaxBins], lbase[kMaxBins];
    for (int b = threadIdx.x; b < nbins; b += blockDim.x) lh[b] = 0;
    __syncthreads();
    const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    int bin = 0;
    unsigned local = 0;
    if (i < ncell) {
        bin = w.bins[i];
        local = atomicAdd(&lh[bin], 1u);
    }
    __syncthreads();
    for (int b = threadIdx.x; b < nbins; b += blockDim.x)
        if (lh[b]) lbase[b] = atomicAdd(&w.cursor[b], lh[b]);
    __syncthreads();
    if (i < ncell) w.order[w.offs[bin] + lbase[bin] + local] = (int32_t)i;
}

// One lane integrates one cell at a time. PLAIN: lane l of block b takes cell
// 64 b + l and stops when it is done (one cell per lane, the drop-in form).
// Otherwise the block is a persistent wave draining the binned queue.
template <class TT, class MT, bool PLAIN>
__global__ void __launch_bounds__(64)
chem_lane_kernel(const MechData<TT>* __restrict__ dm, CellMap cm, long ncell,
                 const double* __restrict__ dt, int per_cell, double* __restrict__ out, long ldo,
                 int8_t* __restrict__ status, unsigned long long* __restrict__ first_failed, ChemParams p,
                 unsigned long long* __restrict__ counters, Workspace w, int refill, int jwait, int swait) {
    constexpr int N = TT::NS - 1;
    double* zbase;
    if constexpr (kSlabLds) {
        __shared__ double zlds[kSlab * 64];
        zbase = zlds;
    } else {
        zbase = static_cast<double*>(w.zs) + (long)blockIdx.x * kSlab * 64;
    }
    const int lane = threadIdx.x;
    const long tile = (long)blockIdx.x * N * N * 64;
    constexpr int sv = 16 / sizeof(MT);   // the slot's vector width (Bdf::SV)
    CellRun<TT, MT, kSlabLds ? 3 : 1> run(*dm, zbase + lane, p, static_cast<MT*>(w.jsave) + tile + lane * sv,
                                        static_cast<MT*>(w.lu) + tile + lane * sv);
    BdfCounters tot;
    int max_steps_cell = 0;
    long cell = -1;
    bool active = false;
    // Wave-uniform: the global queue is empty. Idle lanes are refilled
    // together once at least `refill` of them wait (or when the whole wave is
    // idle), so the cost of starting a cell (initial rate and step-size
    // estimate) is paid by many lanes in the same trip, not one at a time.
    bool exhausted = PLAIN;
    // Setup grouping (wave-uniform): a lane whose next trip evaluates a
    // Jacobian waits until a quarter of the active lanes want one too, every
    // active lane does, or jwait trips have passed since the last trip that
    // evaluated Jacobians; a lane whose next trip only re-forms and factors
    // its Newton matrix does the same with swait (and goes along with any
    // Jacobian trip). Waiting shifts a lane's trips in time and changes none
    // of its arithmetic, so each cell's result is independent of its
    // wave-mates. A wait of 0 never holds a lane back.
    int since_jac = jwait, since_setup = swait;
    if constexpr (PLAIN) {
        cell = (long)blockIdx.x * 64 + lane;
        const double* q;
        int64_t plane;
        if (cell < ncell && !cm.locate(cell, q, plane)) {
            status[cell] = HF_OK;                   // not a leaf: no part of the batch
        } else if (cell < ncell) {
            const int st = run.start(q, plane, per_cell ? dt[cell] : dt[0], p.maxsteps);
            if (st == HF_OK) {
                active = true;
            } else {
                status[cell] = (int8_t)st;
                atomicMin(first_failed, (unsigned long long)cell);
            }
        }
    }
    for (;;) {
        const unsigned long long want = __ballot(!active);
        const int idle = __popcll(want);
        if (!exhausted && idle > 0 && (idle >= refill || idle == 64)) {
            const int leader = __ffsll((long long)want) - 1;
            unsigned base = 0;
            if (lane == leader) base = atomicAdd(w.head, (unsigned)idle);
            base = __shfl(base, leader, 64);
            if (base + (unsigned)idle >= (unsigned)ncell) exhausted = true;
            if (!active) {
                const unsigned slot = base + (unsigned)__popcll(want & ((1ull << lane) - 1));
                if (slot < (unsigned)ncell) {
                    cell = w.order[slot];
                    const double h = per_cell ? dt[cell] : dt[0];
                    const double* q;
                    int64_t plane;
                    const int st = cm.locate(cell, q, plane) ? run.start(q, plane, h, p.maxsteps) : -1;
                    if (st == -1) {
                        status[cell] = HF_OK;       // not a leaf: no part of the batch
                    } else if (st == HF_OK) {
                        active = true;
                    } else {
                        status[cell] = (int8_t)st;
                        atomicMin(first_failed, (unsigned long long)cell);
                    }
                }
            }
        }
        if (!__any(active)) {
            if (exhausted) break;
            continue;
        }
        bool go = active;
        if (!PLAIN && (jwait > 0 || swait > 0)) {
            bool wj = false, ws = false;
            if (active) {
                run.bdf.pre();
                ws = run.bdf.wants_setup();
                wj = ws && run.bdf.wants_jac();
            }
            const unsigned long long wmask = __ballot(wj), smask = __ballot(ws), amask = __ballot(active);
            const int na = __popcll(amask);
            bool jtrip = true;
            if (jwait > 0 && wmask) {
                jtrip = since_jac >= jwait || 4 * __popcll(wmask) >= na || wmask == amask;
                if (!jtrip) go = go && !wj;
            }
            if (wmask && jtrip) since_jac = 0;
            if (swait > 0 && smask) {
                const bool strip = (wmask && jtrip) || since_setup >= swait || 4 * __popcll(smask) >= na ||
                                   smask == amask;
                if (strip) since_setup = 0;
                else go = go && !ws;
            }
            ++since_jac;
            ++since_setup;
        }
        if (go) {
            const int r = run.step();
            if (r != BDF_RUNNING) {
                const double* q;
                int64_t plane;
                cm.locate(cell, q, plane);
                int st = r == BDF_OK ? run.finish(q, plane, out, cell, ldo) : HF_ECONV;
                status[cell] = (int8_t)st;
                if (st != HF_OK) atomicMin(first_failed, (unsigned long long)cell);
                tot.steps += run.bdf.c.steps;
                tot.rhs += run.bdf.c.rhs;
                tot.jac += run.bdf.c.jac;
                tot.lu += run.bdf.c.lu;
                tot.newton += run.bdf.c.newton;
                tot.ncf += run.bdf.c.ncf;
                tot.netf += run.bdf.c.netf;
                max_steps_cell = max(max_steps_cell, run.bdf.c.steps);
                active = false;
            }
        }
    }
    reduce_counters(tot, max_steps_cell, counters);
}

// Wave-per-cell batched form for large mechanisms: each wave of the
// persistent grid takes the next cell of the binned queue and integrates it
// with all 64 lanes (hf_wave.hpp); the post-step checks and the write of
// the species planes follow Chemistry::advance (chemistry.cpp:68-75).
constexpr bool kWaveChem = Topo::NS > 16;

// Two waves per SIMD (at most 256 registers per lane; about 360 B of scratch
// spills): measured 1.6x faster than one wave per SIMD at 332 registers, and
// faster than three waves per SIMD, where the spills grow.
template <class TT, class MT>
__global__ void __launch_bounds__(64, 2)
chem_wave_kernel(const MechData<TT>* __restrict__ dm, CellMap cm, long ncell,
                 const double* __restrict__ dt, int per_cell, double* __restrict__ out, long ldo,
                 int8_t* __restrict__ status, unsigned long long* __restrict__ first_failed, ChemParams p,
                 unsigned long long* __restrict__ counters, Workspace w) {
    constexpr int NS = TT::NS, N = NS - 1, NR = TT::NR;
    __shared__ double scratch[3 * NS + NR];
    __shared__ double coef[13];
    __shared__ __attribute__((aligned(16))) MT lu[NS * 64];  // 16-byte aligned for the blocked LU
    __shared__ int inc[2 * NR * TT::NRMAX];
    const int lane = threadIdx.x;
    // this wave's Jacobian slot: N x 64 entries inside the per-lane Jacobian
    // region, which holds N x N entries for each of at least 64 x (the waves
    // per CU of the launch, or 4) lanes per CU: room for every wave of the grid
    MT* jslot = static_cast<MT*>(w.jsave) + (long)blockIdx.x * N * 64;
    // this wave's reaction-derivative scratch for the analytic Jacobian, in
    // the LU-factor region (the wave form factors in LDS), which holds N x N
    // doubles per lane: (2 NRMAX + 2) x NR doubles per wave fit many times
    double* jd = static_cast<double*>(w.lu) + (long)blockIdx.x * (2 * TT::NRMAX + 2) * NR;
    WaveBdf<TT, MT> bdf(*dm, scratch, inc, jslot, lu, jd, coef, p.rtol, p.atol);
    WaveGas<TT>& gas = bdf.g;
    unsigned long long tot[7] = {0, 0, 0, 0, 0, 0, 0};
    int max_steps_cell = 0;
    for (;;) {
        unsigned slot = 0;
        if (lane == 0) slot = atomicAdd(w.head, 1u);
        slot = (unsigned)__builtin_amdgcn_readfirstlane((int)slot);
        if (slot >= (unsigned)ncell) break;
        const long cell = w.order[slot];
        const double* q;
        int64_t plane;
        if (!cm.locate(cell, q, plane)) {
            if (lane == 0) status[cell] = HF_OK;      // not a leaf: no part of the batch
            continue;
        }
        const double rhoY = lane < NS ? q[lane * plane] : 0.0;
        const double mx = q[NS * plane], my = q[(NS + 1) * plane], E = q[(NS + 2) * plane];
        const double h = per_cell ? dt[cell] : dt[0];
        int st = HF_OK;
        double tt = 1000.0;
        if (!(h > 0) || !finite(h)) {
            st = HF_EINVAL;
        } else if (!gas.properties(rhoY, mx, my, E, tt)) {
            st = HF_EDOMAIN;
        } else {
            const double rho = wave_sum(lane < NS ? rhoY : 0.0);
            const double ux = mx / rho, vy = my / rho;
            gas.rho = rho;
            gas.e = E / rho - 0.5 * (ux * ux + vy * vy);
            gas.t = tt;
            double y = lane < N ? rhoY / rho : 0.0;
            if (!bdf.integrate(y, h, p.maxsteps)) {
                st = HF_ECONV;
            } else if (__any(lane < N && !(finite(y) && y >= -1e-13))) {
                st = HF_EDOMAIN;
            } else {
                double r = lane < N ? rho * fmax(0.0, y) : 0.0;
                const double sum = wave_sum(r);
                if (lane == N) r = rho - sum;
                double t2 = gas.t;
                if (!gas.properties(r, mx, my, E, t2)) st = HF_EDOMAIN;
                else if (lane < NS) out[lane * ldo + cell] = r;
            }
            const WaveCounters& c = bdf.c;
            tot[0] += c.steps;
            tot[1] += c.rhs;
            tot[2] += c.jac;
            tot[3] += c.lu;
            tot[4] += c.newton;
            tot[5] += c.ncf;
            tot[6] += c.netf;
            max_steps_cell = max(max_steps_cell, c.steps);
        }
        if (lane == 0) {
            status[cell] = (int8_t)st;
            if (st != HF_OK) atomicMin(first_failed, (unsigned long long)cell);
        }
    }
    if (lane == 0) {
        #pragma unroll
        for (int k = 0; k < 7; ++k)
            if (tot[k]) atomicAdd(&counters[k], tot[k]);
        atomicMax(&counters[7], (unsigned long long)max_steps_cell);
    }
}

// All-or-nothing commit on the device: the advanced species planes replace
// the caller's only when no cell failed (chemistry_batch.cpp:40-49).
__global__ void __launch_bounds__(256) chem_commit_kernel(const double* __restrict__ out, long ldo, CellMap cm,
                                                          long ncell,
                                                          const unsigned long long* __restrict__ first_failed) {
    if (*first_failed != ~0ull) return;
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const double* q;
    int64_t plane;
    if (i >= ncell || !cm.locate(i, q, plane)) return;
    double* u = const_cast<double*>(q);
    #pragma unroll
    for (int k = 0; k < Topo::NS; ++k) u[k * plane] = out[k * ldo + i];
}

// The batch's cells: the block table, or the flat planes at U (stride ld).
CellMap cell_map(const ChemLaunch& L, const double* U, int64_t ld) {
    return CellMap{U, ld, L.blocks, L.block_start, L.nblock};
}

int chem_commit(const ChemLaunch& L) {
    if (!L.commit) return HF_OK;
    hipLaunchKernelGGL(chem_commit_kernel, dim3((unsigned)((L.ncell + 255) / 256)), dim3(256), 0, L.stream, L.out,
                       (long)L.ldo, cell_map(L, L.commit, L.commit_ld), (long)L.ncell, L.first_failed);
    return hipGetLastError() == hipSuccess ? HF_OK : HF_EDEVICE;
}

int chem(const ChemLaunch& L) {
    if (L.ncell == 0) return HF_OK;
    {
        const ChemParams p{L.rtol, L.atol, L.maxsteps};
        const MD* dm = static_cast<const MD*>(L.mech);
        const CellMap cm = cell_map(L, L.U, L.ld);
        const int64_t lanes = slot_lanes(L.ncell, L.num_cus, L.waves_per_cu);
        if (!L.workspace || L.workspace_bytes < chem_workspace(L.ncell, L.nbins, L.num_cus, L.waves_per_cu))
            return HF_EINVAL;
        if (L.mode == HF_CHEM_BATCHED) {
            const Workspace w = carve(L.workspace, L.ncell, lanes);
            const int nbins = L.nbins > 0 ? min(L.nbins, kMaxBins) : 32;
            const int wpc = L.waves_per_cu > 0 ? L.waves_per_cu : 4;
            // queue chunking: idle lanes refilled together
            const int chunk = L.chunk > 0 ? min(L.chunk, 64) : 8;
            const int jwait = L.jac_wait > 0 ? L.jac_wait : (L.jac_wait == 0 ? 16 : 0);
            const int swait = L.setup_wait > 0 ? L.setup_wait : (L.setup_wait == 0 ? 2 : 0);
            const unsigned g256 = (unsigned)((L.ncell + 255) / 256);
            hipLaunchKernelGGL(chem_init_kernel, dim3(1), dim3(256), 0, L.stream, w, L.first_failed, L.counters);
            hipLaunchKernelGGL(chem_estimate_kernel<Topo>, dim3(g256), dim3(256), 0, L.stream, dm, cm,
                               (long)L.ncell, L.dt, L.dt_per_cell, p, nbins, w);
            hipLaunchKernelGGL(chem_scan_kernel, dim3(1), dim3(64), 0, L.stream, w, nbins);
            hipLaunchKernelGGL(chem_scatter_kernel, dim3(g256), dim3(256), 0, L.stream, (long)L.ncell, nbins, w);
            StreamTimer timer(L.kernel_ms, L.stream);
            if constexpr (kWaveChem) {
                // one cell per wave: more resident waves per CU (LDS ~12-22 KB per wave)
                const int wwpc = L.waves_per_cu > 0 ? L.waves_per_cu : 8;
                if (L.precision == HF_PREC_MIXED)
                    hipLaunchKernelGGL((chem_wave_kernel<Topo, float>), dim3((unsigned)(L.num_cus * wwpc)), dim3(64), 0,
                                       L.stream, dm, cm, (long)L.ncell, L.dt, L.dt_per_cell, L.out,
                                       (long)L.ldo, L.status, L.first_failed, p, L.counters, w);
                else
                    hipLaunchKernelGGL((chem_wave_kernel<Topo, double>), dim3((unsigned)(L.num_cus * wwpc)), dim3(64), 0,
                                       L.stream, dm, cm, (long)L.ncell, L.dt, L.dt_per_cell, L.out,
                                       (long)L.ldo, L.status, L.first_failed, p, L.counters, w);
                if (hipGetLastError() != hipSuccess) return HF_EDEVICE;
                timer.stop();
                return chem_commit(L);
            } else {
                if (L.precision == HF_PREC_MIXED)
                    hipLaunchKernelGGL((chem_lane_kernel<Topo, float, false>), dim3((unsigned)(L.num_cus * wpc)),
                                       dim3(64), 0, L.stream, dm, cm, (long)L.ncell, L.dt, L.dt_per_cell,
                                       L.out, (long)L.ldo, L.status, L.first_failed, p, L.counters, w, chunk, jwait, swait);
                else
                    hipLaunchKernelGGL((chem_lane_kernel<Topo, double, false>), dim3((unsigned)(L.num_cus * wpc)),
                                       dim3(64), 0, L.stream, dm, cm, (long)L.ncell, L.dt, L.dt_per_cell,
                                       L.out, (long)L.ldo, L.status, L.first_failed, p, L.counters, w, chunk, jwait, swait);
                if (hipGetLastError() != hipSuccess) return HF_EDEVICE;
                timer.stop();
                return chem_commit(L);
            }
        }
        // plain: FP64 throughout, one cell per lane, cells in index order
        const Workspace none = carve(L.workspace, L.ncell, lanes);
        hipLaunchKernelGGL(chem_zero_kernel, dim3(1), dim3(64), 0, L.stream, L.first_failed, L.counters);
        StreamTimer timer(L.kernel_ms, L.stream);
        const unsigned grid = (unsigned)((L.ncell + 63) / 64);
        hipLaunchKernelGGL((chem_lane_kernel<Topo, double, true>), dim3(grid), dim3(64), 0, L.stream, dm, cm,
                           (long)L.ncell, L.dt, L.dt_per_cell, L.out, (long)L.ldo, L.status,
                           L.first_failed, p, L.counters, none, 64, 0, 0);
        if (hipGetLastError() != hipSuccess) return HF_EDEVICE;
        timer.stop();
        return chem_commit(L);
    }
}

// ------------------------------------------------------------------ flux stage
// The tile is staged in LDS: (TX+4)(TY+4) cells of NV doubles and their warm
// start temperatures, plus the face fluxes that cross waves. 15 x 15 (one
// 256-lane workgroup) for small mechanisms; 7 x 7 (one wave) for large ones;
// 1-D rows of 255 or 63 cells.
constexpr bool kLargeState = Topo::NS > 16;
constexpr int TX2 = kLargeState ? 7 : 15, TY2 = TX2, TX1 = kLargeState ? 63 : 255;
template <int TX, int TY, bool TWO_D>
constexpr size_t flux_lds_bytes() {
    using FT = FluxTile<Topo, TX, TY, TWO_D>;
    return sizeof(double) * (FT::LDS_IN + FT::LDS_FLUX + FT::GX * FT::GY);
}
static_assert(flux_lds_bytes<TX2, TY2, true>() <= 65536 && flux_lds_bytes<TX1, 1, false>() <= 65536,
              "flux tile exceeds 64 KB of LDS");

// Other 2-D tile shapes of 256 lanes for the tuner (small mechanisms only):
// 31 x 7 (wide rows) and 7 x 31 (tall columns); every shape computes each
// face and cell update from the same staged values, so the results do not
// depend on the shape.
constexpr int NSHAPE = kLargeState ? 1 : 3;
constexpr int SHAPE_TX[3] = {TX2, 31, 7}, SHAPE_TY[3] = {TY2, 7, 31};
static_assert(kLargeState || (flux_lds_bytes<31, 7, true>() <= 65536 && flux_lds_bytes<7, 31, true>() <= 65536),
              "flux tile exceeds 64 KB of LDS");

template <int TX, int TY, bool TWO_D>
void launch_flux(const FluxLaunch& L, unsigned grid) {
    hipLaunchKernelGGL((flux_stage_kernel<Topo, TX, TY, TWO_D>), dim3(grid),
                       dim3(FluxTile<Topo, TX, TY, TWO_D>::THREADS), 0, L.stream, static_cast<const MD*>(L.mech),
                       L.boxes, L.tiles, L.ntiles, L.Uin, L.Uout, L.first_order, !L.skip_validation,
                       L.first_failed, L.counters,
                       L.fluxes, L.in_ptrs, L.out_ptrs);
}

int flux(const FluxLaunch& L) {
    if (L.ntiles == 0) return HF_OK;
    {
    StreamTimer timer(L.kernel_ms, L.stream);
    const unsigned grid = (unsigned)((L.ntiles + 7) / 8) * 8;   // XCD remap needs a multiple of 8
    if (L.one_d) {
        launch_flux<TX1, 1, false>(L, grid);
    } else if constexpr (!kLargeState) {
        if (L.shape == 1) launch_flux<31, 7, true>(L, grid);
        else if (L.shape == 2) launch_flux<7, 31, true>(L, grid);
        else launch_flux<TX2, TY2, true>(L, grid);
    } else {
        launch_flux<TX2, TY2, true>(L, grid);
    }
    const hipError_t err = hipGetLastError();
    if (err != hipSuccess) return HF_EDEVICE;
    timer.stop();
    return HF_OK;
    }
}

// ------------------------------------------------------------------ properties / CFL rate
__global__ void __launch_bounds__(256)
props_kernel(const MD* __restrict__ dm, const double* __restrict__ U, long ncell, long ld, double* __restrict__ p,
             double* __restrict__ a, double* __restrict__ T, unsigned long long* __restrict__ first_failed, double dx,
             double dy, unsigned long long* __restrict__ rate_max) {
    constexpr int NS = Topo::NS, NV = NS + 3;
    const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    double rate = 0;
    if (i < ncell) {
        double u[NV];
        #pragma unroll
        for (int n = 0; n < NV; ++n) u[n] = U[n * ld + i];
        Props pr;
        if (!properties(*dm, u, pr, 1000.0)) {
            atomicMin(first_failed, (unsigned long long)i);
            pr.p = pr.a = pr.T = NAN;
        } else if (dx > 0) {
            double rho = 0;
            #pragma unroll
            for (int k = 0; k < NS; ++k) rho += u[k];
            rate = (fabs(u[NS] / rho) + pr.a) / dx + (dy > 0 ? (fabs(u[NS + 1] / rho) + pr.a) / dy : 0.0);
        }
        if (p) p[i] = pr.p;
        if (a) a[i] = pr.a;
        if (T) T[i] = pr.T;
    }
    if (rate_max) {
        #pragma unroll
        for (int off = 32; off > 0; off >>= 1) rate = fmax(rate, __shfl_xor(rate, off, 64));
        if ((threadIdx.x & 63) == 0 && rate > 0) atomicMax(rate_max, (unsigned long long)__double_as_longlong(rate));
    }
}

int props(const PropsLaunch& L) {
    if (L.ncell == 0) return HF_OK;
    const int block = 256;
    const long grid = (L.ncell + block - 1) / block;
    hipLaunchKernelGGL(props_kernel, dim3((unsigned)grid), dim3(block), 0, L.stream, static_cast<const MD*>(L.mech),
                       L.U, (long)L.ncell, (long)L.ld, L.p, L.a, L.T, L.first_failed, L.dx, L.dy,
                       reinterpret_cast<unsigned long long*>(L.rate_max));
    return hipGetLastError() == hipSuccess ? HF_OK : HF_EDEVICE;
}

// ------------------------------------------------------------------ single-box driver operations
// Ghost fill, roundoff repair and validation over one ghosted SoA box, for a
// driver that advances a single-level box (BlockAMR::fill / repairRoundoff /
// validate / stableDt, BlockAMR.cpp:137-152,286-346).
__device__ __forceinline__ double& box_at(double* U, const hf_box& B, int n, int i, int j) {
    return U[B.in_offset + n * B.in_plane + (int64_t)(j + 2) * B.in_pitch + (i + 2)];
}

// Source index of ghost g (outside [0, len)) for a boundary code:
// HF_BC_TRANSMISSIVE copies the nearest valid cell (AMReX foextrap),
// HF_BC_REFLECTING mirrors about the face (reflect_even / reflect_odd),
// HF_BC_PERIODIC wraps.
__device__ __forceinline__ int ghost_source(int g, int len, int bc) {
    if (bc == HF_BC_PERIODIC) return ((g % len) + len) % len;
    if (bc == HF_BC_REFLECTING) return g < 0 ? -g - 1 : 2 * len - g - 1;
    return g < 0 ? 0 : len - 1;
}

// The fixed inflow state of FlowBoundary (FlowBoundary.hpp:8), passed by value.
struct InletState {
    double q[Topo::NS + 3];
};

// pass 0: x ghosts of the valid rows; pass 1: y ghosts of every column
// (including the x ghosts, so corners come from the filled columns, as AMReX's
// filcc applies the y rule after the x rule). FlowBoundary (FlowBoundary.hpp:
// 12-36) then overrides: a ghost beyond any inflow side is the inlet state; a
// wedge-bottom ghost copies row 0 before the tip and mirrors (negated y
// momentum) from column wedge_i on, reading the clamped valid column. Both
// read only valid cells or pass-0 ghosts, so no pass-1 thread reads another's
// output.
__global__ void __launch_bounds__(256) ghost_kernel(double* __restrict__ U, hf_box B, int4 bc, InletState inlet,
                                                    int pass) {
    constexpr int NV = Topo::NS + 3;
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const int len = pass == 0 ? B.nx : B.ny;
    const int across = pass == 0 ? B.ny : B.nx + 4;
    if (t >= (long)NV * across * 4) return;
    const int g4 = (int)(t % 4);
    const int a = (int)((t / 4) % across);
    const int n = (int)(t / (4L * across));
    const int g = g4 < 2 ? g4 - 2 : len + g4 - 2;          // -2, -1, len, len+1
    const int code = g < 0 ? (pass == 0 ? bc.x : bc.z) : (pass == 0 ? bc.y : bc.w);
    if (pass == 0) {
        if (code == HF_BC_KEEP) return;
        if (code == HF_BC_INFLOW) {
            box_at(U, B, n, g, a) = inlet.q[n];
            return;
        }
        const int s = ghost_source(g, len, code);
        const bool odd = code == HF_BC_REFLECTING && n == Topo::NS;
        const double v = box_at(U, B, n, s, a);
        box_at(U, B, n, g, a) = odd ? -v : v;
        return;
    }
    const int i = a - 2;
    if (code == HF_BC_KEEP) {
        // an in-domain y ghost row: only its cells beyond a physical x side
        // are filled, by that side's rule from the kept cell beside them
        const int xcode = i < 0 ? bc.x : (i >= B.nx ? bc.y : HF_BC_KEEP);
        if (xcode == HF_BC_KEEP) return;
        if (xcode == HF_BC_INFLOW) {
            box_at(U, B, n, i, g) = inlet.q[n];
            return;
        }
        const int s = ghost_source(i, B.nx, xcode);
        const double v = box_at(U, B, n, s, g);
        box_at(U, B, n, i, g) = xcode == HF_BC_REFLECTING && n == Topo::NS ? -v : v;
        return;
    }
    const bool beyond_x_inflow = (i < 0 && bc.x == HF_BC_INFLOW) || (i >= B.nx && bc.y == HF_BC_INFLOW);
    if (beyond_x_inflow || code == HF_BC_INFLOW) {
        box_at(U, B, n, i, g) = inlet.q[n];
        return;
    }
    if (code == HF_BC_WEDGE) {                             // g < 0 here: y-lo only
        const bool wall = i >= B.wedge_i;
        // the column is clamped to the domain: only at a physical x side
        const int ci = (i < 0 && bc.x != HF_BC_KEEP) ? 0 : ((i >= B.nx && bc.y != HF_BC_KEEP) ? B.nx - 1 : i);
        const double v = box_at(U, B, n, ci, wall ? -g - 1 : 0);
        box_at(U, B, n, i, g) = wall && n == Topo::NS + 1 ? -v : v;
        return;
    }
    const int s = ghost_source(g, len, code);
    const bool odd = code == HF_BC_REFLECTING && n == Topo::NS + 1;
    const double v = box_at(U, B, n, i, s);
    box_at(U, B, n, i, g) = odd ? -v : v;
}

// ------------------------------------------------------------------ coarse / fine transfer
__device__ __forceinline__ double& patch_at(double* p, const hf_patch& P, int n, int i, int j) {
    return p[P.offset + n * P.plane + (long)j * P.pitch + i];
}

// One child of ConservativeInterp.cpp:37,47 without contraction:
// q + theta * ((+-0.25 sx) + (+-0.25 sy)).
__device__ __forceinline__ double interp_child(double q, double theta, double cx, double sx, double cy, double sy) {
#pragma clang fp contract(off)
    return q + theta * (cx * sx + cy * sy);
}

__device__ __forceinline__ double minmod1(double x, double y) {
    return x * y > 0 ? copysign(fmin(fabs(x), fabs(y)), x) : 0.0;
}

// Coarse-to-fine conservative interpolation at ratio (2, 2)
// (ConservativeInterp::interp, ConservativeInterp.cpp:27-50): one lane per
// coarse cell of coarsen(region), the theta search over all four children,
// then the children inside region. scal[0] counts limited cells.
// A list of regions (regions non-null: region blockIdx.y) shares the
// patches; with a fine mask, a child is written only inside the mask's
// extent where the mask is 0 (not a valid fine cell), and a parent with no
// such child does nothing.
__global__ void __launch_bounds__(256) interp_kernel(const MD* __restrict__ dm, const double* __restrict__ C,
                                                     hf_patch cp, double* __restrict__ F, hf_patch fp, int4 region0,
                                                     const int4* __restrict__ regions,
                                                     const uint8_t* __restrict__ fmask, long fmask_pitch, int fnx,
                                                     int fny, unsigned long long* __restrict__ scal,
                                                     unsigned long long* __restrict__ first_failed) {
    constexpr int NV = Topo::NS + 3;
    const int4 region = regions ? regions[blockIdx.y] : region0;
    const int cilo = region.x >> 1, cjlo = region.y >> 1;  // floor division (coarsen)
    const int cnx = (region.z >> 1) - cilo + 1, cny = (region.w >> 1) - cjlo + 1;
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)cnx * cny) return;
    const int i = cilo + (int)(t % cnx), j = cjlo + (int)(t / cnx);
    auto writable = [&](int fi, int fj) {
        if (fi < region.x || fi > region.z || fj < region.y || fj > region.w) return false;
        if (!fmask) return true;
        return fi >= 0 && fi < fnx && fj >= 0 && fj < fny && fmask[(long)fj * fmask_pitch + fi] == 0;
    };
    if (fmask) {
        bool any = false;
        for (int k = 0; k < 4; ++k) any = any || writable(2 * i + (k & 1), 2 * j + (k >> 1));
        if (!any) return;
    }
    double* c = const_cast<double*>(C);
    double q[NV], sx[NV], sy[NV];
    #pragma unroll
    for (int n = 0; n < NV; ++n) {
        q[n] = patch_at(c, cp, n, i, j);
        sx[n] = minmod1(q[n] - patch_at(c, cp, n, i - 1, j), patch_at(c, cp, n, i + 1, j) - q[n]);
        sy[n] = minmod1(q[n] - patch_at(c, cp, n, i, j - 1), patch_at(c, cp, n, i, j + 1) - q[n]);
    }
    double theta = 1.0;
    bool accepted = false;
    // Each temperature inversion starts from the last converged one, as the
    // reference's per-thread phase does; only admissibility is used.
    double t_guess = 1000.0;
    for (int attempt = 0; attempt < 48 && !accepted; ++attempt) {
        bool good = true;
        for (int k = 0; k < 4 && good; ++k) {              // (a, b) = (0,0), (1,0), (0,1), (1,1)
            const double cx = (k & 1) ? 0.25 : -0.25, cy = (k & 2) ? 0.25 : -0.25;
            double child[NV];
            #pragma unroll
            for (int n = 0; n < NV; ++n) child[n] = interp_child(q[n], theta, cx, sx[n], cy, sy[n]);
            Props p;
            good = properties(*dm, child, p, t_guess);
            if (good) t_guess = p.T;
        }
        if (good)
            accepted = true;
        else
            theta *= 0.5;
    }
    if (!accepted) {
        theta = 0.0;
        Props p;
        if (!properties(*dm, q, p, 1000.0)) {
            atomicMin(first_failed, regions ? ((unsigned long long)blockIdx.y << 32) | (unsigned long long)t
                                            : (unsigned long long)t);
            return;
        }
    }
    if (theta < 1.0) atomicAdd(&scal[0], 1ull);
    for (int k = 0; k < 4; ++k) {
        const int fi = 2 * i + (k & 1), fj = 2 * j + (k >> 1);
        if (!writable(fi, fj)) continue;
        const double cx = (k & 1) ? 0.25 : -0.25, cy = (k & 2) ? 0.25 : -0.25;
        #pragma unroll
        for (int n = 0; n < NV; ++n) patch_at(F, fp, n, fi, fj) = interp_child(q[n], theta, cx, sx[n], cy, sy[n]);
    }
}

// Flux-register correction of one fine box at ratio (2, 2) (BlockAMR.cpp:
// 247-256): one lane per (coarse face on a selected side, component). The
// four sides' outside cells are distinct, so no two lanes write one value.
__device__ __forceinline__ double reflux_register(double mc, double fc, double mf, double f0, double f1) {
#pragma clang fp contract(off)
    double reg = mc * fc;
    reg += mf * f0;
    reg += mf * f1;
    return reg;
}

__global__ void __launch_bounds__(256) reflux_kernel(double* __restrict__ Uc, hf_patch cp, hf_face_fluxes C,
                                                     hf_face_fluxes F, int4 fb, int sides, double h, double dxc,
                                                     double dyc) {
#pragma clang fp contract(off)
    constexpr int NV = Topo::NS + 3;
    const int ci0 = fb.x >> 1, cj0 = fb.y >> 1, ci1 = fb.z >> 1, cj1 = fb.w >> 1;
    const int ny = cj1 - cj0 + 1, nx = ci1 - ci0 + 1;
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long per = 2L * ny + 2L * nx;
    if (t >= per * NV) return;
    const int n = (int)(t / per);
    long k = t % per;
    int side;
    if (k < ny) side = 0;
    else if (k < 2L * ny) { side = 1; k -= ny; }
    else if (k < 2L * ny + nx) { side = 2; k -= 2L * ny; }
    else { side = 3; k -= 2L * ny + nx; }
    if (!((sides >> side) & 1)) return;
    const double vol = dxc * dyc;
    double reg;
    int oi, oj;
    if (side < 2) {                                        // x faces: column I, coarse row j
        const int j = cj0 + (int)k;
        const int I = side == 0 ? ci0 : ci1 + 1;           // coarse face index
        const int FI = side == 0 ? fb.x : fb.z + 1;        // fine face index
        const double fc = C.fx[C.fx_offset + n * C.fx_plane + (int64_t)j * C.fx_pitch + I];
        const double f0 = F.fx[F.fx_offset + n * F.fx_plane + (int64_t)(2 * j) * F.fx_pitch + FI];
        const double f1 = F.fx[F.fx_offset + n * F.fx_plane + (int64_t)(2 * j + 1) * F.fx_pitch + FI];
        reg = reflux_register(-h * dyc, fc, h * (0.5 * dyc), f0, f1);
        oi = side == 0 ? ci0 - 1 : ci1 + 1;
        oj = j;
    } else {                                               // y faces: row J, coarse column i
        const int i = ci0 + (int)k;
        const int J = side == 2 ? cj0 : cj1 + 1;
        const int FJ = side == 2 ? fb.y : fb.w + 1;
        const double fc = C.fy[C.fy_offset + n * C.fy_plane + (int64_t)J * C.fy_pitch + i];
        const double f0 = F.fy[F.fy_offset + n * F.fy_plane + (int64_t)FJ * F.fy_pitch + 2 * i];
        const double f1 = F.fy[F.fy_offset + n * F.fy_plane + (int64_t)FJ * F.fy_pitch + 2 * i + 1];
        reg = reflux_register(-h * dxc, fc, h * (0.5 * dxc), f0, f1);
        oi = i;
        oj = side == 2 ? cj0 - 1 : cj1 + 1;
    }
    const double sgn = (side & 1) ? 1.0 : -1.0;
    patch_at(Uc, cp, n, oi, oj) += (sgn * reg) / vol;
}

// Refinement tags (BlockAMR::ErrorEst, BlockAMR.cpp:178-198): one lane per
// valid cell of a ghosted box. The neighbours' inversions warm-start from the
// cell's own temperature. scal[0] counts tagged cells.
__global__ void __launch_bounds__(256) tag_kernel(const MD* __restrict__ dm, const double* __restrict__ U, hf_box B,
                                                  hf_tag_params P, uint8_t* __restrict__ tags,
                                                  unsigned long long* __restrict__ scal,
                                                  unsigned long long* __restrict__ first_failed) {
    constexpr int NS = Topo::NS, NV = NS + 3;
    const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
    bool marked = false;
    if (c < (long)B.nx * B.ny) {
        const int i = (int)(c % B.nx), j = (int)(c / B.nx);
        double* u0 = const_cast<double*>(U);
        const double x = P.x0 + (i + 0.5) * P.dx - P.velocity[0] * P.time;
        const double y = P.y0 + (j + 0.5) * P.dy - P.velocity[1] * P.time;
        marked = P.has_box && x >= P.box[0] && x <= P.box[1] && y >= P.box[2] && y <= P.box[3];
        bool ok = true;
        if (P.physics) {
            double q[NV];
            #pragma unroll
            for (int n = 0; n < NV; ++n) q[n] = box_at(u0, B, n, i, j);
            Props p;
            ok = properties(*dm, q, p, 1000.0);
            for (int dd = 0; dd < 2 && ok; ++dd) {
                const int di = dd == 0, dj = dd == 1;
                double ql[NV], qr[NV];
                #pragma unroll
                for (int n = 0; n < NV; ++n) {
                    ql[n] = box_at(u0, B, n, i - di, j - dj);
                    qr[n] = box_at(u0, B, n, i + di, j + dj);
                }
                Props pl, pr;
                ok = properties(*dm, ql, pl, p.T) && properties(*dm, qr, pr, p.T);
                if (!ok) break;
                const double jump = fmax(fabs(pl.p - p.p) / fmin(pl.p, p.p), fabs(pr.p - p.p) / fmin(pr.p, p.p));
                const double temp = fmax(fabs(pl.T - p.T) / fmin(pl.T, p.T), fabs(pr.T - p.T) / fmin(pr.T, p.T));
                marked = marked || jump > P.pressure_tag || temp > P.temperature_tag;
            }
            if (ok && P.chemistry) {
                double C[NS], w[NS], rho = 0;
                #pragma unroll
                for (int k = 0; k < NS; ++k) {
                    C[k] = q[k] * dm->invW[k];
                    rho += q[k];
                }
                production_rates(*dm, p.T, C, w);
                double wf = 0, Wf = 0;
                #pragma unroll
                for (int k = 0; k < NS; ++k)
                    if (k == P.fuel) {
                        wf = w[k];
                        Wf = dm->W[k];
                    }
                marked = marked || -wf * Wf / rho > P.fuel_rate_tag;
            }
        }
        if (!ok) {
            atomicMin(first_failed, (unsigned long long)c);
            marked = false;
        }
        tags[c] = marked ? 1 : 0;
    }
    unsigned long long n = marked ? 1ull : 0ull;
    #pragma unroll
    for (int off = 32; off > 0; off >>= 1) n += __shfl_xor(n, off, 64);
    if ((threadIdx.x & 63) == 0 && n) atomicAdd(&scal[0], n);
}

// Clears one counter and the failure index (a kernel, not a memset, so the
// sequence stays valid inside a captured graph).
__global__ void counter_init_kernel(unsigned long long* __restrict__ counter, unsigned long long* __restrict__ first_failed) {
    if (threadIdx.x == 0) {
        *counter = 0;
        *first_failed = ~0ull;
    }
}

// Average-down at ratio (2, 2): coarse = 0.25 * (((0 + f00) + f10) + f01) + f11),
// one lane per (component, coarse cell); i fastest for coalesced rows.
// With a coarse mask, only coarse cells marked 2 (covered by the fine level) are written.
__global__ void __launch_bounds__(256) avgdown_kernel(const double* __restrict__ F, hf_patch fp, double* __restrict__ C,
                                                      hf_patch cp, int4 region, const uint8_t* __restrict__ cmask,
                                                      long cmask_pitch) {
    constexpr int NV = Topo::NS + 3;
    const int cnx = region.z - region.x + 1, cny = region.w - region.y + 1;
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)NV * cnx * cny) return;
    const int i = region.x + (int)(t % cnx);
    const int j = region.y + (int)((t / cnx) % cny);
    const int n = (int)(t / ((long)cnx * cny));
    if (cmask && cmask[(long)j * cmask_pitch + i] != 2) return;
    double* f = const_cast<double*>(F);
    double sum = 0.0;
    sum += patch_at(f, fp, n, 2 * i, 2 * j);
    sum += patch_at(f, fp, n, 2 * i + 1, 2 * j);
    sum += patch_at(f, fp, n, 2 * i, 2 * j + 1);
    sum += patch_at(f, fp, n, 2 * i + 1, 2 * j + 1);
    patch_at(C, cp, n, i, j) = 0.25 * sum;
}

// Roundoff repair (BlockAMR::repairRoundoff, BlockAMR.cpp:299-334) on one
// level of equal cell volumes. scal layout: [0, NS) peak bits, [NS, 2NS)
// donor index, [2NS, 3NS) deficit in fixed point (units of 64 eps peak /
// 2^40), 3NS repairs, 3NS+1 failure.
__global__ void __launch_bounds__(256) repair_peak_kernel(const double* __restrict__ U, hf_box B,
                                                          unsigned long long* __restrict__ scal) {
    constexpr int NS = Topo::NS;
    __shared__ unsigned long long part[4][NS];
    const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
    // peak starts at 0 (the reference's zero-initialised array): only positive
    // values compete; positive doubles order like their bit patterns, so the
    // workgroup's maximum is taken first and one atomic per species follows
    unsigned long long pv[NS];
    #pragma unroll
    for (int n = 0; n < NS; ++n) pv[n] = 0;
    if (c < (long)B.nx * B.ny) {
        const int i = (int)(c % B.nx), j = (int)(c / B.nx);
        #pragma unroll
        for (int n = 0; n < NS; ++n) {
            const double v = box_at(const_cast<double*>(U), B, n, i, j);
            if (v > 0) pv[n] = (unsigned long long)__double_as_longlong(v);
        }
    }
    #pragma unroll
    for (int n = 0; n < NS; ++n) {
        #pragma unroll
        for (int off = 32; off > 0; off >>= 1) pv[n] = max(pv[n], (unsigned long long)__shfl_xor(pv[n], off, 64));
        if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6][n] = pv[n];
    }
    __syncthreads();
    if (threadIdx.x < NS) {
        const int n = threadIdx.x;
        unsigned long long m = 0;
        for (int w = 0; w < (int)(blockDim.x + 63) / 64; ++w) m = max(m, part[w][n]);
        if (m) atomicMax(&scal[n], m);
    }
}
// donor: the first cell in BoxIterator order (i fastest) holding the peak mass
__global__ void __launch_bounds__(256) repair_donor_kernel(const double* __restrict__ U, hf_box B,
                                                           unsigned long long* __restrict__ scal) {
    constexpr int NS = Topo::NS;
    const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= (long)B.nx * B.ny) return;
    const int i = (int)(c % B.nx), j = (int)(c / B.nx);
    #pragma unroll
    for (int n = 0; n < NS; ++n) {
        const double v = box_at(const_cast<double*>(U), B, n, i, j);
        if (v > 0 && (unsigned long long)__double_as_longlong(v) == scal[n]) atomicMin(&scal[NS + n], (unsigned long long)c);
    }
}
__global__ void __launch_bounds__(256) repair_zero_kernel(double* __restrict__ U, hf_box B,
                                                          unsigned long long* __restrict__ scal,
                                                          unsigned long long* __restrict__ first_failed) {
    constexpr int NS = Topo::NS;
    const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= (long)B.nx * B.ny) return;
    const int i = (int)(c % B.nx), j = (int)(c / B.nx);
    unsigned long long repairs = 0;
    #pragma unroll
    for (int n = 0; n < NS; ++n) {
        double& v = box_at(U, B, n, i, j);
        if (v >= 0) continue;
        const double peak = __longlong_as_double((long long)scal[n]);
        const double limit = 64.0 * 2.220446049250313e-16 * peak;
        if (!finite(v) || -v > limit) {
            atomicMin(first_failed, (unsigned long long)c);
            continue;
        }
        // the deficit in fixed point, in units of limit / 2^40 (|v| <= limit):
        // integer sums do not depend on the order the cells add in
        atomicAdd(&scal[2 * NS + n], (unsigned long long)llrint(-v / limit * 1099511627776.0));
        v = 0.0;
        ++repairs;
    }
    if (repairs) atomicAdd(&scal[3 * NS], repairs);
}
__global__ void repair_donate_kernel(double* __restrict__ U, hf_box B, unsigned long long* __restrict__ scal,
                                     unsigned long long* __restrict__ first_failed) {
    constexpr int NS = Topo::NS;
    const int n = threadIdx.x;
    if (n >= NS) return;
    const double peak = __longlong_as_double((long long)scal[n]);
    const double deficit = (double)scal[2 * NS + n] * (64.0 * 2.220446049250313e-16 * peak / 1099511627776.0);
    if (!(deficit > 0)) return;
    const unsigned long long d = scal[NS + n];
    if (!(peak > deficit) || d == ~0ull) {
        atomicMin(first_failed, d == ~0ull ? 0ull : d);
        scal[3 * NS + 1] = 1;
        return;
    }
    box_at(U, B, n, (int)(d % B.nx), (int)(d / B.nx)) -= deficit;
}
__global__ void repair_init_kernel(unsigned long long* __restrict__ scal, unsigned long long* __restrict__ first_failed) {
    constexpr int NS = Topo::NS;
    const int t = threadIdx.x;
    if (t < NS) {
        scal[t] = 0;                 // peak 0.0
        scal[NS + t] = ~0ull;        // no donor
        scal[2 * NS + t] = 0;        // deficit 0 (fixed point)
    }
    if (t == 0) {
        scal[3 * NS] = 0;
        scal[3 * NS + 1] = 0;
        *first_failed = ~0ull;
    }
}

// validate + stableDt over the box interior: properties of every valid cell
// (lowest failing index in first_failed) and the max CFL rate.
__global__ void __launch_bounds__(256) box_props_kernel(const MD* __restrict__ dm, const double* __restrict__ U,
                                                        hf_box B, double dx, double dy,
                                                        unsigned long long* __restrict__ first_failed,
                                                        unsigned long long* __restrict__ rate_max) {
    constexpr int NS = Topo::NS, NV = NS + 3;
    const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
    double rate = 0;
    if (c < (long)B.nx * B.ny) {
        const int i = (int)(c % B.nx), j = (int)(c / B.nx);
        double u[NV];
        #pragma unroll
        for (int n = 0; n < NV; ++n) u[n] = box_at(const_cast<double*>(U), B, n, i, j);
        Props p;
        if (!properties(*dm, u, p, 1000.0)) {
            atomicMin(first_failed, (unsigned long long)c);
        } else {
            double rho = 0;
            #pragma unroll
            for (int k = 0; k < NS; ++k) rho += u[k];
            rate = (fabs(u[NS] / rho) + p.a) / dx + (dy > 0 ? (fabs(u[NS + 1] / rho) + p.a) / dy : 0.0);
        }
    }
    #pragma unroll
    for (int off = 32; off > 0; off >>= 1) rate = fmax(rate, __shfl_xor(rate, off, 64));
    if ((threadIdx.x & 63) == 0 && rate > 0) atomicMax(rate_max, (unsigned long long)__double_as_longlong(rate));
}

// ------------------------------------------------------------------ operations over cell blocks
// The valid cells of every level of an AMR hierarchy (hf_cells blocks with
// masks: 0 not valid, 1 leaf, 2 covered), addressed through CellMap's block
// form; cell ids run over the blocks in order, then j, then i.
__device__ __forceinline__ double& cell_at(const hf_cells& B, int n, int i, int j) {
    return B.U[B.offset + n * B.plane + (int64_t)j * B.pitch + i];
}
__device__ __forceinline__ int cell_kind(const hf_cells& B, int i, int j) {
    return B.mask ? B.mask[(int64_t)j * B.mask_pitch + i] : 1;
}

// BlockAMR::validate over every valid cell (BlockAMR.cpp:286-298: all
// levels, no mask) and stableDt's rate over the leaf cells (:335-346).
__global__ void __launch_bounds__(256) cells_props_kernel(const MD* __restrict__ dm, CellMap S, long total,
                                                          unsigned long long* __restrict__ first_failed,
                                                          unsigned long long* __restrict__ rate_max) {
    constexpr int NS = Topo::NS, NV = NS + 3;
    const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
    double rate = 0;
    if (c < total) {
        int b, i, j;
        const hf_cells& B = S.find(c, b, i, j);
        const int kind = cell_kind(B, i, j);
        if (kind != 0) {
            double u[NV];
            #pragma unroll
            for (int n = 0; n < NV; ++n) u[n] = cell_at(B, n, i, j);
            Props p;
            if (!properties(*dm, u, p, 1000.0)) {
                atomicMin(first_failed, ((unsigned long long)b << 32) | (unsigned long long)((long)j * B.nx + i));
            } else if (kind == 1) {
                double rho = 0;
                #pragma unroll
                for (int k = 0; k < NS; ++k) rho += u[k];
                rate = (fabs(u[NS] / rho) + p.a) / B.dx + (B.dy > 0 ? (fabs(u[NS + 1] / rho) + p.a) / B.dy : 0.0);
            }
        }
    }
    #pragma unroll
    for (int off = 32; off > 0; off >>= 1) rate = fmax(rate, __shfl_xor(rate, off, 64));
    if ((threadIdx.x & 63) == 0 && rate > 0) atomicMax(rate_max, (unsigned long long)__double_as_longlong(rate));
}

// BlockAMR::repairRoundoff (BlockAMR.cpp:299-334) over the leaf cells of all
// blocks, cell volume dx*dy per block. scal: [0, NS) peak value bits,
// [NS, 2NS) largest mass (value * volume) bits, [2NS, 3NS) donor cell id (the
// first cell holding the largest mass), 3NS repairs, 3NS+1 donor failure.
// The deficits are summed per workgroup in a fixed order (partials) and then
// over the workgroups in a fixed order, so the result does not depend on timing.
__global__ void cells_repair_init_kernel(unsigned long long* __restrict__ scal,
                                         unsigned long long* __restrict__ first_failed) {
    constexpr int NS = Topo::NS;
    const int t = threadIdx.x;
    if (t < NS) {
        scal[t] = 0;
        scal[NS + t] = 0;
        scal[2 * NS + t] = ~0ull;
    }
    if (t == 0) {
        scal[3 * NS] = 0;
        scal[3 * NS + 1] = 0;
        *first_failed = ~0ull;
    }
}
// Largest value and largest mass per species: reduced within the workgroup
// first (positive doubles order like their bit patterns; a maximum does not
// depend on the order it is taken in), so each workgroup issues one atomic
// per species and quantity instead of one per cell.
__global__ void __launch_bounds__(256) cells_repair_peak_kernel(CellMap S, long total,
                                                                unsigned long long* __restrict__ scal) {
    constexpr int NS = Topo::NS;
    __shared__ unsigned long long part[4][2 * NS];
    const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
    unsigned long long pv[NS], pm[NS];
    #pragma unroll
    for (int n = 0; n < NS; ++n) pv[n] = pm[n] = 0;
    if (c < total) {
        int b, i, j;
        const hf_cells& B = S.find(c, b, i, j);
        if (cell_kind(B, i, j) == 1) {
            const double vol = B.dy > 0 ? B.dx * B.dy : B.dx;
            #pragma unroll
            for (int n = 0; n < NS; ++n) {
                const double v = cell_at(B, n, i, j);
                if (v > 0) {
                    pv[n] = (unsigned long long)__double_as_longlong(v);
                    pm[n] = (unsigned long long)__double_as_longlong(v * vol);
                }
            }
        }
    }
    #pragma unroll
    for (int n = 0; n < NS; ++n) {
        #pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            pv[n] = max(pv[n], (unsigned long long)__shfl_xor(pv[n], off, 64));
            pm[n] = max(pm[n], (unsigned long long)__shfl_xor(pm[n], off, 64));
        }
        if ((threadIdx.x & 63) == 0) {
            part[threadIdx.x >> 6][n] = pv[n];
            part[threadIdx.x >> 6][NS + n] = pm[n];
        }
    }
    __syncthreads();
    if (threadIdx.x < 2 * NS) {
        const int q = threadIdx.x;
        unsigned long long m = 0;
        for (int w = 0; w < (int)(blockDim.x + 63) / 64; ++w) m = max(m, part[w][q]);
        if (m) atomicMax(&scal[q], m);
    }
}
__global__ void __launch_bounds__(256) cells_repair_donor_kernel(CellMap S, long total,
                                                                 unsigned long long* __restrict__ scal) {
    constexpr int NS = Topo::NS;
    const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= total) return;
    int b, i, j;
    const hf_cells& B = S.find(c, b, i, j);
    if (cell_kind(B, i, j) != 1) return;
    const double vol = B.dy > 0 ? B.dx * B.dy : B.dx;
    #pragma unroll
    for (int n = 0; n < NS; ++n) {
        const double v = cell_at(B, n, i, j);
        if (v > 0 && (unsigned long long)__double_as_longlong(v * vol) == scal[NS + n])
            atomicMin(&scal[2 * NS + n], (unsigned long long)c);
    }
}
__global__ void __launch_bounds__(256) cells_repair_zero_kernel(CellMap S, long total,
                                                                unsigned long long* __restrict__ scal,
                                                                double* __restrict__ partials,
                                                                unsigned long long* __restrict__ first_failed) {
    constexpr int NS = Topo::NS;
    __shared__ double part[4][NS];
    const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
    double d[NS];
    #pragma unroll
    for (int n = 0; n < NS; ++n) d[n] = 0.0;
    unsigned long long repairs = 0;
    if (c < total) {
        int b, i, j;
        const hf_cells& B = S.find(c, b, i, j);
        if (cell_kind(B, i, j) == 1) {
            const double vol = B.dy > 0 ? B.dx * B.dy : B.dx;
            #pragma unroll
            for (int n = 0; n < NS; ++n) {
                double& v = cell_at(B, n, i, j);
                if (v >= 0) continue;
                const double peak = __longlong_as_double((long long)scal[n]);
                if (!finite(v) || -v > 64.0 * 2.220446049250313e-16 * peak) {
                    atomicMin(first_failed, ((unsigned long long)b << 32) | (unsigned long long)((long)j * B.nx + i));
                    continue;
                }
                d[n] = -v * vol;
                v = 0.0;
                ++repairs;
            }
        }
    }
    // fixed-order sums: butterfly within the wave, then the four waves in order
    #pragma unroll
    for (int n = 0; n < NS; ++n) {
        double x = d[n];
        #pragma unroll
        for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
        if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6][n] = x;
    }
    #pragma unroll
    for (int off = 32; off > 0; off >>= 1) repairs += __shfl_xor(repairs, off, 64);
    if ((threadIdx.x & 63) == 0 && repairs) atomicAdd(&scal[3 * NS], repairs);
    __syncthreads();
    if (threadIdx.x < NS) {
        const int n = threadIdx.x;
        partials[(long)blockIdx.x * NS + n] = ((part[0][n] + part[1][n]) + part[2][n]) + part[3][n];
    }
}
// One workgroup per species: the deficit is the sum of the workgroup partials
// in a fixed order (a strided pass per lane, then a fixed butterfly and the
// four wave sums in order), so it does not depend on timing; lane 0 then
// charges it to the donor.
__global__ void __launch_bounds__(256) cells_repair_donate_kernel(CellMap S, int nparts,
                                                                  unsigned long long* __restrict__ scal,
                                                                  const double* __restrict__ partials,
                                                                  unsigned long long* __restrict__ first_failed) {
    constexpr int NS = Topo::NS;
    __shared__ double wsum[4];
    const int n = blockIdx.x;
    if (*first_failed != ~0ull) return;
    double x = 0.0;
    for (int g = threadIdx.x; g < nparts; g += 256) x += partials[(long)g * NS + n];
    #pragma unroll
    for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
    if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = x;
    __syncthreads();
    if (threadIdx.x != 0) return;
    const double deficit = ((wsum[0] + wsum[1]) + wsum[2]) + wsum[3];
    if (!(deficit > 0)) return;
    const double mass = __longlong_as_double((long long)scal[NS + n]);
    const unsigned long long d = scal[2 * NS + n];
    if (!(mass > deficit) || d == ~0ull) {
        scal[3 * NS + 1] = 1;
        return;
    }
    int b, i, j;
    const hf_cells& B = S.find((int64_t)d, b, i, j);
    const double vol = B.dy > 0 ? B.dx * B.dy : B.dx;
    cell_at(B, n, i, j) -= deficit / vol;
}

// Flux-register corrections of one level from its finer level (BlockAMR.cpp:
// 247-256), one lane per coarse cell of region: a leaf cell (mask 1) takes,
// for each face whose other side is covered (mask 2), in the order x-lo,
// x-hi, y-lo, y-hi, the register (-h dy_c) Fc + (h dy_f) Ff0 + (h dy_f) Ff1
// (dx for y faces) over its volume: + on the cell's lo faces (the fine box
// lies below it), - on its hi faces. Coarse and fine face fluxes are
// addressed in their levels' index spaces. No fused multiply-add.
__global__ void __launch_bounds__(256) reflux_cells_kernel(double* __restrict__ Uc, hf_patch cp, hf_face_fluxes C,
                                                           hf_face_fluxes F, int4 region,
                                                           const uint8_t* __restrict__ cmask, long cmask_pitch,
                                                           double h_arg, const double* __restrict__ hp, double dxc,
                                                           double dyc) {
#pragma clang fp contract(off)
    constexpr int NV = Topo::NS + 3;
    const double h = hp ? *hp : h_arg;
    const int nx = region.z - region.x + 1;
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)nx * (region.w - region.y + 1)) return;
    const int i = region.x + (int)(t % nx), j = region.y + (int)(t / nx);
    auto kind = [&](int a, int b) -> int {
        if (a < region.x || a > region.z || b < region.y || b > region.w) return 0;
        return cmask[(long)b * cmask_pitch + a];
    };
    if (kind(i, j) != 1) return;
    const bool xlo = kind(i - 1, j) == 2, xhi = kind(i + 1, j) == 2;
    const bool ylo = kind(i, j - 1) == 2, yhi = kind(i, j + 1) == 2;
    if (!(xlo || xhi || ylo || yhi)) return;
    const double vol = dxc * dyc;
    for (int n = 0; n < NV; ++n) {
        double u = patch_at(Uc, cp, n, i, j);
        if (xlo) {          // the fine box's x-hi side at coarse face i, fine face 2i
            const double fc = C.fx[C.fx_offset + n * C.fx_plane + (int64_t)j * C.fx_pitch + i];
            const double f0 = F.fx[F.fx_offset + n * F.fx_plane + (int64_t)(2 * j) * F.fx_pitch + 2 * i];
            const double f1 = F.fx[F.fx_offset + n * F.fx_plane + (int64_t)(2 * j + 1) * F.fx_pitch + 2 * i];
            u += reflux_register(-h * dyc, fc, h * (0.5 * dyc), f0, f1) / vol;
        }
        if (xhi) {          // the fine box's x-lo side at coarse face i+1
            const double fc = C.fx[C.fx_offset + n * C.fx_plane + (int64_t)j * C.fx_pitch + i + 1];
            const double f0 = F.fx[F.fx_offset + n * F.fx_plane + (int64_t)(2 * j) * F.fx_pitch + 2 * i + 2];
            const double f1 = F.fx[F.fx_offset + n * F.fx_plane + (int64_t)(2 * j + 1) * F.fx_pitch + 2 * i + 2];
            u += (-reflux_register(-h * dyc, fc, h * (0.5 * dyc), f0, f1)) / vol;
        }
        if (ylo) {
            const double fc = C.fy[C.fy_offset + n * C.fy_plane + (int64_t)j * C.fy_pitch + i];
            const double f0 = F.fy[F.fy_offset + n * F.fy_plane + (int64_t)(2 * j) * F.fy_pitch + 2 * i];
            const double f1 = F.fy[F.fy_offset + n * F.fy_plane + (int64_t)(2 * j) * F.fy_pitch + 2 * i + 1];
            u += reflux_register(-h * dxc, fc, h * (0.5 * dxc), f0, f1) / vol;
        }
        if (yhi) {
            const double fc = C.fy[C.fy_offset + n * C.fy_plane + (int64_t)(j + 1) * C.fy_pitch + i];
            const double f0 = F.fy[F.fy_offset + n * F.fy_plane + (int64_t)(2 * j + 2) * F.fy_pitch + 2 * i];
            const double f1 = F.fy[F.fy_offset + n * F.fy_plane + (int64_t)(2 * j + 2) * F.fy_pitch + 2 * i + 1];
            u += (-reflux_register(-h * dxc, fc, h * (0.5 * dxc), f0, f1)) / vol;
        }
        patch_at(Uc, cp, n, i, j) = u;
    }
}

int box_ops(const BoxLaunch& L) {
    const long cells = (long)L.box.nx * L.box.ny;
    const unsigned g256 = (unsigned)((cells + 255) / 256);
    const MD* dm = static_cast<const MD*>(L.mech);
    switch (L.op) {
    case BoxLaunch::FILL: {
        constexpr int NV = Topo::NS + 3;
        const int4 bc = make_int4(L.bc[0], L.bc[1], L.bc[2], L.bc[3]);
        InletState inlet{};
        if (L.inlet)
            for (int n = 0; n < NV; ++n) inlet.q[n] = L.inlet[n];
        const long t0 = (long)NV * L.box.ny * 4;
        hipLaunchKernelGGL(ghost_kernel, dim3((unsigned)((t0 + 255) / 256)), dim3(256), 0, L.stream, L.U, L.box, bc,
                           inlet, 0);
        if (L.two_d) {
            const long t1 = (long)NV * (L.box.nx + 4) * 4;
            hipLaunchKernelGGL(ghost_kernel, dim3((unsigned)((t1 + 255) / 256)), dim3(256), 0, L.stream, L.U, L.box, bc,
                               inlet, 1);
        }
        break;
    }
    case BoxLaunch::INTERP: {
        const int4 r = make_int4(L.region[0], L.region[1], L.region[2], L.region[3]);
        // a list: L.total is the largest region's parent count
        const long cells = L.regions ? (long)L.total
                                     : (long)((r.z >> 1) - (r.x >> 1) + 1) * ((r.w >> 1) - (r.y >> 1) + 1);
        hipLaunchKernelGGL(counter_init_kernel, dim3(1), dim3(64), 0, L.stream, L.scal, L.first_failed);
        hipLaunchKernelGGL(interp_kernel, dim3((unsigned)((cells + 255) / 256), L.regions ? (unsigned)L.nregion : 1u),
                           dim3(256), 0, L.stream, dm, L.src, L.sp, L.dst, L.dp, r,
                           reinterpret_cast<const int4*>(L.regions), L.mask, (long)L.mask_pitch, L.mask_nx, L.mask_ny,
                           L.scal, L.first_failed);
        break;
    }
    case BoxLaunch::REFLUX_CELLS: {
        const int4 r = make_int4(L.region[0], L.region[1], L.region[2], L.region[3]);
        const long t = (long)(r.z - r.x + 1) * (r.w - r.y + 1);
        hipLaunchKernelGGL(reflux_cells_kernel, dim3((unsigned)((t + 255) / 256)), dim3(256), 0, L.stream, L.dst, L.dp,
                           L.cflux, L.fflux, r, L.mask, (long)L.mask_pitch, L.h, L.hp, L.dx, L.dy);
        break;
    }
    case BoxLaunch::CELLS_PROPS: {
        const CellMap S{nullptr, 0, L.blocks, L.block_start, L.nblock};
        const unsigned g = (unsigned)((L.total + 255) / 256);
        hipLaunchKernelGGL(cells_props_kernel, dim3(g), dim3(256), 0, L.stream, dm, S, (long)L.total, L.first_failed,
                           L.scal);
        break;
    }
    case BoxLaunch::CELLS_REPAIR: {
        const CellMap S{nullptr, 0, L.blocks, L.block_start, L.nblock};
        const unsigned g = (unsigned)((L.total + 255) / 256);
        hipLaunchKernelGGL(cells_repair_init_kernel, dim3(1), dim3(64), 0, L.stream, L.scal, L.first_failed);
        hipLaunchKernelGGL(cells_repair_peak_kernel, dim3(g), dim3(256), 0, L.stream, S, (long)L.total, L.scal);
        hipLaunchKernelGGL(cells_repair_donor_kernel, dim3(g), dim3(256), 0, L.stream, S, (long)L.total, L.scal);
        hipLaunchKernelGGL(cells_repair_zero_kernel, dim3(g), dim3(256), 0, L.stream, S, (long)L.total, L.scal,
                           L.partials, L.first_failed);
        hipLaunchKernelGGL(cells_repair_donate_kernel, dim3(Topo::NS), dim3(256), 0, L.stream, S, (int)g, L.scal,
                           L.partials, L.first_failed);
        break;
    }
    case BoxLaunch::REFLUX: {
        constexpr int NV = Topo::NS + 3;
        const int4 r = make_int4(L.region[0], L.region[1], L.region[2], L.region[3]);
        const long per = 2L * ((r.w >> 1) - (r.y >> 1) + 1) + 2L * ((r.z >> 1) - (r.x >> 1) + 1);
        const long t = per * NV;
        hipLaunchKernelGGL(reflux_kernel, dim3((unsigned)((t + 255) / 256)), dim3(256), 0, L.stream, L.dst, L.dp,
                           L.cflux, L.fflux, r, L.sides, L.h, L.dx, L.dy);
        break;
    }
    case BoxLaunch::TAG: {
        hipLaunchKernelGGL(counter_init_kernel, dim3(1), dim3(64), 0, L.stream, L.scal, L.first_failed);
        hipLaunchKernelGGL(tag_kernel, dim3(g256), dim3(256), 0, L.stream, dm, L.U, L.box, L.tag, L.tags, L.scal,
                           L.first_failed);
        break;
    }
    case BoxLaunch::AVGDOWN: {
        constexpr int NV = Topo::NS + 3;
        const int4 r = make_int4(L.region[0], L.region[1], L.region[2], L.region[3]);
        const long t = (long)NV * (r.z - r.x + 1) * (r.w - r.y + 1);
        hipLaunchKernelGGL(avgdown_kernel, dim3((unsigned)((t + 255) / 256)), dim3(256), 0, L.stream, L.src, L.sp, L.dst,
                           L.dp, r, L.mask, (long)L.mask_pitch);
        break;
    }
    case BoxLaunch::REPAIR:
        hipLaunchKernelGGL(repair_init_kernel, dim3(1), dim3(64), 0, L.stream, L.scal, L.first_failed);
        hipLaunchKernelGGL(repair_peak_kernel, dim3(g256), dim3(256), 0, L.stream, L.U, L.box, L.scal);
        hipLaunchKernelGGL(repair_donor_kernel, dim3(g256), dim3(256), 0, L.stream, L.U, L.box, L.scal);
        hipLaunchKernelGGL(repair_zero_kernel, dim3(g256), dim3(256), 0, L.stream, L.U, L.box, L.scal,
                           L.first_failed);
        hipLaunchKernelGGL(repair_donate_kernel, dim3(1), dim3(64), 0, L.stream, L.U, L.box, L.scal, L.first_failed);
        break;
    case BoxLaunch::PROPS:
        hipLaunchKernelGGL(box_props_kernel, dim3(g256), dim3(256), 0, L.stream, dm, L.U, L.box, L.dx, L.dy,
                           L.first_failed, L.scal);
        break;
    }
    return hipGetLastError() == hipSuccess ? HF_OK : HF_EDEVICE;
}

// ------------------------------------------------------------------ sensitivities
// Fixed-step backward-Euler chemistry with its discrete adjoint (reverse
// sweep over checkpointed segments) and its forward tangent, one lane per
// cell. Checkpoint and trajectory slots are [slot][element][cell]: element j
// (y_0..y_{N-1}, then T) of slot m for cell c at [(m (N+1) + j) ncell + c].
struct SensArgs {
    long ncell;
    const double *rho, *T0, *Y0;
    double h;
    int nsteps, ckpt;
    const double* shift;
    double *J, *g_lnA, *g_T0, *g_Y0;
    const double *v_lnA, *v_T0, *v_Y0;
    double* dJ;
    double *ck, *traj;
    unsigned long long* first_failed;
};

// e, cv and the species energies e_k of the initial state (T0, Y0), Y0 as given.
template <class TT>
__device__ __forceinline__ void initial_energy(const MechData<TT>& d, double T0, const double* Y0, double& e,
                                               double& cv, double* ek) {
    e = cv = 0;
    #pragma unroll
    for (int k = 0; k < TT::NS; ++k) {
        const Nasa7 na = nasa_select(d, k, T0);
        const double* a = na.a;
        const double cpr = fma(T0, fma(T0, fma(T0, fma(T0, a[4], a[3]), a[2]), a[1]), a[0]);
        const double hrt = fma(T0, fma(T0, fma(T0, fma(T0, a[4] * 0.2, a[3] * 0.25), a[2] * (1.0 / 3.0)), a[1] * 0.5),
                               a[0]) + a[5] / T0;
        ek[k] = (hrt - 1.0) * kR * T0 * d.invW[k];
        e = fma(Y0[k], ek[k], e);
        cv = fma(Y0[k], (cpr - 1.0) * kR * d.invW[k], cv);
    }
}

template <class TT, bool TANGENT>
__global__ void __launch_bounds__(64) sens_kernel(const MechData<TT>* __restrict__ dm, SensArgs a) {
    constexpr int NS = TT::NS, N = NS - 1, NR = TT::NR;
    const long c = (long)blockIdx.x * 64 + threadIdx.x;
    if (c >= a.ncell) return;
    const MechData<TT>& d = *dm;
    const double rho = a.rho[c], h = a.h;
    double Y0[NS], ek0[NS], e, cv0;
    #pragma unroll
    for (int k = 0; k < NS; ++k) Y0[k] = a.Y0[c * NS + k];
    initial_energy(d, a.T0[c], Y0, e, cv0, ek0);
    double y[N], t = a.T0[c];
    #pragma unroll
    for (int k = 0; k < N; ++k) y[k] = Y0[k];
    BeState<TT> s;
    const size_t nc = (size_t)a.ncell;
    auto ck = [&](int m, int j) HF_AI -> double& { return a.ck[((size_t)m * (N + 1) + j) * nc + c]; };
    auto tr = [&](int m, int j) HF_AI -> double& { return a.traj[((size_t)m * (N + 1) + j) * nc + c]; };
    // tangent state
    double dy[N], de0 = 0, dJ = 0;
    if constexpr (TANGENT) {
        de0 = cv0 * a.v_T0[c];
        #pragma unroll
        for (int k = 0; k < NS; ++k) de0 = fma(ek0[k], a.v_Y0[c * NS + k], de0);
        #pragma unroll
        for (int k = 0; k < N; ++k) dy[k] = a.v_Y0[c * NS + k];
    }
    double Jv = 0;
    bool ok = true;
    for (int n = 0; n < a.nsteps && ok; ++n) {
        if constexpr (!TANGENT) {
            if (n % a.ckpt == 0) {
                #pragma unroll
                for (int j = 0; j < N; ++j) ck(n / a.ckpt, j) = y[j];
                ck(n / a.ckpt, N) = t;
            }
        }
        double yold[N];
        #pragma unroll
        for (int k = 0; k < N; ++k) yold[k] = y[k];
        ok = be_step(d, rho, e, a.shift, h, yold, y, t, s);
        if (!ok) break;
        Jv += h * be_qdot(s, rho);
        if constexpr (TANGENT) {
            // (I - hJ) dy' = dy + h (df/dtheta v_theta + fT/cv de0)
            double rq[N];
            #pragma unroll
            for (int k = 0; k < N; ++k) rq[k] = 0;
            double dq_theta = 0;   // dqdot/dtheta . v_theta
            static_for<0, NR>([&](auto ic) HF_AI {
                constexpr int I = decltype(ic)::value;
                const double qv = s.q[I] * a.v_lnA[(size_t)c * NR + I];
                static_for<0, TT::NRMAX>([&](auto jc) HF_AI {
                    constexpr int r = TT::reac[I][decltype(jc)::value], p = TT::prod[I][decltype(jc)::value];
                    if constexpr (r >= 0 && r < N) rq[r] -= qv;
                    if constexpr (p >= 0 && p < N) rq[p] += qv;
                });
                dq_theta -= qv * nu_w_dot<TT, I>(d, s.de);
            });
            double b[N];
            #pragma unroll
            for (int k = 0; k < N; ++k) b[k] = dy[k] + h * (d.W[k] / rho * rq[k] + s.fT[k] / s.cv * de0);
            double dqdy[N], dqde;
            be_dqdot(s, rho, dqdy, dqde);
            double M[N * N];
            be_matrix(s, h, M);
            ok = be_factor<N>(M);
            be_solve<N>(M, b);
            double acc = dqde * de0 + dq_theta;
            #pragma unroll
            for (int k = 0; k < N; ++k) {
                dy[k] = b[k];
                acc = fma(dqdy[k], b[k], acc);
            }
            dJ += h * acc;
        }
    }
    if (ok) {
        Jv += s.t;
        if constexpr (TANGENT) {
            double acc = de0 / s.cv;
            #pragma unroll
            for (int k = 0; k < N; ++k) acc -= s.de[k] / s.cv * dy[k];
            dJ += acc;
        }
    }
    if constexpr (TANGENT) {
        a.J[c] = ok ? Jv : NAN;
        a.dJ[c] = ok ? dJ : NAN;
        if (!ok) atomicMin(a.first_failed, (unsigned long long)c);
        return;
    } else {
        double lam[N], gth[NR], ge = 0;
        #pragma unroll
        for (int i = 0; i < NR; ++i) gth[i] = 0;
        #pragma unroll
        for (int k = 0; k < N; ++k) lam[k] = 0;
        const int K = a.ckpt, nseg = (a.nsteps + K - 1) / K;
        bool first = true;
        for (int g = nseg - 1; g >= 0 && ok; --g) {
            const int n0 = g * K, n1 = min(a.nsteps, n0 + K);
            double yy[N], tt = ck(g, N);
            #pragma unroll
            for (int j = 0; j < N; ++j) yy[j] = ck(g, j);
            for (int n = n0; n < n1 && ok; ++n) {
                double yold[N];
                #pragma unroll
                for (int k = 0; k < N; ++k) yold[k] = yy[k];
                ok = be_step(d, rho, e, a.shift, h, yold, yy, tt, s);
                #pragma unroll
                for (int j = 0; j < N; ++j) tr(n - n0, j) = yy[j];
                tr(n - n0, N) = tt;
            }
            for (int n = n1; n > n0 && ok; --n) {
                double yn[N], tn = tr(n - n0 - 1, N);
                #pragma unroll
                for (int j = 0; j < N; ++j) yn[j] = tr(n - n0 - 1, j);
                ok = be_eval(d, rho, e, a.shift, yn, tn, s);
                if (!ok) break;
                if (first) {
                    // terminal term T(y_S): dT/dy_j = -(e_j - e_N)/cv, dT/de = 1/cv
                    #pragma unroll
                    for (int j = 0; j < N; ++j) lam[j] = -s.de[j] / s.cv;
                    ge = 1.0 / s.cv;
                    first = false;
                }
                double dqdy[N], dqde, mu[N];
                be_dqdot(s, rho, dqdy, dqde);
                #pragma unroll
                for (int j = 0; j < N; ++j) mu[j] = fma(h, dqdy[j], lam[j]);
                double muft = 0, wv[N];
                {
                    double M[N * N];
                    be_matrix(s, h, M);
                    ok = be_factor<N>(M);
                    be_solve_t<N>(M, mu);
                }
                #pragma unroll
                for (int k = 0; k < N; ++k) {
                    wv[k] = mu[k] / rho - s.de[k];
                    muft = fma(mu[k], s.fT[k], muft);
                    lam[k] = mu[k];
                }
                static_for<0, NR>([&](auto ic) HF_AI {
                    constexpr int I = decltype(ic)::value;
                    gth[I] = fma(h * s.q[I], nu_w_dot<TT, I>(d, wv), gth[I]);
                });
                ge += h * (muft / s.cv + dqde);
            }
        }
        a.J[c] = ok ? Jv : NAN;
        #pragma unroll
        for (int i = 0; i < NR; ++i) a.g_lnA[(size_t)c * NR + i] = ok ? gth[i] : NAN;
        a.g_T0[c] = ok ? ge * cv0 : NAN;
        #pragma unroll
        for (int k = 0; k < NS; ++k) a.g_Y0[(size_t)c * NS + k] = ok ? (k < N ? lam[k] : 0.0) + ge * ek0[k] : NAN;
        if (!ok) atomicMin(a.first_failed, (unsigned long long)c);
    }
}

// Workspace of the adjoint: checkpoints plus one segment of trajectory.
size_t sens_workspace(int64_t ncell, int nsteps, int ckpt) {
    constexpr int N = Topo::NS - 1;
    const int64_t nseg = (nsteps + ckpt - 1) / ckpt;
    return sizeof(double) * (size_t)(N + 1) * (size_t)ncell * (size_t)(nseg + ckpt);
}

// Lane-resident sensitivities (the N x N Jacobian lives in VGPRs): compiled
// for small mechanisms only.
constexpr bool kSens = Topo::NS <= 16;

int sens(const AdjointLaunch& L, bool tangent) {
    if (L.ncell == 0) return HF_OK;
    if constexpr (!kSens) {
        return HF_EINVAL;
    } else {
    if (!(L.h > 0) || L.nsteps < 1 || (!tangent && L.ckpt_every < 1)) return HF_EINVAL;
    constexpr int N = Topo::NS - 1;
    SensArgs a{};
    a.ncell = (long)L.ncell;
    a.rho = L.rho;
    a.T0 = L.T0;
    a.Y0 = L.Y0;
    a.h = L.h;
    a.nsteps = L.nsteps;
    a.ckpt = tangent ? 1 : L.ckpt_every;
    a.shift = L.lnA_shift;
    a.J = L.J;
    a.g_lnA = L.grad_lnA;
    a.g_T0 = L.grad_T0;
    a.g_Y0 = L.grad_Y0;
    a.v_lnA = L.v_lnA;
    a.v_T0 = L.v_T0;
    a.v_Y0 = L.v_Y0;
    a.dJ = L.dJ;
    a.first_failed = L.first_failed;
    if (!tangent) {
        if (!L.workspace || L.workspace_bytes < sens_workspace(L.ncell, L.nsteps, L.ckpt_every)) return HF_EINVAL;
        const int64_t nseg = (L.nsteps + L.ckpt_every - 1) / L.ckpt_every;
        a.ck = L.workspace;
        a.traj = L.workspace + (size_t)nseg * (N + 1) * (size_t)L.ncell;
    }
    hipLaunchKernelGGL(chem_zero_kernel, dim3(1), dim3(64), 0, L.stream, L.first_failed,
                       static_cast<unsigned long long*>(nullptr));
    const unsigned grid = (unsigned)((L.ncell + 63) / 64);
    if (tangent)
        hipLaunchKernelGGL((sens_kernel<Topo, true>), dim3(grid), dim3(64), 0, L.stream, static_cast<const MD*>(L.mech), a);
    else
        hipLaunchKernelGGL((sens_kernel<Topo, false>), dim3(grid), dim3(64), 0, L.stream, static_cast<const MD*>(L.mech), a);
    return hipGetLastError() == hipSuccess ? HF_OK : HF_EDEVICE;
    }
}

int adjoint(const AdjointLaunch& L) { return sens(L, false); }
int tangent(const AdjointLaunch& L) { return sens(L, true); }

}  // namespace

// The table of host entry points exists in the host compilation only.
#ifndef __HIP_DEVICE_COMPILE__
#define HF_OPS_NAME2(x) ops_##x
#define HF_OPS_NAME(x) HF_OPS_NAME2(x)
extern const TopoOps HF_OPS_NAME(HF_TOPO);
const TopoOps HF_OPS_NAME(HF_TOPO) = {HF_STR(HF_TOPO), Topo::topology, Topo::NS, Topo::NR, sizeof(MD), fill, chem,
                                      chem_workspace, flux, props, adjoint, tangent, sens_workspace,
                                      box_ops, TX2, TY2, TX1, NSHAPE, {SHAPE_TX[0], SHAPE_TX[1], SHAPE_TX[2]},
                                      {SHAPE_TY[0], SHAPE_TY[1], SHAPE_TY[2]}};
#endif

}  // namespace hf
