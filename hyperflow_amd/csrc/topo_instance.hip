// Kernels and launchers for one mechanism topology. Built once per mechanism
// with -DHF_TOPO=<struct name> -DHF_TOPO_INC="<generated header>"; the
// generated header fixes the topology (see mechgen.cpp), the coefficients are
// uploaded at run time.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>

#include "hf_chem.hpp"
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
}

// ------------------------------------------------------------------ chemistry (plain)
template <class TT>
__global__ void __launch_bounds__(64)
chem_plain_kernel(const MechData<TT>* __restrict__ dm, const double* __restrict__ U, long ncell, long ld,
                  const double* __restrict__ dt, int per_cell, double* __restrict__ out, long ldo,
                  int8_t* __restrict__ status, unsigned long long* __restrict__ first_failed, ChemParams p,
                  unsigned long long* __restrict__ counters) {
    constexpr int N = TT::NS - 1;
    __shared__ double zlds[6 * N * 64];
    const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    BdfCounters c;
    if (i < ncell) {
        const double h = per_cell ? dt[i] : dt[0];
        const int st = advance_cell<TT>(*dm, U, i, ld, h, out, ldo, p, c, zlds + threadIdx.x);
        status[i] = (int8_t)st;
        if (st != HF_OK) atomicMin(first_failed, (unsigned long long)i);
    }
    reduce_counters(c, counters);
}

// Register-resident per-lane integration is compiled for small mechanisms
// (the Newton matrix of N = NS-1 unknowns lives in VGPRs).
constexpr bool kLaneChem = Topo::NS <= 16;

int chem(const ChemLaunch& L) {
    if (L.ncell == 0) return HF_OK;
    if constexpr (!kLaneChem) {
        return HF_EINVAL;
    } else {
    const ChemParams p{L.rtol, L.atol, L.maxsteps};
    StreamTimer timer(L.kernel_ms, L.stream);
    const int block = 64;
    const long grid = (L.ncell + block - 1) / block;
    hipLaunchKernelGGL(chem_plain_kernel<Topo>, dim3((unsigned)grid), dim3(block), 0, L.stream,
                       static_cast<const MD*>(L.mech), L.U, (long)L.ncell, (long)L.ld, L.dt, L.dt_per_cell, L.out,
                       (long)L.ldo, L.status, L.first_failed, p, L.counters);
    if (hipGetLastError() != hipSuccess) return HF_EDEVICE;
    timer.stop();
    return HF_OK;
    }
}

size_t chem_workspace(int64_t, int, int, int) { return 0; }

// ------------------------------------------------------------------ flux stage
constexpr int TX2 = 15, TY2 = 15, TX1 = 255;

// The LDS-staged tiles hold (TX+4)(TY+4) cells of NV doubles: built for small mechanisms.
constexpr bool kFlux = Topo::NS <= 16;

int flux(const FluxLaunch& L) {
    if (L.ntiles == 0) return HF_OK;
    if constexpr (!kFlux) {
        return HF_EINVAL;
    } else {
    StreamTimer timer(L.kernel_ms, L.stream);
    if (L.one_d) {
        hipLaunchKernelGGL((flux_stage_kernel<Topo, TX1, 1, false>), dim3(L.ntiles),
                           dim3(FluxTile<Topo, TX1, 1, false>::THREADS), 0, L.stream, static_cast<const MD*>(L.mech),
                           L.boxes, L.tiles, L.ntiles, L.Uin, L.Uout, L.first_order, L.first_failed, L.counters);
    } else {
        hipLaunchKernelGGL((flux_stage_kernel<Topo, TX2, TY2, true>), dim3(L.ntiles),
                           dim3(FluxTile<Topo, TX2, TY2, true>::THREADS), 0, L.stream, static_cast<const MD*>(L.mech),
                           L.boxes, L.tiles, L.ntiles, L.Uin, L.Uout, L.first_order, L.first_failed, L.counters);
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

int adjoint(const AdjointLaunch&) { return HF_EINVAL; }
int tangent(const AdjointLaunch&) { return HF_EINVAL; }

}  // namespace

// The table of host entry points exists in the host compilation only.
#ifndef __HIP_DEVICE_COMPILE__
#define HF_OPS_NAME2(x) ops_##x
#define HF_OPS_NAME(x) HF_OPS_NAME2(x)
extern const TopoOps HF_OPS_NAME(HF_TOPO);
const TopoOps HF_OPS_NAME(HF_TOPO) = {HF_STR(HF_TOPO), Topo::topology, Topo::NS, Topo::NR, sizeof(MD), fill, chem,
                                      chem_workspace, flux, props, adjoint, tangent};
#endif

}  // namespace hf
