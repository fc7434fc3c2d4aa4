// Flux-sweep kernel: one Euler stage of BlockAMR::euler's face loop and
// update (HyperFlow_V3/src/BlockAMR.cpp:210-243) over a table of tiles of
// ragged ghosted SoA boxes.
//
// A workgroup owns a TX x TY tile of cells. It stages the ghosted tile
// (TX+4) x (TY+4) cells x NV components into LDS once, then runs the two
// directions in turn: every lane computes exactly one face (NND/minmod
// reconstruction of both sides, admissibility + EOS by one energy->T Newton
// per side, HLLC, wall zeroing) into an LDS flux tile, and every cell lane
// accumulates its flux difference in registers. The updated cell is
// validated (properties) before it is written. With TX = TY = 15 the
// workgroup is 16 x 16 = 256 lanes and no lane is idle in a face phase.
#pragma once
#include "hf_device.hpp"
#include "hyperflow.h"

namespace hf {

template <class T, int TX, int TY, bool TWO_D>
struct FluxTile {
    static constexpr int NS = T::NS, NV = NS + 3;
    static constexpr int GX = TX + 4, GY = TWO_D ? TY + 4 : TY;  // staged rows
    static constexpr int BX = TX + 1, BY = TWO_D ? TY + 1 : TY;  // lanes
    static constexpr int THREADS = BX * BY;
    static constexpr int LDS_IN = NV * GX * GY;
    static constexpr int LDS_FLUX = NV * BX * BY;
};

// Reconstructed face state and its properties; falls back to the cell value
// (first order) when the reconstruction is inadmissible (BlockAMR.cpp:221-222).
// Returns false when even the cell value is inadmissible. t_hint: the
// temperature of the adjacent cell, the warm start of the inversion.
template <class T>
__device__ __forceinline__ bool face_side(const MechData<T>& d, double* s, const double* cell, Props& pr,
                                          int& fallbacks, double t_hint) {
    constexpr int NV = T::NS + 3;
    if (properties(d, s, pr, t_hint)) return true;
    ++fallbacks;
    #pragma unroll
    for (int c = 0; c < NV; ++c) s[c] = cell[c];
    return properties(d, s, pr, t_hint);
}

template <class T, int TX, int TY, bool TWO_D>
__global__ void __launch_bounds__((TX + 1) * (TWO_D ? TY + 1 : TY))
flux_stage_kernel(const MechData<T>* __restrict__ dm, const hf_box* __restrict__ boxes,
                  const int32_t* __restrict__ tiles, int ntiles, const double* __restrict__ Uin,
                  double* __restrict__ Uout, int first_order, unsigned long long* __restrict__ first_failed,
                  unsigned long long* __restrict__ counters, const hf_face_fluxes* __restrict__ fout,
                  const double* const* __restrict__ in_ptrs, double* const* __restrict__ out_ptrs) {
    using FT = FluxTile<T, TX, TY, TWO_D>;
    constexpr int NS = T::NS, NV = NS + 3, E = NS + 2;
    constexpr int GX = FT::GX, GY = FT::GY, BX = FT::BX;
    __shared__ double s_in[FT::LDS_IN];
    __shared__ double s_fx[FT::LDS_FLUX];
    __shared__ double s_t[GX * GY];   // temperature of each staged cell (warm starts)
    const MechData<T>& d = *dm;

    // XCD-aware tile order: workgroups are dealt round-robin to the 8 XCDs,
    // so give each XCD a contiguous run of tiles (neighbours share L2 halos).
    // The grid is a multiple of 8, so (blockIdx % 8, blockIdx / 8) covers
    // every tile index below 8 * per exactly once.
    const int per = gridDim.x / 8;
    const int t = (blockIdx.x % 8) * per + blockIdx.x / 8;
    if (t >= ntiles) return;
    const int bi = tiles[3 * t], ti = tiles[3 * t + 1], tj = tiles[3 * t + 2];
    const hf_box B = boxes[bi];
    // the box's own arrays when the caller passed a pointer per box
    const double* __restrict__ src = in_ptrs ? in_ptrs[bi] : Uin;
    double* __restrict__ dst = out_ptrs ? out_ptrs[bi] : Uout;
    const int i0 = ti * TX, j0 = tj * TY;
    const int lane = threadIdx.x;
    const int lx = lane % BX, ly = lane / BX;

    // ---- stage the ghosted tile: rows j0-2 .. j0+TY+1 (2-D) or j0 .. j0+TY-1 (1-D)
    // Eight loads per lane are issued before their LDS stores: one load per
    // trip left every trip waiting out a full HBM round trip.
    const int jlo = TWO_D ? j0 - 2 : j0;
    constexpr int SB = 8;
    for (int e0 = lane; e0 < NV * GX * GY; e0 += SB * FT::THREADS) {
        double v[SB];
        #pragma unroll
        for (int k = 0; k < SB; ++k) {
            const int e = e0 + k * FT::THREADS;
            const int n = e / (GX * GY);
            const int r = e - n * GX * GY;
            const int gj = r / GX, gi = r - gj * GX;
            const int ci = i0 - 2 + gi, cj = jlo + gj;
            v[k] = 0;
            if (e < NV * GX * GY && ci < B.nx + 2 && cj < B.ny + 2)
                v[k] = src[B.in_offset + n * B.in_plane + (int64_t)(cj + 2) * B.in_pitch + (ci + 2)];
        }
        #pragma unroll
        for (int k = 0; k < SB; ++k)
            if (e0 + k * FT::THREADS < NV * GX * GY) s_in[e0 + k * FT::THREADS] = v[k];
    }
    __syncthreads();

    auto cellv = [&](int gi, int gj, int n) HF_AI -> double { return s_in[n * GX * GY + gj * GX + gi]; };
    for (int e = lane; e < GX * GY; e += FT::THREADS) {
        double u[NV];
        #pragma unroll
        for (int c = 0; c < NV; ++c) u[c] = s_in[c * GX * GY + e];
        s_t[e] = state_temperature(d, u, 1000.0);
    }
    __syncthreads();
    int fallbacks = 0, faces = 0;
    bool fail = false;
    unsigned long long fail_idx = ~0ull;
    double acc[NV];
    const int ci = i0 + lx, cj = j0 + ly;             // this lane's cell
    const bool owns_cell = lx < TX && ly < TY && ci < B.nx && cj < B.ny;
    const int gyc = TWO_D ? ly + 2 : ly;               // staged row of this lane's cell

    // ---- one face: stencil cells (a, b | c, d) along direction DIR
    auto face = [&](auto dirc, int fi, int fj, bool active) HF_AI {
        constexpr int DIR = decltype(dirc)::value;
        constexpr int NRM = DIR == 0 ? NS : NS + 1, TAN = DIR == 0 ? NS + 1 : NS;
        double f[NV];
        #pragma unroll
        for (int c = 0; c < NV; ++c) f[c] = 0;
        if (active) {
            // staged coordinates of cell c (the face's right/upper cell)
            const int gi = (fi - i0) + 2, gj = TWO_D ? (fj - jlo) : (fj - j0);
            const int di = DIR == 0, dj = DIR == 1;
            double qb[NV], qc[NV], L[NV], R[NV];
            #pragma unroll
            for (int c = 0; c < NV; ++c) {
                const double a = cellv(gi - 2 * di, gj - 2 * dj, c);
                const double b = cellv(gi - di, gj - dj, c);
                const double cc = cellv(gi, gj, c);
                const double dd = cellv(gi + di, gj + dj, c);
                qb[c] = b;
                qc[c] = cc;
                if (first_order) {
                    L[c] = b;
                    R[c] = cc;
                } else {
                    L[c] = b + 0.5 * minmod(b - a, cc - b);
                    R[c] = cc - 0.5 * minmod(cc - b, dd - cc);
                }
            }
            Props pl, pr;
            const bool okl = face_side(d, L, qb, pl, fallbacks, s_t[(gj - dj) * GX + gi - di]);
            const bool okr = face_side(d, R, qc, pr, fallbacks, s_t[gj * GX + gi]);
            if (!okl || !okr) {
                fail = true;
                const int bad_i = okl ? fi : fi - di, bad_j = okl ? fj : fj - dj;
                const int ii = min(max(bad_i, 0), B.nx - 1), jj = min(max(bad_j, 0), B.ny - 1);
                fail_idx = min(fail_idx, ((unsigned long long)bi << 32) | (unsigned)(jj * B.nx + ii));
            } else {
                hllc<NS>(L, R, NRM, pl, pr, f);
                const int coord = DIR == 0 ? fi : fj, len = DIR == 0 ? B.nx : B.ny;
                const bool lo = (B.walls >> (2 * DIR)) & 1, hi = (B.walls >> (2 * DIR + 1)) & 1;
                const bool wall = (coord == 0 && lo && (DIR == 0 || fi >= B.wedge_i)) || (coord == len && hi);
                if (wall) {
                    #pragma unroll
                    for (int k = 0; k < NS; ++k) f[k] = 0;
                    f[TAN] = 0;
                    f[E] = 0;
                }
            }
            ++faces;
            // face-flux output for the flux registers (a face two tiles share is
            // computed by both from the same staged cells: identical values)
            if (fout) {
                const hf_face_fluxes& O = fout[bi];
                double* dst = DIR == 0 ? O.fx : O.fy;
                const int64_t off = DIR == 0 ? O.fx_offset : O.fy_offset;
                const int64_t pitch = DIR == 0 ? O.fx_pitch : O.fy_pitch;
                const int64_t plane = DIR == 0 ? O.fx_plane : O.fy_plane;
                if (dst) {
                    #pragma unroll
                    for (int c = 0; c < NV; ++c) dst[off + c * plane + (int64_t)fj * pitch + fi] = f[c];
                }
            }
        }
        #pragma unroll
        for (int c = 0; c < NV; ++c) s_fx[c * FT::THREADS + lane] = f[c];
    };

    // ---- x faces: lane (lx, ly) -> face lx of row ly
    {
        const int fi = i0 + lx, fj = j0 + ly;
        face(std::integral_constant<int, 0>{}, fi, fj, ly < TY && fi <= B.nx && fj < B.ny && lx <= TX);
    }
    __syncthreads();
    if (owns_cell) {
        #pragma unroll
        for (int c = 0; c < NV; ++c) {
            const double q = cellv(lx + 2, gyc, c);
            acc[c] = q - B.dtdx * (s_fx[c * FT::THREADS + lane + 1] - s_fx[c * FT::THREADS + lane]);
        }
    }
    if constexpr (TWO_D) {
        __syncthreads();
        // ---- y faces: lane (lx, ly) -> face ly of column lx
        const int fi = i0 + lx, fj = j0 + ly;
        face(std::integral_constant<int, 1>{}, fi, fj, lx < TX && fi < B.nx && fj <= B.ny && ly <= TY);
        __syncthreads();
        if (owns_cell) {
            #pragma unroll
            for (int c = 0; c < NV; ++c)
                acc[c] = acc[c] - B.dtdy * (s_fx[c * FT::THREADS + lane + BX] - s_fx[c * FT::THREADS + lane]);
        }
    }
    if (owns_cell) {
        #pragma unroll
        for (int c = 0; c < NV; ++c)
            dst[B.out_offset + c * B.out_plane + (int64_t)cj * B.out_pitch + ci] = acc[c];
        Props p;
        if (!properties(d, acc, p, s_t[gyc * GX + lx + 2])) {
            fail = true;
            fail_idx = min(fail_idx, ((unsigned long long)bi << 32) | (unsigned)(cj * B.nx + ci));
        }
    }
    if (fail) atomicMin(first_failed, fail_idx);
    // counters: faces computed, first-order sides
    unsigned long long v0 = faces, v1 = fallbacks;
    #pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        v0 += __shfl_xor(v0, off, 64);
        v1 += __shfl_xor(v1, off, 64);
    }
    if ((lane & 63) == 0) {
        atomicAdd(&counters[0], v0);
        if (v1) atomicAdd(&counters[1], v1);
    }
}

}  // namespace hf
