// Flux-sweep kernel: one Euler stage of BlockAMR::euler's face loop and
// update (HyperFlow_V3/src/BlockAMR.cpp:210-243) over a table of tiles of
// ragged ghosted SoA boxes.
//
// A workgroup owns a TX x TY tile of cells. It stages the ghosted tile
// (TX+4) x (TY+4) cells x NV components into LDS once (each lane loads all
// components of its cells, so the address arithmetic is paid once per cell),
// and the temperature of every staged cell that warm-starts an inversion.
// Then it runs the two directions in turn: every lane computes exactly one
// face (NND/minmod reconstruction of both sides, admissibility + EOS by one
// energy->T Newton per side, HLLC, wall zeroing), and every cell lane takes
// the flux of its other face from the neighbouring lane: along x that lane
// is in the same 16/32/8-lane row of the wave (a cross-lane shuffle), along y
// it is BX lanes on, in the same wave except for a wave's last row, whose
// faces come through a small LDS buffer. The flux differences accumulate in
// registers into the cell update, which is written once and (optionally)
// validated. With TX = TY = 15 the workgroup is 16 x 16 = 256 lanes and no
// lane is idle in a face phase. The 1-D form (TX x 1, one row of 256 or 64
// lanes) passes its face fluxes through LDS.
//
// The stage is bound by issuing FP64 instructions (the inversions), so the
// code is kept small: the two directions and the four inversions of a face
// run as loops over one copy of the face code, and the kernel carries the
// mechanism's thermodynamic form only (the mixture polynomial when every
// species shares Tmid), not both behind a run-time test.
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
    // 2-D: a wave holds RPW whole lane rows (BX divides 64); the y faces of
    // the first row of every wave but the first pass through LDS
    static constexpr int RPW = TWO_D ? 64 / BX : 1;
    static constexpr int LDS_FLUX = TWO_D ? NV * BX * (BY / RPW) : NV * THREADS;
    // staged cells whose temperature warm-starts an inversion: rows 2..TY+1 x
    // columns 1..TX+2 (x faces, validation) and rows 1 and TY+2 x columns
    // 2..TX+1 (y faces); the 1-D form: columns 1..TX+2
    static constexpr int NT_A = TWO_D ? TY * (TX + 2) : TX + 2;
    static constexpr int NT = TWO_D ? NT_A + 2 * TX : NT_A;
    static_assert(!TWO_D || (64 % BX == 0 && BY % RPW == 0), "2-D lane rows must tile the waves");
};

// One tile of the stage; TF selects the thermodynamic form the code carries
// (the kernel below picks the one the mechanism needs), so the loop bodies
// hold a single copy of the inversion each: the stage is bound by issuing
// its instructions, and a kernel that outgrows the instruction cache stalls
// on fetches.
template <class T, int TX, int TY, bool TWO_D, int TF>
__device__ __forceinline__ void flux_tile(const MechData<T>& d, const hf_box* __restrict__ boxes,
                                          const int32_t* __restrict__ tiles, int t, const double* __restrict__ Uin,
                                          double* __restrict__ Uout, int first_order, int validate,
                                          unsigned long long* __restrict__ first_failed,
                                          unsigned long long* __restrict__ counters,
                                          const hf_face_fluxes* __restrict__ fout,
                                          const double* const* __restrict__ in_ptrs,
                                          double* const* __restrict__ out_ptrs, double* s_in, double* s_fx,
                                          double* s_t) {
    using FT = FluxTile<T, TX, TY, TWO_D>;
    constexpr int NS = T::NS, NV = NS + 3, E = NS + 2;
    constexpr int GX = FT::GX, GY = FT::GY, BX = FT::BX, RPW = FT::RPW;
    const int bi = tiles[3 * t], ti = tiles[3 * t + 1], tj = tiles[3 * t + 2];
    const hf_box B = boxes[bi];
    // the box's own arrays when the caller passed a pointer per box, addressed
    // in the global address space (global, not flat, loads and stores)
    typedef const __attribute__((address_space(1))) double GIn;
    typedef __attribute__((address_space(1))) double GOut;
    const GIn* __restrict__ src = (const GIn*)(in_ptrs ? in_ptrs[bi] : Uin);
    GOut* __restrict__ dst = (GOut*)(out_ptrs ? out_ptrs[bi] : Uout);
    const int i0 = ti * TX, j0 = tj * TY;
    const int lane = threadIdx.x;
    const int lx = lane % BX, ly = lane / BX;

    // ---- stage the ghosted tile: rows j0-2 .. j0+TY+1 (2-D) or j0 .. j0+TY-1 (1-D).
    // A lane loads every component of its staged cell before storing any, so
    // NV loads are in flight per trip.
    const int jlo = TWO_D ? j0 - 2 : j0;
    for (int e = lane; e < GX * GY; e += FT::THREADS) {
        const int gj = e / GX, gi = e - gj * GX;
        const int ci = i0 - 2 + gi, cj = jlo + gj;
        const bool in = ci < B.nx + 2 && cj < B.ny + 2;
        // cells past the box's last ghost layer stage zeros; their loads read
        // the last ghost cell instead (an address inside the array), so every
        // lane issues the same unconditional loads
        const int cil = min(ci, B.nx + 1), cjl = min(cj, B.ny + 1);
        const GIn* __restrict__ q = src + B.in_offset + (int64_t)(cjl + 2) * B.in_pitch + (cil + 2);
        double v[NV];
        #pragma unroll
        for (int n = 0; n < NV; ++n) {
            const double x = q[n * B.in_plane];
            v[n] = in ? x : 0.0;
        }
        #pragma unroll
        for (int n = 0; n < NV; ++n) s_in[n * GX * GY + e] = v[n];
    }
    __syncthreads();

    auto cellv = [&](int gi, int gj, int n) HF_AI -> double { return s_in[n * GX * GY + gj * GX + gi]; };
    for (int e = lane; e < FT::NT; e += FT::THREADS) {
        int gi, gj;
        if (!TWO_D) {
            gi = 1 + e;
            gj = 0;
        } else if (e < FT::NT_A) {
            gj = 2 + e / (TX + 2);
            gi = 1 + e % (TX + 2);
        } else {
            const int r = e - FT::NT_A;
            gj = r < TX ? 1 : TY + 2;
            gi = 2 + (r < TX ? r : r - TX);
        }
        double u[NV];
        #pragma unroll
        for (int c = 0; c < NV; ++c) u[c] = cellv(gi, gj, c);
        s_t[gj * GX + gi] = state_temperature<T, TF>(d, u, 1000.0);
    }
    __syncthreads();
    int fallbacks = 0, faces = 0;
    bool fail = false;
    unsigned long long fail_idx = ~0ull;
    const int ci = i0 + lx, cj = j0 + ly;             // this lane's cell
    const bool owns_cell = lx < TX && ly < TY && ci < B.nx && cj < B.ny;
    const int gyc = TWO_D ? ly + 2 : ly;               // staged row of this lane's cell
    double acc[NV];
    #pragma unroll
    for (int c = 0; c < NV; ++c) acc[c] = cellv(lx + 2, gyc, c);

    // ---- the faces, x then y (a loop, so the face code exists once): lane
    // (lx, ly) computes x face lx of row ly, then y face ly of column lx.
    // The y face is solved in the face-normal frame (the two momenta swapped
    // on the way into HLLC and back), which is the same arithmetic.
    #pragma nounroll
    for (int dir = 0; dir < (TWO_D ? 2 : 1); ++dir) {
        const int di = dir == 0, dj = dir;
        const int fi = i0 + lx, fj = j0 + ly;
        const bool active = dir == 0 ? ly < TY && fi <= B.nx && fj < B.ny && lx <= TX
                                     : lx < TX && fi < B.nx && fj <= B.ny && ly <= TY;
        double f[NV];
        #pragma unroll
        for (int c = 0; c < NV; ++c) f[c] = 0;
        if (active) {
            // staged cells of the stencil (a, b | c, d): b and c beside the face
            const int gi = lx + 2, gj = TWO_D ? ly + 2 : ly;
            const int sb = (gj - dj) * GX + gi - di, sc = gj * GX + gi, step = dir == 0 ? 1 : GX;
            const double tl = s_t[sb], tr = s_t[sc];
            double L[NV], R[NV];
            Props pl, pr;
            bool okl = false, okr = false;
            // four inversions through one copy of the code: the left side, its
            // first-order retry (BlockAMR.cpp:221), the right side, its retry
            #pragma nounroll
            for (int task = 0; task < 4; ++task) {
                const bool right = task >= 2, retry = task & 1;
                if (retry && (right ? okr : okl)) continue;
                double S[NV];
                if (right) {
                    #pragma unroll
                    for (int c = 0; c < NV; ++c) {
                        const double* w = s_in + c * GX * GY;
                        const double b = w[sb], cc = w[sc], dd = w[sc + step];
                        S[c] = retry || first_order ? cc : cc - 0.5 * minmod(cc - b, dd - cc);
                    }
                } else {
                    #pragma unroll
                    for (int c = 0; c < NV; ++c) {
                        const double* w = s_in + c * GX * GY;
                        const double a = w[sb - step], b = w[sb], cc = w[sc];
                        S[c] = retry || first_order ? b : b + 0.5 * minmod(b - a, cc - b);
                    }
                }
                Props ps;
                const bool ok = properties<T, TF>(d, S, ps, right ? tr : tl);
                if (retry) ++fallbacks;
                if (right) {
                    okr = ok;
                    pr = ps;
                    #pragma unroll
                    for (int c = 0; c < NV; ++c) R[c] = S[c];
                } else {
                    okl = ok;
                    pl = ps;
                    #pragma unroll
                    for (int c = 0; c < NV; ++c) L[c] = S[c];
                }
            }
            if (!okl || !okr) {
                fail = true;
                const int bad_i = okl ? fi : fi - di, bad_j = okl ? fj : fj - dj;
                const int ii = min(max(bad_i, 0), B.nx - 1), jj = min(max(bad_j, 0), B.ny - 1);
                fail_idx = min(fail_idx, ((unsigned long long)bi << 32) | (unsigned)(jj * B.nx + ii));
            } else {
                if (dir) {
                    const double l0 = L[NS], r0 = R[NS];
                    L[NS] = L[NS + 1];
                    L[NS + 1] = l0;
                    R[NS] = R[NS + 1];
                    R[NS + 1] = r0;
                }
                hllc<NS>(L, R, NS, pl, pr, f);
                const int coord = dir == 0 ? fi : fj, len = dir == 0 ? B.nx : B.ny;
                const bool lo = (B.walls >> (2 * dir)) & 1, hi = (B.walls >> (2 * dir + 1)) & 1;
                const bool wall = (coord == 0 && lo && (dir == 0 || fi >= B.wedge_i)) || (coord == len && hi);
                if (wall) {
                    #pragma unroll
                    for (int k = 0; k < NS; ++k) f[k] = 0;
                    f[NS + 1] = 0;             // tangential momentum (face-normal frame)
                    f[E] = 0;
                }
                if (dir) {
                    const double f0 = f[NS];
                    f[NS] = f[NS + 1];
                    f[NS + 1] = f0;
                }
            }
            ++faces;
            // face-flux output for the flux registers (a face two tiles share is
            // computed by both from the same staged cells: identical values)
            if (fout) {
                const hf_face_fluxes& O = fout[bi];
                GOut* fdst = (GOut*)(dir == 0 ? O.fx : O.fy);
                const int64_t off = dir == 0 ? O.fx_offset : O.fy_offset;
                const int64_t pitch = dir == 0 ? O.fx_pitch : O.fy_pitch;
                const int64_t plane = dir == 0 ? O.fx_plane : O.fy_plane;
                if (fdst) {
                    #pragma unroll
                    for (int c = 0; c < NV; ++c) fdst[off + c * plane + (int64_t)fj * pitch + fi] = f[c];
                }
            }
        }
        // ---- the cell's other face: along x the next lane's (one row of the
        // wave); along y the lane BX on, through LDS when that lane is in the
        // next wave. Every lane runs the shuffles.
        if (!TWO_D) {
            #pragma unroll
            for (int c = 0; c < NV; ++c) s_fx[c * FT::THREADS + lane] = f[c];
            __syncthreads();
            #pragma unroll
            for (int c = 0; c < NV; ++c)
                acc[c] = acc[c] - B.dtdx * (s_fx[c * FT::THREADS + lane + 1] - f[c]);
        } else if (dir == 0) {
            #pragma unroll
            for (int c = 0; c < NV; ++c) {
                const double fr = __shfl_down(f[c], 1, BX);
                acc[c] = acc[c] - B.dtdx * (fr - f[c]);
            }
        } else {
            const int wrow = ly % RPW, slot = ly / RPW;
            if (wrow == 0 && slot > 0) {
                #pragma unroll
                for (int c = 0; c < NV; ++c) s_fx[(c * (FT::BY / RPW) + slot - 1) * BX + lx] = f[c];
            }
            __syncthreads();
            #pragma unroll
            for (int c = 0; c < NV; ++c) {
                double gu = __shfl_down(f[c], BX, 64);
                if (wrow == RPW - 1) gu = s_fx[(c * (FT::BY / RPW) + slot) * BX + lx];
                acc[c] = acc[c] - B.dtdy * (gu - f[c]);
            }
        }
    }
    if (owns_cell) {
        #pragma unroll
        for (int c = 0; c < NV; ++c)
            dst[B.out_offset + c * B.out_plane + (int64_t)cj * B.out_pitch + ci] = acc[c];
        Props p;
        if (validate && !properties<T, TF>(d, acc, p, s_t[gyc * GX + lx + 2])) {
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

// Three waves per SIMD (at most 168 VGPRs) for the 256-lane 2-D tiles: the
// stage is a long run of dependent FP64 chains, and a third resident wave
// keeps the FP64 pipe issuing (three workgroups of ~44 KB fit a CU's LDS).
// Measured on the 4096 x 1024 stage: 1.01 ms against 1.25 ms at two waves.
template <class T, int TX, int TY, bool TWO_D>
__global__ void __launch_bounds__((TX + 1) * (TWO_D ? TY + 1 : TY))
__attribute__((amdgpu_waves_per_eu(TWO_D && (TX + 1) * (TY + 1) == 256 ? 3 : 1)))
flux_stage_kernel(const MechData<T>* __restrict__ dm, const hf_box* __restrict__ boxes,
                  const int32_t* __restrict__ tiles, int ntiles, const double* __restrict__ Uin,
                  double* __restrict__ Uout, int first_order, int validate,
                  unsigned long long* __restrict__ first_failed,
                  unsigned long long* __restrict__ counters, const hf_face_fluxes* __restrict__ fout,
                  const double* const* __restrict__ in_ptrs, double* const* __restrict__ out_ptrs) {
    using FT = FluxTile<T, TX, TY, TWO_D>;
    __shared__ double s_in[FT::LDS_IN];
    __shared__ double s_fx[FT::LDS_FLUX];
    __shared__ double s_t[FT::GX * FT::GY];   // temperature of the staged cells that warm-start inversions
    // XCD-aware tile order: workgroups are dealt round-robin to the 8 XCDs,
    // so give each XCD a contiguous run of tiles (neighbours share L2 halos).
    // The grid is a multiple of 8, so (blockIdx % 8, blockIdx / 8) covers
    // every tile index below 8 * per exactly once.
    const int per = gridDim.x / 8;
    const int t = (blockIdx.x % 8) * per + blockIdx.x / 8;
    if (t >= ntiles) return;
    if (dm->common_tmid)
        flux_tile<T, TX, TY, TWO_D, kThermoPoly>(*dm, boxes, tiles, t, Uin, Uout, first_order, validate,
                                                 first_failed, counters, fout, in_ptrs, out_ptrs, s_in, s_fx, s_t);
    else
        flux_tile<T, TX, TY, TWO_D, kThermoSpecies>(*dm, boxes, tiles, t, Uin, Uout, first_order, validate,
                                                    first_failed, counters, fout, in_ptrs, out_ptrs, s_in, s_fx,
                                                    s_t);
}

}  // namespace hf
