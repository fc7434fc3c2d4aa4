// Interface between the host C-ABI (capi.cpp) and the kernels compiled for
// one mechanism topology (topo_instance.hip, built once per mechanism).
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <string>

#include "hyperflow.h"
#include "mechanism.hpp"

namespace hf {

// Where cell c of a chemistry batch lives. Flat form (blocks null): cell c
// of SoA planes U[n*ld + c]. Block form: the cells of blocks[0..nblock) in
// block order, then j, then i (the order BlockAMR::react gathers them,
// BlockAMR.cpp:271-281); start[b] is the first cell id of block b (device
// prefix sums, start[nblock] = total). A block cell whose mask is not 1 (not
// a leaf) is no part of the batch.
struct CellMap {
    const double* U;
    int64_t ld;
    const hf_cells* blocks;
    const int64_t* start;
    int nblock;
    // component 0 of cell c and the plane stride; false for a non-leaf cell
    __device__ __forceinline__ bool locate(int64_t c, const double*& q, int64_t& plane) const {
        if (!blocks) {
            q = U + c;
            plane = ld;
            return true;
        }
        int lo = 0, hi = nblock - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (start[mid] <= c) lo = mid;
            else hi = mid - 1;
        }
        const hf_cells& B = blocks[lo];
        const int64_t k = c - start[lo];
        const int64_t j = k / B.nx, i = k - j * B.nx;
        if (B.mask && B.mask[j * B.mask_pitch + i] != 1) return false;
        q = B.U + B.offset + j * B.pitch + i;
        plane = B.plane;
        return true;
    }
    // block form: the block of cell c and its position in it
    __device__ __forceinline__ const hf_cells& find(int64_t c, int& b, int& i, int& j) const {
        int lo = 0, hi = nblock - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (start[mid] <= c) lo = mid;
            else hi = mid - 1;
        }
        const hf_cells& B = blocks[lo];
        const int64_t k = c - start[lo];
        b = lo;
        j = (int)(k / B.nx);
        i = (int)(k - (int64_t)j * B.nx);
        return B;
    }
};

struct ChemLaunch {
    const void* mech;               // device MechData<T>
    const double* U;                // device, nv planes, stride ld
    int64_t ncell, ld;
    const hf_cells* blocks;         // optional device block table (CellMap); U, ld unused then
    const int64_t* block_start;
    int nblock;
    const double* dt;               // device: 1 or ncell values
    int dt_per_cell;
    double* out;                    // device, ns planes, stride ldo
    int64_t ldo;
    int8_t* status;                 // device, ncell
    unsigned long long* first_failed;  // device scalar (atomicMin)
    unsigned long long* counters;   // device, 8 values
    double rtol, atol;
    int maxsteps;
    int mode;                       // HF_CHEM_PLAIN / HF_CHEM_BATCHED
    int precision;
    // batched form
    int num_cus, waves_per_cu, chunk, nbins;
    int jac_wait;                   // lane form: trips a Jacobian request may wait for others (0: none)
    int setup_wait;                 // lane form: the same for a Newton-matrix setup without a Jacobian
    void* workspace;                // device scratch for the batched form
    size_t workspace_bytes;
    hipStream_t stream;
    float* kernel_ms;               // optional: measured with events around the integration kernel
    double* commit;                 // optional device U: species planes written back (stride commit_ld)
    int64_t commit_ld;              //   only when no cell failed (all-or-nothing on the device)
};

struct FluxLaunch {
    const void* mech;
    const hf_box* boxes;            // device copy of the box table
    int nbox;
    const int32_t* tiles;           // device tile table: (box, tile_i, tile_j) triples
    int ntiles;
    const double* Uin;
    double* Uout;
    int first_order;
    // 1: the stage does not validate the updated cells (a caller that repairs
    // and validates them with its own kernels); first_failed then reports only
    // inadmissible cell values met by the face loop
    int skip_validation;
    int one_d;                      // x-only stage: tiles are flux_tx1 x 1, else flux_tx2 x flux_ty2
    int shape;                      // 2-D tile shape index (flux_shape_tx/ty), 0 the default
    unsigned long long* first_failed;  // device scalar
    unsigned long long* counters;   // device: [faces, first-order sides]
    const hf_face_fluxes* fluxes;   // device copy of the per-box face-flux outputs, or null
    // device tables of per-box base pointers (box b reads in_ptrs[b] + in_offset ...), or null for Uin/Uout
    const double* const* in_ptrs;
    double* const* out_ptrs;
    hipStream_t stream;
    float* kernel_ms;
};

struct PropsLaunch {
    const void* mech;
    const double* U;
    int64_t ncell, ld;
    double *p, *a, *T;
    unsigned long long* first_failed;
    double dx, dy;                  // for the CFL rate (dx > 0)
    double* rate_max;               // device scalar (max of the CFL rate), optional
    hipStream_t stream;
};

struct AdjointLaunch {
    const void* mech;
    int64_t ncell;
    const double *rho, *T0, *Y0;    // device
    double h;
    int nsteps, ckpt_every;
    const double* lnA_shift;        // device, nr (or null)
    double* J;
    double *grad_lnA, *grad_T0, *grad_Y0;
    const double *v_lnA, *v_T0, *v_Y0;  // tangent directions (tangent mode)
    double* dJ;
    double* workspace;              // checkpoints + one segment of trajectory (adjoint)
    size_t workspace_bytes;
    unsigned long long* first_failed;   // device scalar
    hipStream_t stream;
};

// Single-box driver operations (ghost fill, roundoff repair, validation and
// CFL rate) on a ghosted box in device memory.
struct BoxLaunch {
    enum Op { FILL, REPAIR, PROPS, INTERP, AVGDOWN, REFLUX, TAG, CELLS_PROPS, CELLS_REPAIR, REFLUX_CELLS };
    const void* mech;
    double* U;
    hf_box box;
    int op;
    int bc[4];                      // x-lo, x-hi, y-lo, y-hi boundary codes (y < 0: no y ghosts)
    int two_d;
    const double* inlet;            // FILL: host inflow state (ns+3 values), null without HF_BC_INFLOW
    double dx, dy;
    // INTERP (coarse -> fine) and AVGDOWN (fine -> coarse): source and target patches
    const double* src;
    double* dst;
    hf_patch sp, dp;
    int region[4];                  // INTERP, REFLUX: fine cells; AVGDOWN: coarse cells (ilo, jlo, ihi, jhi)
    hf_face_fluxes cflux, fflux;    // REFLUX: coarse and fine face fluxes (level index spaces)
    int sides;                      // REFLUX: bit0 x-lo, bit1 x-hi, bit2 y-lo, bit3 y-hi
    double h;                       // REFLUX: stage time step (dx, dy: coarse cell size)
    const double* hp;               // REFLUX_CELLS: device h, read in place of h when set
    hf_tag_params tag;              // TAG: criteria
    uint8_t* tags;                  // TAG: device output, nx*ny
    // CELLS_PROPS, CELLS_REPAIR: a device block table (CellMap's block form), total cells
    const hf_cells* blocks;
    const int64_t* block_start;
    int nblock;
    int64_t total;
    double* partials;               // CELLS_REPAIR: device, one NS-vector per workgroup
    // AVGDOWN, REFLUX_CELLS: the coarse level's mask (2: covered by the fine level);
    // INTERP: the fine level's mask (children are written only where it is 0)
    const uint8_t* mask;
    int64_t mask_pitch;
    int mask_nx, mask_ny;           // the mask's extent (cells outside it are not written / not scanned)
    const int32_t* regions;         // INTERP: device list of nregion fine regions (ilo, jlo, ihi, jhi), or null
    int nregion;
    unsigned long long* scal;       // device scratch (repair: 3 NS + 2 words; props: rate bits;
                                    // interp: limited-cell count)
    unsigned long long* first_failed;
    hipStream_t stream;
};

struct TopoOps {
    const char* name;
    const char* topology;
    int ns, nr;
    size_t mech_bytes;
    void (*fill)(const Mechanism& m, void* host);   // writes a MechData<T>
    int (*chem)(const ChemLaunch& L);
    size_t (*chem_workspace)(int64_t ncell, int nbins, int num_cus, int waves_per_cu);
    int (*flux)(const FluxLaunch& L);
    int (*props)(const PropsLaunch& L);
    int (*adjoint)(const AdjointLaunch& L);
    int (*tangent)(const AdjointLaunch& L);
    size_t (*sens_workspace)(int64_t ncell, int nsteps, int ckpt_every);
    int (*box_ops)(const BoxLaunch& L);
    int flux_tx2, flux_ty2, flux_tx1;               // flux tile shapes: 2-D TX x TY, 1-D TX x 1
    int flux_nshape;                                // 2-D tile shapes available to the tuner (shape 0 the default)
    int flux_shape_tx[3], flux_shape_ty[3];
};

// All compiled topologies (null-terminated).
const TopoOps* const* compiled_topologies();

// Device time of the work enqueued between construction and stop() on one
// stream, measured with HIP events on that stream. A null target disables it.
struct StreamTimer {
    hipEvent_t e0 = nullptr, e1 = nullptr;
    float* target;
    hipStream_t s;
    StreamTimer(float* t, hipStream_t stream) : target(t), s(stream) {
        if (target && (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess ||
                       hipEventRecord(e0, s) != hipSuccess))
            target = nullptr;
    }
    void stop() {
        if (!target) return;
        float ms = 0;
        if (hipEventRecord(e1, s) == hipSuccess && hipEventSynchronize(e1) == hipSuccess &&
            hipEventElapsedTime(&ms, e0, e1) == hipSuccess)
            *target = ms;
        target = nullptr;
    }
    ~StreamTimer() {
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
    }
};

}  // namespace hf
