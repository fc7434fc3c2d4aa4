/* hyperflow.h — C ABI of the MI355X-native HyperFlow V3 hot path.
 *
 * The reference's two hot entry points are C++ member functions:
 *   ChemistryBatch::advance(Field&, double dt, size_t count)
 *       HyperFlow_V3/src/chemistry_batch.hpp:8-17, chemistry_batch.cpp:21-50,
 *       called from BlockAMR::react (HyperFlow_V3/src/BlockAMR.cpp:282);
 *   the per-box face loop + update of BlockAMR::euler
 *       HyperFlow_V3/src/BlockAMR.cpp:210-243, calling properties / minmod /
 *       admissible / hllc (HyperFlow_V3/src/solver.hpp:27-30, physics.cpp:4-48).
 * Behind this ABI they run as HIP kernels for gfx950. Signatures carry plain
 * pointers and sizes only.
 *
 * Conventions (each mirrors the reference):
 *  - State layout: species-first conserved vector [rhoY_0..rhoY_{ns-1}, rho u,
 *    rho v, rho E] (solver.hpp:17-18), stored SoA: component n of cell i at
 *    U[n*ld + i].
 *  - All-or-nothing: on any failed cell the caller's field is unchanged and
 *    the LOWEST failing cell index is reported (chemistry_batch.cpp:26,40-49).
 *  - Inadmissible states return HF_EDOMAIN (std::domain_error in physics.cpp:7,13,
 *    chemistry.cpp:71); integrator failure returns HF_ECONV (CVODE flag < 0 /
 *    incomplete interval, chemistry.cpp:7,67-68); bad arguments HF_EINVAL
 *    (chemistry.cpp:56, chemistry_batch.cpp:23-24).
 *  - One caller at a time per context: a concurrent or nested call returns
 *    HF_EBUSY (chemistry_batch.hpp:6-7, chemistry_batch.cpp:22).
 *  - Pointers may be device (HBM) or host memory; host buffers are staged
 *    through device memory for the call. The library keeps no pointer after
 *    a call returns.
 */
#ifndef HYPERFLOW_H
#define HYPERFLOW_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    HF_OK = 0,
    HF_EINVAL = 1,
    HF_EDOMAIN = 2,
    HF_ECONV = 3,
    HF_ENOMEM = 4,
    HF_EDEVICE = 5,
    HF_EBUSY = 6
};

/* Chemistry forms. PLAIN: one cell per lane, FP64 (the drop-in).
 * BATCHED: persistent kernel over a device work queue of stiffness-binned
 * cells, lanes refilled as their cells finish. */
enum { HF_CHEM_PLAIN = 0, HF_CHEM_BATCHED = 1 };

/* Newton-matrix precision: FP64 throughout, or the Jacobian and its LU in
 * FP32 with FP64 residuals, state and error control. */
enum { HF_PREC_FP64 = 0, HF_PREC_MIXED = 1 };

typedef struct hf_ctx hf_ctx;

typedef struct hf_options {
    double rtol;        /* relative tolerance; reference: 1e-8 (chemistry.cpp:19) */
    double atol;        /* absolute tolerance; reference: 1e-15 (chemistry.cpp:19) */
    int max_steps;      /* per cell; reference: 5000 (chemistry.cpp:25) */
    int precision;      /* HF_PREC_FP64 or HF_PREC_MIXED */
    int device;         /* HIP device ordinal */
    /* batched-form tuning. When all five tuning values (waves_per_cu, chunk,
     * nbins, jac_wait, setup_wait) are 0 the library tunes them online per
     * size class (coordinate descent over candidates on the first calls;
     * hf_chem_stats reports the values used); otherwise a 0 selects that
     * value's default (waves_per_cu 4 lanes / 8 waves, chunk 8, nbins 32,
     * jac_wait 16, setup_wait 2). Results do not depend on any of them. */
    int waves_per_cu;   /* resident waves per CU of the persistent kernel */
    int chunk;          /* idle lanes of a wave refilled together from the queue */
    int nbins;          /* stiffness bins */
    /* 0: a device-resident hf_chem_advance call runs as a captured HIP graph,
     * replayed while calls repeat their arguments and buffers; 1: launch the
     * kernels one by one */
    int eager;
    /* batched lane form: a lane whose next trip evaluates a Jacobian may wait
     * up to this many trips for other lanes of its wave to need one too, so
     * that the evaluations fall into few trips; waiting never changes a
     * cell's arithmetic. 0 selects the default (16), a negative value no
     * waiting. */
    int jac_wait;
    /* the same for a lane whose next trip re-forms and factors its Newton
     * matrix from its saved Jacobian (0: default 2; negative: no waiting) */
    int setup_wait;
} hf_options;

typedef struct hf_chem_stats {
    int64_t cells;          /* cells advanced (BlockAMR.cpp:282 chemistry_cell_calls) */
    int64_t steps, rhs, jac, lu, newton, conv_failures, err_failures;
    int64_t max_cell_steps; /* largest step count of any cell */
    int team;               /* lanes of work (ChemistryBatch::lastTeamSize analogue) */
    double kernel_ms;       /* device time of the integration kernel(s) */
    int chunk;              /* batched form: idle lanes refilled per queue claim in this call */
    int graph;              /* 0 kernels launched one by one, 1 cached HIP graph replayed,
                               2 graph captured by this call and launched */
    /* batched form: the values this call ran with (the online tuner's when
     * every hf_options tuning value is 0; no value changes any result) */
    int waves_per_cu, nbins, jac_wait, setup_wait;
    int tuned;              /* 0 fixed by the options, 1 a tuning trial, 2 the tuned configuration */
} hf_chem_stats;

/* One ghosted SoA box for a flux stage. Element (component n, cell i, j) of
 * the input, with -2 <= i < nx+2 and -2 <= j < ny+2 (two ghost layers,
 * BlockAMR.cpp:206), is Uin[in_offset + n*in_plane + (j+2)*in_pitch + (i+2)];
 * the updated valid cell (0 <= i < nx, 0 <= j < ny) goes to
 * Uout[out_offset + n*out_plane + j*out_pitch + i]. */
typedef struct hf_box {
    int32_t nx, ny;
    int32_t walls;      /* bit0 x-lo, bit1 x-hi, bit2 y-lo, bit3 y-hi: the box face lies on a
                           reflecting wall (species, tangential momentum and energy flux zeroed,
                           BlockAMR.cpp:224-229) */
    int32_t wedge_i;    /* the y-lo wall applies only to faces with i >= wedge_i (0: all) */
    int64_t in_offset, in_pitch, in_plane;
    int64_t out_offset, out_pitch, out_plane;
    double dtdx, dtdy;  /* h/dx, h/dy; dtdy == 0 gives an x-only (1-D) stage */
} hf_box;

typedef struct hf_flux_stats {
    int64_t cells, faces;
    int64_t first_order_sides;  /* face sides that fell back to first order (BlockAMR.cpp:221-222) */
    double kernel_ms;
    int tile_x, tile_y;         /* the tile shape this call ran with (2-D: chosen online per size
                                   class among 15 x 15, 31 x 7 and 7 x 31 for small mechanisms;
                                   the results do not depend on it) */
} hf_flux_stats;

/* A rectangle of cells in device memory: one box of a level, or a whole
 * level array. Cell (i, j), 0 <= i < nx, 0 <= j < ny, component n is at
 * U[offset + n*plane + j*pitch + i]. mask (optional, device, one byte per
 * cell at mask[j*mask_pitch + i]) says what the cell is in the AMR
 * hierarchy: 0 not a valid cell of its level, 1 a leaf (valid and not
 * covered by a finer level: BlockAMR::mask, BlockAMR.cpp:165-168), 2 valid
 * and covered. NULL: every cell is a leaf. dx, dy: the level's cell size
 * (the cell volume is dx*dy; dy = 0 for a 1-D block). */
typedef struct hf_cells {
    double *U;
    int64_t offset, pitch, plane;
    int32_t nx, ny;
    const uint8_t *mask;
    int64_t mask_pitch;
    double dx, dy;
} hf_cells;

/* Create a context: load the Cantera-YAML mechanism (replaces
 * Cantera::newSolution(mechanism, "", "none"), chemistry_batch.cpp:17-18) and
 * upload its tables. opt may be NULL for the reference defaults.
 * Returns HF_EINVAL for an unreadable mechanism or one whose topology has no
 * compiled kernels (hf_last_error names it). */
int hf_create(const char *mechanism_yaml, const hf_options *opt, hf_ctx **out);
void hf_destroy(hf_ctx *ctx);
/* Message of the last failure on this context (e.g. "chemistry cell 17: invalid accepted species"). */
const char *hf_last_error(const hf_ctx *ctx);
/* Message of the last hf_create failure (no context exists then). */
const char *hf_create_error(void);

int hf_num_species(const hf_ctx *ctx);
int hf_num_reactions(const hf_ctx *ctx);
const char *hf_species_name(const hf_ctx *ctx, int k);
/* Molecular weights (kg/kmol), ns values. */
int hf_molecular_weights(const hf_ctx *ctx, double *W);
int hf_set_options(hf_ctx *ctx, const hf_options *opt);
int hf_get_options(const hf_ctx *ctx, hf_options *opt);

/* ChemistryBatch::advance (chemistry_batch.cpp:21-50) over SoA planes U[n*ld + i],
 * n < ns+3, i < ncell. dt: one value, or one per cell when dt_per_cell != 0.
 * Each cell is advanced over its dt at constant density and internal energy
 * (Chemistry::advance, chemistry.cpp:55-76); only the species planes change.
 * mode: HF_CHEM_PLAIN or HF_CHEM_BATCHED. first_failed (optional) receives the
 * lowest failing cell index, -1 on success. */
int hf_chem_advance(hf_ctx *ctx, double *U, int64_t ncell, int64_t ld, const double *dt, int dt_per_cell,
                    int mode, hf_chem_stats *stats, int64_t *first_failed);

/* One Euler stage of the flux sweep (BlockAMR.cpp:210-243) over nbox boxes:
 * NND/minmod reconstruction of the conserved variables, per-side first-order
 * fallback when a reconstructed state is inadmissible, HLLC flux, wall-face
 * zeroing, conservative update Uout = U - dtdx dFx - dtdy dFy. first_order != 0
 * forces first-order face states (BlockAMR.cpp:220). Returns HF_EDOMAIN with
 * first_failed = (box index << 32) | (j*nx + i) when an updated cell is
 * inadmissible, so the driver can redo the stage first-order
 * (BlockAMR.cpp:257-263). boxes is a host array. */
int hf_flux_stage(hf_ctx *ctx, const hf_box *boxes, int nbox, const double *Uin, double *Uout, int first_order,
                  hf_flux_stats *stats, int64_t *first_failed);

/* Face fluxes of one box, as the face-centred flux arrays BlockAMR::euler
 * keeps for the flux registers (BlockAMR.cpp:208, 230): x face (i, j) lies
 * between cells i-1 and i, 0 <= i <= nx, 0 <= j < ny, component n at
 * fx[fx_offset + n*fx_plane + j*fx_pitch + i]; y face (i, j), 0 <= i < nx,
 * 0 <= j <= ny, at fy[fy_offset + n*fy_plane + j*fy_pitch + i]. The values are
 * the HLLC fluxes after wall zeroing (per unit area and time). Device memory;
 * fy is unused by a 1-D stage. */
typedef struct hf_face_fluxes {
    double *fx;
    int64_t fx_offset, fx_pitch, fx_plane;
    double *fy;
    int64_t fy_offset, fy_pitch, fy_plane;
} hf_face_fluxes;

/* hf_flux_stage that also stores every box's face fluxes (fluxes: nbox host
 * descriptors, in box order), for refluxing across levels. The update of
 * Uout is bit-identical to hf_flux_stage's. */
int hf_flux_stage_fluxes(hf_ctx *ctx, const hf_box *boxes, int nbox, const double *Uin, double *Uout, int first_order,
                         const hf_face_fluxes *fluxes, hf_flux_stats *stats, int64_t *first_failed);

/* hf_flux_stage_fluxes over boxes that live in separate allocations (an
 * AMReX MultiFab's FArrayBoxes, BlockAMR.cpp:210-243): box b reads
 * in[b] + in_offset + ... and writes out[b] + out_offset + ..., with its
 * hf_box offsets relative to its own pointers (typically in_offset 0 for the
 * ghosted FAB and the valid box's corner for the output FAB). in and out are
 * host arrays of nbox device pointers; each box must lie inside its own
 * allocations (HF_EINVAL otherwise). fluxes may be NULL. The update is
 * bit-identical to hf_flux_stage's on the same data. */
int hf_flux_stage_boxes(hf_ctx *ctx, const hf_box *boxes, int nbox, const double *const *in, double *const *out,
                        int first_order, const hf_face_fluxes *fluxes, hf_flux_stats *stats, int64_t *first_failed);

/* hf_flux_stage_fluxes (fluxes may be NULL: then hf_flux_stage) that
 * allocates nothing. No counterpart in the reference: it is for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc). The box table, face-flux
 * descriptors, tile table and counters live in a static device array of
 * 1.4 MB that each GPU holds until the process ends; at most 4096 boxes and
 * 65536 tiles per call (15 x 15 cells per 2-D tile, 255 per 1-D tile), else
 * HF_EINVAL. Uin, Uout and the face fluxes must be device memory (HF_EINVAL
 * otherwise), and the context's device must have been bound by hf_create
 * (a GPU was visible then; HF_EDEVICE otherwise). stats->kernel_ms is 0 (no
 * timing events are created). Every call on a GPU shares that GPU's array:
 * calls that overlap on one GPU race. Results are bit-identical to
 * hf_flux_stage_fluxes. */
int hf_flux_stage_noalloc(hf_ctx *ctx, const hf_box *boxes, int nbox, const double *Uin, double *Uout,
                          int first_order, const hf_face_fluxes *fluxes, hf_flux_stats *stats, int64_t *first_failed);

/* properties(u) (physics.cpp:4-15) on SoA planes: p, a, T per cell; status
 * per cell (HF_OK / HF_EDOMAIN). Returns HF_EDOMAIN when any cell fails
 * (BlockAMR::validate, BlockAMR.cpp:286-298), first_failed = lowest index. */
int hf_properties(hf_ctx *ctx, const double *U, int64_t ncell, int64_t ld, double *p, double *a, double *T,
                  int64_t *first_failed);

/* CFL time step of BlockAMR::stableDt (BlockAMR.cpp:335-346) over SoA cells:
 * returns cfl / max((|u|+a)/dx + (|v|+a)/dy) in *dt. */
int hf_stable_dt(hf_ctx *ctx, const double *U, int64_t ncell, int64_t ld, double dx, double dy, double cfl,
                 double *dt);

/* Discrete adjoint of a fixed-step backward-Euler chemistry integration (no
 * reference counterpart; rate-parameter sensitivities). Per cell c: density
 * rho[c], initial temperature T0[c] and mass fractions Y0[c*ns + k]; e is held
 * at e(T0, Y0). nsteps steps of size h; checkpoints every ckpt_every steps are
 * recomputed in reverse. Objective J = T(t_end) + sum_n h qdot(y_{n+1}),
 * qdot = -sum_k u_k wdot_k (u_k molar internal energy), in J[c].
 * Gradients: grad_lnA[c*nr + i] = dJ/d ln A_i (A_i scales k and k0 of
 * reaction i), grad_T0[c], grad_Y0[c*ns + k]. lnA_shift (optional, nr values)
 * perturbs every cell's rate constants. All arrays host or device. Each step
 * is solved by Newton to a correction below 1e-15; a cell whose iteration
 * fails gets NaN outputs and the call returns HF_ECONV naming the lowest such
 * cell in hf_last_error. Built for mechanisms of at most 16 species (the
 * per-cell backward-Euler matrix is held in registers); a larger mechanism
 * returns HF_EINVAL. */
int hf_chem_adjoint(hf_ctx *ctx, int64_t ncell, const double *rho, const double *T0, const double *Y0, double h,
                    int nsteps, int ckpt_every, const double *lnA_shift, double *J, double *grad_lnA,
                    double *grad_T0, double *grad_Y0);

/* Forward tangent of the same integration along one direction per cell:
 * dJ = dJ/dlnA . v_lnA + dJ/dT0 v_T0 + dJ/dY0 . v_Y0 (v arrays shaped like the
 * gradients of hf_chem_adjoint). */
int hf_chem_tangent(hf_ctx *ctx, int64_t ncell, const double *rho, const double *T0, const double *Y0, double h,
                    int nsteps, const double *lnA_shift, const double *v_lnA, const double *v_T0,
                    const double *v_Y0, double *J, double *dJ);

/* ------------------------------------------------------------------ single-box driver operations
 * Device-memory helpers for a driver that advances one single-level ghosted
 * box (the hf_box input layout; the output fields of the box are unused).
 * All pointers are device pointers (HF_EINVAL otherwise). */
enum { HF_BC_TRANSMISSIVE = 0, HF_BC_REFLECTING = 1, HF_BC_PERIODIC = 2, HF_BC_INFLOW = 3, HF_BC_WEDGE = 4,
       HF_BC_KEEP = 5 };

/* Fill the two ghost layers of U (BlockAMR::fill, BlockAMR.cpp:137-152, with
 * the boundary types of BlockAMR.cpp:48-59): transmissive copies the nearest
 * valid cell (AMReX foextrap); reflecting mirrors about the face, negating the
 * normal momentum (reflect_even / reflect_odd); periodic wraps. bc = {x-lo,
 * x-hi, y-lo, y-hi}; y-lo < 0 fills x ghosts only (a 1-D box). At a corner
 * the y rule is applied to the filled x ghost column (AMReX filcc order).
 * Inflow and wedge codes need hf_fill_ghosts_flow (HF_EINVAL here). */
int hf_fill_ghosts(hf_ctx *ctx, double *U, const hf_box *box, const int32_t bc[4]);

/* hf_fill_ghosts plus the two conditions of FlowBoundary (FlowBoundary.hpp:
 * 7-37, filled after the rules above as GpuBndryFuncFab does, BlockAMR.cpp:
 * 137-152): HF_BC_INFLOW (ext_dir) sets every ghost cell that lies beyond an
 * inflow side, corners included, to inlet (host, ns+3 conserved values: the
 * reference's premixed `fresh` state, BlockAMR.cpp:132-136); HF_BC_WEDGE (y-lo
 * only, 2-D, ny >= 2) makes the bottom ghost rows an outflow before the wedge
 * tip (copy of row 0) and a slip wall from column box->wedge_i on (mirror,
 * negated y momentum), reading the valid cell whose column is the ghost's
 * clamped to [0, nx). A corner beyond an inflow side is inlet even on a wedge
 * bottom. inlet may be NULL when no side is HF_BC_INFLOW. HF_BC_KEEP marks a
 * side inside the domain (a fine box's side whose ghosts came from the coarse
 * level): its ghosts are kept, except that a ghost beyond a physical x side
 * but beside a kept y side takes that x side's rule (FillPatchTwoLevels fills
 * the in-domain ghosts first, then the physical ones). */
int hf_fill_ghosts_flow(hf_ctx *ctx, double *U, const hf_box *box, const int32_t bc[4], const double *inlet);

/* One SoA array of a level patch: element (component n, cell i, j), i and j
 * in the level's own index space, is at base[offset + n*plane + j*pitch + i]. */
typedef struct hf_patch {
    int64_t offset, pitch, plane;
} hf_patch;

/* Coarse-to-fine conservative interpolation at refinement ratio (2, 2)
 * (ConservativeInterp::interp, ConservativeInterp.cpp:18-51). region =
 * {ilo, jlo, ihi, jhi}: the fine cells to write, inclusive, in the fine
 * index space. For every coarse cell (i, j) of coarsen(region): minmod slopes
 * sx, sy from its four neighbours (all must be readable in coarse:
 * grow(coarsen(region), 1)); the first theta in 1, 1/2, ..., 2^-47 for which
 * all four children q + theta (+-sx/4 +- sy/4) are admissible, else 0; the
 * children inside region are written. No fused multiply-add in the child
 * arithmetic. limited (optional) receives the number of coarse cells with
 * theta < 1. A parent that is itself inadmissible returns HF_EDOMAIN with the
 * lowest such coarse cell (row-major over coarsen(region), i fastest) in
 * first_failed (optional); fine cells are then written for the other parents. */
int hf_interp_coarse_to_fine(hf_ctx *ctx, const double *coarse, const hf_patch *cp, double *fine,
                             const hf_patch *fp, const int32_t region[4], int64_t *limited,
                             int64_t *first_failed);

/* Average-down at ratio (2, 2) (BlockAMR::average, BlockAMR.cpp:157-159, over
 * AMReX average_down, amrex @2d8c2e7e): every coarse cell of cregion = {ilo,
 * jlo, ihi, jhi} (inclusive, coarse index space) becomes 0.25 times the sum of
 * its four fine children, summed from 0 in the order (2i, 2j), (2i+1, 2j),
 * (2i, 2j+1), (2i+1, 2j+1). */
int hf_average_down(hf_ctx *ctx, const double *fine, const hf_patch *fp, double *coarse, const hf_patch *cp,
                    const int32_t cregion[4]);

/* The flux-register correction of one fine box at ratio (2, 2) (BlockAMR.cpp:
 * 247-256: FluxRegister CrseInit with -h dy_c (x faces) / -h dx_c (y faces),
 * FineAdd with h dy_f / h dx_f, Reflux with scale 1; Cartesian volumes).
 * fine_box = {ilo, jlo, ihi, jhi}: the fine box's cells, inclusive, in the fine
 * index space (ilo, jlo even, ihi, jhi odd). coarse addresses the coarse
 * stage's face fluxes and fine the fine box's, each with its offset taken in
 * its level's index space (face (i, j) of the coarse level at
 * fx[fx_offset + n*fx_plane + j*fx_pitch + i]). sides (bit0 x-lo, bit1 x-hi,
 * bit2 y-lo, bit3 y-hi) selects the box sides that face coarse cells: leave out
 * a side on the domain boundary or shared with another fine box
 * (ClearInternalBorders). For each coarse face on a selected side:
 *   reg = (-h dy_c) Fc;  reg += (h dy_f) Ff(2j);  reg += (h dy_f) Ff(2j+1)
 * (x faces; dx for y faces), and the coarse cell just outside the box gets
 * -reg / (dx_c dy_c) on a lo side, +reg / (dx_c dy_c) on a hi side, in Uc
 * (addressed by cp, coarse index space). No fused multiply-add. */
int hf_reflux(hf_ctx *ctx, double *Uc, const hf_patch *cp, const hf_face_fluxes *coarse, const hf_face_fluxes *fine,
              const int32_t fine_box[4], int32_t sides, double h, double dxc, double dyc);

/* ------------------------------------------------------------------ AMR hierarchy operations
 * A level of an AMR hierarchy may be held as one array over its whole
 * domain (plus two ghost layers) in which its boxes are index rectangles;
 * a per-level hf_cells mask then says which cells are valid (in a box),
 * which of those are leaves and which are covered by the next finer level.
 * Boxes of one level read each other's valid cells as ghosts directly. */

/* BlockAMR::react's chemistry (BlockAMR.cpp:271-285): ChemistryBatch::advance
 * (chemistry_batch.cpp:21-50) over every leaf cell (mask 1, or every cell of
 * a block without mask) of every block, in block order, then j, then i, all
 * in one batch read and written in place through the block table (no gather
 * or scatter). dt: one interval (host or device). All-or-nothing over the
 * whole batch; first_failed = (block << 32) | (j*nx + i) of the lowest
 * failing cell. stats->cells counts every block cell, leaves or not. Device
 * memory only. */
int hf_chem_advance_cells(hf_ctx *ctx, const hf_cells *blocks, int nblock, const double *dt, int mode,
                          hf_chem_stats *stats, int64_t *first_failed);

/* BlockAMR::validate over every valid cell of every block (mask != 0;
 * BlockAMR.cpp:286-298) and stableDt's rate max((|u|+a)/dx + (|v|+a)/dy)
 * over the leaf cells (:335-346) into *rate. HF_EDOMAIN with first_failed =
 * (block << 32) | (j*nx + i) of the lowest inadmissible cell. */
int hf_validate_cells(hf_ctx *ctx, const hf_cells *blocks, int nblock, double *rate, int64_t *first_failed);

/* BlockAMR::repairRoundoff (BlockAMR.cpp:299-334) over the leaf cells of all
 * blocks, cell volume dx*dy: per species the peak over all leaves and the
 * donor (the first leaf, in block/j/i order, holding the largest value *
 * volume); negative values within 64 eps of the peak are zeroed and their
 * volume-weighted total is taken from the donor. HF_EDOMAIN with the lowest
 * materially negative cell ((block << 32) | (j*nx + i)), or when the donor's
 * mass does not exceed the deficit. repairs receives the zeroed count. */
int hf_repair_cells(hf_ctx *ctx, const hf_cells *blocks, int nblock, int64_t *repairs, int64_t *first_failed);

/* hf_interp_coarse_to_fine over nregion fine regions (regions: host, 4
 * values each, ilo jlo ihi jhi) in one launch. With fine_mask (device, the
 * fine level's hf_cells mask, extent fine_extent = {nx, ny} of the fine
 * level, row pitch fine_mask_pitch), a child is written only inside that
 * extent where the mask is 0: the ghost cells of a level's boxes that no box
 * of the level holds and that lie in the domain (FillPatchTwoLevels'
 * coarse-filled part). first_failed = (region << 32) | parent index. */
int hf_interp_regions(hf_ctx *ctx, const double *coarse, const hf_patch *cp, double *fine, const hf_patch *fp,
                      const int32_t *regions, int nregion, const uint8_t *fine_mask, int64_t fine_mask_pitch,
                      const int32_t fine_extent[2], int64_t *limited, int64_t *first_failed);

/* hf_average_down restricted to the coarse cells of cregion whose coarse
 * mask (device, row pitch coarse_mask_pitch, coarse level index space) is 2
 * (covered by the fine level): BlockAMR::average over a whole level pair. */
int hf_average_down_masked(hf_ctx *ctx, const double *fine, const hf_patch *fp, double *coarse, const hf_patch *cp,
                           const uint8_t *coarse_mask, int64_t coarse_mask_pitch, const int32_t cregion[4]);

/* The flux-register correction of a level from every box of its finer level
 * (BlockAMR.cpp:247-256, ClearInternalBorders and Reflux over a whole
 * BoxArray): each leaf coarse cell of cregion (mask 1) with a covered face
 * neighbour (mask 2; neighbours outside cregion count as neither) takes, per
 * such face in the order x-lo, x-hi, y-lo, y-hi, the register of hf_reflux
 * over its volume, + on its lo faces and - on its hi faces. Faces shared by
 * two fine boxes, and faces on the domain boundary, carry no correction.
 * No fused multiply-add; bit-identical to hf_reflux for one box. */
int hf_reflux_masked(hf_ctx *ctx, double *Uc, const hf_patch *cp, const hf_face_fluxes *coarse,
                     const hf_face_fluxes *fine, const uint8_t *coarse_mask, int64_t coarse_mask_pitch,
                     const int32_t cregion[4], double h, double dxc, double dyc);

/* Refinement criteria of BlockAMR::ErrorEst (BlockAMR.cpp:170-200). */
typedef struct hf_tag_params {
    double pressure_tag;     /* relative pressure jump to a face neighbour (reference default 0.08) */
    double temperature_tag;  /* relative temperature jump to a face neighbour (default 0.08) */
    double fuel_rate_tag;    /* fuel consumption rate -wdot_f W_f / rho in 1/s (default 50) */
    int32_t physics;         /* 0: geometry only (tag_mode "box") */
    int32_t chemistry;       /* nonzero: apply the fuel-rate criterion (config "chemistry") */
    int32_t fuel;            /* fuel species index (the reference's species 0, H2) */
    int32_t has_box;         /* nonzero: tag_box is given */
    double box[4];           /* tag_box: x0, x1, y0, y1 (inclusive) */
    double velocity[2];      /* tag_velocity: the box moves with it */
    double x0, y0;           /* coordinates of the box's lower-left corner (cell (0, 0)'s low faces) */
    double dx, dy, time;     /* cell size and the time the tags are taken at */
} hf_tag_params;

/* Tags of the valid cells of one ghosted box (the hf_box input layout; the
 * first ghost layer must be filled, e.g. by hf_fill_ghosts_flow): cell (i, j)
 * is tagged when its centre, moved back by velocity * time, lies in the box;
 * or (physics) when max(|p_n - p| / min(p_n, p)) over the two face neighbours
 * of either direction exceeds pressure_tag, or the same for T exceeds
 * temperature_tag; or (chemistry) when -wdot_fuel W_fuel / rho exceeds
 * fuel_rate_tag. tags: device, nx*ny bytes, tags[j*nx + i] = 1 or 0. ntagged
 * (optional) receives the count. An inadmissible cell or neighbour (properties
 * throws in the reference) returns HF_EDOMAIN with the lowest such valid cell
 * (j*nx + i) in first_failed. The tag buffer and the clustering into boxes
 * belong to AMReX's regrid and are not part of this call. */
int hf_tag_cells(hf_ctx *ctx, const double *U, const hf_box *box, const hf_tag_params *prm, uint8_t *tags,
                 int64_t *ntagged, int64_t *first_failed);

/* One Euler stage of BlockAMR::euler (BlockAMR.cpp:202-265) on one
 * single-level ghosted box in a single call: fill Uin's ghosts
 * (hf_fill_ghosts_flow), the flux stage into Uout's valid cells (both arrays
 * in the box's input layout; out_* must address Uout's valid cells), roundoff
 * repair and validation of Uout; when the repair or the validation rejects
 * the stage, the same with first-order face states (BlockAMR.cpp:257-263).
 * One host synchronisation per attempt. repairs (optional) receives the
 * values the accepted attempt zeroed, first_order (optional) 1 when the
 * first-order redo ran. HF_EDOMAIN with first_failed = the lowest cell
 * (j*nx + i) that the repair or the validation rejected, when the
 * first-order stage is rejected too. Device pointers only. */
int hf_euler_stage(hf_ctx *ctx, double *Uin, double *Uout, const hf_box *box, const int32_t bc[4], const double *inlet,
                   double dx, double dy, int64_t *repairs, int32_t *first_order, int64_t *first_failed);

/* One whole Strang step of a single-level box (BlockAMR::run's loop body,
 * BlockAMR.cpp:400-425) as one HIP graph, captured on the first call and
 * replayed while the arrays, box, boundary, options and context buffers
 * repeat: dt = cfl / (the CFL rate of U) (capped at time_left when > 0),
 * flow(dt/2), react(dt), flow(dt/2), where flow is the Heun average of two
 * hf_euler_stage stages (ghost fill, flux stage, roundoff repair,
 * validation) written into U, and react one chemistry batch over the box's
 * cells (chem_mode) in place. U, U1, U2, old, backup: device arrays in the
 * box's input layout (the box's out_* fields are ignored); dt and h live in
 * device memory, so no host value enters the graph. Every decision of the
 * step stays on the device except one: when any stage is rejected by the
 * repair or the validation (the reference then redoes it first-order), or
 * the chemistry fails, U is restored to the step's start and *taken is 0, so
 * that the caller takes the step with host decisions (hf_euler_stage, the
 * dt halving of BlockAMR.cpp:416). Otherwise *taken is 1 (replayed), 2
 * (captured and launched) or 3 (hf_options.eager: kernels launched one by
 * one), *dt_taken the step and *repairs the zeroed values. The results are
 * bit-identical to the same step taken call by call. */
int hf_box_strang_step(hf_ctx *ctx, double *U, double *U1, double *U2, double *old, double *backup,
                       const hf_box *box, const int32_t bc[4], const double *inlet, double dx, double dy, double cfl,
                       int chem_mode, double time_left, double *dt_taken, int32_t *taken, int64_t *repairs);

/* BlockAMR::totals (BlockAMR.cpp:347-355): the integral of every conserved
 * component (|value| when absolute != 0) times the cell volume dx*dy (dx for
 * a 1-D block) over the leaf cells (mask 1, or every cell without a mask) of
 * a block table. The reference sums in long double cell by cell; here each
 * workgroup sums its cells and the workgroups are summed in a fixed order, so
 * the result does not depend on timing. totals: nv = ns + 3 values (host). */
int hf_totals_cells(hf_ctx *ctx, const hf_cells *blocks, int nblock, int absolute, double *totals);

/* One level of a block-structured hierarchy at refinement ratio 2 (the
 * levels of BlockAMR, BlockAMR.hpp). The level is one ghosted array over its
 * whole index space: component n of cell (i, j) at
 * n (nx + 4)(ny + 4) + (j + 2)(nx + 4) + i + 2, with five such arrays (the
 * state U and the step's work arrays). fx, fy: the level's face fluxes,
 * nv planes of ny x (nx + 1) and (ny + 1) x nx. mask: ny x nx bytes, 0 no
 * box, 1 leaf, 2 covered by the next finer level. boxes: (i0, j0, i1, j1)
 * per box, inclusive (level 0: one box over the level). ghosts: the fine
 * ghost strips (i0, j0, i1, j1) interpolated from the coarser level where
 * the mask is 0 (hf_interp_regions). wedge_i: the level's first wedge-wall
 * column. Host tables; device arrays. */
typedef struct hf_level {
    double *U, *U1, *U2, *old, *backup;
    double *fx, *fy;
    const uint8_t *mask;
    int32_t nx, ny;
    double dx, dy;
    const int32_t *boxes;
    int32_t nbox;
    const int32_t *ghosts;
    int32_t nghost;
    int32_t wedge_i;
} hf_level;

/* One whole Strang step of a hierarchy (BlockAMR::run's loop body,
 * BlockAMR.cpp:400-425, over every level) as one HIP graph, captured on the
 * first call and replayed while the levels, boxes, arrays, boundary,
 * options and context buffers repeat: dt = cfl / (the CFL rate of the leaf
 * cells) (capped at time_left when > 0), flow(dt/2), react(dt) over every
 * leaf cell (chemistry != 0, chem_mode), flow(dt/2). flow is the Heun
 * average of two composite Euler stages (BlockAMR::euler, BlockAMR.cpp:202-
 * 298): per level, coarse to fine, the ghost strips interpolated from the
 * coarser level and the physical ghost fill; one flux stage over every box
 * of every level with face fluxes; reflux and average-down, finest to
 * coarsest; roundoff repair over the leaves; average-down; validation.
 * dt and h live in device memory, so no host value enters the graph, and
 * the step makes the decisions of the stage-by-stage calls on the device
 * except one: when any interpolation, repair, validation or chemistry batch
 * fails (the reference then redoes a stage first-order or halves dt), every
 * level's U is restored to the step's start and *taken is 0, for the
 * caller's step with host decisions. Otherwise *taken is 1 (replayed), 2
 * (captured and launched) or 3 (hf_options.eager), *dt_taken the step and
 * *repairs the zeroed values. The step's copies and Heun average touch the
 * valid cells only (every ghost cell is refilled before it is read, so no
 * side may be HF_BC_KEEP). The results are bit-identical to the same step
 * taken with hf_interp_regions, hf_fill_ghosts_flow,
 * hf_flux_stage_fluxes, hf_reflux_masked, hf_average_down_masked,
 * hf_repair_cells, hf_validate_cells and hf_chem_advance_cells. */
int hf_amr_strang_step(hf_ctx *ctx, const hf_level *levels, int nlevel, const int32_t bc[4], const double *inlet,
                       double cfl, int chemistry, int chem_mode, double time_left, double *dt_taken, int32_t *taken,
                       int64_t *repairs);

/* BlockAMR::repairRoundoff (BlockAMR.cpp:299-334) on the valid cells of one
 * box of equal cell volumes: negative species within 64 eps of the species'
 * peak are zeroed and their total is taken from the first cell holding the
 * peak; a larger negative value, a non-finite one or a missing donor returns
 * HF_EDOMAIN with the cell (j*nx + i) in first_failed. repairs (optional)
 * receives the number of zeroed values. */
int hf_repair_roundoff(hf_ctx *ctx, double *U, const hf_box *box, int64_t *repairs, int64_t *first_failed);

/* BlockAMR::validate and stableDt (BlockAMR.cpp:286-298,335-346) over the
 * valid cells of one box: HF_EDOMAIN with the lowest inadmissible cell in
 * first_failed, else *rate = max((|u|+a)/dx + (|v|+a)/dy) (dy = 0: x only). */
int hf_validate_box(hf_ctx *ctx, const double *U, const hf_box *box, double dx, double dy, double *rate,
                    int64_t *first_failed);

/* y = a y + b x over n doubles (MultiFab::LinComb of the Heun average,
 * BlockAMR.cpp:268). */
int hf_lincomb(hf_ctx *ctx, double *y, const double *x, double a, double b, int64_t n);

/* Synthetic inputs (no reference counterpart). Seeded splitmix64/xoshiro256**
 * cells for the batched-chemistry workloads. Per cell, in draw order:
 * T, ln p, phi, f, ln dt (uniform doubles u = (x >> 11) 2^-53).
 * Y = (1-f) Y_fresh(phi) + f Y_eq(T, p, phi), rho = p Wbar / (R T), e = e(T, Y),
 * u = v = 0. which: 0 = H2/air (T in [1100,2400] K, p log-uniform
 * [0.5,50] atm, phi in [0.5,2], progress f in [0,1] toward TP equilibrium,
 * dt log-uniform [1e-8,1e-6] s); 1 = DME/air (T in [700,1800] K, p log-uniform
 * [1,40] atm, phi in [0.6,1.5]); 2 = H2/air adjoint cells (T in [1000,1500] K,
 * p in [1,10] atm, phi in [0.5,2], fresh mixture). Writes host SoA planes
 * U[n*ncell + i] (ns+3 planes) and dt[i] (which 0/1). */
int hf_synth_cells(hf_ctx *ctx, int which, uint64_t seed, int64_t ncell, double *U, double *dt);

/* TP equilibrium mass fractions for element amounts given by a reactant
 * composition (mole fractions X, ns values) at (T, p), by element-potential
 * Newton on the ideal-gas Gibbs minimum. Y out (ns values). Host arrays. */
int hf_equilibrate_tp(hf_ctx *ctx, double T, double p, const double *X, double *Y);

/* Conserved state (SoA, nv values) of a gas at (T, p, Y, u, v): replaces the
 * reference's primitive() initial-condition helper (BlockAMR.cpp:20-24). */
int hf_state_from_TPY(hf_ctx *ctx, double T, double p, const double *Y, double u, double v, double *q);

#ifdef __cplusplus
}
#endif

#endif /* HYPERFLOW_H */
