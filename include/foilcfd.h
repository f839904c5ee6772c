/*
 * FoilCFD MI355X solver core -- C ABI.
 *
 * The drop-in boundary behind the reference's src/sim solver interface
 * (class LBMSolver, src/sim/lbm_solver.h:133-514), its unit conversions
 * (src/sim/units.h), the flag builders it is fed from (src/geom/voxelizer.h,
 * src/geom/airfoil.h) and the particle entry points (src/render/particles.cuh).
 * Plain C: opaque handle, raw host pointers and sizes, integer status codes,
 * no exceptions across the boundary, no HIP or torch types in the signatures.
 *
 * Index convention for every host array: cell = x + nx*(y + ny*z), x fastest
 * (reference GridDims, src/sim/lbm_core.cuh:137-158); population arrays are
 * canonical SoA f[q*cells + cell] in the frozen D3Q19 ordering
 * (src/sim/lbm_core.cuh:20-68), independent of the device storage format.
 */
#ifndef FOILCFD_H
#define FOILCFD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
#define FC_OK 0
#define FC_ERR_INVALID 1   /* bad argument (reference: cudaErrorInvalidValue) */
#define FC_ERR_ALLOC 2     /* device / pinned allocation failed */
#define FC_ERR_DEVICE 3    /* a HIP runtime call or launch failed */
#define FC_ERR_STATE 4     /* call not valid in the current state (e.g. before init) */
#define FC_ERR_IO 5        /* snapshot file could not be written / read or was rejected */

/* ---- population storage formats --------------------------------------- */
#define FC_STORAGE_F32 0   /* FP32 shifted deviations f - w */
#define FC_STORAGE_F16 1   /* IEEE FP16 of 2^15 (f - w); FP32 arithmetic */
#define FC_STORAGE_BF16 2  /* BF16 of (f - w); FP32 arithmetic */
#define FC_STORAGE_F64 3   /* FP64 storage and arithmetic (reference-quality runs) */
#define FC_STORAGE_F16C 4  /* 16-bit s1e4m11 float of (f - w), normals 2^-14..2; FP32 arithmetic */

/* ---- plain data ---------------------------------------------------------- */
typedef struct fc_dims { int nx, ny, nz; } fc_dims;  /* GridDims, lbm_core.cuh:139-158 */

/* LatticeScaling, units.h:115-174 */
typedef struct fc_scaling {
    float chord_m, airspeed_ms, nu_m2s;
    int chord_cells;
    float u_lat;
    float dx, dt;
    float nu_lat_target, nu_lat, tau;
    int tau_clamped;
    float re_target, re_effective;
} fc_scaling;

/* DomainLayout, voxelizer.h:57-85 */
typedef struct fc_layout {
    fc_dims dims;
    int chord_cells;
    float anchor_x_frac, anchor_y_frac;
    float anchor_x_cells, anchor_y_cells;
    int anchor_absolute;
} fc_layout;

/* PatchBox, lbm_refine.cuh:54-61 (coarse cells, [x0,x1) x [y0,y1) x full z) */
typedef struct fc_patch_box { int x0, y0, x1, y1; } fc_patch_box;

/* ForceReadout, lbm_solver.h:28-52 */
typedef struct fc_force_readout {
    float cl, cd, lift_to_drag;
    int valid;
    float flow_throughs;
    float cl_avg, cl_min, cl_max, cl_median;
    float cd_avg, cd_min, cd_max, cd_median;
    float ld_avg, ld_min, ld_max, ld_median;
} fc_force_readout;

/* SolverPerfStats, lbm_solver.h:68-74 (+ bytes moved per step) */
typedef struct fc_perf_stats {
    int last_steps_per_frame;
    double last_step_ms;
    double mlups;
    double population_bytes_per_step;
} fc_perf_stats;

/* Options of fc_init (extensions over the reference's init, which has
 * neither a storage choice nor periodic x / y). */
typedef struct fc_init_options {
    int storage;     /* FC_STORAGE_* */
    int periodic_x;  /* wrap x (fully periodic boxes; reference: never) */
    int periodic_y;  /* wrap y */
} fc_init_options;

/* ---- library ------------------------------------------------------------- */
const char* fc_version(void);
/* Number of HIP devices visible (0 without a GPU); no other side effect. */
int fc_device_count(void);

/* ---- units (src/sim/units.h) ----------------------------------------------- */
/* computeScaling, units.h:192-219 */
int fc_compute_scaling(float chord_m, float airspeed_ms, float nu_m2s, int chord_cells,
                       float u_lat, fc_scaling* out);
/* maxEffectiveRe, units.h:224-227 */
float fc_max_effective_re(int chord_cells, float u_lat);
/* refinedScaling, units.h:242-254 */
int fc_refined_scaling(const fc_scaling* parent, int factor, fc_scaling* out);
/* fineTauFor, lbm_refine.cuh:99-101 */
float fc_fine_tau_for(float tau_coarse, int factor);
/* rampedTau / rampedU, units.h:274-283 / :310-318 */
float fc_ramped_tau(const fc_scaling* s, long long step, int nx);
float fc_ramped_u(const fc_scaling* s, long long step, int nx);
/* LatticeScaling::forceToCoefficient / flowThroughSteps, units.h:158-173 */
float fc_force_to_coefficient(const fc_scaling* s, float force_lat, int span_cells);
float fc_flow_through_steps(const fc_scaling* s, int nx);

/* ---- geometry -> flags (src/geom/airfoil.h, src/geom/voxelizer.h) -------- */
/* generateNACA4, airfoil.h:73. Writes up to cap (x,y) pairs into xy; returns
 * the point count (2*n-1), or -1 with a reason in err. */
int fc_generate_naca4(const char* digits, int points_per_surface, float* xy, int cap,
                      int* le_index, char* err, int errlen);
/* buildBoundaryFlags, voxelizer.h:97 (flags: dims cells bytes) */
int fc_build_boundary_flags(fc_dims dims, uint8_t* flags);
/* voxelizeAirfoil, voxelizer.h:115 (xy: npts (x,y) pairs of a Selig loop) */
int fc_voxelize_airfoil(const float* xy, int npts, float aoa_deg, const fc_layout* layout,
                        uint8_t* flags);
/* closeTrailingEdgeGaps, voxelizer.h:125; returns the closed-cell count */
int fc_close_trailing_edge_gaps(fc_dims dims, uint8_t* flags);
/* buildCleanFoilFlags, voxelizer.h:135 */
int fc_build_clean_foil_flags(const float* xy, int npts, float aoa_deg, const fc_layout* layout,
                              uint8_t* flags);
/* makeFineLayout / fineDimsFor / stampInterfaceShell / buildFinePatchFlags /
 * derivePatchBox, voxelizer.h:155-199, lbm_refine.cuh:87-94 */
int fc_make_fine_layout(const fc_layout* coarse, const fc_patch_box* box, int factor,
                        fc_layout* out);
int fc_fine_dims(const fc_patch_box* box, fc_dims coarse, int factor, fc_dims* out);
int fc_stamp_interface_shell(fc_dims dims, uint8_t* flags);
int fc_build_fine_patch_flags(const float* xy, int npts, float aoa_deg, const fc_layout* coarse,
                              const fc_patch_box* box, int factor, uint8_t* flags);
int fc_derive_patch_box(fc_dims dims, const uint8_t* flags, int margin_x0, int margin_x1,
                        int margin_y0, int margin_y1, int clearance, fc_patch_box* out);

/* ---- vortex generators and q-links (src/geom/vg.h, airfoil.h:105) -------- */
/* VGParams, vg.h:24-36. type: 0 single vane, 1 counter-rotating pair,
 * 2 co-rotating array, 3 ramp. */
typedef struct fc_vg_params {
    int type;
    float x_c, height_c, length_h, beta_deg, pitch_c, gap_h;
    int count;
    int common_flow_down;
    int enabled;
} fc_vg_params;

/* VaneSlab, voxelizer.h:47-55 (oriented box of one vane, lattice cells). */
typedef struct fc_vane_slab {
    float center[3], n_axis[3], d_axis[3], b_axis[3];
    float half_height, half_length, half_thick;
} fc_vane_slab;

/* surfaceFrameAt, airfoil.h:105: out6 = point(2), tangent(2), normal(2);
 * returns 1 when valid. */
int fc_surface_frame_at(const float* xy, int npts, int le_index, float x_c, int upper, float* out6);
/* buildFlagsWithVGs, vg.h (clean flags + every VG, then TE closure). slabs
 * (may be NULL) receive up to slab_cap vane boxes; *nslabs the total. */
int fc_build_flags_with_vgs(const float* xy, int npts, int le_index, float aoa_deg,
                            const fc_layout* layout, const fc_vg_params* vgs, int nvg,
                            const uint8_t* clean_flags, uint8_t* flags_out,
                            fc_vane_slab* slabs, int slab_cap, int* nslabs);
/* buildVaneQLinks + densifyQLinks, vg.h: dense cut bytes qfrac[cell*19+q]
 * and two-node mask ffmask[cell] for every fluid->vane link; returns the
 * link count. */
int fc_build_vane_qlinks(fc_dims dims, const uint8_t* active, const uint8_t* clean,
                         const fc_vane_slab* slabs, int nslabs, uint8_t* qfrac, uint32_t* ffmask);
/* rayOBBEntry, vg.cpp:272-296 (entry parameter in [0,1] or -1). */
float fc_ray_obb_entry(const float* origin3, const float* dir3, const fc_vane_slab* slab);

/* ---- solver lifecycle (LBMSolver, lbm_solver.h:133-514) ------------------ */
typedef struct fc_solver fc_solver;

fc_solver* fc_create(void);                 /* LBMSolver::LBMSolver */
/* Message of the last failing call on this solver (empty if none); no
 * reference counterpart (the reference logs to stderr). */
int fc_last_error(fc_solver* s, char* buf, int cap);
void fc_destroy(fc_solver* s);              /* LBMSolver::~LBMSolver */
/* LBMSolver::init (lbm_solver.h:152-154): allocate, upload flags, reset.
 * flags: dims cells bytes (CellFlag values). options may be NULL (FP32,
 * z-periodic only). Fails with a reason in err, leaving nothing allocated. */
int fc_init(fc_solver* s, fc_dims dims, const fc_scaling* scaling, const uint8_t* flags,
            const fc_init_options* options, char* err, int errlen);
void fc_shutdown(fc_solver* s);             /* LBMSolver::shutdown */
int fc_reset(fc_solver* s);                 /* LBMSolver::reset */
int fc_set_flags(fc_solver* s, const uint8_t* flags);  /* LBMSolver::setFlags */
int fc_set_startup_ramp(fc_solver* s, int enabled);    /* setStartupRampEnabled, :269 */
int fc_set_force_ema_window(fc_solver* s, float flow_throughs);  /* setForceEmaWindow, :490 */
/* Smagorinsky constant (reference: fixed 0.1733 inside stepN, lbm_solver.cpp:1148);
 * 0 disables the eddy-viscosity term. Default 0.1733. */
int fc_set_smagorinsky_cs(fc_solver* s, float cs);

/* Step kernel selection (no reference counterpart: the reference has one
 * kernel). FC_STEP_AUTO (default) is the library's choice: one cell per
 * lane with 64-bit population addressing (the fastest measured on MI355X);
 * on a factor-2 patch level without a nested level, the last fine sub-step
 * of each coarse step also writes the restriction's child-average records.
 * FC_STEP_SINGLE runs the same cell kernel on every sub-step and leaves the
 * records to the separate restriction pass. FC_STEP_BUF runs the one-cell
 * kernel with 32-bit buffer addressing (population arrays under 2 GiB, the
 * default kernel beyond) and the separate restriction pass. FC_STEP_WIDE is
 * FC_STEP_AUTO with the 64-bit population index that grids of 2^32 / 19
 * cells or more use for wall-adjacent cells, forced on any grid (to test
 * that path). All four give the same cell update, bit for bit. */
enum { FC_STEP_AUTO = 0, FC_STEP_SINGLE = 1, FC_STEP_BUF = 2, FC_STEP_WIDE = 3 };
int fc_set_step_variant(fc_solver* s, int variant);

/* Rounding of the 16-bit population stores of the step (FP16C, FP16, BF16;
 * no counterpart in the reference, whose populations are FP32). Enabled
 * (the default): stochastic rounding, each stored value rounds up with
 * probability equal to the dropped fraction, so the storage error has zero
 * mean; the random bits are a hash of (cell, direction, step), so results
 * stay deterministic, identical across step variants and group stepping,
 * and exact snapshots restore the sequence. Disabled: round to nearest even.
 * Round to nearest leaves a systematic error in steady regions (the same
 * value rounds the same way every step): on the refined 23.6 M-cell wing at
 * chord Re ~1000, FP16C drag came out 4.5 % low against FP64 with round to
 * nearest and within 0.05 % with stochastic rounding. 32- and 64-bit
 * storages ignore the setting. */
int fc_set_stochastic_rounding(fc_solver* s, int enabled);
int fc_get_stochastic_rounding(const fc_solver* s);

/* ---- stepping ------------------------------------------------------------ */
/* LBMSolver::stepN (lbm_solver.h:354): n fused steps with the startup ramps,
 * watchdog every 200 steps, one force sample + wall-model update per call.
 * Asynchronous on the solver's stream. */
int fc_step(fc_solver* s, int n);
/* Wait for the solver's stream and fold finished readbacks. */
int fc_synchronize(fc_solver* s);

/* ---- readouts ------------------------------------------------------------ */
int fc_forces(fc_solver* s, fc_force_readout* out);   /* LBMSolver::forces, :371 */
/* Last folded raw momentum-exchange sample (lattice units, FP64 sum) and the
 * step it was taken at. */
int fc_force_raw(fc_solver* s, double* out3, long long* at_step);
/* Synchronous momentum-exchange force of the current state (launchForceReduction). */
int fc_force_now(fc_solver* s, double* out3);
/* fc_force_now without any device allocation (no reference counterpart;
 * for callers that cannot allocate: code around HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc). The reduction's per-block
 * partials and its result live in the library's static device scratch
 * (192 KiB per GPU, held until the process ends); the same kernels as
 * fc_force_now, so the result is bit-identical. At most 8192 * 256 force
 * cells (FC_ERR_INVALID beyond). Calls that overlap on one GPU (solvers on
 * different streams) share the scratch and race. */
int fc_force_now_noalloc(fc_solver* s, double* out3);
/* The force fold of fc_forces on caller-supplied samples (the reference
 * keeps it inside LBMSolver: EMA fold lbm_solver.cpp:633-654, statistics
 * lbm_solver.cpp:1414-1509, 256-sample ring lbm_solver.cpp:120). The solver
 * runs the same fold on its own samples; callers that take theirs with
 * fc_force_now (or from another solver) fold them here. add: the EMA weight
 * is 1 - exp(-d_steps / max(1, window_steps)), the first sample after a
 * create / clear seeds it, every sample enters the ring. readout: the EMA
 * coefficients (scaling, span_cells) and, when gate_open, mean / min / max /
 * median of Cl, Cd and L/D over the ring samples taken at step >=
 * one_flow_through_steps; flow_throughs is left 0. */
typedef struct fc_force_fold fc_force_fold;
fc_force_fold* fc_force_fold_create(void);
void fc_force_fold_destroy(fc_force_fold* f);
int fc_force_fold_clear(fc_force_fold* f);
int fc_force_fold_add(fc_force_fold* f, long long at_step, float d_steps, float window_steps, double fx,
                      double fy);
int fc_force_fold_readout(const fc_force_fold* f, const fc_scaling* scaling, int span_cells,
                          long long one_flow_through_steps, int gate_open, fc_force_readout* out);
/* Synchronous FP64 global mass sum over Fluid cells of sum_q f_q. */
int fc_mass(fc_solver* s, double* out);
/* Host copies of the macroscopic fields (velocityField/deviceRho, :425-429);
 * written on the last step of each fc_step call. Any pointer may be NULL. */
int fc_read_macro(fc_solver* s, float* rho, float* u, float* v, float* w);
/* Host copy of the cell flags (hostFlags, :485). */
int fc_read_flags(fc_solver* s, uint8_t* flags);
/* Canonical post-collision populations (the unpadded FP32/FP64 form of
 * activeDeviceF, :454-458). Populations of cells that do not own their
 * storage (solid / slip / interface) are unspecified; slots that left a
 * non-periodic domain read NaN. */
int fc_download_f(fc_solver* s, void* host, int as_f64);
/* Replace the state (tests' uploadF, tests/lbm_test_util.h:116-147) and run
 * the exact-restore bookkeeping (notifySnapshotRestored(true), :465). */
int fc_upload_f(fc_solver* s, const void* host, int is_f64);
int fc_notify_snapshot_restored(fc_solver* s, int full_state);  /* :1750-1771 */
/* applyEditedFlags (lbm_solver.h:174, lbm_solver.cpp:1086-1126; fixup kernel
 * lbm_core.cu:644-717): new flags on a developed field. Cells whose solid
 * state flips are fixed in place (newly solid: f = 0; newly fluid: f =
 * feq(mean rho, mean u) over the neighbours that are Fluid and unedited,
 * rest state when none); all other populations are kept. No viscosity ramp;
 * the force gate reopens after half a flow-through. */
int fc_apply_edited_flags(fc_solver* s, const uint8_t* flags);
long long fc_last_edit_count(fc_solver* s);  /* cells fixed by the last call */
/* restoreFromMacroscopic (lbm_solver.h:478, lbm_solver.cpp:1773-1803): the
 * four fields (dims cells each) become the solver's macro arrays and f =
 * feq(rho, u) (solid cells at rest), then notifySnapshotRestored(false). An
 * active fine patch (and the level nested in it) is re-seeded from it. */
int fc_restore_from_macroscopic(fc_solver* s, const float* rho, const float* u, const float* v,
                                const float* w);
int fc_nan_detected(fc_solver* s);          /* nanDetected, :411 (1 = latched) */
int fc_nan_kind(fc_solver* s);              /* watchdog verdict of the trip */
int fc_get_perf_stats(fc_solver* s, fc_perf_stats* out);   /* perfStats, :419 */
long long fc_step_count(fc_solver* s);      /* stepCount, :501 */
float fc_current_tau(fc_solver* s);         /* currentTau, :507 */
float fc_flow_throughs_completed(fc_solver* s);  /* flowThroughsCompleted, :504 */
int fc_get_dims(fc_solver* s, fc_dims* out);     /* dims, :495 */
int fc_get_scaling(fc_solver* s, fc_scaling* out);  /* scaling, :498 */
int fc_storage(fc_solver* s);
/* Device pointers of the macro fields and flags (velocityField, deviceRho,
 * deviceFlags): valid until the next init/shutdown. */
int fc_device_fields(fc_solver* s, void** rho, void** u, void** v, void** w, void** flags);
/* Bytes of device memory owned by the handle. */
long long fc_device_bytes(fc_solver* s);

/* ---- two-level refinement (src/sim/lbm_refine.cuh, lbm_solver.h:192-258) ---- */
/* RefinementInfo, lbm_solver.h:105-114 (coarse/fine level only). */
typedef struct fc_refinement_info {
    int active;               /* fine level allocated and stepping */
    int factor;               /* refinement factor m (2..4) */
    fc_patch_box box;         /* patch in coarse cells */
    fc_dims fine_dims;        /* (m W, m H, m nz) */
    fc_scaling fine_scaling;  /* refinedScaling(coarse, m) */
    double device_bytes;      /* fine populations + kind bytes + coupling band */
    int forces_from_fine;     /* momentum exchange on the fine grid */
} fc_refinement_info;

/* LBMSolver::initRefinement (lbm_solver.h:192-194): allocate the fine level
 * over `box` (coarse cells, full span) at `factor` with its flags (Fluid /
 * Solid / Interface, 2-cell Interface shell on the x / y faces, as
 * buildFinePatchFlags makes them), seed it from the current coarse state and
 * step it m times per coarse step from then on. The box must keep 2 coarse
 * cells from the x / y domain faces (the reference: 1). */
int fc_init_refinement(fc_solver* s, const fc_patch_box* box, int factor, const uint8_t* fine_flags,
                       char* err, int errlen);
int fc_shutdown_refinement(fc_solver* s);                          /* shutdownRefinement, :199 */
int fc_set_refined_flags(fc_solver* s, const uint8_t* fine_flags);  /* setRefinedFlags, :205 */
int fc_get_refinement_info(fc_solver* s, fc_refinement_info* out);  /* refinementInfo, :246 */
/* Nested third level (initFinerRefinement, lbm_solver.h:223-225; sequencing
 * lbm_solver.cpp:1209-1267): a patch inside the fine grid (box in fine cells,
 * at least 3 fine cells from the fine x / y faces), factor 2..4, full fine z.
 * Flags hold Fluid, Solid and Interface (an Interface shell on its x / y
 * faces, as for the fine level). Each fine sub-step then runs factor finer
 * sub-steps and the finer-to-fine restriction. Re-initialising or shutting
 * down the fine level drops it. The info struct reports the finer box,
 * dims and scaling in its fine_* fields. */
int fc_init_finer_refinement(fc_solver* s, const fc_patch_box* box, int factor, const uint8_t* finer_flags,
                             char* err, int errlen);
int fc_shutdown_finer_refinement(fc_solver* s);
int fc_get_finer_refinement_info(fc_solver* s, fc_refinement_info* out);
int fc_download_finer_f(fc_solver* s, void* host, int as_f64);
int fc_upload_finer_f(fc_solver* s, const void* host, int is_f64);
/* Fine-level populations in canonical SoA (no reference counterpart: the
 * reference exposes the padded fine buffer only to its tests). */
int fc_download_fine_f(fc_solver* s, void* host, int as_f64);
int fc_upload_fine_f(fc_solver* s, const void* host, int is_f64);

/* ---- iMEM wall model (lbm_wallmodel.cuh, wall_model.h) -------------------- */
#define FC_WALL_OFF 0
#define FC_WALL_NEWTON 1   /* Reichardt inversion by Newton (the reference) */
#define FC_WALL_MLP 2      /* learned u_tau network, Newton outside its envelope */

/* WallModelReadout, lbm_solver.h:82-91 (+ mode, listed and network-served cells) */
typedef struct fc_wall_model_readout {
    int enabled, mode;
    int listed, cells;
    float mean_yplus, max_yplus, clamped_frac;
    int excluded_vg, degenerate;
    int mlp_cells;
    int from_fine;   /* the readout is of a patch level (the finest with a wall list) */
} fc_wall_model_readout;

/* Learned u_tau network: 3 -> 32 -> 32 -> 1 tanh MLP on normalised
 * (ln y_s, ln |u_t|, ln nu), output ln(u_tau/|u_t|) (de-normalised by
 * out_mean/out_std); BF16 weight bits, FP32 biases; envelope bounds on
 * (ln y_s, ln |u_t|, ln nu, ln Re_y). No reference counterpart. */
typedef struct fc_mlp_params {
    float in_mean[4];
    float in_inv_std[4];
    float env_lo[4];
    float env_hi[4];
    float out_mean, out_std;
    uint16_t w1[32 * 3];
    float b1[32];
    uint16_t w2[32 * 32];
    float b2[32];
    uint16_t w3[32];
    float b3;
} fc_mlp_params;

/* buildWallCellList, wall_model.h:107-111 (host). Returns the listed count;
 * arrays (each cap entries, may be NULL to query the count) receive the SoA
 * list; stats3 (may be NULL) receives listed / excludedVG / degenerate. */
int fc_build_wall_cell_list(fc_dims dims, const uint8_t* active, const uint8_t* clean,
                            int sample_cells, long long* cell_idx, long long* sample_idx,
                            float* nx, float* ny, float* nz, float* sample_dist,
                            uint32_t* link_mask, int cap, int* stats3);
/* setWallModelEnabled, lbm_solver.h:290 (mode FC_WALL_*) */
int fc_set_wall_model(fc_solver* s, int mode);
/* setSurfaceReference, lbm_solver.h:280 (clean VG-free flags; NULL = live flags) */
int fc_set_surface_reference(fc_solver* s, const uint8_t* clean_flags);
/* wallModelReadout, lbm_solver.h:307 */
int fc_get_wall_model_readout(fc_solver* s, fc_wall_model_readout* out);
/* Slip field (half bits, cells each) and per-list-entry u_tau; NULL skips. */
int fc_read_wall_slip(fc_solver* s, uint16_t* uwx, uint16_t* uwy, uint16_t* uwz, float* utau);
/* Run one wall-model update on the current state now (fc_step runs one per call). */
int fc_wall_update_now(fc_solver* s);
int fc_wall_cell_count(fc_solver* s);
/* Per level (0 coarse, 1 fine patch, 2 nested finer patch): wall-list size
 * (-1 when the level is absent) and the level's slip field (its cells) and
 * u_tau. Patch levels list their wall cells at a sample distance of 3 cells
 * and update at their own viscosity (lbm_solver.cpp:425-455, :1368-1389). */
int fc_wall_cell_count_level(fc_solver* s, int level);
int fc_read_wall_slip_level(fc_solver* s, int level, uint16_t* uwx, uint16_t* uwy, uint16_t* uwz, float* utau);
/* setRefinedSurfaceReference (lbm_solver.h:302) and its finer analog
 * (lbm_solver.cpp:2022): clean VG-free flags of the patch level (NULL =
 * the level's live flags). */
int fc_set_refined_surface_reference(fc_solver* s, const uint8_t* clean_fine_flags);
int fc_set_refined_finer_surface_reference(fc_solver* s, const uint8_t* clean_finer_flags);
/* Harvest training samples (y_s, |u_t|, nu, Newton u_tau; u_tau < 0 = none)
 * from every update while enabled; fc_wall_harvest copies the last update's
 * samples (4 floats each) of the readout level (the finest level with wall
 * physics, as fc_get_wall_model_readout) and returns the count (out NULL:
 * count only). No reference counterpart: training data of the learned wall
 * function. */
int fc_set_wall_harvest(fc_solver* s, int enabled);
int fc_wall_harvest(fc_solver* s, float* out, int cap);
/* Load (or with NULL, drop) the learned u_tau network used by FC_WALL_MLP. */
int fc_set_wall_mlp(fc_solver* s, const fc_mlp_params* params);

/* ---- learned u_tau network: GPU training (no reference counterpart) ------- */
/* The trainer holds harvested samples on the GPU and fits the network there
 * with the library's own forward / backward / Adam kernels (BF16 matrix-core
 * tiles, FP32 accumulate and FP32 master parameters). Master parameters are
 * FC_MLP_NPARAMS floats: W1[32][3], b1[32], W2[32][32] (row = output unit),
 * b2[32], w3[32], b3. Inputs are the normalised (ln y_s, ln |u_t|, ln nu),
 * the target the normalised ln(u_tau / |u_t|). */
#define FC_MLP_NPARAMS 1217
typedef struct fc_mlp_trainer fc_mlp_trainer;
/* samples: count x 4 floats (y_s, |u_t|, nu, u_tau) as fc_wall_harvest writes
 * them; entries with a non-positive or non-finite field are skipped. Sets the
 * normalisation and the envelope from the kept samples and initialises the
 * weights (Glorot-uniform, deterministic in seed). NULL on error (text in err). */
fc_mlp_trainer* fc_mlp_trainer_create(const float* samples, long long count, uint32_t seed, char* err,
                                      int errlen);
void fc_mlp_trainer_destroy(fc_mlp_trainer* t);
long long fc_mlp_trainer_samples(fc_mlp_trainer* t);
/* steps Adam steps on minibatches of `batch` samples (a multiple of 32) drawn
 * on the GPU; losses (may be NULL, steps entries) receive each step's batch
 * mean squared error. The step counter persists across calls. */
int fc_mlp_trainer_train(fc_mlp_trainer* t, int steps, int batch, float lr, float* losses);
/* Gradient of the batch-mean squared error over samples idx[0..batch) (batch
 * a multiple of 32) at params (NULL = the current master parameters) into
 * grad[FC_MLP_NPARAMS]; *sse = sum of squared errors. */
int fc_mlp_trainer_gradient(fc_mlp_trainer* t, const float* params, const int* idx, int batch, float* grad,
                            float* sse);
int fc_mlp_trainer_get_params(fc_mlp_trainer* t, float* params);
/* Replaces the master parameters and restarts Adam (moments and counter). */
int fc_mlp_trainer_set_params(fc_mlp_trainer* t, const float* params);
/* The normalised training inputs (SoA, x[3][n]) and targets (t[n]); NULL skips. */
int fc_mlp_trainer_data(fc_mlp_trainer* t, float* x, float* target);
/* The network for fc_set_wall_mlp: BF16 (round to nearest even) weights,
 * FP32 biases, normalisation and envelope. */
int fc_mlp_trainer_export(fc_mlp_trainer* t, fc_mlp_params* out);
/* Batched inference on the GPU with the wall model's own fused routine:
 * out[i] = network u_tau of samples[4i..4i+2] (y_s, |u_t|, nu), or -1
 * outside the envelope. */
int fc_mlp_infer(const fc_mlp_params* params, const float* samples, int count, float* out);

/* ---- lattice groups: batched stepping (no reference counterpart) ---------- */
/* A group steps several initialized solvers (e.g. the cases of an angle-of-
 * attack polar sweep, each with its own grid) with one batched step launch
 * per step; the launches of a batch are captured into a HIP graph and
 * replayed. fc_group_step(g, n) is equivalent to fc_step(member, n) for
 * every member (same cell updates, ramps, macro store, force sample and
 * wall-model update per batch; the NaN watchdog samples at the batch end
 * once 200 or more steps have passed). Members must share the storage format
 * (16- or 32-bit, population array under 2 GiB), have x / y not periodic, no
 * refinement patch and no q-LIBB links; up to 64 members. */
typedef struct fc_group fc_group;
typedef struct fc_group_info {
    int members;
    long long cells;
    int waves_per_block;  /* launch shape: 1, 2, 4 or 8 wavefronts per workgroup */
    int slab_order;       /* 1: each XCD sweeps a contiguous slab of the task range */
    int use_graph;
    long long cache_budget; /* sub-batch footprint budget in bytes (0: all members per launch) */
    int sub_batches;
    double last_step_ms;  /* batched step time of the last timed batch */
    double glups;         /* aggregate: all members' cells / last_step_ms */
} fc_group_info;
fc_group* fc_group_create(fc_solver* const* members, int count, char* err, int errlen);
void fc_group_destroy(fc_group* g);
int fc_group_step(fc_group* g, int n);
int fc_group_set_launch(fc_group* g, int waves_per_block, int slab_order, int use_graph);
/* Splits the members into consecutive sub-batches whose populations (plus
 * kind bytes) fit `bytes`; each sub-batch runs its steps before the next, so
 * a footprint that fits the last-level cache stays resident. 0 = one batch,
 * 1 = one member per sub-batch. Results do not depend on it. */
int fc_group_set_cache_budget(fc_group* g, long long bytes);
/* Waits for the last batch and reports its time per step (also stored for
 * fc_group_get_info). */
int fc_group_last_step_ms(fc_group* g, double* ms);
int fc_group_get_info(fc_group* g, fc_group_info* out);
int fc_group_last_error(fc_group* g, char* buf, int cap);
/* Times the sub-batch budgets (0, 1, 128 / 256 / 512 MiB) and then every
 * launch shape (waves per workgroup x task order), each over two batches of
 * steps_per_trial real steps (the members advance 26 x 2 x steps_per_trial
 * steps; results do not depend on the choices), keeps the fastest and
 * reports it. */
int fc_group_autotune(fc_group* g, int steps_per_trial, fc_group_info* best);

/* ---- interpolated bounce-back (q-LIBB), lbm_solver.h:311-346 -------------- */
/* setQLIBBEnabled, :315 (off frees the link field; every link reverts to half-way). */
int fc_set_qlibb_enabled(fc_solver* s, int enabled);
/* setCoarseQLinks, :329-331: dense cut bytes qfrac[cell*19 + q] (0 = no link,
 * b -> q = b/255) and the per-cell two-node mask ffmask[cell] (bit q); links /
 * fallback are the readout counts (QLIBBReadout, lbm_solver.h:97-101). */
int fc_set_qlinks(fc_solver* s, const uint8_t* qfrac, const uint32_t* ffmask, int links, int fallback);
/* setFineQLinks / setFinerQLinks, :332-337 (lbm_solver.cpp:1049-1063): the
 * same fields over the fine patch (fine cells) or the nested finer patch
 * (finer cells); FC_ERR_STATE when that level is not active. The link field
 * is freed with its level. The pre-pass of a level runs just before each of
 * its sub-steps, after that sub-step's shell fill and band snapshots. */
int fc_set_fine_qlinks(fc_solver* s, const uint8_t* qfrac, const uint32_t* ffmask, int links, int fallback);
int fc_set_finer_qlinks(fc_solver* s, const uint8_t* qfrac, const uint32_t* ffmask, int links, int fallback);
/* qlibbReadout, :346 (+ active_links: links the kernel interpolates). */
int fc_qlibb_readout(fc_solver* s, int* enabled, int* links, int* fallback, int* active_links);

/* ---- tracer particles (src/render/particles.cuh:29-82, lbm_core.cuh:403-407) */
/* Particles are 4 floats each: x, y, z in lattice-cell space, age in [0,1]. */
/* launchParticleSeed: (re)allocate count particles, uniform box seed (lowbias32). */
int fc_particles_seed(fc_solver* s, int count, unsigned seed);
/* launchParticleAdvectRK2 on the solver-owned particles through the macro
 * velocity field of the last fc_step call (dt_steps lattice steps; per-call seed). */
int fc_particles_advect(fc_solver* s, float dt_steps, unsigned seed);
/* launchAdvectParticles: same advection on a caller-owned device buffer. */
int fc_advect_particles(fc_solver* s, void* device_pos4, int count, float dt_steps, unsigned seed);
int fc_particles_read(fc_solver* s, float* pos4);
int fc_particles_write(fc_solver* s, const float* pos4);
int fc_particles_count(fc_solver* s);
/* ParticleAdvectParams ageRate / inletSpawnFraction (defaults 0.002 / 0.3). */
int fc_particles_set_params(fc_solver* s, float age_rate, float inlet_fraction);

/* ---- warm-start snapshots (src/sim/snapshot.h, snapshot.cpp) ------------- */
#define FC_SOLVER_VERSION 1  /* kSolverVersion, snapshot.h:22 */
/* SnapshotKey (snapshot.h:33-64). geometry_id: "naca:2412", "dat:<file>",
 * "stl:<16-hex hash>" (NUL-terminated, at most 4096 bytes for files). */
typedef struct fc_snapshot_key {
    const char* geometry_id;
    int aoa_bucket;       /* fc_snapshot_aoa_bucket(aoa_deg) = lround(2 aoa) */
    float u_lat;
    fc_dims dims;
    int solver_version;   /* FC_SOLVER_VERSION */
} fc_snapshot_key;
int fc_snapshot_aoa_bucket(float aoa_deg);                        /* bucketForAoA, :45 */
uint64_t fc_snapshot_key_hash(const fc_snapshot_key* key);        /* hash, snapshot.cpp:89 */
int fc_snapshot_key_string(const fc_snapshot_key* key, char* out, int cap);  /* toString, :101 */
/* "<16 hex digits of the hash>.fcs" (cacheFileName, snapshot.cpp:76-81) */
int fc_snapshot_key_file_name(const fc_snapshot_key* key, char* out, int cap);
int fc_snapshot_key_equal(const fc_snapshot_key* a, const fc_snapshot_key* b);  /* operator==, :52 */
uint64_t fc_hash_flags(const uint8_t* flags, uint64_t count);     /* hashFlags, FNV-1a 64 */

/* VramSnapshot (snapshot.h:77-122): a device-resident exact copy of the
 * populations (coarse f, outlet feed planes, and the fine / finer levels'
 * f when active) with their layout parity. Capture and restore are
 * device-to-device copies on the solver's stream; restore needs the same
 * dims, storage and level shapes and runs notifySnapshotRestored(true).
 * Each object is its own slot. */
typedef struct fc_vram_snapshot fc_vram_snapshot;
fc_vram_snapshot* fc_vram_snapshot_create(void);
void fc_vram_snapshot_destroy(fc_vram_snapshot* snap);
int fc_vram_snapshot_capture(fc_vram_snapshot* snap, fc_solver* s, const fc_snapshot_key* key,
                             char* err, int errlen);
int fc_vram_snapshot_restore(fc_vram_snapshot* snap, fc_solver* s);
int fc_vram_snapshot_has_data(const fc_vram_snapshot* snap);
/* The stored key; geometry_id points into id_buf (id_cap bytes). */
int fc_vram_snapshot_get_key(const fc_vram_snapshot* snap, fc_snapshot_key* out, char* id_buf, int id_cap);
uint64_t fc_vram_snapshot_flags_hash(const fc_vram_snapshot* snap);
long long fc_vram_snapshot_bytes(const fc_vram_snapshot* snap);
void fc_vram_snapshot_release(fc_vram_snapshot* snap);

/* CompactSnapshot (snapshot.h:124-180): host copy of rho, u, v, w (the macro
 * fields of the last stepped batch), key, flags hash and scaling; restore =
 * fc_restore_from_macroscopic. Files are the FCS1 container (snapshot.cpp:
 * 287-424): u32 magic 0x31534346, u32 version 1, u32 id length + id bytes,
 * i32 aoa bucket, f32 u_lat, i32 nx ny nz, i32 solver version, u64 flags
 * hash, scaling (f32 chord_m airspeed_ms nu_m2s, i32 chord_cells, f32 u_lat
 * dx dt nu_lat_target nu_lat tau, u8 tau_clamped, f32 re_target
 * re_effective), u32 compression 0, u64 cells, then the four FP32 arrays;
 * little-endian, unpadded. */
typedef struct fc_compact_snapshot fc_compact_snapshot;
fc_compact_snapshot* fc_compact_snapshot_create(void);
void fc_compact_snapshot_destroy(fc_compact_snapshot* snap);
int fc_compact_snapshot_capture(fc_compact_snapshot* snap, fc_solver* s, const fc_snapshot_key* key,
                                char* err, int errlen);
/* Fill from host arrays (a field computed elsewhere; no solver needed). */
int fc_compact_snapshot_set(fc_compact_snapshot* snap, const fc_snapshot_key* key, uint64_t flags_hash,
                            const fc_scaling* scaling, const float* rho, const float* u, const float* v,
                            const float* w, long long ncells);
int fc_compact_snapshot_restore(fc_compact_snapshot* snap, fc_solver* s);
int fc_compact_snapshot_save(const fc_compact_snapshot* snap, const char* path, char* err, int errlen);
int fc_compact_snapshot_load(fc_compact_snapshot* snap, const char* path, char* err, int errlen);
int fc_compact_snapshot_has_data(const fc_compact_snapshot* snap);
int fc_compact_snapshot_get_key(const fc_compact_snapshot* snap, fc_snapshot_key* out, char* id_buf,
                                int id_cap);
uint64_t fc_compact_snapshot_flags_hash(const fc_compact_snapshot* snap);
int fc_compact_snapshot_scaling(const fc_compact_snapshot* snap, fc_scaling* out);
/* Copies of the held fields (cells floats each; any pointer may be NULL). */
int fc_compact_snapshot_fields(const fc_compact_snapshot* snap, float* rho, float* u, float* v, float* w);

/* DiskSnapshotCache (snapshot.h:182-219): a directory of FCS1 files named by
 * key hash, evicted oldest-modification-first past a byte budget (the file
 * just stored is never evicted); a hit refreshes the file's time. */
typedef struct fc_snapshot_cache fc_snapshot_cache;
fc_snapshot_cache* fc_snapshot_cache_open(const char* dir, uint64_t budget_bytes);
void fc_snapshot_cache_close(fc_snapshot_cache* c);  /* queued writes finish first */
/* 1 = hit (path written to path_out when non-NULL), 0 = miss, -1 = bad argument. */
int fc_snapshot_cache_find(fc_snapshot_cache* c, const fc_snapshot_key* key, char* path_out, int cap);
int fc_snapshot_cache_store(fc_snapshot_cache* c, const fc_compact_snapshot* snap, char* err, int errlen);
/* storeAsync: written by the cache's worker thread; the snapshot's data is
 * shared with the queue, so the handle may be reused or destroyed at once. */
int fc_snapshot_cache_store_async(fc_snapshot_cache* c, const fc_compact_snapshot* snap);
/* Waits until the queue is empty; returns the count of failed async stores. */
long long fc_snapshot_cache_flush(fc_snapshot_cache* c);
int fc_snapshot_cache_set_budget(fc_snapshot_cache* c, uint64_t budget_bytes);
uint64_t fc_snapshot_cache_used_bytes(const fc_snapshot_cache* c);

/* adaptiveStepsForBudget (lbm_solver.h:363, lbm_solver.cpp:1392-1408): steps
 * for the next frame within budget_ms (reference default 10 ms) from the
 * measured per-step time: 1 until measured, at most 2x the last choice,
 * at most 4096, 0 while the NaN watchdog is latched. */
int fc_adaptive_steps_for_budget(fc_solver* s, double budget_ms);

/* ---- mesh sequencing (lbm_solver.h:258, src/main.cpp:756-817) ------------- */
/* seedFromCoarse (lbm_solver.cpp:2062-2102): trilinear upsample of the
 * presolver's macro fields (lbm_refine.cu:437-501) into s, f = feq(rho, u),
 * compact-restore bookkeeping, startup ramp re-armed, patch levels re-seeded. */
int fc_seed_from_coarse(fc_solver* s, fc_solver* presolver);
/* runPreconverge: a presolver on (pre_dims, pre_scaling, pre_flags) - the
 * same geometry voxelized at max(48, chord / 4) cells per chord - runs two
 * flow-throughs in batches of 1024 steps (stopping with FC_ERR_STATE if it
 * diverges), then seeds s. steps_run (may be NULL) receives its step count. */
int fc_preconverge(fc_solver* s, fc_dims pre_dims, const fc_scaling* pre_scaling, const uint8_t* pre_flags,
                   const fc_init_options* options, long long* steps_run, char* err, int errlen);

/* ---- boundary-layer readouts and the VG vortex audit ---------------------- */
/* (lbm_solver.h:58-65, :379-412; src/geom/vg_audit.h). The readouts work on
 * the macro fields of the last fc_step call at mid-span (z = nz / 2); the
 * surface is the clean-foil reference when set (fc_set_surface_reference),
 * else the live flags. */
typedef struct fc_delta99_sample {   /* Delta99Sample */
    float x_c, delta99_c, ue_edge;
    int valid;
} fc_delta99_sample;
typedef struct fc_plane_circulation { float pos, neg; } fc_plane_circulation;  /* PlaneCirculation */
typedef struct fc_band_circulation {  /* BandCirculation */
    fc_plane_circulation zone, ambient;
    int zone_cols, ambient_cols;
} fc_band_circulation;
typedef struct fc_vg_audit_readout {  /* VGAuditReadout */
    int valid;
    float gamma_measured, gamma_predicted, ratio, plane_xc;
} fc_vg_audit_readout;
/* extractSuctionDelta99 (lbm_solver.cpp:1511-1584): one sample per station. */
int fc_extract_suction_delta99(fc_solver* s, const float* stations_xc, int count, fc_delta99_sample* out);
/* separationOnsetXc (:1587-1617): onset x/c, or -1 when attached. */
int fc_separation_onset_xc(fc_solver* s, float* out);
/* downloadCrossflowPlane (:1619-1640): v, w at column ix, ny*nz each (y + ny*z). */
int fc_download_crossflow_plane(fc_solver* s, int ix, float* v, float* w);
int fc_suction_surface_y(fc_solver* s, int ix);                 /* suctionSurfaceY, :1642; -1 if none */
int fc_lattice_x_for_chord_station(fc_solver* s, float xc);    /* latticeXForChordStation, :1648 */
/* integrateBandCirculation of the live field's column ix computed on the
 * device (FP64 sums); same clamps and empty-band result as the host form. */
int fc_band_circulation_live(fc_solver* s, int ix, int y_min, int y_max, int z0, int z1,
                             fc_band_circulation* out);
/* Host math of vg_audit.h:44-105 on host planes (index y + ny*z). */
float fc_wendt_circulation_lattice(float beta_rad, float ue_lat, float h_cells, float vane_chord_cells,
                                   float delta99_cells);
int fc_integrate_streamwise_circulation(const float* v, const float* w, int ny, int nz, int y_min, int y_max,
                                        fc_plane_circulation* out);
int fc_integrate_band_circulation(const float* v, const float* w, int ny, int nz, int y_min, int y_max, int z0,
                                  int z1, fc_band_circulation* out);
/* auditVGVortexStrength (vg_audit.cpp:103-187), band integral on the device. */
int fc_audit_vg_vortex_strength(fc_solver* s, const fc_vg_params* vg, int chord_cells, float ue_lat,
                                float delta99_cells, fc_vg_audit_readout* out);

/* ---- measurement --------------------------------------------------------- */
/* Per-launch HIP-event timing of the stream-collide kernel (on the solver's
 * stream); fc_kernel_timing returns launches and summed milliseconds since
 * the last enable, after synchronizing. */
int fc_set_kernel_timing(fc_solver* s, int enabled);
int fc_kernel_timing(fc_solver* s, int* launches, double* total_ms);
/* The same per level: step launches of level 0 (coarse), 1 (fine patch) or
 * 2 (nested finer patch) and their summed milliseconds. */
int fc_kernel_timing_level(fc_solver* s, int level, int* launches, double* total_ms);
/* Wall time of n fc_step calls of one step each, bracketed by HIP events on
 * the solver's stream (ms). */
int fc_time_steps(fc_solver* s, int n, int steps_per_call, double* ms);

/* Traffic-only probes of the HBM ceiling (no counterpart in the reference;
 * measurement infrastructure for the stream-collide roofline). The AA probe
 * allocates a 19-slab array of dims cells with elem_bytes (2 or 4) per
 * population, then times `reps` launches of a kernel with the step kernels'
 * decomposition that moves exactly one in-place step's bytes (19 reads and 19
 * writes per cell): pattern = 0 is the local S->L pattern, 1 the L->S pattern
 * with neighbour reads and writes, 2 the esoteric-twist pattern (slot q read
 * and written at cell + max(e_q, 0), the same for every step); `work` FP32
 * FMAs per population (19
 * independent chains) sit between the read and the write (0: none). The
 * copy probe times a plain device copy of `bytes`. Both free what they
 * allocate; ms per launch. */
int fc_probe_aa_traffic(fc_dims dims, int elem_bytes, int pattern, int work, int reps, double* ms_per_launch);
/* The step kernels' store codec of a 16-bit storage (FC_STORAGE_F16C, _F16 or
 * _BF16) run on the device over n given values (the codec's input, the scaled
 * deviation K (f - w)) with the given 32-bit random words: stochastic = 1
 * rounds as fc_set_stochastic_rounding's default, 0 to nearest even. Writes
 * the n 16-bit codes (host arrays; allocates and frees its own device
 * buffers). For checking the codec against its restatement. */
int fc_probe_store_rounding(int storage, const float* values, const unsigned* random_bits, int stochastic,
                            long long n, unsigned short* codes);
int fc_probe_copy(size_t bytes, int reps, double* ms_per_launch);

#ifdef __cplusplus
}
#endif

#endif /* FOILCFD_H */
