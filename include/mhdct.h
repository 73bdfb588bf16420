/*
 * mhdct.h -- C ABI of the MI355X MHD-CT time step (libmhdct.so).
 *
 * The entry points replace the reference's GodunovImplemBase<dim,device_t> boundary
 * (scheme/GodunovImplemBase.h:45-170) and the two solver-side kernels next to it:
 *
 *   mhdct_create        GodunovImplemBase ctor + SolverGodunovMHD::create_godunov_implem
 *                       (scheme/GodunovImplemBase.h:54-73, SolverGodunovMHD.h:326-340)
 *   mhdct_set_mesh      GodunovImplemV0::resize_auxiliary_data (scheme/GodunovImplemV0.cpp:40-75)
 *                       after the host octree changed (SolverGodunovMHD.cpp:307)
 *   mhdct_mem_bytes     GodunovImplemV0::total_mem_size_in_bytes (scheme/GodunovImplemV0.cpp:79-97)
 *   mhdct_do_time_step  GodunovImplemV0::do_time_step (scheme/GodunovImplemBase.h:85-90,
 *                       scheme/GodunovImplemV0.cpp:101-207)
 *   mhdct_compute_dt    SolverGodunovMHD::compute_dt_local (SolverGodunovMHD.cpp:377-403,
 *                       scheme/ComputeDtMHDFunctor.cpp:44-151)
 *   mhdct_fill_outside  SolverGodunovMHD::fill_outside_quadrants (SolverGodunovMHD.cpp:1068-1088,
 *                       border_conditions/FillOutside.cpp:47-395)
 *
 * The mesh entry points stand in for the p4est / kalypsso::core AMR manager the reference
 * keeps on the host (refine, coarsen and 2:1 balance stay host-side).
 *
 * Data layout (both sides of the ABI, the reference's DataArrayBlock / FaceDataArrayBlock):
 *   U[block][var][cell]      var = rho, E, rho*u, rho*v, rho*w; cell = i + b*(j + b*k)
 *   B[block][face]           per block: the x-face grid (b+1) x b x b, then the y-face grid
 *                            b x (b+1) x b, then the z-face grid b x b x (b+1), each x-fastest;
 *                            in 2D: (b+1) x b, b x (b+1) and the cell-centred Bz b x b.
 *   blocks are the owned leaves in mesh order followed by the outside blocks.
 * All pointers passed to the step functions are device pointers (FP64).  Streams are
 * hipStream_t passed as void* (NULL = the default stream).
 *
 * Every function returns 0 on success and a negative code on error; the message is in
 * mhdct_last_error / mhdct_mesh_last_error.  Nothing aborts the process.
 */
#ifndef MHDCT_H
#define MHDCT_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct mhdct_ctx  mhdct_ctx;
typedef struct mhdct_mesh mhdct_mesh;

/* [hydro] / [mhd] settings the hot path reads (models/HydroSettings.cpp:19-30). */
typedef struct mhdct_settings
{
  double gamma0;         /* [hydro] gamma0, default 1.4 */
  double cfl;            /* [hydro] cfl, default 0.5 */
  double slope_type;     /* [hydro] slope_type: 1 minmod, 2 MC, default 1 */
  double smallr;         /* [hydro] smallr, default 1e-10 */
  double smallc;         /* [hydro] smallc, default 1e-10 (smallp = smallc^2/gamma0) */
  double scaling_factor; /* [mesh] scaling_factor, default 1 */
  int    riemann;        /* RiemannSolverType (models/riemann_solver_types.h): 5 = HLLD, the only one supported */
  int    prolongation;   /* [amr] cell_centered_prolongation_type of the primitive conversion's ghost cells
                            beside a coarser block (ConvertToPrimitivesVariablesFunctor.cpp:697-710):
                            0 = SIMPLE_COPY, 1 = EXTRAPOLATE_LINEAR_MINMOD */
  int    abort_when_negative_eint; /* [hydro] abort_when_negative_eint: report eint < 0 as an error */
  int    time_integrator; /* [run] time_integrator: MHDCT_TIME_HANCOCK or MHDCT_TIME_RK2_SSP */
  int    gravity_enabled; /* [gravity] enabled: a uniform gravity field (AddGravitySourceTerm.cpp) */
  double gravity[3];      /* [gravity] uniform_gx, uniform_gy, uniform_gz */
  int    viscosity_enabled; /* [viscosity] enabled: viscous fluxes after the hydro ones
                               (GodunovImplemV0.cpp:179-192); the FP64 forms (plain, fused),
                               one context per step (not in a batch) */
  double viscosity_mu;      /* [viscosity] mu, the dynamic viscosity (>= 0) */
  int    viscosity_hancock_predictor; /* [viscosity] hancock_predictor_enabled: the viscous term of
                                         the half-step predictor (ComputeUpdatedPrimitiveVariablesFunctor.cpp:380-383) */
} mhdct_settings;

/* [run] time_integrator values (SolverGodunovMHD.cpp:494, 551; RK3_SSP is rejected, as the
 * reference aborts on it, SolverGodunovMHD.cpp:174-177).  Under RK2_SSP, do_time_step is one
 * forward-Euler stage: no half-step predictor (ComputeUpdatedPrimitiveVariablesFunctor.cpp:396-407)
 * and no face-B source in the edge states (ComputeEmfAndStoreFunctor.cpp:189-193, 297-301). */
#define MHDCT_TIME_HANCOCK 0
#define MHDCT_TIME_RK2_SSP 1

/* implementation_version values ([hydro] implementation_version, SolverGodunovMHD.cpp:76) */
#define MHDCT_IMPL_PLAIN 10 /* one kernel per stage, drop-in for GodunovImplemV0 */
#define MHDCT_IMPL_FUSED 11 /* fused stages, FP64 */
#define MHDCT_IMPL_FUSED_FP32 12 /* fused stages, FP32 slopes/HLLD/EMF, FP64 primitive recovery, accumulation and dt */

/* ---- host block octree (p4est brick forest replacement) ---------------------------- */

/* Uniform forest at `level` over an nbrick[0] x nbrick[1] (x nbrick[2]) brick of trees. */
int mhdct_mesh_create(int dim, int block_size, const int nbrick[3], const int periodic[3], int level,
                      mhdct_mesh ** out);
/* Border-condition type per face (XMIN,XMAX,YMIN,YMAX,ZMIN,ZMAX), BC_MHD enum values
 * (border_conditions/FillOutside.h:55-61): 2 ZERO_GRADIENT, 3 WALL.  Only read for
 * non-periodic directions. */
int mhdct_mesh_set_bc(mhdct_mesh * mesh, const int bc[6]);
/* Per-leaf flags: +1 refine, -1 coarsen (applied when all siblings ask), 0 keep. */
int mhdct_mesh_adapt(mhdct_mesh * mesh, const int8_t * flags, int nleaves);
/* 2:1 balance across faces, edges and corners. */
int mhdct_mesh_balance(mhdct_mesh * mesh);
/* Replace the leaves by n given leaves (levels[n], coords[3n] in blocks at each leaf's level)
 * that tile the brick exactly: the leaf set of a checkpoint, in place of reading the p4est file
 * (init/init.cpp:123-127).  Call mhdct_mesh_finalize after it. */
int mhdct_mesh_set_leaves(mhdct_mesh * mesh, int n, const int * levels, const int * coords);
/* Outside blocks and neighbour tables; call after the last adapt/balance. */
int mhdct_mesh_finalize(mhdct_mesh * mesh);
int mhdct_mesh_num_leaves(const mhdct_mesh * mesh);
int mhdct_mesh_num_outside(const mhdct_mesh * mesh);
/* levels[n] and coords[3n] (block units at that level) of owned leaves then outside blocks */
int mhdct_mesh_leaves(const mhdct_mesh * mesh, int * levels, int * coords);
/* source leaf and outward direction (3 ints) of every outside block */
int mhdct_mesh_outside(const mhdct_mesh * mesh, int * src, int * dirs);
const char * mhdct_mesh_last_error(const mhdct_mesh * mesh);
void         mhdct_mesh_destroy(mhdct_mesh * mesh);

/* ---- GodunovImplemBase ---------------------------------------------------------------- */

/* -4: unknown impl_version; -5: settings this build does not run (a Riemann solver other than
 * HLLD, a prolongation other than SIMPLE_COPY or EXTRAPOLATE_LINEAR_MINMOD, an unknown time integrator, viscosity in the
 * FP32-bulk form or with mu < 0); -6: unequal block sizes. */
int      mhdct_create(const mhdct_settings * settings, int dim, const int block_size[3], int impl_version,
                      mhdct_ctx ** out);
/* Takes a finalized mesh: uploads its tables and sizes the scratch.  The context keeps no
 * reference to `mesh` afterwards.  Synchronises the device first (work in flight may still
 * read the previous mesh's tables, which are reused when the new mesh fits them). */
int      mhdct_set_mesh(mhdct_ctx * ctx, const mhdct_mesh * mesh);
uint64_t mhdct_mem_bytes(const mhdct_ctx * ctx);
/* U2 == U and B2 == B on entry (the solver's contract); on exit U2, B2 hold t + dt for the
 * owned blocks.  Asynchronous on `stream`. */
int mhdct_do_time_step(mhdct_ctx * ctx, const double * U, double * U2, const double * B, double * B2, double dt,
                       void * stream);
/* dt = cfl / max over owned cells of sum_d (|v_d| + c_fast,d) / dx.  Synchronises `stream`;
 * returns -3 if a NaN state was met and -2 if eint < 0 was flagged by a previous step with
 * abort_when_negative_eint set. */
int mhdct_compute_dt(mhdct_ctx * ctx, const double * U, const double * B, double * dt_out, void * stream);
/* Fill the outside blocks from their sources by the border conditions (no-op when periodic). */
int mhdct_fill_outside(mhdct_ctx * ctx, double * U, double * B, void * stream);
/* U_RK = c0 * U + c1 * U_RK and B_RK = c0 * B + c1 * B_RK over every block (owned and outside),
 * each value as two products and a sum.  Replaces LinearCombination<dim>::apply of the RK2-SSP
 * driver (SolverGodunovMHD.cpp:583-588, 617-622), whose stage coefficients are (0, 1) and
 * (1/2, 1/2).  Asynchronous on `stream`. */
int mhdct_linear_combination(mhdct_ctx * ctx, const double * U, double * U_RK, const double * B, double * B_RK,
                             double c0, double c1, void * stream);
const char * mhdct_last_error(const mhdct_ctx * ctx);
void         mhdct_destroy(mhdct_ctx * ctx);

/* ---- AMR cycle (SolverGodunovMHD::do_amr_cycle, SolverGodunovMHD.cpp:263-319) -----------
 * The reference computes the refinement flags on the device (ComputeRefineFlags), adapts and
 * balances the forest on the host (p4est) and remaps U and Bface on the device
 * (UserDataRemapper); the first and the last are kalypsso::core, absent from its tree, and are
 * restated here (amr.hip): Loehner's unsplit second-derivative indicator with the noise filter
 * epsilon_lohner, SIMPLE_COPY / mean cell prolongation and restriction, divergence-free face-B
 * prolongation and flux-conserving face-B restriction. */
typedef struct mhdct_amr_params
{
  int    level_min;       /* [amr] level_min */
  int    level_max;       /* [amr] level_max */
  double epsilon_refine;  /* [amr] epsilon_refine, default 0.01 */
  double epsilon_coarsen; /* [amr] epsilon_coarsen, default 0.00125 */
  double epsilon_lohner;  /* [amr] epsilon_lohner, default 0.02 */
  int    ivar;            /* the refinement variable when nvars == 0: a variable of U, 0 rho */
  int    criterion;       /* [amr] refine_criterion: MHDCT_REFINE_LOHNER_UNSPLIT or MHDCT_REFINE_SIMPLE_GRADIENT */
  int    nvars;           /* [amr] variable_used_to_compute_refine_criterion: 1-4 entries of vars (0: ivar) */
  int    vars[4];         /* 0-4 a variable of U (rho, e_tot, rho_vx, rho_vy, rho_vz) or MHDCT_RVAR_* */
} mhdct_amr_params;

/* [amr] refine_criterion values (kalypsso::core's RefineIndicatorData, absent from the reference
 * tree): Loehner's unsplit second-derivative estimate with the noise filter epsilon_lohner, or the
 * largest normalised jump |q_n - q_c| / max(|q_n|, |q_c|) to a face neighbour. */
#define MHDCT_REFINE_LOHNER_UNSPLIT 0
#define MHDCT_REFINE_SIMPLE_GRADIENT 1
/* derived refinement variables (DERIVED_QUANTITY, utils/ComputeDerivedQuantities.h:44-53), from the
 * cell's primitives; zero in outside blocks, as the reference's derived array is only computed
 * for the owned blocks (:104, 109-111) */
#define MHDCT_RVAR_THERMAL_PRESSURE 10
#define MHDCT_RVAR_SPEED_OF_SOUND 11
#define MHDCT_RVAR_SPECIFIC_EKIN 12
#define MHDCT_RVAR_MAGNETIC_PRESSURE 13
#define MHDCT_RVAR_LOCAL_MACH_NUMBER 14
#define MHDCT_RVAR_VX 15
#define MHDCT_RVAR_VY 16
#define MHDCT_RVAR_VZ 17

/* Replaces compute_refine_flags (SolverGodunovMHD.cpp:1094-1185): the indicator of every owned
 * block (the maximum over its cells and over the refinement variables, each read with one ghost
 * layer, so the outside blocks must be filled) and its flag, +1 refine (indicator >
 * epsilon_refine and level < level_max, or level < level_min), -1 coarsen (indicator <
 * epsilon_coarsen and level > level_min, or level > level_max), 0 keep, into host arrays of
 * num_leaves entries (`indicator` may be NULL).  With several variables a block is refined when
 * any asks it and coarsened when all do.  B (the face field) is read for derived variables only
 * and may be NULL otherwise.  Synchronises `stream`. */
int mhdct_refine_flags(mhdct_ctx * ctx, const double * U, const double * B, const mhdct_amr_params * params,
                       int8_t * flags, double * indicator, void * stream);
/* No counterpart in the reference.  mhdct_refine_flags for callers that cannot allocate device
 * memory (an AMR cycle inside a serving loop or another allocation-free host loop): the same
 * results as mhdct_refine_flags, with its device buffers (per-block maxima, indicators, flags)
 * in a file-scope device array of the library sized for at most MHDCT_NOALLOC_MAX_BLOCKS owned
 * blocks (4.25 MiB, one copy per GPU, held until the process ends); a larger mesh is refused
 * with -8 before any device work.  The array keeps the last call's contents; calls that overlap
 * on one GPU (several contexts or streams) race on it.  The context's device must be current.
 * Synchronises `stream`. */
#define MHDCT_NOALLOC_MAX_BLOCKS 262144
int mhdct_refine_flags_noalloc(mhdct_ctx * ctx, const double * U, const double * B, const mhdct_amr_params * params,
                               int8_t * flags, double * indicator, void * stream);
/* A copy of a mesh (the leaf set before an adaptation, kept for the remap). */
int mhdct_mesh_clone(const mhdct_mesh * mesh, mhdct_mesh ** out);

/* Replaces UserDataRemapper (map_userdata_after_adapt, SolverGodunovMHD.cpp:1191-1269): the
 * owned leaves of new_mesh from those of old_mesh, each a copy of an old leaf, a prolongation
 * of its old parent, or a restriction of its old children (one level of change per cycle).
 * The plan is built on the host; counts[3] = copied, prolonged, restricted leaves. */
typedef struct mhdct_remap mhdct_remap;
int mhdct_remap_create(const mhdct_mesh * old_mesh, const mhdct_mesh * new_mesh, mhdct_remap ** out);
int mhdct_remap_counts(const mhdct_remap * remap, int counts[3]);
/* U_new, B_new (the new mesh's layout) for its owned leaves from U_old, B_old (the old mesh's);
 * the outside blocks of U_new, B_new are left for mhdct_fill_outside.  Asynchronous on `stream`. */
int mhdct_remap_apply(mhdct_remap * remap, const double * U_old, const double * B_old, double * U_new,
                      double * B_new, void * stream);
/* the message of the last failed mhdct_remap_create (remap NULL) or mhdct_remap_apply */
const char * mhdct_remap_last_error(const mhdct_remap * remap);
void         mhdct_remap_destroy(mhdct_remap * remap);

/* ---- batches ---------------------------------------------------------------------------
 * No counterpart in the reference, which advances one mesh per process
 * (SolverGodunovMHD::next_iteration_impl, SolverGodunovMHD.cpp:409-484): a batch advances
 * many fused-form contexts (impl 11 or 12, one dimension) together, each with its own CFL dt
 * (the ComputeDtMHD reduction and dt = cfl / inv_dt on the device, SolverGodunovMHD.cpp:377-403),
 * through one kernel per stage over the concatenated work list of every slot and block size.
 * A slot's results are the same alone and in any batch.
 *
 * Each slot's state ping-pongs between two caller-owned buffer pairs: it starts in (U, B);
 * after n batch steps it is in (U, B) if n is even, else in (U2, B2).  A context lent to a
 * batch must not be given a new mesh or destroyed before it is removed. */
typedef struct mhdct_batch mhdct_batch;

int mhdct_batch_create(int dim, int impl_version, mhdct_batch ** out);
/* Add a context (mesh set, same dim and impl) with its state buffers; the slot starts at the
 * next batch step.  A batch holds contexts of one time integrator: HANCOCK, or RK2_SSP, whose
 * step (SolverGodunovMHD.cpp:551-647) runs both stages through a stage buffer the batch
 * allocates for the slot.  Synchronises the device. */
int mhdct_batch_add(mhdct_batch * batch, mhdct_ctx * ctx, double * U, double * U2, double * B, double * B2,
                    int * slot_id);
/* Remove a slot after the steps enqueued so far: waits for the stream of the batch's last steps
 * (and, for an RK2_SSP slot, for the device). */
int mhdct_batch_remove(mhdct_batch * batch, int slot_id);
/* Enqueue `nsteps` steps of every slot on `stream` (outside fill, dt, stages).  A batch's steps
 * go to one stream at a time: the calls that read its clocks (remove, slot_state) wait for the
 * stream of its last steps only, so that batches on other streams keep running. */
int mhdct_batch_step(mhdct_batch * batch, int nsteps, void * stream);
/* The same as mhdct_batch_step without graphs, with HIP events recorded around each stage on
 * the stream it runs on: stage_ms[MHDCT_NUM_STAGES] receives the average milliseconds per step
 * of each stage (names from mhdct_stage_name; with the overlap on, k_edge runs beside
 * k_border_flux and k_flux_tile; with k_front selected, stage 2 is k_front and stages 3 and 4
 * launch nothing).  HANCOCK batches only (-1 for an RK2_SSP batch).
 * Synchronises `stream`. */
#define MHDCT_NUM_STAGES 10
int          mhdct_batch_step_timed(mhdct_batch * batch, int nsteps, void * stream, double * stage_ms);
const char * mhdct_stage_name(int stage);
/* Simulated time, steps taken and which buffer pair holds the state (0: U/B, 1: U2/B2).
 * Synchronises the device. */
int mhdct_batch_slot_state(mhdct_batch * batch, int slot_id, int * current, double * time, int64_t * steps);
/* Error flags of every slot (eint < 0 with abort_when_negative_eint: -2; NaN speed: -3, with
 * the slot in bad_slot).  Synchronises the device. */
int mhdct_batch_check(mhdct_batch * batch, int * bad_slot);
/* Capture each step (per ping-pong parity) in a HIP graph and replay it. */
int mhdct_batch_set_graph(mhdct_batch * batch, int enable);
/* Run k_edge on a second stream beside the flux stages (default off: k_edge fills every CU's
 * registers, so on MI355X the two branches measured no overlap), joined before the EMF
 * correction; the results do not depend on it. */
int mhdct_batch_set_overlap(mhdct_batch * batch, int enable);
/* Launch tunables: at most grid_cap workgroups per element kernel and tile_cap per flux-tile
 * kernel (multiples of 8; below the work size the kernels loop), and the 3D flux-tile depth
 * tile_z: 8 (8x8x8 cells, 576-thread workgroups) or 4 (8x8x4, 320 threads). */
int mhdct_batch_set_tune(mhdct_batch * batch, int grid_cap, int tile_cap, int tile_z);
/* A batch of several block sizes: enable = 1 (the default) launches every stage once per
 * block-size class, over that class's slots, with the kernels specialised to its block size; 0
 * launches each stage once over all slots with kernels that dispatch on each slot's block size
 * (three times the code).  The results are the same bit for bit.  No counterpart in the reference
 * (a launch choice of this library). */
int mhdct_batch_set_class_launch(mhdct_batch * batch, int enable);
/* How the primitive, electric-field and cell stages run: enable = 0 (the default) as three
 * kernels (k_prim, k_efield, k_cell) through HBM scratch; 1 as one workgroup per 8 x 8 x 4 cell
 * sub-tile (8 x 8 x 2 in 8^3 blocks; 8 x 8 in 2D), k_front, holding the sub-tile's primitive
 * state, face field and edge electric field in LDS.  The results are the same bit for bit; on
 * MI355X k_front is the slower of the two (it recomputes the halo of every sub-tile).  No
 * counterpart in the reference: a launch choice of this library (GodunovImplemV0.cpp:128-156 runs
 * these stages as separate device loops). */
int mhdct_batch_set_front(mhdct_batch * batch, int enable);
/* Workgroup sizes of the cell stage (k_cell: 256 or 512 threads) and of the edge stage
 * (k_edge: 256 or 512); a larger workgroup puts more consecutive wavefronts, which read the
 * same cells, on one CU.  Same results bit for bit.  No counterpart in the reference (a launch
 * choice). */
int mhdct_batch_set_stage_threads(mhdct_batch * batch, int cell_threads, int edge_threads);
/* Time each candidate (grid_cap, tile_cap, tile_z) over `reps` steps that leave the slots'
 * states and clocks unchanged, keep the fastest and return it. */
int mhdct_batch_autotune(mhdct_batch * batch, int reps, void * stream, int * grid_cap, int * tile_cap,
                         int * tile_z);
int          mhdct_batch_num_slots(const mhdct_batch * batch);
const char * mhdct_batch_last_error(const mhdct_batch * batch);
void         mhdct_batch_destroy(mhdct_batch * batch);

/* ---- .ini settings ---------------------------------------------------------------- */

/* The run's .ini file, read as the reference reads it through kalypsso::core's ConfigMap
 * (absent from the reference tree; its calls are at models/HydroSettings.cpp:19-30,
 * SolverGodunovMHD.cpp:76, 1107-1113, app/solver_godunov_mhd_ct.cpp:72, 87, init/init.cpp:40,
 * 116-117).  Lines are `[section]`, `key = value`, blank, or comments starting with '#' or ';'.
 * Section and key names are case-insensitive; a repeated key keeps its last value. */
typedef struct mhdct_config mhdct_config;
/* 0; a positive line number (the first line of none of those forms; *out stays NULL); -1. */
int mhdct_config_parse(const char * text, mhdct_config ** out);
/* As mhdct_config_parse on a file's text; -2 when the file cannot be read. */
int mhdct_config_load(const char * path, mhdct_config ** out);
/* The value of section/key as written (blanks around it dropped), or NULL when absent. */
const char * mhdct_config_get(const mhdct_config * cfg, const char * section, const char * key);
/* mhdct_settings from the keys the time step reads, with the reference's defaults:
 * [hydro] gamma0 1.4, cfl 0.5, slope_type 1.0, smallr 1e-10, smallc 1e-10, riemann (no default:
 * a missing or unknown name gives riemann 0, UNKNOWN, which mhdct_create refuses, as the
 * reference fails on it, RiemannSolvers_MHD.h:586), abort_when_negative_eint false;
 * [mesh] scaling_factor 1; [amr] cell_centered_prolongation_type SIMPLE_COPY;
 * time_integrator HANCOCK, read from [run] (or from [amr], where the shipped .ini files write it;
 * RK3_SSP is refused, SolverGodunovMHD.cpp:174-177); [gravity] enabled false, uniform_gx,
 * uniform_gy, uniform_gz 0 (z in 3D only). */
int mhdct_config_settings(mhdct_config * cfg, mhdct_settings * out);
/* mhdct_amr_params from [amr] level_min and level_max (required), epsilon_refine 0.01,
 * epsilon_coarsen 0.00125, epsilon_lohner 0.02 (SolverGodunovMHD.cpp:1107-1110),
 * variable_used_to_compute_refine_criterion [rho] (:1112-1113; one to four comma-separated names:
 * a derived quantity in any case, else a variable of U, else rho, as at :1117-1160) and
 * refine_criterion LOHNER_UNSPLIT (default) or SIMPLE_GRADIENT. */
int mhdct_config_amr_params(mhdct_config * cfg, mhdct_amr_params * out);
/* The mesh the run starts from: [run] dimension (required), [amr] bx = by (= bz) (required),
 * uniform at [amr] level_min, over [p4est_connectivity] nbrick_{x,y,z} (default 1) with
 * periodic_{x,y,z} (default 1), [mesh] boundary_type_{x,y,z}{min,max} (NONE, PERIODIC,
 * ZERO_GRADIENT or WALL; default ZERO_GRADIENT).  Not finalized. */
int mhdct_config_mesh(mhdct_config * cfg, mhdct_mesh ** out);
const char * mhdct_config_last_error(const mhdct_config * cfg);
void         mhdct_config_destroy(mhdct_config * cfg);

#ifdef __cplusplus
}
#endif

#endif /* MHDCT_H */
