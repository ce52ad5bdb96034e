/*
 * meso_hip.h — C ABI of the MI355X (gfx950) DUGKS/CDUGKS time step.
 *
 * The step replaces meso's per-step update, i.e. the body of
 *   CDUGKS_SHAKHOV::do_step()  src/solver/boltzmann/cdugks_shakhov.cpp:724-764
 *   CDUGKS::do_step()          src/solver/boltzmann/cdugks.cpp:337-371
 *   DUGKS::do_step()           src/solver/boltzmann/dugks.cpp:346-383
 * (reconstruct() + fvm_update() + residual bookkeeping), called from the
 * time loop of handle_solver<T>() (src/meso/exec.cpp:81-165).
 *
 * Conventions
 *   - Every entry point returns 0 on success and a negative MESO_E_* code on
 *     failure; meso_hip_last_error()/meso_case_last_error() give the text.
 *     No C++ exception crosses this boundary.
 *   - Host arrays passed in are borrowed for the duration of the call only.
 *   - A context owns its device memory and one HIP stream (default priority);
 *     calls on one context are stream-ordered and not thread-safe.
 *   - Distributions on the host are FP64. Layout VELOCITY_MAJOR is meso's own
 *     DistributionFunction layout f[p][cell] (include/solver/solver.h:55);
 *     CELL_MAJOR is f[cell][p]. For two distributions (Shakhov g, h) the
 *     buffer holds all of g followed by all of h.
 */
#ifndef MESO_HIP_H
#define MESO_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
#define MESO_OK 0
#define MESO_E_INVALID (-1)  /* bad argument / unsupported case        */
#define MESO_E_IO (-2)       /* file could not be read or parsed       */
#define MESO_E_DEVICE (-3)   /* HIP runtime error                      */
#define MESO_E_NUMERIC (-4)  /* NaN density or T<0 detected in a step  */

/* ---- enumerations ------------------------------------------------------ */
/* solver selected by the case key `solver` (src/meso/main.cpp:47-57) */
#define MESO_MODEL_CDUGKS_BGK 1     /* "cdugks@incompressible" */
#define MESO_MODEL_CDUGKS_SHAKHOV 2 /* "cdugks@shakhov"        */
#define MESO_MODEL_DUGKS_BGK 3      /* "dugks@incompressible"  */

#define MESO_STORAGE_FP64 0 /* f, fluxes, gradients in FP64 */
#define MESO_STORAGE_FP32 1 /* f, fluxes, gradients in FP32; moments, macros, residual in FP64 */
/* f as FP32 plus an 8-bit extension of its mantissa (a 31-bit significand of
 * the deviation from the rest Maxwellian, 5 bytes per value); the step runs
 * the FP64 kernels (face values, fluxes and equilibria in FP64) */
#define MESO_STORAGE_FP32X 2
/* f in the FP32X format (31-bit significand, 5 bytes per value); the step runs
 * the FP32 kernels (face values and fluxes handed over in FP32, stochastically
 * rounded; equilibria in FP32; transport arithmetic, moments and macros in FP64) */
#define MESO_STORAGE_FP32E 3

#define MESO_LAYOUT_VELOCITY_MAJOR 0
#define MESO_LAYOUT_CELL_MAJOR 1

/* Boundary kinds: the values of meso's BoundaryType (include/solver/config.h:8-18). */
#define MESO_BC_FLUID_INTERIOR 0
#define MESO_BC_FARFIELD_INLET 1
#define MESO_BC_PRESSURE_INLET 2
#define MESO_BC_FREESTREAM_INLET 3
#define MESO_BC_FARFIELD_OUTLET 4
#define MESO_BC_PRESSURE_OUTLET 5
#define MESO_BC_WALL 6
#define MESO_BC_SLIP_WALL 7
#define MESO_BC_SYMMETRY 8

/* Patch kinds of a mark's "temperature" key (include/solver/config.h:20-30). */
#define MESO_PATCH_ZERO_GRADIENT 0
#define MESO_PATCH_FIXED_VALUE 1

/* ---- plain-data descriptors -------------------------------------------- */

/* One [mark] block of the case file (include/solver/config.h:66-72). */
typedef struct meso_mark_desc {
    int32_t type;      /* MESO_BC_*                                    */
    int32_t T_kind;    /* MESO_PATCH_* of key "temperature"            */
    int32_t direction; /* mirror axis (0,1,2) for symmetry / slip-wall */
    int32_t reserved;
    double velocity[3];
    double temperature;
    double density;
    double pressure;
    double alpha_u; /* slip-wall-alpha-u */
    double alpha_T; /* slip-wall-alpha-T */
} meso_mark_desc;

/* Mesh topology + geometry as built by fvmMesh::load_gambit
 * (src/mesh/object.cpp:38-172, src/mesh/geom.cpp:165-211). */
typedef struct meso_mesh_desc {
    int32_t dim;
    int32_t n_cells;
    int32_t n_faces;
    int32_t n_marks;
    int32_t n_nbr;        /* entries of the neighbour CSR      */
    int32_t n_cell_faces; /* entries of the cell->face CSR     */
    double total_volume;
    double min_cell_size;
    const double *cell_pos;       /* [n_cells][3] node-average centroid        */
    const double *cell_vol;       /* [n_cells]                                 */
    const int32_t *face_cells;    /* [n_faces][2] owner, neighbour (=owner on a boundary) */
    const double *face_pos;       /* [n_faces][3] node-average face centre     */
    const double *face_area;      /* [n_faces]                                 */
    const double *face_normal;    /* [n_faces][3] unit normal out of the owner */
    const int32_t *cell_face_ptr; /* [n_cells+1]                               */
    const int32_t *cell_face_ids; /* [n_cell_faces] faces of each cell, local order */
    const int32_t *nbr_ptr;       /* [n_cells+1] least-squares neighbours      */
    const int32_t *nbr_ids;       /* [n_nbr]                                   */
    const double *lsq_dr;         /* [n_nbr][3] x_nbr - x_cell                 */
    const double *lsq_w;          /* [n_nbr] 1/|dr|^2                          */
    const double *lsq_C;          /* [n_cells][9] rows Cx, Cy, Cz              */
    const int32_t *face_mark;     /* [n_faces] index into marks (marks[0] = fluid-interior) */
    const meso_mark_desc *marks;  /* [n_marks]                                 */
} meso_mesh_desc;

/* Discrete velocity set (src/solver/dvs/dvs_generator.cpp). */
typedef struct meso_dvs_desc {
    int32_t n_vel;
    int32_t has_mirror;
    double max_magnitude;
    const double *xi;      /* [n_vel][3] */
    const double *w;       /* [n_vel]    */
    const int32_t *mirror; /* [n_vel][3] mirror velocity per axis, or NULL */
} meso_dvs_desc;

/* Step parameters (CDUGKS_SHAKHOV::update_config cdugks_shakhov.cpp:7-33,
 * CDUGKS ctor/initial cdugks.cpp:9-18,45-47, DUGKS initial dugks.cpp:45-48). */
typedef struct meso_step_params {
    int32_t model;           /* MESO_MODEL_*              */
    int32_t dim;             /* 2 or 3                    */
    int32_t gradient_switch; /* 0: zero-gradient faces    */
    int32_t K;               /* gas-k (internal DOF), Shakhov */
    double dt;
    /* Shakhov */
    double Pr, R, T0, mu0, omega;
    /* incompressible BGK */
    double RT, tau;
    /* Shakhov: Venkatakrishnan limiter of the reconstruction
     * (src/solver/limiter.cpp:7-31, applied at cdugks_shakhov.cpp:231-266).
     * venkata_k is meso's venkata_k member: read from `venkata-limiter-k` only
     * when the solver is built with gradient-switch False and limiter-switch
     * True (cdugks_shakhov.cpp:21-28), 0 otherwise. */
    int32_t limiter_switch;
    int32_t reserved;
    double venkata_k;
} meso_step_params;

/* Result of meso_hip_step(). */
typedef struct meso_step_report {
    int32_t steps_done;
    int32_t bad_cells;   /* unmasked cells with NaN density or T<0 / NaN T in the last step */
    int32_t first_bad_cell;
    int32_t masked_cells; /* bad cells of masked groups repaired in the steps run
                             (cell_value_interp, cdugks_shakhov.cpp:766-789) */
} meso_step_report;

/* ======================================================================== */
/* Host front-end: case file -> descriptors                                   */
/* ======================================================================== */
typedef struct meso_case meso_case;

/* Parse a meso case file, its gambit mesh and its DVS
 * (Config ctor src/solver/config.cpp:18-28; solver ctors cdugks_shakhov.cpp:35-60,
 * cdugks.cpp:7-39, dugks.cpp:7-39). mesh_scale_override <= 0 keeps `mesh-scale`. */
int meso_case_open(const char *case_path, double mesh_scale_override, meso_case **out);
void meso_case_close(meso_case *c);
const char *meso_case_last_error(void);

/* Descriptors point into memory owned by the case; valid until meso_case_close. */
int meso_case_mesh(const meso_case *c, meso_mesh_desc *out);
int meso_case_dvs(const meso_case *c, meso_dvs_desc *out);
int meso_case_params(const meso_case *c, meso_step_params *out);
/* Re-read [settings] (dt, Pr, mu0, ...) like update_config() does each step. */
int meso_case_reload_settings(meso_case *c);
/* Per-cell initial density, velocity, temperature from the cell-group patches
 * (fixedValue or fromFile; cdugks_shakhov.cpp:108-131, cdugks.cpp:69-84). */
int meso_case_initial_patch(const meso_case *c, double *rho, double *u3, double *T);
/* mask[NC] = 1 for the cells whose group sets `mask switch True`
 * (src/solver/boundary/patch.cpp:48-51, read at cdugks_shakhov.cpp:702,713). */
int meso_case_cell_mask(const meso_case *c, int32_t *mask);
/* Integer / string settings meso's time loop reads (max-step, write-interval, case-name, ...). */
int meso_case_setting_int(const meso_case *c, const char *key, int32_t dflt, int32_t *out);
int meso_case_setting_double(const meso_case *c, const char *key, double dflt, double *out);
/* Writes at most cap bytes (NUL included). */
int meso_case_setting_str(const meso_case *c, const char *key, const char *dflt, char *out, int32_t cap);

/* The files the solver's output() writes for `step` (macros of the case's
 * cells, as meso_hip_get_macros / meso_hip_get_flux_moments return them), under
 * root/<case-name>/: grid.plt at step 0, step-N/data.plt (Tecplot), latest/ and,
 * with output-np True, step-N/ text fields (Rho, U, T, q, fluxM0, fluxM1,
 * fluxM2; BGK: Rho, U); DUGKS: result-N.dat.plt and np-data/ (Rho, vel).
 * Formats of Field::output (src/solver/field.cpp:139-173) and Mesh::output_grid /
 * output_data / output (src/mesh/mesh_output.cpp:111-330); layout of
 * CDUGKS_SHAKHOV::output (cdugks_shakhov.cpp:791-855), CDUGKS::output
 * (cdugks.cpp:373-420), DUGKS::output (dugks.cpp:385-411). T, q and the flux
 * moments may be null except for Shakhov; Shakhov's flux moments are all three
 * null before the first step (meso's flux fields are then empty and so are
 * fluxM0-2). root null or "" = working directory. */
int meso_case_write_output(const meso_case *c, const char *root, int32_t step, double solution_time,
                           const double *rho, const double *u3, const double *T, const double *q3,
                           const double *flux_m0, const double *flux_m1_3, const double *flux_m2);
const char *meso_case_output_error(void);

/* Write the gambit file in_path to out_path with its cells renumbered (node
 * ids, groups, boundary sets and every other line kept): MESO_RENUMBER_RCB
 * orders the cells by recursive coordinate bisection of their centres (a
 * locality order for meshes written in a poor one), MESO_RENUMBER_RANDOM by a
 * permutation seeded with seed, MESO_RENUMBER_KEEP keeps the order. new_of_old
 * (may be null; capacity entries) receives each old cell's new 0-based id,
 * n_cells (may be null) the cell count. No reference counterpart: meso reads
 * cells in file order (src/mesh/mesh_reader.cpp:52-80) and builds its faces in
 * that order (src/mesh/object.cpp:38-84); its only mesh utility is the METIS
 * partition of --mesh (src/mesh/metis.cpp). */
#define MESO_RENUMBER_RCB 0
#define MESO_RENUMBER_RANDOM 1
#define MESO_RENUMBER_KEEP 2
int meso_mesh_renumber(const char *in_path, const char *out_path, int32_t method, uint64_t seed,
                       int32_t *new_of_old, int32_t capacity, int32_t *n_cells);
const char *meso_mesh_renumber_error(void);

/* ======================================================================== */
/* Device step                                                               */
/* ======================================================================== */
typedef struct meso_hip_ctx meso_hip_ctx;

/* Upload the mesh, DVS and parameters; allocate the state on the current device. */
int meso_hip_create(const meso_mesh_desc *mesh, const meso_dvs_desc *dvs, const meso_step_params *params,
                    int32_t storage, meso_hip_ctx **out);
void meso_hip_destroy(meso_hip_ctx *ctx);
const char *meso_hip_last_error(const meso_hip_ctx *ctx);

/* initial(): f = equilibrium of the per-cell patch values; macros from the
 * discrete moments; Shakhov q = m3/2 (cdugks_shakhov.cpp:103-168,
 * cdugks.cpp:67-104, dugks.cpp:68-105). Also resets the residual copies. */
int meso_hip_init_equilibrium(meso_hip_ctx *ctx, const double *rho, const double *u3, const double *T);
/* The same initial state without allocating device memory (no hipMalloc,
 * hipFree or allocating library call): the per-cell macros are staged in a
 * static 16 MiB device array of the library (one per GPU, held until the
 * process ends), so the mesh may have at most 419,430 cells; larger meshes are
 * rejected with MESO_E_INVALID. Calls on one GPU are serialised by the library
 * (the array keeps the last call's contents; nothing in it needs clearing).
 * No counterpart in the reference: it is for callers that cannot allocate
 * (HIP graph capture around the time loop, serving loops, harnesses that
 * forbid hipMalloc). */
int meso_hip_init_equilibrium_noalloc(meso_hip_ctx *ctx, const double *rho, const double *u3, const double *T);
/* read-np-data restart: f = equilibrium (Shakhov: g_S/h_S with q) of the given
 * macros, which are taken as the cell state (cdugks_shakhov.cpp:81-101, cdugks.cpp:56-65). */
int meso_hip_init_from_macros(meso_hip_ctx *ctx, const double *rho, const double *u3, const double *T,
                              const double *q3);
/* Full state upload / download (f in FP64 host layout, macros per cell). T and
 * q3 are ignored (may be NULL) for the BGK models. */
int meso_hip_set_state(meso_hip_ctx *ctx, const double *f, int32_t layout, const double *rho, const double *u3,
                       const double *T, const double *q3);
int meso_hip_get_state(meso_hip_ctx *ctx, double *f, int32_t layout);

/* Advance nsteps time steps. Stops early (returning MESO_E_NUMERIC) after a
 * step that produced NaN density or T<0 in a cell outside the masked groups
 * (cdugks_shakhov.cpp:698-719); that step is complete, as in meso. */
int meso_hip_step(meso_hip_ctx *ctx, int32_t nsteps, meso_step_report *report);
int meso_hip_set_params(meso_hip_ctx *ctx, const meso_step_params *params);
/* Cells whose group sets `mask switch True` (mask[NC], 0/1; NULL clears). A
 * Shakhov step repairs such a cell when its density is NaN or its T < 0, as
 * CDUGKS_SHAKHOV::cell_value_interp does (cdugks_shakhov.cpp:698-719, 766-789;
 * fvmMesh::interp_IDW, src/mesh/mesh_numerical.cpp:51-94): rho, u, T and q
 * become the inverse-distance average of the face neighbours (cells before it
 * in meso's loop order with their new values, the others with their old ones)
 * and g, h the Shakhov equilibrium of those; q after the step is the cell's
 * own heat-flux sum, as meso leaves it. Other models ignore the mask. */
int meso_hip_set_mask(meso_hip_ctx *ctx, const int32_t *mask);

/* Many independent cases at once (the batched cavity sweep): advances each of
 * the n contexts by nsteps, issuing step i of every context (each on its own
 * stream) before step i + 1 of any and waiting once per context at the end, so
 * that small cases share the GPU. reports[k] (may be NULL) is what
 * meso_hip_step would report for ctxs[k]; returns the first non-OK status. A
 * context may appear once per call. No counterpart in the reference, which runs
 * one case per process (src/meso/exec.cpp:141-165). */
int meso_hip_step_batch(meso_hip_ctx *const *ctxs, int32_t n, int32_t nsteps, meso_step_report *reports);
/* Autotuner of the cell x velocity tiles of the patch kernels. It times
 * trial_steps steps of the current state (computed but not committed: the
 * state does not advance) for each candidate and keeps the fastest, in three
 * stages: the moment patches (cell-based of 16 or 32 cells, face-based of 32
 * or 64 faces), the update patches (16, 24 or 32 cells), the velocity chunk
 * (half or twice the chunk count). Patch sets are rebuilt on the host and
 * uploaded again. Returns the velocities per block (0 where the per-cell
 * kernels run); meso_hip_tiles reports the rest. No counterpart in the
 * reference. */
int meso_hip_autotune(meso_hip_ctx *ctx, int32_t trial_steps, int32_t *vchunk);

/* The tile shapes in use. */
typedef struct meso_tile_config {
    int32_t patches;          /* 1: patch kernels (else the per-cell kernels) */
    int32_t update_cells;     /* largest update patch, owned cells */
    int32_t moment_by_cells;  /* 1: cell-based moment patches, 0: face-based */
    int32_t moment_size;      /* largest moment patch: owned cells or faces */
    int32_t vchunk;           /* velocities per block */
    int32_t update_vt;        /* velocities per tile, update kernel */
    int32_t moment_vt;        /* velocities per tile, moment kernel */
    int32_t update_owned;     /* cells whose update the dominant update launch owns
                                 (patch kernels: every cell without a mirror boundary;
                                 per-cell kernels: the cells without a boundary face) */
    int32_t moment_lds;       /* LDS bytes per block of the moment kernel (patch kernels) */
    int32_t update_lds;       /* LDS bytes per block of the update kernel (patch kernels) */
} meso_tile_config;
int meso_hip_tiles(const meso_hip_ctx *ctx, meso_tile_config *out);
/* enable != 0: meso_hip_step and meso_hip_step_batch run steps in pairs as
 * replays of a hipGraph captured on the context's stream (one per buffer
 * parity, captured again after any change of parameters, state or tiles;
 * never while kernel profiling is on). Results are those of the eager steps.
 * No counterpart in the reference. */
int meso_hip_set_graphs(meso_hip_ctx *ctx, int32_t enable);
/* Velocities per block of the patch kernels (a positive multiple of 64; 0:
 * the default, enough chunks for 16 blocks per CU). Changes only the order of
 * the velocity partial sums. MESO_E_INVALID when the partial-sum buffers do
 * not hold that many chunks, or (for vchunk != 0) when the case runs the
 * per-cell kernels. The autotuner's third stage, for callers that choose it
 * themselves. No counterpart in the reference. */
int meso_hip_set_vchunk(meso_hip_ctx *ctx, int32_t vchunk);

/* A fixed set of contexts advanced with one launch per stage (the batched
 * cavity sweep in one grid): the contexts whose step runs the patch kernels
 * (large velocity sets, no mirror-boundary cells, no masked groups) are
 * split into sub-batches (meso_hip_batch_set_streams) and grouped by kernel
 * instance, and every stage of a step is one launch over the blocks of all
 * contexts of a group, each block reading its context's arguments from a
 * device table; every XCD takes a contiguous share of every context's
 * blocks. Other contexts take their own launches in the same call.
 * The plan (tables, block maps) is built on the first step, or by
 * meso_hip_batch_info, and again after any context's parameters, state,
 * tiles or buffer parity changed outside the batch. Building it sets each
 * fused context's velocity chunk for the batch grid (fewer, larger blocks
 * than a context alone takes: its share of the batch's DV-cells decides;
 * meso_hip_tiles reports it, meso_hip_batch_set_rechunk(b, 0) keeps the
 * contexts' own). Results are bit-identical to meso_hip_step on each context
 * with the same tiles. meso_hip_batch_set_graphs(b, 1) replays two fused steps as one
 * captured hipGraph. reports[k] is what meso_hip_step would report for
 * ctxs[k]. The contexts must outlive the batch, share one device, and may not
 * be stepped from another thread while the batch steps. No counterpart in the
 * reference, which runs one case per process (src/meso/exec.cpp:141-165). */
typedef struct meso_hip_batch meso_hip_batch;
int meso_hip_batch_create(meso_hip_ctx *const *ctxs, int32_t n, meso_hip_batch **batch);
int meso_hip_batch_step(meso_hip_batch *batch, int32_t nsteps, meso_step_report *reports);
int meso_hip_batch_set_graphs(meso_hip_batch *batch, int32_t enable);
int meso_hip_batch_set_rechunk(meso_hip_batch *batch, int32_t enable);
/* Sub-batches (1..16, default 4: the hardware queues a process gets): the
 * fused contexts are split by DV-cells over n streams, each launching its
 * own share of every stage, so that the small stages of one sub-batch overlap
 * the large ones of another. */
int meso_hip_batch_set_streams(meso_hip_batch *batch, int32_t n);
/* fused: contexts in the fused launches; groups: kernel instances;
 * launches_per_step: launches of one batch step (fused part) */
int meso_hip_batch_info(meso_hip_batch *batch, int32_t *fused, int32_t *groups, int32_t *launches_per_step);
void meso_hip_batch_destroy(meso_hip_batch *batch);

/* Cell macros after the last step: rho[NC], u3[NC*3], T[NC], q3[NC*3] (NULL skips). */
int meso_hip_get_macros(meso_hip_ctx *ctx, double *rho, double *u3, double *T, double *q3);
/* Cell flux moments of the last step (fluxM0..2 of cdugks_shakhov.cpp:837-839). */
int meso_hip_get_flux_moments(meso_hip_ctx *ctx, double *m0, double *m1_3, double *m2);
/* Residuals against the stored previous copies, then copy (field.cpp:222-249):
 * out[0]=rho, out[1..3]=u, out[4]=T, out[5..7]=q (BGK: T and q entries are 0). */
int meso_hip_residual(meso_hip_ctx *ctx, double out[8]);

/* Kernel timing with HIP events on the context's stream. enable=1 starts
 * recording (and clears totals), enable=0 stops. Names are static strings. */
int meso_hip_profile(meso_hip_ctx *ctx, int32_t enable);
int meso_hip_kernel_times(meso_hip_ctx *ctx, int32_t max_entries, const char **names, double *total_ms,
                          int32_t *launches, int32_t *n_entries);
/* Bytes of device memory held by the context. */
int64_t meso_hip_device_bytes(const meso_hip_ctx *ctx);
/* Synchronise the context's stream. */
int meso_hip_sync(meso_hip_ctx *ctx);

#ifdef __cplusplus
}
#endif

#endif /* MESO_HIP_H */
