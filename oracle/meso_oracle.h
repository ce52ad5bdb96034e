/*
 * meso_oracle.h — FP64 CPU restatement of meso's DUGKS/CDUGKS time step.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the checker: it is loaded by
 * tests/, by __graft_entry__.smoke() and by the cpu_baseline leg of bench.py,
 * and by nothing else. The product (libmeso_hip.so) neither links nor calls it.
 *
 * It consumes the same plain descriptors as the device library
 * (include/meso_hip.h) and stores distributions the way meso does,
 * velocity-major f[p][cell].
 */
#ifndef MESO_ORACLE_H
#define MESO_ORACLE_H

#include "meso_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct oracle_ctx oracle_ctx;

int oracle_create(const meso_mesh_desc *mesh, const meso_dvs_desc *dvs, const meso_step_params *params,
                  oracle_ctx **out);
void oracle_destroy(oracle_ctx *ctx);
const char *oracle_last_error(void);
/* number of OpenMP threads used by the step loops (<=0: library default) */
void oracle_set_threads(int32_t n);
int32_t oracle_get_threads(void);

int oracle_set_params(oracle_ctx *ctx, const meso_step_params *params);
/* cells of groups with `mask switch True` (NULL clears); Shakhov repairs them */
int oracle_set_mask(oracle_ctx *ctx, const int32_t *mask);
int oracle_init_equilibrium(oracle_ctx *ctx, const double *rho, const double *u3, const double *T);
int oracle_init_from_macros(oracle_ctx *ctx, const double *rho, const double *u3, const double *T,
                            const double *q3);
int oracle_set_state(oracle_ctx *ctx, const double *f, int32_t layout, const double *rho, const double *u3,
                     const double *T, const double *q3);
int oracle_get_state(oracle_ctx *ctx, double *f, int32_t layout);
int oracle_step(oracle_ctx *ctx, int32_t nsteps, meso_step_report *report);
int oracle_get_macros(oracle_ctx *ctx, double *rho, double *u3, double *T, double *q3);
int oracle_get_flux_moments(oracle_ctx *ctx, double *m0, double *m1_3, double *m2);
int oracle_residual(oracle_ctx *ctx, double out[8]);

/* Stage probes for unit tests (no state change):
 * LSQ gradient of an arbitrary cell field (mesh_numerical.cpp:24-49): grad[NC*3]. */
int oracle_grad(const oracle_ctx *ctx, const double *field, double *grad);
/* Face values g_face/h_face (velocity-major [n_d][NV][NF]) produced by the
 * last step's reconstruct(), i.e. after equilibrium mixing and boundaries. */
int oracle_face_values(const oracle_ctx *ctx, double *f_face);
/* Half-step face reconstruction f-bar+ alone (no mixing, no boundaries) of a
 * velocity-major field cell_f[NV][NC] into face_f[NV][NF]
 * (cdugks_shakhov.cpp:213-299); for checks of the upwind extrapolation. */
int oracle_fbar(oracle_ctx *ctx, const double *cell_f, double *face_f);

#ifdef __cplusplus
}
#endif

#endif
