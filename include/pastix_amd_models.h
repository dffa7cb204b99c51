/*
 * pastix_amd_models.h -- per-kernel cost models and the kernel timing log.
 *
 * The polynomial models, their file format and their defaults follow the
 * reference (common/models.h:20-87, common/models.c:41-593); the engine adds
 * the timing log of its own kernels (the reference's model.csv,
 * kernels/kernels_trace.c:222-249), a least-squares fit of the polynomials to
 * such a log, a writer of model files, and a learned model (an MLP trained on
 * the GPU) that is compared with the polynomials.  Included by pastix.h.
 */
#ifndef PASTIX_AMD_MODELS_H
#define PASTIX_AMD_MODELS_H

#include "pastix.h"
#include "pastix/kernels_enums.h"
#include "pastix_amd_device.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Unset coefficient marker (common/models.c:497-507). */
#define PASTIX_MODEL_UNSET 3735928559.0   /* 0xdeadbeef */

/* Model of one computational unit: coefficients[arith][ktype][0..7], arith
 * 0..3 = s, d, c, z (common/models.h:24-27). */
typedef struct pastix_model_s {
    char  *name;
    double coefficients[4][PastixKernelLvl1Nbr][8];
} pastix_model_t;

/* Time (s) of a kernel with 1, 2 or 3 size parameters, clamped at 0:
 *   a3 N^3 + a2 N^2 + a1 N + a0                       (common/models.h:38-44)
 *   a5 M N^2 + a4 M N + a3 N^2 + a2 M + a1 N + a0      (common/models.h:53-59)
 *   a7 MNK + a6 MK + a5 KN + a4 MN + a3 M + a2 N + a1 K + a0
 *                                                      (common/models.h:71-83) */
double modelsGetCost1Param( const double *coefs, pastix_int_t N );
double modelsGetCost2Param( const double *coefs, pastix_int_t M, pastix_int_t N );
double modelsGetCost3Param( const double *coefs, pastix_int_t M, pastix_int_t N, pastix_int_t K );

/* Kernel id of a model-file kernel name and its number of coefficients
 * (4, 6 or 8); -1 for an unknown name (common/models.c:41-84). */
int  modelsGetKernelId( const char *kernelstr, int *nbcoef );
/* Copy the coefficients of (arithm, kernelid) to the unset entries of the
 * related kernels and arithmetics (common/models.c:100-201). */
void modelsPropagate( pastix_model_t *model, int arithm, pastix_ktype_t kernelid );
/* Read a model file: '#' lines skipped, then the model name, then lines
 * "arith;kernel;c0;c1;..." (common/models.c:219-317).  0 or -1. */
int  modelsRead( pastix_model_t *model, const char *modelfilename );
/* Built-in defaults (common/models.c:337-474). */
int  modelsInitDefaultCPU( pastix_model_t *model );
int  modelsInitDefaultGPU( pastix_model_t *model );

pastix_model_t *pastixModelsNew( void );              /* common/models.c:490-511 */
void            pastixModelsFree( pastix_model_t *model );   /* :520-531 */
/* Load the CPU and GPU models of an instance: PASTIX_MODELS_CPU /
 * PASTIX_MODELS_GPU name model files, the defaults otherwise (:549-593).
 * pastixInit calls it. */
void            pastixModelsLoad( pastix_data_t *pastix_data );

/* --- engine additions (no counterpart in the reference) --------------- */

/* The CPU (gpu = 0) or GPU (gpu = 1) model of an instance. */
pastix_model_t *pastix_amd_models_get( pastix_data_t *pd, int gpu );
/* Time of one kernel call under a model: the 1-, 2- or 3-parameter form
 * that the kernel type uses (panels: m = n of the panel). */
double pastix_amd_model_cost( const pastix_model_t *model, int arith, int ktype,
                              pastix_int_t m, pastix_int_t n, pastix_int_t k );
/* Records of the kernel timing log of the last factorization (the option
 * "model_log" set to 1 before pastix_task_numfact); *count gets the total. */
int pastix_amd_model_log( pastix_data_t *pd, pxd_model_entry_t *entries, pastix_int_t max,
                          pastix_int_t *count );
/* Write that log as the reference's model.csv (kernels/kernels_trace.c:
 * 222-249): "# GPU Model data", "# Precision: ...", "Kernel;M;N;K;Time". */
int pastix_amd_model_write_csv( pastix_data_t *pd, const char *filename );
/* Least-squares fit (relative error) of each kernel type present in the
 * records to its polynomial form, stored for arithmetic arith and
 * propagated.  The model is renamed to name when name is not NULL. */
int pastix_amd_models_fit( pastix_model_t *model, const char *name, int arith,
                           const pxd_model_entry_t *entries, pastix_int_t count );
/* Write a model in the format modelsRead reads (set coefficients only). */
int pastix_amd_models_write( const pastix_model_t *model, const char *filename );

/*
 * Learned model: the GPU multilayer perceptron of pastix_amd_device.h
 * (pxd_mlp_*) fitted to the natural log of the times of timing-log records.
 * Its 10 inputs, standardised over the training records, are the kernel
 * family as five 0/1 values (diagonal block, TRSM, and the update on each of
 * its three kernel classes: 64 x 64 tiles, tall tiles, one wavefront, from
 * the record's variant), ln(1+m), ln(1+n), ln(1+k), ln(1+flops) of the
 * record's sizes and the arithmetic.
 * Training runs steps Adam steps, the rate lr divided by 3 after each
 * quarter; *loss gets the final mean squared error of the standardised log
 * times.
 */
typedef struct pastix_amd_learned_s pastix_amd_learned_t;
int  pastix_amd_learned_train  ( pastix_amd_learned_t **lm, int device, int arith,
                                 const pxd_model_entry_t *entries, pastix_int_t count, int steps,
                                 double lr, uint64_t seed, double *loss );
/* Predicted times (s) of count records. */
int  pastix_amd_learned_predict( pastix_amd_learned_t *lm, int arith, const pxd_model_entry_t *entries,
                                 pastix_int_t count, double *time );
void pastix_amd_learned_free   ( pastix_amd_learned_t *lm );

/*
 * Kernel class of every blok update chosen by the smallest predicted cost:
 * with the learned model lm, else with poly[0..2], the polynomial models
 * (GEMMCblk2d2d entries) of the three classes (64 x 64 tiles, tall tiles,
 * one wavefront); with neither, the engine's cut-offs are restored.  The
 * cost is the predicted time alone on the device times the fraction of the
 * device's CUs the class's workgroups fill (the updates of a level share
 * the device).  The
 * reference's schedulers consult the same kind of prediction per update
 * (modelsGetCost3Param, sopalin/starpu/codelet_cblk_zgemmsp.c:137-170).
 * *predicted (optional) gets the predicted total cost of the choice.
 * Call after pastix_subtask_bcsc2ctab; applies from the next factorization.
 */
int  pastix_amd_choose_update_kernels( pastix_data_t *pd, pastix_amd_learned_t *lm, pastix_model_t *const *poly,
                                       int arith, double *predicted );

#ifdef __cplusplus
}
#endif

#endif /* PASTIX_AMD_MODELS_H */
