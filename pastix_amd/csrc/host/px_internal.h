/*
 * px_internal.h -- host-side data structures of the engine (not installed).
 *
 * symbol / solver structures follow the layout contract of the reference
 * (symbol/symbol.h:45-91, blend/solver.h:118-273): a column block (cblk) is a
 * range of consecutive columns in the permuted numbering with a contiguous list
 * of blocks (bloks), the diagonal blok first; every blok is a column-major
 * rows x ncols tile (ld = rows) stored at coefind = rowoffset * ncols inside
 * the cblk buffer (2D tile layout, blend/solver_matrix_gen_utils.h:90).
 */
#ifndef PX_INTERNAL_H
#define PX_REF_CACHE 16   /* refinement device buffers kept per instance */
#define PX_INTERNAL_H

#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "pastix.h"
#include "pastix_amd_device.h"

#define STEP_INIT      (1 << 0)
#define STEP_ORDERING  (1 << 1)
#define STEP_SYMBFACT  (1 << 2)
#define STEP_ANALYSE   (1 << 3)
#define STEP_CSC2BCSC  (1 << 4)
#define STEP_BCSC2CTAB (1 << 5)
#define STEP_NUMFACT   (1 << 6)
#define STEP_SOLVE     (1 << 7)
#define STEP_REFINE    (1 << 8)

/* cblktype flags (reference blend/solver.h:48-54) */
#define CBLK_FANIN      (1 << 0)
#define CBLK_LAYOUT_2D  (1 << 1)
#define CBLK_TASKS_2D   (1 << 2)
#define CBLK_COMPRESSED (1 << 3)
#define CBLK_IN_SCHUR   (1 << 4)
#define CBLK_IN_LAST    (1 << 5)
#define CBLK_RECV       (1 << 6)

static inline pastix_int_t px_imin( pastix_int_t a, pastix_int_t b ) { return a < b ? a : b; }
static inline pastix_int_t px_imax( pastix_int_t a, pastix_int_t b ) { return a > b ? a : b; }
static inline pastix_int_t px_iceil( pastix_int_t a, pastix_int_t b ) { return ( a + b - 1 ) / b; }

void  px_print_error( const char *fmt, ... );
void  px_print_warning( const char *fmt, ... );
void *px_malloc( size_t size );
void *px_calloc( size_t n, size_t size );
double px_wtime( void );

/* ----------------------------------------------------------------------- */
/* Graph: symmetric adjacency without the diagonal, CSC, 0-based, sorted.   */
typedef struct px_graph_s {
    pastix_int_t  n;
    pastix_int_t *colptr;
    pastix_int_t *rows;
} px_graph_t;

int  px_graph_prepare( const spmatrix_t *spm, px_graph_t *graph );
void px_graph_exit( px_graph_t *graph );

/* ----------------------------------------------------------------------- */
/* Order helpers (order.c).                                                 */
int px_order_nd( const px_graph_t *graph, pastix_order_t *order );
void px_order_fundamental_supernodes( const px_graph_t *g, pastix_order_t *ord );
int px_order_apply_level_order( pastix_order_t *order, pastix_int_t level_tasks2d,
                                pastix_int_t width_tasks2d );

/* ----------------------------------------------------------------------- */
/* Symbol matrix.                                                           */
typedef struct px_symcblk_s {
    pastix_int_t fcolnum;
    pastix_int_t lcolnum;
    pastix_int_t bloknum;
    pastix_int_t brownum;
    int8_t       selevtx;
} px_symcblk_t;

typedef struct px_symblok_s {
    pastix_int_t frownum;
    pastix_int_t lrownum;
    pastix_int_t lcblknm;
    pastix_int_t fcblknm;
} px_symblok_t;

typedef struct px_symbol_s {
    pastix_int_t  cblknbr;
    pastix_int_t  bloknbr;
    pastix_int_t  nodenbr;
    pastix_int_t  schurfcol;
    pastix_int_t  browmax;
    pastix_int_t  dof;
    px_symcblk_t *cblktab;   /* cblknbr+1 entries, last is a sentinel */
    px_symblok_t *bloktab;   /* bloknbr entries                        */
    pastix_int_t *browtab;   /* bloknbr-cblknbr entries                */
} px_symbol_t;

int  px_symbol_fax_direct( px_symbol_t *symb, const px_graph_t *graph, const pastix_order_t *order );
void px_symbol_build_rowtab( px_symbol_t *symb );
void px_symbol_exit( px_symbol_t *symb );
int  px_symbol_check( const px_symbol_t *symb );
pastix_int_t px_symbol_nnz( const px_symbol_t *symb );
void px_symbol_flops( const px_symbol_t *symb, int flttype, int factotype,
                      double *thflops, double *rlflops );
void px_symbol_reordering( px_symbol_t *symb, pastix_order_t *order,
                           pastix_int_t split_level, pastix_int_t stop_criterion,
                           int nthreads );
void px_symbol_split( px_symbol_t *symb, pastix_int_t blcolmin, pastix_int_t blcolmax );

/* ----------------------------------------------------------------------- */
/* Solver matrix (single process, 2D layout everywhere).                    */
typedef struct px_cblk_s {
    pastix_int_t fcolnum;
    pastix_int_t lcolnum;
    pastix_int_t fblok;      /* index of the diagonal blok in bloktab      */
    pastix_int_t stride;     /* total number of rows of the cblk           */
    pastix_int_t coefoff;    /* offset of the cblk buffer in the coeftab   */
    pastix_int_t brownum;    /* first entry in browtab                     */
    pastix_int_t brown2d;
    pastix_int_t lcolidx;    /* offset of the cblk rows in a RHS vector     */
    pastix_int_t priority;
    int8_t       cblktype;
} px_cblk_t;

typedef struct px_blok_s {
    pastix_int_t frownum;
    pastix_int_t lrownum;
    pastix_int_t lcblknm;
    pastix_int_t fcblknm;
    pastix_int_t coefind;    /* offset of the tile inside the cblk buffer   */
    pastix_int_t browind;
} px_blok_t;

typedef struct px_solver_s {
    pastix_int_t  cblknbr;
    pastix_int_t  bloknbr;
    pastix_int_t  nodenbr;
    pastix_int_t  coefnbr;
    pastix_int_t  brownbr;
    pastix_int_t  colmax;
    pastix_int_t  offdmax;
    pastix_int_t  blokmax;
    pastix_int_t  nb2dcblk;
    pastix_int_t  nb2dblok;
    px_cblk_t    *cblktab;   /* cblknbr+1 (sentinel) */
    px_blok_t    *bloktab;   /* bloknbr+1 (sentinel) */
    pastix_int_t *browtab;
    /* static task list: tasks = cblks, sorted for the static schedule */
    pastix_int_t *tasktab;
    double        diagthreshold;
    int           nbpivots;
} px_solver_t;

int  px_solver_gen( px_solver_t *solv, const px_symbol_t *symb,
                    pastix_int_t width_tasks2d, pastix_int_t level_tasks2d );
void px_solver_exit( px_solver_t *solv );
int  px_solver_check( const px_solver_t *solv );

/* ----------------------------------------------------------------------- */
/* bcsc: the matrix permuted into the solver numbering, full pattern, CSC.  */
typedef struct px_bcsc_s {
    pastix_int_t   n;
    pastix_int_t   nnz;
    int            flttype;  /* of the values below                       */
    int            mtxtype;
    pastix_int_t  *colptr;   /* n+1                                        */
    pastix_int_t  *rowtab;   /* nnz, sorted inside each column             */
    void          *Lvalues;  /* nnz values of P A P^T (CSC)                */
    pastix_int_t  *trowptr;  /* CSR (= CSC of the transpose) for general   */
    pastix_int_t  *tcoltab;
    void          *Uvalues;
    double         Anorm_f;  /* ||A||_F */
    double         Anorm_inf;
} px_bcsc_t;

int  px_bcsc_init( px_bcsc_t *bcsc, const spmatrix_t *spm, const pastix_order_t *ord );
void px_bcsc_exit( px_bcsc_t *bcsc );

/* ----------------------------------------------------------------------- */
struct pastix_data_s {
    pastix_int_t    *iparm;
    double          *dparm;
    int              steps;
    int              procnum;
    PASTIX_Comm      comm;
    const spmatrix_t *csc;
    px_graph_t      *graph;
    pastix_order_t  *ordemesh;
    px_symbol_t     *symbmtx;
    px_solver_t     *solvmatr;
    px_bcsc_t       *bcsc;
    pxd_ctx_t       *dev;       /* device engine context (NULL until first use) */
    int              factored_flttype;   /* precision of the device factor    */
    int              nthreads;
    char            *dir_global;
    double           berr_target;        /* >0: refinement stops on backward error */
    int              refine_nrhs_batch;   /* 1: reference behaviour (one RHS at a time) */
    void            *ref_cache[PX_REF_CACHE];       /* device buffers kept for the refinement */
    size_t           ref_cache_bytes[PX_REF_CACHE];
    pastix_int_t     schur_n;             /* Schur unknowns (pastixSetSchurUnknownList)       */
    pastix_int_t    *schur_list;
    pastix_model_t  *cpu_models;          /* kernel cost models (common/pastixdata.h:106-107) */
    pastix_model_t  *gpu_models;
};

int px_device_ready( pastix_data_t *pd );

/* schur.c */
int px_order_with_schur( pastix_data_t *pd, const spmatrix_t *spm, pastix_order_t *ord,
                         int ( *order_fn )( const px_graph_t *, pastix_order_t * ) );
int px_order_schur_boundary( pastix_data_t *pd, const spmatrix_t *spm, pastix_order_t *ord );

/* io.c */
const char *px_gendirectories( pastix_data_t *pd );
int  px_symbol_load_default( px_symbol_t *symb );
int  px_symbol_save_default( pastix_data_t *pd, const px_symbol_t *symb );
void px_symbol_expand( px_symbol_t *symb, pastix_int_t dof );
int  px_spm_expand( spmatrix_t *spm );

#endif /* PX_INTERNAL_H */
