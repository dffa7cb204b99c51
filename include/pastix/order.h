/*
 * pastix/order.h -- ordering structure and routines of the PaStiX C API.
 *
 * pastix_order_t keeps the field names, meanings and order of the reference
 * structure (reference include/pastix/order.h:47-59) because users fill it in
 * for PastixOrderPersonal and read it back after pastix_subtask_order.
 */
#ifndef PASTIX_AMD_ORDER_H
#define PASTIX_AMD_ORDER_H

#include <stdint.h>
#include <stdio.h>
#include "spm.h"

#ifdef __cplusplus
extern "C" {
#endif

#ifndef PASTIX_AMD_INT_DEFINED
#define PASTIX_AMD_INT_DEFINED
typedef int64_t  pastix_int_t;
typedef uint64_t pastix_uint_t;
#define PASTIX_INT_MAX INT64_MAX
#endif

typedef struct pastix_order_s {
    pastix_int_t  baseval;     /* base value used for numbering                 */
    pastix_int_t  vertnbr;     /* number of vertices                            */
    pastix_int_t  cblknbr;     /* number of column blocks (supernodes)          */
    pastix_int_t *permtab;     /* permutation, size vertnbr                     */
    pastix_int_t *peritab;     /* inverse permutation, size vertnbr             */
    pastix_int_t *rangtab;     /* supernode ranges, size cblknbr+1              */
    pastix_int_t *treetab;     /* supernode father, size cblknbr (-1 for roots) */
    int8_t       *selevtx;     /* unused by this engine                         */
    pastix_int_t  sndenbr;     /* number of supernodes before splitting         */
    pastix_int_t *sndetab;     /* supernode ranges before splitting             */
    pastix_int_t *peritab_exp; /* unused by this engine                         */
} pastix_order_t;

int  pastixOrderInit ( pastix_order_t *ordeptr, pastix_int_t baseval,
                       pastix_int_t vertnbr, pastix_int_t cblknbr,
                       pastix_int_t *permtab, pastix_int_t *peritab,
                       pastix_int_t *rangtab, pastix_int_t *treetab );
int  pastixOrderAlloc( pastix_order_t *ordeptr, pastix_int_t vertnbr, pastix_int_t cblknbr );
void pastixOrderExit ( pastix_order_t *ordeptr );
void pastixOrderBase ( pastix_order_t *ordeptr, pastix_int_t baseval );
int  pastixOrderCheck( const pastix_order_t *ordeptr );
int  pastixOrderCopy ( pastix_order_t *ordedst, const pastix_order_t *ordesrc );
/* Expand a node ordering to the spm->dof unknowns of each node (reference
 * order/order.c:396-469; constant dof only). */
void pastixOrderExpand( pastix_order_t *ordeptr, const spmatrix_t *spm );

/* Nested dissection of an nx x ny x nz grid (reference order/order_grids.c:306). */
int  pastixOrderGrid ( pastix_order_t **myorder, pastix_int_t nx, pastix_int_t ny, pastix_int_t nz );

#ifndef PASTIX_AMD_DATA_DEFINED
#define PASTIX_AMD_DATA_DEFINED
typedef struct pastix_data_s pastix_data_t;
typedef uintptr_t PASTIX_Comm;          /* no MPI (reference include/pastix.h:40) */
#endif

/* Identity ordering with one supernode (reference order/order.c:123-149). */
int  pastixOrderAllocId( pastix_order_t *ordeptr, pastix_int_t vertnbr );
/* The ordering of an instance (reference order/order.c:649-656). */
pastix_order_t *pastixOrderGet( const pastix_data_t *pastix_data );
/* One process: nothing to broadcast (reference order/order.c:679-). */
void pastixOrderBcast( pastix_order_t *ordemesh, int root, PASTIX_Comm pastix_comm );

/*
 * Ordering files in the reference's format (reference order/order_io.c:
 * 132-175, 293-337): version, cblknbr, vertnbr, rangtab, permtab and treetab.
 * pastixOrderLoad reads the file named by the environment variable
 * PASTIX_FILE_ORDER, else "ordername" in the current directory;
 * pastixOrderSave writes PASTIX_FILE_ORDER (default "ordergen") into the
 * instance's output directory "<PASTIX_OUTPUT_DIR or pastix>-XXXXXX".
 */
int  pastixOrderLoad( const pastix_data_t *pastix_data, pastix_order_t *ordeptr );
int  pastixOrderSave( pastix_data_t *pastix_data, const pastix_order_t *ordeptr );

#ifdef __cplusplus
}
#endif

#endif /* PASTIX_AMD_ORDER_H */
