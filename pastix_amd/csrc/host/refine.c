/*
 * refine.c -- iterative refinement on the device, preconditioned by the
 * factorization (of possibly lower precision).
 *
 * GMRES(m) follows z_gmres_smp (reference refinement/z_refine_gmres.c:48-371):
 * right preconditioning, W_i = M^{-1} V_i kept to form x at restart, modified
 * Gram-Schmidt, Givens rotations on the host, stop when the estimated
 * ||r||_2 / ||b||_2 <= eps (eps = DPARM_EPSILON_REFINEMENT, 1e-12 by default
 * for double, refinement/pastix_task_refine.c:132-140).  The reference refines
 * one right-hand side at a time (:150-161); here all columns run together so
 * that each preconditioner application streams the factor once for every RHS,
 * each column keeping its own Krylov basis, rotations and stopping point.
 *
 * With the engine option "berr_target" the stopping rule is the componentwise
 * backward error ||b-Ax||_inf / (||A||_inf ||x||_inf + ||b||_inf) of every
 * column, checked on the true residual at every restart.
 *
 * The Krylov vectors, the residual, the SpMV and all dot products stay in
 * double (complex double) on the GPU; only the preconditioner may be single.
 */
#include <complex.h>
#include <math.h>
#include "px_internal.h"

typedef double complex zc;

typedef struct ref_ctx_s {
    pastix_data_t *pd;
    pxd_ctx_t     *dev;
    pastix_int_t   n, nrhs;
    int            cplx;
    size_t         es;        /* bytes per vector element */
    void          *dscal;     /* device scalars (nrhs) */
    double        *dnrm;      /* device reals (nrhs) */
} ref_ctx_t;

/*
 * Device buffers of the refinement are kept by the instance between calls
 * (PX_REF_CACHE slots, freed by pastixFinalize): a call takes the smallest
 * cached buffer that is large enough, else allocates, and hands its buffers
 * back when it returns.  Repeated refinements then allocate nothing.
 */
static int
ref_alloc( ref_ctx_t *r, size_t bytes, void **p )
{
    pastix_data_t *pd   = r->pd;
    int            best = -1;
    for ( int i = 0; i < PX_REF_CACHE; i++ ) {
        if ( pd->ref_cache[i] && pd->ref_cache_bytes[i] >= bytes &&
             ( best < 0 || pd->ref_cache_bytes[i] < pd->ref_cache_bytes[best] ) ) {
            best = i;
        }
    }
    if ( best >= 0 ) {
        *p                  = pd->ref_cache[best];
        pd->ref_cache[best] = NULL;
        return 0;
    }
    return pxd_malloc( r->dev, bytes, p );
}

static void
ref_release( ref_ctx_t *r, void *p, size_t bytes )
{
    pastix_data_t *pd = r->pd;
    if ( p == NULL ) {
        return;
    }
    for ( int i = 0; i < PX_REF_CACHE; i++ ) {
        if ( pd->ref_cache[i] == NULL ) {
            pd->ref_cache[i]       = p;
            pd->ref_cache_bytes[i] = bytes;
            return;
        }
    }
    pxd_free( r->dev, p );
}

static void
put_scalars( ref_ctx_t *r, const zc *h, void *dptr, pastix_int_t cnt )
{
    if ( r->cplx ) {
        pxd_h2d( r->dev, dptr, h, cnt * sizeof( zc ) );
    }
    else {
        double *t = px_malloc( cnt * sizeof( double ) );
        for ( pastix_int_t i = 0; i < cnt; i++ ) {
            t[i] = creal( h[i] );
        }
        pxd_h2d( r->dev, dptr, t, cnt * sizeof( double ) );
        free( t );
    }
}

static void
get_scalars( ref_ctx_t *r, zc *h, const void *dptr, pastix_int_t cnt )
{
    if ( r->cplx ) {
        pxd_d2h( r->dev, h, dptr, cnt * sizeof( zc ) );
    }
    else {
        double *t = px_malloc( cnt * sizeof( double ) );
        pxd_d2h( r->dev, t, dptr, cnt * sizeof( double ) );
        for ( pastix_int_t i = 0; i < cnt; i++ ) {
            h[i] = t[i];
        }
        free( t );
    }
}

static void
norms2( ref_ctx_t *r, const void *x, double *out )
{
    pxd_vec_nrm2( r->dev, r->n, r->nrhs, x, r->dnrm );
    pxd_d2h( r->dev, out, r->dnrm, r->nrhs * sizeof( double ) );
}

static void
normsinf( ref_ctx_t *r, const void *x, double *out )
{
    pxd_vec_maxabs( r->dev, r->n, r->nrhs, x, r->dnrm );
    pxd_d2h( r->dev, out, r->dnrm, r->nrhs * sizeof( double ) );
}

static void
scale_cols( ref_ctx_t *r, const zc *alpha, void *x )
{
    put_scalars( r, alpha, r->dscal, r->nrhs );
    pxd_vec_scal( r->dev, r->n, r->nrhs, r->dscal, 0, x );
}

/* per-column stopping thresholds on ||r||_2; returns the number of columns
 * already converged (their flag set in conv[]) */
static int
stop_targets( ref_ctx_t *r, const void *db, const void *dx, const void *dres, double eps, double berr,
              double normA, const double *normb2, double *target, int *conv )
{
    pastix_int_t nr = r->nrhs;
    double      *rn = px_malloc( nr * sizeof( double ) );
    int          nconv = 0;
    if ( berr > 0. ) {
        double *xi = px_malloc( nr * sizeof( double ) );
        double *bi = px_malloc( nr * sizeof( double ) );
        double *ri = px_malloc( nr * sizeof( double ) );
        normsinf( r, dx, xi );
        normsinf( r, db, bi );
        normsinf( r, dres, ri );
        for ( pastix_int_t c = 0; c < nr; c++ ) {
            double den = normA * xi[c] + bi[c];
            target[c]  = berr * den;
            conv[c]    = ( ri[c] <= berr * den );
        }
        free( xi );
        free( bi );
        free( ri );
    }
    else {
        norms2( r, dres, rn );
        for ( pastix_int_t c = 0; c < nr; c++ ) {
            target[c] = eps * normb2[c];
            conv[c]   = ( rn[c] <= target[c] );
        }
    }
    for ( pastix_int_t c = 0; c < nr; c++ ) {
        nconv += conv[c];
    }
    free( rn );
    return nconv;
}

static int
gmres_run( ref_ctx_t *r, void *db, void *dx, double eps, double berr, pastix_int_t im,
           pastix_int_t itermax, pastix_int_t *iters_out )
{
    pastix_int_t n = r->n, nr = r->nrhs, im1 = im + 1;
    size_t       blk = (size_t)n * nr * r->es;
    void        *V, *W, *dh, *dy;
    zc          *H, *cs, *sn, *g, *h, *alpha, *y;
    double      *normb, *target, *res, normA = 0.;
    int         *conv, *active;
    pastix_int_t *mcol, iters = 0;
    int          rc = PASTIX_SUCCESS;

    V = W = dh = dy = NULL;
    if ( ref_alloc( r, blk * im1, &V ) || ref_alloc( r, blk * im, &W ) ||
         ref_alloc( r, im1 * nr * r->es, &dh ) || ref_alloc( r, im * nr * r->es, &dy ) ) {
        /* hand back what was taken before the failure */
        ref_release( r, V, blk * im1 );
        ref_release( r, W, blk * im );
        ref_release( r, dh, im1 * nr * r->es );
        ref_release( r, dy, im * nr * r->es );
        return PASTIX_ERR_OUTOFMEMORY;
    }
    H      = px_calloc( (size_t)nr * im1 * im, sizeof( zc ) );
    cs     = px_calloc( (size_t)nr * im, sizeof( zc ) );
    sn     = px_calloc( (size_t)nr * im, sizeof( zc ) );
    g      = px_calloc( (size_t)nr * im1, sizeof( zc ) );
    h      = px_calloc( (size_t)nr * im1, sizeof( zc ) );
    y      = px_calloc( (size_t)nr * im, sizeof( zc ) );
    alpha  = px_calloc( nr, sizeof( zc ) );
    normb  = px_calloc( nr, sizeof( double ) );
    target = px_calloc( nr, sizeof( double ) );
    res    = px_calloc( nr, sizeof( double ) );
    conv   = px_calloc( nr, sizeof( int ) );
    active = px_calloc( nr, sizeof( int ) );
    mcol   = px_calloc( nr, sizeof( pastix_int_t ) );

    pxd_matrix_norm_inf( r->dev, &normA );
    norms2( r, db, normb );
    for ( pastix_int_t c = 0; c < nr; c++ ) {
        if ( normb[c] == 0. ) {
            normb[c] = 1.;
        }
    }

    for ( ;; ) {
        pastix_int_t i;
        int          nact = 0;
        /* V_0 = b - A x */
        pxd_vec_copy( r->dev, n, nr, db, V );
        pxd_spmv( r->dev, -1., dx, 1., V, nr, n );
        if ( stop_targets( r, db, dx, V, eps, berr, normA, normb, target, conv ) == nr ) {
            break;
        }
        norms2( r, V, res );
        for ( pastix_int_t c = 0; c < nr; c++ ) {
            active[c] = !conv[c];
            mcol[c]   = 0;
            nact += active[c];
            alpha[c]  = active[c] ? 1. / res[c] : 0.;
            g[c * im1] = res[c];
            for ( pastix_int_t k = 1; k < im1; k++ ) {
                g[c * im1 + k] = 0.;
            }
        }
        scale_cols( r, alpha, V );

        for ( i = 0; i < im && nact > 0 && iters < itermax; i++ ) {
            char *Vi  = (char *)V + i * blk;
            char *Vi1 = (char *)V + ( i + 1 ) * blk;
            char *Wi  = (char *)W + i * blk;

            pxd_vec_copy( r->dev, n, nr, Vi, Wi );
            if ( ( rc = pxd_solve( r->dev, nr, Wi, n ) ) != PASTIX_SUCCESS ) {
                goto done;
            }
            pxd_spmv( r->dev, 1., Wi, 0., Vi1, nr, n );
            for ( pastix_int_t j = 0; j <= i; j++ ) {
                char *hj = (char *)dh + j * nr * r->es;
                pxd_vec_dot( r->dev, n, nr, (char *)V + j * blk, Vi1, hj );
                pxd_vec_axpy( r->dev, n, nr, -1., hj, (char *)V + j * blk, Vi1 );
            }
            norms2( r, Vi1, res );
            get_scalars( r, h, dh, ( i + 1 ) * nr );
            for ( pastix_int_t c = 0; c < nr; c++ ) {
                alpha[c] = ( active[c] && res[c] > 1e-50 ) ? 1. / res[c] : ( active[c] ? 1. : 0. );
            }
            scale_cols( r, alpha, Vi1 );
            iters++;

            for ( pastix_int_t c = 0; c < nr; c++ ) {
                zc *Hc, *Hi, tmp;
                if ( !active[c] ) {
                    continue;
                }
                Hc = H + (size_t)c * im1 * im;
                Hi = Hc + i * im1;
                for ( pastix_int_t j = 0; j <= i; j++ ) {
                    Hi[j] = h[j * nr + c];
                }
                Hi[i + 1] = res[c];
                for ( pastix_int_t j = 0; j < i; j++ ) {
                    tmp       = Hi[j];
                    Hi[j]     = cs[c * im + j] * tmp + sn[c * im + j] * Hi[j + 1];
                    Hi[j + 1] = cs[c * im + j] * Hi[j + 1] - conj( sn[c * im + j] ) * tmp;
                }
                tmp = csqrt( Hi[i] * Hi[i] + Hi[i + 1] * Hi[i + 1] );
                if ( cabs( tmp ) <= eps ) {
                    tmp = eps;
                }
                cs[c * im + i]     = Hi[i] / tmp;
                sn[c * im + i]     = Hi[i + 1] / tmp;
                g[c * im1 + i + 1] = -sn[c * im + i] * g[c * im1 + i];
                g[c * im1 + i]     = cs[c * im + i] * g[c * im1 + i];
                Hi[i]              = cs[c * im + i] * Hi[i] + sn[c * im + i] * Hi[i + 1];
                mcol[c]            = i + 1;
                if ( cabs( g[c * im1 + i + 1] ) <= target[c] ) {
                    active[c] = 0;
                    nact--;
                }
            }
        }

        /* y = H^{-1} g per column, x += W y */
        for ( pastix_int_t c = 0; c < nr; c++ ) {
            zc          *Hc = H + (size_t)c * im1 * im;
            pastix_int_t m  = mcol[c];
            for ( pastix_int_t k = 0; k < im; k++ ) {
                y[k * nr + c] = 0.;
            }
            for ( pastix_int_t k = m - 1; k >= 0; k-- ) {
                zc s = g[c * im1 + k];
                for ( pastix_int_t l = k + 1; l < m; l++ ) {
                    s -= Hc[l * im1 + k] * y[l * nr + c];
                }
                y[k * nr + c] = s / Hc[k * im1 + k];
            }
        }
        put_scalars( r, y, dy, im * nr );
        pxd_vec_gemv( r->dev, n, nr, i, W, dy, dx );
        if ( iters >= itermax ) {
            break;
        }
    }
done:
    pxd_sync( r->dev );
    *iters_out = iters;
    ref_release( r, V, blk * im1 );
    ref_release( r, W, blk * im );
    ref_release( r, dh, im1 * nr * r->es );
    ref_release( r, dy, im * nr * r->es );
    free( H ); free( cs ); free( sn ); free( g ); free( h ); free( y ); free( alpha );
    free( normb ); free( target ); free( res ); free( conv ); free( active ); free( mcol );
    return rc;
}

/* Simple iterative refinement: x += M^{-1} (b - A x) until converged. */
static int
sr_run( ref_ctx_t *r, void *db, void *dx, double eps, double berr, pastix_int_t itermax,
        pastix_int_t *iters_out )
{
    pastix_int_t n = r->n, nr = r->nrhs, iters = 0;
    size_t       blk = (size_t)n * nr * r->es;
    void        *R;
    double      *normb = px_calloc( nr, sizeof( double ) ), *target = px_calloc( nr, sizeof( double ) );
    int         *conv = px_calloc( nr, sizeof( int ) );
    zc          *one = px_calloc( nr, sizeof( zc ) );
    double       normA = 0.;
    int          rc = PASTIX_SUCCESS;

    if ( ref_alloc( r, blk, &R ) ) {
        free( normb );
        free( target );
        free( conv );
        free( one );
        return PASTIX_ERR_OUTOFMEMORY;
    }
    pxd_matrix_norm_inf( r->dev, &normA );
    norms2( r, db, normb );
    for ( pastix_int_t c = 0; c < nr; c++ ) {
        if ( normb[c] == 0. ) normb[c] = 1.;
    }
    for ( ; iters < itermax; iters++ ) {
        pxd_vec_copy( r->dev, n, nr, db, R );
        pxd_spmv( r->dev, -1., dx, 1., R, nr, n );
        if ( stop_targets( r, db, dx, R, eps, berr, normA, normb, target, conv ) == nr ) {
            break;
        }
        if ( ( rc = pxd_solve( r->dev, nr, R, n ) ) != PASTIX_SUCCESS ) {
            break;
        }
        for ( pastix_int_t c = 0; c < nr; c++ ) {
            one[c] = conv[c] ? 0. : 1.;
        }
        put_scalars( r, one, r->dscal, nr );
        pxd_vec_axpy( r->dev, n, nr, 1., r->dscal, R, dx );
    }
    pxd_sync( r->dev );
    *iters_out = iters;
    ref_release( r, R, blk );
    free( normb ); free( target ); free( conv ); free( one );
    return rc;
}

/* per-column dot products (x, y) = sum conj(x_i) y_i, on the host */
static void
dots( ref_ctx_t *r, const void *x, const void *y, zc *out )
{
    pxd_vec_dot( r->dev, r->n, r->nrhs, x, y, r->dscal );
    get_scalars( r, out, r->dscal, r->nrhs );
}

/* y += sign * a[c] * x for every column c */
static void
axpy_cols( ref_ctx_t *r, double sign, const zc *a, const void *x, void *y )
{
    put_scalars( r, a, r->dscal, r->nrhs );
    pxd_vec_axpy( r->dev, r->n, r->nrhs, sign, r->dscal, x, y );
}

/*
 * BiCGStab preconditioned by the factorization, following z_bicgstab_smp
 * (refinement/z_refine_bicgstab.c:48-228): y = M^{-1} p, v = A y, alpha =
 * (r, r2) / (v, r2), s = r - alpha v, z = M^{-1} s, t = A z, omega =
 * (M^{-1} t, z) / (M^{-1} t, M^{-1} t), x += alpha y + omega z, r = s - omega
 * t, beta = (r, r2) / (r_old, r2) * alpha / omega, p = r + beta (p - omega v);
 * stop when ||r||_2 / ||b||_2 <= eps.  The scalars are complex for complex
 * matrices (the reference keeps alpha and beta in double variables).  All
 * columns run together, each with its own scalars and stopping point.
 */
static int
bicgstab_run( ref_ctx_t *r, void *db, void *dx, double eps, double berr, pastix_int_t itermax,
              pastix_int_t *iters_out )
{
    pastix_int_t n = r->n, nr = r->nrhs, iters = 0;
    size_t       blk = (size_t)n * nr * r->es;
    void        *R = NULL, *R2 = NULL, *P = NULL, *Yv = NULL, *V = NULL, *S = NULL, *Z = NULL, *Tv = NULL,
                *G2 = NULL;
    void       **vec[] = { &R, &R2, &P, &Yv, &V, &S, &Z, &Tv, &G2 };
    double      *normb = px_calloc( nr, sizeof( double ) ), *target = px_calloc( nr, sizeof( double ) );
    double      *rn = px_calloc( nr, sizeof( double ) );
    int         *conv = px_calloc( nr, sizeof( int ) );
    zc          *alpha = px_calloc( nr, sizeof( zc ) ), *beta = px_calloc( nr, sizeof( zc ) );
    zc          *om = px_calloc( nr, sizeof( zc ) ), *t1 = px_calloc( nr, sizeof( zc ) );
    zc          *t2 = px_calloc( nr, sizeof( zc ) );
    double       normA = 0.;
    int          rc = PASTIX_SUCCESS;

    for ( size_t i = 0; i < sizeof( vec ) / sizeof( vec[0] ); i++ ) {
        if ( ref_alloc( r, blk, vec[i] ) ) {
            for ( size_t j = 0; j < i; j++ ) ref_release( r, *vec[j], blk );
            free( normb ); free( target ); free( rn ); free( conv ); free( alpha ); free( beta );
            free( om ); free( t1 ); free( t2 );
            return PASTIX_ERR_OUTOFMEMORY;
        }
    }
    pxd_matrix_norm_inf( r->dev, &normA );
    norms2( r, db, normb );
    for ( pastix_int_t c = 0; c < nr; c++ ) {
        if ( normb[c] == 0. ) normb[c] = 1.;
    }
    /* r = b - A x, r2 = r, p = r */
    pxd_vec_copy( r->dev, n, nr, db, R );
    pxd_spmv( r->dev, -1., dx, 1., R, nr, n );
    pxd_vec_copy( r->dev, n, nr, R, R2 );
    pxd_vec_copy( r->dev, n, nr, R, P );
    int nconv = stop_targets( r, db, dx, R, eps, berr, normA, normb, target, conv );

    while ( nconv < nr && iters < itermax ) {
        iters++;
        /* y = M^{-1} p, v = A y */
        pxd_vec_copy( r->dev, n, nr, P, Yv );
        if ( ( rc = pxd_solve( r->dev, nr, Yv, n ) ) != PASTIX_SUCCESS ) break;
        pxd_spmv( r->dev, 1., Yv, 0., V, nr, n );
        /* alpha = (r, r2) / (v, r2) */
        dots( r, V, R2, t1 );
        dots( r, R, R2, beta );
        for ( pastix_int_t c = 0; c < nr; c++ ) {
            alpha[c] = ( conv[c] || t1[c] == 0. ) ? 0. : beta[c] / t1[c];
        }
        /* s = r - alpha v;  z = M^{-1} s;  t = A z;  g2 = M^{-1} t */
        pxd_vec_copy( r->dev, n, nr, R, S );
        axpy_cols( r, -1., alpha, V, S );
        pxd_vec_copy( r->dev, n, nr, S, Z );
        if ( ( rc = pxd_solve( r->dev, nr, Z, n ) ) != PASTIX_SUCCESS ) break;
        pxd_spmv( r->dev, 1., Z, 0., Tv, nr, n );
        pxd_vec_copy( r->dev, n, nr, Tv, G2 );
        if ( ( rc = pxd_solve( r->dev, nr, G2, n ) ) != PASTIX_SUCCESS ) break;
        /* omega = (g2, z) / (g2, g2) */
        dots( r, G2, Z, t1 );
        dots( r, G2, G2, t2 );
        for ( pastix_int_t c = 0; c < nr; c++ ) {
            om[c] = ( conv[c] || t2[c] == 0. ) ? 0. : t1[c] / t2[c];
        }
        /* x += alpha y + omega z;  r = s - omega t (frozen columns keep r) */
        axpy_cols( r, 1., alpha, Yv, dx );
        axpy_cols( r, 1., om, Z, dx );
        for ( pastix_int_t c = 0; c < nr; c++ ) t1[c] = conv[c] ? 0. : 1.;
        axpy_cols( r, -1., t1, R, R );        /* r <- 0 on the active columns */
        axpy_cols( r, 1., t1, S, R );         /* r <- s */
        axpy_cols( r, -1., om, Tv, R );       /* r <- s - omega t */
        /* beta = (r, r2) / (r_old, r2) * alpha / omega */
        dots( r, R, R2, t2 );
        for ( pastix_int_t c = 0; c < nr; c++ ) {
            zc b = 0.;
            if ( !conv[c] && beta[c] != 0. && om[c] != 0. ) {
                b = t2[c] / beta[c] * ( alpha[c] / om[c] );
            }
            beta[c] = b;
        }
        /* p = r + beta (p - omega v) */
        axpy_cols( r, -1., om, V, P );
        for ( pastix_int_t c = 0; c < nr; c++ ) t1[c] = conv[c] ? 1. : beta[c];
        put_scalars( r, t1, r->dscal, nr );
        pxd_vec_scal( r->dev, n, nr, r->dscal, 0, P );
        for ( pastix_int_t c = 0; c < nr; c++ ) t1[c] = conv[c] ? 0. : 1.;
        axpy_cols( r, 1., t1, R, P );
        /* stop on ||r|| (recursive residual) or on the backward error of the
         * true residual; a breakdown ends the column */
        if ( berr > 0. ) {
            pxd_vec_copy( r->dev, n, nr, db, S );
            pxd_spmv( r->dev, -1., dx, 1., S, nr, n );
            nconv = stop_targets( r, db, dx, S, eps, berr, normA, normb, target, conv );
        }
        else {
            norms2( r, R, rn );
            nconv = 0;
            for ( pastix_int_t c = 0; c < nr; c++ ) {
                if ( rn[c] <= eps * normb[c] || alpha[c] == 0. || om[c] == 0. ) conv[c] = 1;
                nconv += conv[c];
            }
        }
    }
    pxd_sync( r->dev );
    *iters_out = iters;
    for ( size_t i = 0; i < sizeof( vec ) / sizeof( vec[0] ); i++ ) {
        ref_release( r, *vec[i], blk );
    }
    free( normb ); free( target ); free( rn ); free( conv ); free( alpha ); free( beta );
    free( om ); free( t1 ); free( t2 );
    return rc;
}

/*
 * Preconditioned conjugate gradient following z_grad_smp
 * (refinement/z_refine_grad.c:40-200), for symmetric (Hermitian) positive
 * definite matrices: z = M^{-1} r, p = z; then alpha = (r, z) / (Ap, p),
 * x += alpha p, r -= alpha Ap, z = M^{-1} r, beta = (r_new, z_new) / (r, z),
 * p = z + beta p; stop when ||r||_2 / ||b||_2 <= eps.  Columns run together.
 */
static int
cg_run( ref_ctx_t *r, void *db, void *dx, double eps, double berr, pastix_int_t itermax,
        pastix_int_t *iters_out )
{
    pastix_int_t n = r->n, nr = r->nrhs, iters = 0;
    size_t       blk = (size_t)n * nr * r->es;
    void        *R = NULL, *P = NULL, *Z = NULL, *Q = NULL;
    void       **vec[] = { &R, &P, &Z, &Q };
    double      *normb = px_calloc( nr, sizeof( double ) ), *target = px_calloc( nr, sizeof( double ) );
    double      *rn = px_calloc( nr, sizeof( double ) );
    int         *conv = px_calloc( nr, sizeof( int ) );
    zc          *alpha = px_calloc( nr, sizeof( zc ) ), *rz = px_calloc( nr, sizeof( zc ) );
    zc          *t1 = px_calloc( nr, sizeof( zc ) );
    double       normA = 0.;
    int          rc = PASTIX_SUCCESS;

    for ( size_t i = 0; i < sizeof( vec ) / sizeof( vec[0] ); i++ ) {
        if ( ref_alloc( r, blk, vec[i] ) ) {
            for ( size_t j = 0; j < i; j++ ) ref_release( r, *vec[j], blk );
            free( normb ); free( target ); free( rn ); free( conv ); free( alpha ); free( rz ); free( t1 );
            return PASTIX_ERR_OUTOFMEMORY;
        }
    }
    pxd_matrix_norm_inf( r->dev, &normA );
    norms2( r, db, normb );
    for ( pastix_int_t c = 0; c < nr; c++ ) {
        if ( normb[c] == 0. ) normb[c] = 1.;
    }
    pxd_vec_copy( r->dev, n, nr, db, R );
    pxd_spmv( r->dev, -1., dx, 1., R, nr, n );
    int nconv = stop_targets( r, db, dx, R, eps, berr, normA, normb, target, conv );
    pxd_vec_copy( r->dev, n, nr, R, Z );
    if ( nconv < nr && ( rc = pxd_solve( r->dev, nr, Z, n ) ) == PASTIX_SUCCESS ) {
        pxd_vec_copy( r->dev, n, nr, Z, P );
        dots( r, R, Z, rz );
    }
    while ( rc == PASTIX_SUCCESS && nconv < nr && iters < itermax ) {
        iters++;
        /* q = A p, alpha = (r, z) / (q, p) */
        pxd_spmv( r->dev, 1., P, 0., Q, nr, n );
        dots( r, Q, P, t1 );
        for ( pastix_int_t c = 0; c < nr; c++ ) {
            alpha[c] = ( conv[c] || t1[c] == 0. ) ? 0. : rz[c] / t1[c];
        }
        axpy_cols( r, 1., alpha, P, dx );
        axpy_cols( r, -1., alpha, Q, R );
        pxd_vec_copy( r->dev, n, nr, R, Z );
        if ( ( rc = pxd_solve( r->dev, nr, Z, n ) ) != PASTIX_SUCCESS ) break;
        dots( r, R, Z, t1 );
        for ( pastix_int_t c = 0; c < nr; c++ ) {
            zc beta = ( conv[c] || rz[c] == 0. ) ? 1. : t1[c] / rz[c];
            rz[c]   = conv[c] ? rz[c] : t1[c];
            t1[c]   = beta;
        }
        /* p = z + beta p (frozen columns keep p) */
        put_scalars( r, t1, r->dscal, nr );
        pxd_vec_scal( r->dev, n, nr, r->dscal, 0, P );
        for ( pastix_int_t c = 0; c < nr; c++ ) t1[c] = conv[c] ? 0. : 1.;
        axpy_cols( r, 1., t1, Z, P );
        if ( berr > 0. ) {
            pxd_vec_copy( r->dev, n, nr, db, Q );
            pxd_spmv( r->dev, -1., dx, 1., Q, nr, n );
            nconv = stop_targets( r, db, dx, Q, eps, berr, normA, normb, target, conv );
        }
        else {
            norms2( r, R, rn );
            nconv = 0;
            for ( pastix_int_t c = 0; c < nr; c++ ) {
                if ( rn[c] <= eps * normb[c] || alpha[c] == 0. ) conv[c] = 1;
                nconv += conv[c];
            }
        }
    }
    pxd_sync( r->dev );
    *iters_out = iters;
    for ( size_t i = 0; i < sizeof( vec ) / sizeof( vec[0] ); i++ ) {
        ref_release( r, *vec[i], blk );
    }
    free( normb ); free( target ); free( rn ); free( conv ); free( alpha ); free( rz ); free( t1 );
    return rc;
}

int
px_refine_run( pastix_data_t *pd, pastix_int_t nrhs, void *db, void *dx, pastix_int_t *iters )
{
    ref_ctx_t    r;
    double       eps = pd->dparm[DPARM_EPSILON_REFINEMENT];
    int          rc  = PASTIX_SUCCESS;
    pastix_int_t it = 0, itmax = 0;

    if ( eps < 0. ) {
        eps = 1e-12;   /* double-precision input (refinement/pastix_task_refine.c:132-140) */
        pd->dparm[DPARM_EPSILON_REFINEMENT] = eps;
    }
    r.pd   = pd;
    r.dev  = pd->dev;
    r.n    = pd->bcsc->n;
    r.cplx = ( pd->bcsc->flttype == SpmComplex64 );
    r.es   = r.cplx ? 16 : 8;
    r.dscal = NULL;
    r.dnrm  = NULL;
    if ( ref_alloc( &r, 64 * 16, &r.dscal ) || ref_alloc( &r, 64 * 8, (void **)&r.dnrm ) ) {
        ref_release( &r, r.dscal, 64 * 16 );
        return PASTIX_ERR_OUTOFMEMORY;
    }

    /* columns are refined in groups (all at once in batch mode, one by one
     * otherwise); device scalars hold at most 64 columns */
    {
        pastix_int_t group = pd->refine_nrhs_batch ? 64 : 1;
        for ( pastix_int_t c0 = 0; c0 < nrhs && rc == PASTIX_SUCCESS; c0 += group ) {
            pastix_int_t nc = px_imin( group, nrhs - c0 );
            char        *b  = (char *)db + c0 * r.n * r.es;
            char        *x  = (char *)dx + c0 * r.n * r.es;
            r.nrhs = nc;
            switch ( pd->iparm[IPARM_REFINEMENT] ) {
            case PastixRefineSR:
                rc = sr_run( &r, b, x, eps, pd->berr_target, pd->iparm[IPARM_ITERMAX], &it );
                break;
            case PastixRefineCG:
                rc = cg_run( &r, b, x, eps, pd->berr_target, pd->iparm[IPARM_ITERMAX], &it );
                break;
            case PastixRefineBiCGSTAB:
                rc = bicgstab_run( &r, b, x, eps, pd->berr_target, pd->iparm[IPARM_ITERMAX], &it );
                break;
            case PastixRefineGMRES:
            default:
                rc = gmres_run( &r, b, x, eps, pd->berr_target, pd->iparm[IPARM_GMRES_IM],
                                pd->iparm[IPARM_ITERMAX], &it );
                break;
            }
            itmax = px_imax( itmax, it );
        }
    }
    ref_release( &r, r.dscal, 64 * 16 );
    ref_release( &r, r.dnrm, 64 * 8 );
    *iters = itmax;
    return rc;
}
