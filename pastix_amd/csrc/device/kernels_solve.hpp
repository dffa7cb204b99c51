// kernels_solve.hpp -- supernodal triangular solve sweeps, SpMV and block
// vector kernels of the MI355X engine.
//
// Forward sweep (reference kernels/solve_ztrsmsp.c:300-424), level by level
// from the leaves: x_k = L_kk^{-1} b_k, then b[rows of blok i] -= A_i x_k for
// every off-diagonal blok (atomics: several cblks of a level update the same
// ancestor rows).  Backward sweep (:449-625), levels from the root:
// b_k -= sum_i A_i^T x[rows of i] (a gather of final values), then
// x_k = L_kk^{-T} b_k.  RHS columns are processed together (nrhs <= 8 per
// launch), so the factor is streamed once per sweep for all of them.
#pragma once
#include "kernels_fact.hpp"
#include "mfma.hpp"

#define PXD_MAXRHS 8

// x_k <- L_kk^{-1} x_k  (or L_kk^{-T} when trans) for a list of cblks.
template <typename T>
__global__ void __launch_bounds__( PXD_THREADS )
k_solve_diag( const int32_t *__restrict__ list, const dcblk *__restrict__ cb,
              const T *__restrict__ L, const T *__restrict__ dinv, T *__restrict__ X,
              int64_t ldx, int nrhs, int trans, int unit )
{
    extern __shared__ unsigned char smem_raw[];
    T *sx = reinterpret_cast<T *>( smem_raw );   // [n][nrhs]
    __shared__ T st[PXD_NB][PXD_MAXRHS];

    const int32_t k   = list[blockIdx.x];
    const dcblk   c   = cb[k];
    const int32_t n   = c.lcolnum - c.fcolnum + 1;
    const int     tid = threadIdx.x;
    const T      *A   = L + c.coefoff;   // n x n, ld n
    const T      *DI  = dinv + c.dinvoff;
    T            *x   = X + c.fcolnum;

    for ( int idx = tid; idx < n * nrhs; idx += PXD_THREADS ) {
        int r = idx % n, q = idx / n;
        sx[r * nrhs + q] = x[r + q * ldx];
    }
    __syncthreads();
    const int32_t nblk = ( n + PXD_NB - 1 ) / PXD_NB;
    for ( int32_t jb = 0; jb < nblk; jb++ ) {
        const int32_t J  = trans ? ( nblk - 1 - jb ) * PXD_NB : jb * PXD_NB;
        const int32_t bs = min( PXD_NB, n - J );
        for ( int idx = tid; idx < bs * nrhs; idx += PXD_THREADS ) {
            int cc = idx % bs, q = idx / bs;
            T   s  = sx[( J + cc ) * nrhs + q];
            if ( !trans ) {
                for ( int32_t l = 0; l < J; l++ ) {
                    s -= A[( J + cc ) + (int64_t)l * n] * sx[l * nrhs + q];
                }
            }
            else {
                for ( int32_t l = J + bs; l < n; l++ ) {
                    s -= A[l + (int64_t)( J + cc ) * n] * sx[l * nrhs + q];
                }
            }
            st[cc][q] = s;
        }
        __syncthreads();
        const T *Dj = DI + (int64_t)( J / PXD_NB ) * PXD_NB * PXD_NB;
        for ( int idx = tid; idx < bs * nrhs; idx += PXD_THREADS ) {
            int cc = idx % bs, q = idx / bs;
            T   s  = (T)0;
            if ( unit ) {
                s = st[cc][q];
            }
            else if ( !trans ) {
                for ( int l = 0; l <= cc; l++ ) {
                    s += Dj[cc + l * PXD_NB] * st[l][q];
                }
            }
            else {
                for ( int l = cc; l < bs; l++ ) {
                    s += Dj[l + cc * PXD_NB] * st[l][q];
                }
            }
            sx[( J + cc ) * nrhs + q] = s;
        }
        __syncthreads();
    }
    for ( int idx = tid; idx < n * nrhs; idx += PXD_THREADS ) {
        int r = idx % n, q = idx / n;
        x[r + q * ldx] = sx[r * nrhs + q];
    }
}

// Forward update of a 64-row tile of the off-diagonal rows of a cblk:
// X[grow] -= A(rows, :) * x_k, four wave-groups splitting the columns.
// B columns per load batch: 16 for launches with many workgroups (the
// lower tree, where occupancy matters), every column at once (up to 96
// registers) for launches with few, long workgroups (the upper tree).
template <typename T> struct solve_wide {
    static constexpr int value = 96 * 4 / (int)sizeof( T );
};
template <typename T, typename V, int B>
__global__ void __launch_bounds__( PXD_THREADS )
k_solve_fwd_upd( const row_task *__restrict__ tasks, const dcblk *__restrict__ cb,
                 const dblok *__restrict__ bk, const T *__restrict__ L, V *X,
                 int64_t ldx, int nrhs, const V *Xs, int64_t lds )
{
    extern __shared__ unsigned char smem_raw[];
    V *sx = reinterpret_cast<V *>( smem_raw );    // [n][nrhs]
    __shared__ V sred[4][64][PXD_MAXRHS + 1];

    const row_task t   = tasks[xcd_task( blockIdx.x, gridDim.x )];
    const dcblk    c   = cb[t.cblk];
    const int32_t  n   = c.lcolnum - c.fcolnum + 1;
    const int32_t  lb  = cb[t.cblk + 1].fblok;
    const int      tid = threadIdx.x;
    const int      r   = tid % 64, part = tid / 64;

    // x_k, solved: in Xs (the diagonal solve's output) or in X itself
    for ( int idx = tid; idx < n * nrhs; idx += PXD_THREADS ) {
        int rr = idx % n, q = idx / n;
        sx[rr * nrhs + q] = Xs[c.fcolnum + rr + q * lds];
    }
    __shared__ int32_t skeys[PXD_ROWKEY_CAP];
    stage_row_keys( skeys, bk, c.fblok, lb );
    __syncthreads();

    int32_t s     = t.r0 + r;
    bool    valid = ( s < c.stride );
    int64_t base  = 0;
    int32_t ld = 0, g = 0;
    if ( valid ) {
        int32_t b  = search_row_keys( skeys, bk, c.fblok, lb, s );
        dblok   bb = bk[b];
        base = c.coefoff + (int64_t)bb.rowoff * n + ( s - bb.rowoff );
        ld   = bb.lrownum - bb.frownum + 1;
        g    = bb.frownum + ( s - bb.rowoff );
    }
    V acc[PXD_MAXRHS];
#pragma unroll
    for ( int q = 0; q < PXD_MAXRHS; q++ ) acc[q] = (V)0;
    if ( valid ) {
        // the loads of a batch of B columns are all in flight at once
        const int32_t c0 = ( n * part ) / 4, c1 = ( n * ( part + 1 ) ) / 4;
        for ( int32_t cb0 = c0; cb0 < c1; cb0 += B ) {
            T a[B];
#pragma unroll
            for ( int u = 0; u < B; u++ ) a[u] = ( cb0 + u < c1 ) ? L[base + (int64_t)( cb0 + u ) * ld] : (T)0;
#pragma unroll
            for ( int u = 0; u < B; u++ ) {
                if ( cb0 + u < c1 ) {
                    const V av = px_cvt<V>( a[u] );
#pragma unroll
                    for ( int q = 0; q < PXD_MAXRHS; q++ ) {
                        if ( q < nrhs ) acc[q] += av * sx[( cb0 + u ) * nrhs + q];
                    }
                }
            }
        }
    }
#pragma unroll
    for ( int q = 0; q < PXD_MAXRHS; q++ ) {
        if ( q < nrhs ) sred[part][r][q] = acc[q];
    }
    __syncthreads();
    if ( part == 0 && valid ) {
        for ( int q = 0; q < nrhs; q++ ) {
            V v = sred[0][r][q] + sred[1][r][q] + sred[2][r][q] + sred[3][r][q];
            px_atomic_add( X + g + q * ldx, -v );
        }
    }
}

// Backward gather of a 64-row tile: x_k -= A(rows, :)^T * X[grow], one lane
// per row, wave reductions per column.
template <typename T>
__device__ inline T
wave_sum( T v )
{
#pragma unroll
    for ( int off = 32; off > 0; off >>= 1 ) {
        v += __shfl_xor( v, off, 64 );
    }
    return v;
}

template <typename T>
__global__ void __launch_bounds__( PXD_THREADS )
k_solve_bwd_gather( const row_task *__restrict__ tasks, const dcblk *__restrict__ cb,
                    const dblok *__restrict__ bk, const T *__restrict__ L, T *__restrict__ X,
                    int64_t ldx, int nrhs )
{
    const row_task t    = tasks[blockIdx.x];
    const dcblk    c    = cb[t.cblk];
    const int32_t  n    = c.lcolnum - c.fcolnum + 1;
    const int32_t  lb   = cb[t.cblk + 1].fblok;
    const int      lane = threadIdx.x % 64, wave = threadIdx.x / 64;

    int32_t s     = t.r0 + lane;
    bool    valid = ( s < c.stride );
    int64_t base  = 0;
    int32_t ld = 0;
    T       xv[PXD_MAXRHS];
#pragma unroll
    for ( int q = 0; q < PXD_MAXRHS; q++ ) xv[q] = (T)0;
    if ( valid ) {
        int32_t b  = find_blok_by_row( bk, c.fblok, lb, s );
        dblok   bb = bk[b];
        base = c.coefoff + (int64_t)bb.rowoff * n + ( s - bb.rowoff );
        ld   = bb.lrownum - bb.frownum + 1;
        int32_t g = bb.frownum + ( s - bb.rowoff );
#pragma unroll
        for ( int q = 0; q < PXD_MAXRHS; q++ ) {
            if ( q < nrhs ) xv[q] = X[g + q * ldx];
        }
    }
    for ( int32_t cc = wave; cc < n; cc += PXD_THREADS / 64 ) {
        T a = valid ? L[base + (int64_t)cc * ld] : (T)0;
        for ( int q = 0; q < nrhs; q++ ) {
            T v = wave_sum( a * xv[q] );
            if ( lane == 0 ) {
                px_atomic_add( X + c.fcolnum + cc + q * ldx, -v );
            }
        }
    }
}

// ---------------------------------------------------------------------------
// SpMV on the full-pattern CSC: column j of the stored matrix gives row j of
// the operator (symmetric: A itself; general: the transposed CSC is passed).
// y = alpha * A x + beta * y for nrhs columns (real alpha / beta); the
// column's entries are read once for up to PXD_SPMV_NR right-hand sides.
#define PXD_SPMV_NR 8
// G lanes per row (G a power of two up to 64, about the average entries per
// stored column): the lanes of a group read consecutive entries of the
// column, so a wave's loads of values and row indices are contiguous, and
// the G partial sums meet by shuffles.
__device__ inline double spmv_shfl_xor( double v, int m ) { return __shfl_xor( v, m, 64 ); }
__device__ inline pxcd   spmv_shfl_xor( pxcd v, int m )
{
    return pxcd{ __shfl_xor( v.re, m, 64 ), __shfl_xor( v.im, m, 64 ) };
}

template <typename T, typename TV, int G>
__global__ void __launch_bounds__( 256 )
k_spmv_grp( int64_t n, const int64_t *__restrict__ colptr, const int64_t *__restrict__ rowtab,
            const T *__restrict__ val, const TV *__restrict__ x, TV *__restrict__ y, int64_t ld,
            int nrhs, double alpha, double beta )
{
    typedef typename px_real<TV>::type RV;
    // the workgroups of one XCD take a contiguous range of columns, so that
    // the rows of x that neighbouring columns gather share that XCD's L2
    const int64_t j   = ( (int64_t)xcd_task( blockIdx.x, gridDim.x ) * blockDim.x + threadIdx.x ) / G;
    const int     sub = threadIdx.x % G;
    // rows past n take part in the shuffles with no entries
    int64_t e0 = 0, e1 = 0;
    if ( j < n ) {
        e0 = colptr[j];
        e1 = colptr[j + 1];
    }
    for ( int q0 = 0; q0 < nrhs; q0 += PXD_SPMV_NR ) {
        TV s[PXD_SPMV_NR];
#pragma unroll
        for ( int q = 0; q < PXD_SPMV_NR; q++ ) s[q] = px_zero( TV() );
        for ( int64_t e = e0 + sub; e < e1; e += G ) {
            const T       v = val[e];
            const int64_t r = rowtab[e];
#pragma unroll
            for ( int q = 0; q < PXD_SPMV_NR; q++ ) {
                if ( q0 + q < nrhs ) s[q] = px_fma( v, x[r + ( q0 + q ) * ld], s[q] );
            }
        }
#pragma unroll
        for ( int q = 0; q < PXD_SPMV_NR; q++ ) {
#pragma unroll
            for ( int m = G / 2; m > 0; m >>= 1 ) s[q] += spmv_shfl_xor( s[q], m );
        }
        if ( j < n && sub == 0 ) {
#pragma unroll
            for ( int q = 0; q < PXD_SPMV_NR; q++ ) {
                if ( q0 + q < nrhs ) {
                    TV *yy = y + j + ( q0 + q ) * ld;
                    *yy = ( beta == 0. ) ? s[q] * (RV)alpha : s[q] * (RV)alpha + *yy * (RV)beta;
                }
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Block vector kernels.  Reductions accumulate per column into res[] with
// atomics; res must be cleared first (k_zero).
template <typename T>
__global__ void
k_zero( T *p, int64_t n )
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if ( i < n ) {
        p[i] = px_zero( T() );
    }
}

__device__ inline double px_conj_mul( double a, double b ) { return a * b; }
__device__ inline pxcd   px_conj_mul( pxcd a, pxcd b )
{
    return pxcd{ a.re * b.re + a.im * b.im, a.re * b.im - a.im * b.re };
}
__device__ inline double px_re2( double a ) { return a * a; }
__device__ inline double px_re2( pxcd a ) { return a.re * a.re + a.im * a.im; }

__device__ inline double wave_sum_any( double v )
{
#pragma unroll
    for ( int off = 32; off > 0; off >>= 1 ) v += __shfl_xor( v, off, 64 );
    return v;
}
__device__ inline pxcd wave_sum_any( pxcd v )
{
#pragma unroll
    for ( int off = 32; off > 0; off >>= 1 ) {
        v.re += __shfl_xor( v.re, off, 64 );
        v.im += __shfl_xor( v.im, off, 64 );
    }
    return v;
}

// dot: res[q] += sum_i conj(x_i) y_i ; grid (blocks, nrhs)
template <typename T>
__global__ void __launch_bounds__( 256 )
k_dot( int64_t n, const T *__restrict__ x, const T *__restrict__ y, T *__restrict__ res )
{
    __shared__ T sh[4];
    const int q = blockIdx.y;
    const T  *xx = x + q * n, *yy = y + q * n;
    T         s  = px_zero( T() );
    const int64_t st = (int64_t)gridDim.x * blockDim.x;
    int64_t       i  = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    for ( ; i + 3 * st < n; i += 4 * st ) {
        const T a0 = px_conj_mul( xx[i], yy[i] ), a1 = px_conj_mul( xx[i + st], yy[i + st] );
        const T a2 = px_conj_mul( xx[i + 2 * st], yy[i + 2 * st] ), a3 = px_conj_mul( xx[i + 3 * st], yy[i + 3 * st] );
        s = s + ( ( a0 + a1 ) + ( a2 + a3 ) );
    }
    for ( ; i < n; i += st ) {
        s = s + px_conj_mul( xx[i], yy[i] );
    }
    const int lane = threadIdx.x % 64, wave = threadIdx.x / 64;
    s = wave_sum_any( s );
    if ( lane == 0 ) sh[wave] = s;
    __syncthreads();
    if ( threadIdx.x == 0 ) {
        T r = sh[0] + sh[1] + sh[2] + sh[3];
        px_atomic_add( res + q, r );
    }
}

template <typename T>
__global__ void __launch_bounds__( 256 )
k_nrm2sq( int64_t n, const T *__restrict__ x, double *__restrict__ res )
{
    __shared__ double sh[4];
    const int q  = blockIdx.y;
    const T  *xx = x + q * n;
    double    s  = 0.;
    const int64_t st = (int64_t)gridDim.x * blockDim.x;
    int64_t       i  = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    for ( ; i + 3 * st < n; i += 4 * st ) {
        s += ( px_re2( xx[i] ) + px_re2( xx[i + st] ) ) + ( px_re2( xx[i + 2 * st] ) + px_re2( xx[i + 3 * st] ) );
    }
    for ( ; i < n; i += st ) {
        s += px_re2( xx[i] );
    }
    const int lane = threadIdx.x % 64, wave = threadIdx.x / 64;
    s = wave_sum_any( s );
    if ( lane == 0 ) sh[wave] = s;
    __syncthreads();
    if ( threadIdx.x == 0 ) {
        unsafeAtomicAdd( res + q, sh[0] + sh[1] + sh[2] + sh[3] );
    }
}

// max that keeps a NaN of either operand (fmax drops it)
__device__ inline double
max_nan( double a, double b )
{
    return ( a > b || a != a ) ? a : b;
}

template <typename T>
__global__ void __launch_bounds__( 256 )
k_maxabs( int64_t n, const T *__restrict__ x, unsigned long long *__restrict__ res )
{
    __shared__ double sh[4];
    const int q  = blockIdx.y;
    const T  *xx = x + q * n;
    double    s  = 0.;
    const int64_t st = (int64_t)gridDim.x * blockDim.x;
    int64_t       i  = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    for ( ; i + 3 * st < n; i += 4 * st ) {
        const double a = max_nan( (double)px_abs( xx[i] ), (double)px_abs( xx[i + st] ) );
        const double b = max_nan( (double)px_abs( xx[i + 2 * st] ), (double)px_abs( xx[i + 3 * st] ) );
        s              = max_nan( s, max_nan( a, b ) );
    }
    for ( ; i < n; i += st ) {
        s = max_nan( s, (double)px_abs( xx[i] ) );
    }
    const int lane = threadIdx.x % 64, wave = threadIdx.x / 64;
#pragma unroll
    for ( int off = 32; off > 0; off >>= 1 ) s = max_nan( s, __shfl_xor( s, off, 64 ) );
    if ( lane == 0 ) sh[wave] = s;
    __syncthreads();
    if ( threadIdx.x == 0 ) {
        double m = max_nan( max_nan( sh[0], sh[1] ), max_nan( sh[2], sh[3] ) );
        // non-negative doubles order like their bit patterns, and a positive
        // NaN orders above +inf, so a NaN entry makes the norm NaN
        atomicMax( res + q, (unsigned long long)__double_as_longlong( m ) );
    }
}

// sqrt in place of nrm2 results
__global__ void
k_sqrt_inplace( double *r, int n )
{
    int i = threadIdx.x;
    if ( i < n ) r[i] = sqrt( r[i] );
}

// y(:,q) += sign * alpha[q] * x(:,q)
template <typename T>
__global__ void
k_axpy( int64_t n, int nrhs, double sign, const T *__restrict__ alpha, const T *__restrict__ x,
        T *__restrict__ y )
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if ( i >= n ) return;
    for ( int q = 0; q < nrhs; q++ ) {
        T a = alpha[q] * (typename px_real<T>::type)sign;
        y[i + q * n] = px_fma( a, x[i + q * n], y[i + q * n] );
    }
}

template <typename T>
__device__ inline T px_inv( T a ) { return px_div( (T)1, a ); }
template <>
__device__ inline pxcd px_inv( pxcd a ) { return px_div( pxcd{ 1., 0. }, a ); }

template <typename T>
__global__ void
k_scal( int64_t n, int nrhs, const T *__restrict__ alpha, int inv, T *__restrict__ x )
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if ( i >= n ) return;
    for ( int q = 0; q < nrhs; q++ ) {
        T a = inv ? px_inv( alpha[q] ) : alpha[q];
        x[i + q * n] = a * x[i + q * n];
    }
}

template <typename T>
__global__ void
k_gemv( int64_t n, int nrhs, int m, const T *__restrict__ V, const T *__restrict__ y,
        T *__restrict__ x )
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if ( i >= n ) return;
    for ( int q = 0; q < nrhs; q++ ) {
        T s = x[i + q * n];
        for ( int j = 0; j < m; j++ ) {
            s = px_fma( V[(int64_t)j * n * nrhs + i + q * n], y[j * nrhs + q], s );
        }
        x[i + q * n] = s;
    }
}

// precision conversions between the vector type and the factor type
template <typename TO, typename TI>
__global__ void
k_convert( int64_t n, const TI *__restrict__ in, TO *__restrict__ out )
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if ( i < n ) out[i] = px_from_any<TO>( in[i] );
}

// Y(:, q) <- X(:, q) for the columns of the grid's y dimension
template <typename V>
__global__ void
k_copy_cols( int64_t n, const V *__restrict__ X, int64_t ldx, V *__restrict__ Y, int64_t ldy )
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if ( i < n ) Y[i + blockIdx.y * ldy] = X[i + blockIdx.y * ldx];
}

// Y(:, q) <- 0 for the columns of the grid's y dimension
template <typename V>
__global__ void
k_zero_cols( int64_t n, V *__restrict__ Y, int64_t ldy )
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if ( i < n ) Y[i + blockIdx.y * ldy] = (V)0;
}

// dd[i] = the diagonal entry i of the factor's diagonal tiles (permuted
// numbering), widened to the vector precision (coeftab_zgetdiag,
// sopalin/coeftab_z.c:659-689)
template <typename T, typename V>
__global__ void
k_get_diag( int64_t n, const int32_t *__restrict__ col2cblk, const dcblk *__restrict__ cb,
            const T *__restrict__ L, V *__restrict__ dd )
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if ( i >= n ) return;
    const dcblk   c  = cb[col2cblk[i]];
    const int32_t nk = c.lcolnum - c.fcolnum + 1;
    dd[i] = px_cvt<V>( L[c.coefoff + (int64_t)( i - c.fcolnum ) * ( nk + 1 )] );
}

// LDL^T solve, between the sweeps: y_i <- y_i / D_i, D read from the
// diagonal of the cblk's diagonal tile (solve_cblk_zdiag, kernels/solve_ztrsmsp.c:629-679);
// factor entries T, vectors V (T or its double-precision counterpart)
template <typename T, typename V>
__global__ void
k_diag_scale( int64_t n, const int32_t *__restrict__ col2cblk, const dcblk *__restrict__ cb,
              const T *__restrict__ L, V *__restrict__ Y, int64_t ldy, int nrhs )
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if ( i >= n ) return;
    const dcblk   c  = cb[col2cblk[i]];
    const int32_t nk = c.lcolnum - c.fcolnum + 1;
    const V       d  = px_cvt<V>( L[c.coefoff + (int64_t)( i - c.fcolnum ) * ( nk + 1 )] );
    for ( int q = 0; q < nrhs; q++ ) Y[i + q * ldy] = px_div( Y[i + q * ldy], d );
}
