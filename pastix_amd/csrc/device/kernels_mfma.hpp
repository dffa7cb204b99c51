// kernels_mfma.hpp -- matrix-core factorization kernels and the inverse-based
// solve kernels of the MI355X engine.
//
// The diagonal block L_kk of every cblk is inverted explicitly (Linv, n x n,
// column major, upper part zero): the off-diagonal TRSM of the reference
// (core_ztrsmsp.c:148-183, L_off <- L_off L_kk^{-T}) becomes the product
// L_off Linv^T on the matrix cores, and the diagonal solves of the sweeps
// (solve_ztrsmsp.c:74-108) become Linv b and Linv^T b.
#pragma once
#include "kernels_fact.hpp"
#include "mfma.hpp"

// tasks of the block-column inversion: cblk and first column of the block
struct inv_task {
    int32_t cblk;
    int32_t J;
};

// ---------------------------------------------------------------------------
// Panel: Cholesky of the diagonal tile with static pivoting, right looking by
// PXD_NB columns; the unblocked step updates with unscaled columns so that it
// needs one barrier per column.  The inverses of the PXD_NB x PXD_NB diagonal
// sub-blocks go to the diagonal blocks of Linv.
template <typename T>
__global__ void __launch_bounds__( PXD_THREADS )
k_potrf2( const int32_t *__restrict__ list, const dcblk *__restrict__ cb, T *__restrict__ L,
          T *__restrict__ Linv, double thresh, int *__restrict__ npiv )
{
    __shared__ T sA[PXD_NB][PXD_NB + 1];   // sA[col][row]
    __shared__ T sI[PXD_NB][PXD_NB + 1];   // inverse, sI[col][row]
    __shared__ T sD[PXD_NB];
    __shared__ T sPi[PXD_NB][64 + 1];
    __shared__ T sPj[PXD_NB][64 + 1];

    const int32_t k   = list[blockIdx.x];
    const dcblk   c   = cb[k];
    const int32_t n   = c.lcolnum - c.fcolnum + 1;
    const int     tid = threadIdx.x;
    T *A  = L + c.coefoff;            // n x n, ld n
    T *LI = Linv + c.dinvoff;         // n x n, ld n
    int piv = 0;
    const T tau = (T)thresh;

    for ( int32_t kb = 0; kb < n; kb += PXD_NB ) {
        const int32_t bs = min( PXD_NB, n - kb );

        for ( int idx = tid; idx < PXD_NB * PXD_NB; idx += PXD_THREADS ) {
            int r = idx % PXD_NB, cc = idx / PXD_NB;
            sA[cc][r] = ( r < bs && cc < bs && r >= cc ) ? A[( kb + r ) + (int64_t)( kb + cc ) * n] : (T)0;
            sI[cc][r] = (T)0;
        }
        __syncthreads();

        for ( int cc = 0; cc < bs; cc++ ) {
            T d = sA[cc][cc];
            if ( px_abs( d ) < tau ) {
                d = tau;
                if ( tid == 0 ) {
                    piv++;
                }
            }
            const int m = bs - cc - 1;
            const T   r = (T)1 / d;
            for ( int idx = tid; idx < m * m; idx += PXD_THREADS ) {
                int rr = cc + 1 + idx % m, c2 = cc + 1 + idx / m;
                if ( rr >= c2 ) {
                    sA[c2][rr] -= sA[cc][rr] * sA[cc][c2] * r;
                }
            }
            if ( tid == 0 ) {
                sD[cc] = d;
            }
            __syncthreads();
        }
        // scale the columns by 1/sqrt(pivot)
        for ( int idx = tid; idx < bs * bs; idx += PXD_THREADS ) {
            int rr = idx % bs, cc = idx / bs;
            if ( rr > cc ) {
                sA[cc][rr] = sA[cc][rr] / px_sqrt( sD[cc] );
            }
        }
        __syncthreads();
        if ( tid < bs ) {
            sA[tid][tid] = px_sqrt( sD[tid] );
        }
        __syncthreads();

        // inverse of the lower-triangular block: thread t builds column t
        if ( tid < bs ) {
            const int t = tid;
            sI[t][t] = (T)1 / sA[t][t];
            for ( int i = t + 1; i < bs; i++ ) {
                T s = (T)0;
                for ( int l = t; l < i; l++ ) {
                    s += sA[l][i] * sI[t][l];
                }
                sI[t][i] = -s / sA[i][i];
            }
        }
        __syncthreads();

        for ( int idx = tid; idx < bs * bs; idx += PXD_THREADS ) {
            int rr = idx % bs, cc = idx / bs;
            if ( rr >= cc ) {
                A[( kb + rr ) + (int64_t)( kb + cc ) * n] = sA[cc][rr];
            }
            LI[( kb + rr ) + (int64_t)( kb + cc ) * n] = sI[cc][rr];
        }

        // rows below the block inside the tile: A21 <- A21 * L11^{-T}; each
        // thread keeps its row's results in registers until all are computed
        for ( int32_t r = kb + bs + tid; r < n; r += PXD_THREADS ) {
            T out[PXD_NB];
#pragma unroll
            for ( int c0 = 0; c0 < PXD_NB; c0 += 8 ) {
                T acc[8];
#pragma unroll
                for ( int j = 0; j < 8; j++ ) acc[j] = (T)0;
                const int lend = min( c0 + 8, (int)bs );
                for ( int l = 0; l < lend; l++ ) {
                    const T a = A[r + (int64_t)( kb + l ) * n];
#pragma unroll
                    for ( int j = 0; j < 8; j++ ) {
                        if ( l <= c0 + j ) acc[j] += a * sI[l][c0 + j];
                    }
                }
#pragma unroll
                for ( int j = 0; j < 8; j++ ) out[c0 + j] = acc[j];
            }
#pragma unroll
            for ( int cc = 0; cc < PXD_NB; cc++ ) {
                if ( cc < bs ) A[r + (int64_t)( kb + cc ) * n] = out[cc];
            }
        }
        __syncthreads();

        // trailing update A22 -= A21 * A21^T (lower part, 64 x 64 macro tiles)
        const int32_t m0 = kb + bs;
        const int32_t m  = n - m0;
        const int32_t nt = ( m + 63 ) / 64;
        const int     tx = tid % 16, ty = tid / 16;
        for ( int32_t ti = 0; ti < nt; ti++ ) {
            for ( int32_t tj = 0; tj <= ti; tj++ ) {
                for ( int idx = tid; idx < 64 * PXD_NB; idx += PXD_THREADS ) {
                    int r = idx % 64, l = idx / 64;
                    int32_t ri = ti * 64 + r, rj = tj * 64 + r;
                    sPi[l][r] = ( ri < m && l < bs ) ? A[( m0 + ri ) + (int64_t)( kb + l ) * n] : (T)0;
                    sPj[l][r] = ( rj < m && l < bs ) ? A[( m0 + rj ) + (int64_t)( kb + l ) * n] : (T)0;
                }
                __syncthreads();
                T acc[4][4];
#pragma unroll
                for ( int a = 0; a < 4; a++ )
#pragma unroll
                    for ( int b = 0; b < 4; b++ ) acc[a][b] = (T)0;
                for ( int l = 0; l < bs; l++ ) {
                    T ai[4], aj[4];
#pragma unroll
                    for ( int a = 0; a < 4; a++ ) {
                        ai[a] = sPi[l][tx + 16 * a];
                        aj[a] = sPj[l][ty + 16 * a];
                    }
#pragma unroll
                    for ( int a = 0; a < 4; a++ )
#pragma unroll
                        for ( int b = 0; b < 4; b++ ) acc[a][b] += ai[a] * aj[b];
                }
#pragma unroll
                for ( int a = 0; a < 4; a++ ) {
#pragma unroll
                    for ( int b = 0; b < 4; b++ ) {
                        int32_t r = ti * 64 + tx + 16 * a, cc = tj * 64 + ty + 16 * b;
                        if ( r < m && cc < m && r >= cc ) {
                            A[( m0 + r ) + (int64_t)( m0 + cc ) * n] -= acc[a][b];
                        }
                    }
                }
                __syncthreads();
            }
        }
        __syncthreads();
    }
    if ( tid == 0 && piv > 0 ) {
        atomicAdd( npiv, piv );
    }
}

// ---------------------------------------------------------------------------
// Inverse of the lower-triangular L_kk, one workgroup per (cblk, block column
// J): Linv[I,J] = -Linv[I,I] * sum_{K=J}^{I-1} L[I,K] Linv[K,J] for I > J, the
// diagonal blocks Linv[I,I] written by the panel.  Block column J of Linv is
// produced top-down and read back from global memory by the same workgroup.
template <typename T>
__global__ void __launch_bounds__( PXD_THREADS )
k_trtri( const inv_task *__restrict__ tasks, const dcblk *__restrict__ cb, const T *__restrict__ L,
         T *__restrict__ Linv )
{
    __shared__ T sL[PXD_NB][PXD_NB + 1];   // L[I,K]        sL[col][row]
    __shared__ T sX[PXD_NB][PXD_NB + 1];   // Linv[K,J]     sX[col][row]
    __shared__ T sT[PXD_NB][PXD_NB + 1];   // accumulated T sT[col][row]

    const inv_task t   = tasks[blockIdx.x];
    const dcblk    c   = cb[t.cblk];
    const int32_t  n   = c.lcolnum - c.fcolnum + 1;
    const int      tid = threadIdx.x;
    const T       *A   = L + c.coefoff;
    T             *LI  = Linv + c.dinvoff;
    const int32_t  J   = t.J;
    const int32_t  bj  = min( PXD_NB, n - J );

    for ( int32_t I = J + PXD_NB; I < n; I += PXD_NB ) {
        const int32_t bi = min( PXD_NB, n - I );
        T             acc[4];
#pragma unroll
        for ( int q = 0; q < 4; q++ ) acc[q] = (T)0;
        // thread owns (row r, cols cq..): r = tid % 32, cols tid/32 + 8q
        const int r = tid % PXD_NB, c0 = tid / PXD_NB;
        for ( int32_t K = J; K < I; K += PXD_NB ) {
            const int32_t bk = min( PXD_NB, n - K );
            for ( int idx = tid; idx < PXD_NB * PXD_NB; idx += PXD_THREADS ) {
                int rr = idx % PXD_NB, cc = idx / PXD_NB;
                sL[cc][rr] = ( rr < bi && cc < bk ) ? A[( I + rr ) + (int64_t)( K + cc ) * n] : (T)0;
                sX[cc][rr] = ( rr < bk && cc < bj ) ? LI[( K + rr ) + (int64_t)( J + cc ) * n] : (T)0;
            }
            __syncthreads();
#pragma unroll 8
            for ( int l = 0; l < PXD_NB; l++ ) {
                T a = sL[l][r];
#pragma unroll
                for ( int q = 0; q < 4; q++ ) acc[q] += a * sX[c0 + 8 * q][l];
            }
            __syncthreads();
        }
#pragma unroll
        for ( int q = 0; q < 4; q++ ) sT[c0 + 8 * q][r] = acc[q];
        // Linv[I,I] into sL
        for ( int idx = tid; idx < PXD_NB * PXD_NB; idx += PXD_THREADS ) {
            int rr = idx % PXD_NB, cc = idx / PXD_NB;
            sL[cc][rr] = ( rr < bi && cc < bi ) ? LI[( I + rr ) + (int64_t)( I + cc ) * n] : (T)0;
        }
        __syncthreads();
#pragma unroll
        for ( int q = 0; q < 4; q++ ) {
            const int cc = c0 + 8 * q;
            T         s  = (T)0;
            for ( int l = 0; l <= r; l++ ) {
                s += sL[l][r] * sT[cc][l];
            }
            if ( r < bi && cc < bj ) {
                LI[( I + r ) + (int64_t)( J + cc ) * n] = -s;
            }
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// Off-diagonal TRSM on the matrix cores: X = B * Linv^T for a tile of 64
// off-diagonal rows and all n columns of the cblk, in place (the workgroup
// owns complete rows, so every read of B precedes every write of X).
// Wave w computes rows 16w..16w+15 against every 16-column tile; column tiles
// that only meet the zero upper part of Linv are skipped.
#define PXD_TRSM_NT 22   /* up to 352 columns */
template <typename T>
__global__ void __launch_bounds__( PXD_THREADS )
k_trsm_mfma( const row_task *__restrict__ tasks, const dcblk *__restrict__ cb,
             const dblok *__restrict__ bk, T *__restrict__ L, const T *__restrict__ Linv )
{
    typedef mfma16<T>          M;
    typedef typename M::acc_t  acc_t;
    constexpr int BK = 16;
    __shared__ T       sB[BK][64 + 1];
    __shared__ T       sLi[BK][PXD_TRSM_NT * 16 + 1];
    __shared__ int64_t sBase[64];
    __shared__ int32_t sLd[64];

    const row_task t    = tasks[blockIdx.x];
    const dcblk    c    = cb[t.cblk];
    const int32_t  n    = c.lcolnum - c.fcolnum + 1;
    const int32_t  lb   = cb[t.cblk + 1].fblok;
    const int      tid  = threadIdx.x;
    const int      lane = tid & 63, w = tid >> 6;
    const T       *LI   = Linv + c.dinvoff;
    const int      nt   = ( n + 15 ) >> 4;

    if ( tid < 64 ) {
        int32_t s = t.r0 + tid;
        if ( s < c.stride ) {
            int32_t b  = find_blok_by_row( bk, c.fblok, lb, s );
            dblok   bb = bk[b];
            sBase[tid] = c.coefoff + (int64_t)bb.rowoff * n + ( s - bb.rowoff );
            sLd[tid]   = bb.lrownum - bb.frownum + 1;
        }
        else {
            sBase[tid] = -1;
            sLd[tid]   = 0;
        }
    }
    __syncthreads();

    acc_t acc[PXD_TRSM_NT];
#pragma unroll
    for ( int j = 0; j < PXD_TRSM_NT; j++ ) acc[j] = M::zero();

    const int     lr    = tid & 63;          // row this thread loads
    const int64_t lbase = sBase[lr];
    const int32_t lld   = sLd[lr];

    for ( int32_t l0 = 0; l0 < n; l0 += BK ) {
        for ( int kk = tid >> 6; kk < BK; kk += 4 ) {
            int32_t l = l0 + kk;
            sB[kk][lr] = ( lbase >= 0 && l < n ) ? L[lbase + (int64_t)l * lld] : (T)0;
        }
        for ( int idx = tid; idx < BK * PXD_TRSM_NT * 16; idx += PXD_THREADS ) {
            int cc = idx % ( PXD_TRSM_NT * 16 ), kk = idx / ( PXD_TRSM_NT * 16 );
            int32_t l = l0 + kk;
            sLi[kk][cc] = ( cc < n && l < n ) ? LI[cc + (int64_t)l * n] : (T)0;
        }
        __syncthreads();
#pragma unroll
        for ( int k4 = 0; k4 < BK; k4 += 4 ) {
            const T a = sB[k4 + ( lane >> 4 )][16 * w + ( lane & 15 )];
#pragma unroll
            for ( int j = 0; j < PXD_TRSM_NT; j++ ) {
                if ( j < nt && 16 * j + 15 >= l0 + k4 ) {
                    const T b = sLi[k4 + ( lane >> 4 )][16 * j + ( lane & 15 )];
                    acc[j] = M::mma( a, b, acc[j] );
                }
            }
        }
        __syncthreads();
    }
#pragma unroll
    for ( int j = 0; j < PXD_TRSM_NT; j++ ) {
        if ( j < nt ) {
            const int cc = 16 * j + ( lane & 15 );
#pragma unroll
            for ( int v = 0; v < 4; v++ ) {
                const int r = 16 * w + M::row( lane, v );
                if ( sBase[r] >= 0 && cc < n ) {
                    L[sBase[r] + (int64_t)cc * sLd[r]] = acc[j][v];
                }
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Supernodal update on the matrix cores: C[f(i)] -= A_i * A_b^T for one
// 64 x 64 output tile (four waves of 32 x 32), fused row remap into the
// facing cblk in the epilogue (atomic adds when another source of the same
// level may hit the same columns).
template <typename T>
__global__ void __launch_bounds__( PXD_THREADS )
k_update_mfma( const upd_task *__restrict__ tasks, const dcblk *__restrict__ cb,
               const dblok *__restrict__ bk, T *__restrict__ L )
{
    typedef mfma16<T>         M;
    typedef typename M::acc_t acc_t;
    constexpr int BK = 16;
    __shared__ T       sA[BK][64 + 4];
    __shared__ T       sB[BK][64 + 4];
    __shared__ int64_t sSrc[64];
    __shared__ int32_t sSrcLd[64];
    __shared__ int64_t sDst[64];
    __shared__ int32_t sDstLd[64];

    const upd_task t    = tasks[blockIdx.x];
    const int      tid  = threadIdx.x;
    const int      lane = tid & 63, w = tid >> 6;
    const dcblk    c    = cb[t.cblk];
    const int32_t  K    = c.lcolnum - c.fcolnum + 1;
    const int32_t  lb   = cb[t.cblk + 1].fblok;
    const dblok    bb   = bk[t.blok];
    const int32_t  N    = bb.lrownum - bb.frownum + 1;
    const int32_t  Mr   = c.stride - bb.rowoff;
    const int32_t  F    = bb.fcblknm;
    const dcblk    cf   = cb[F];
    const int32_t  Fn   = cf.lcolnum - cf.fcolnum + 1;
    const int32_t  flb  = cb[F + 1].fblok;
    const int32_t  colF0 = bb.frownum - cf.fcolnum + t.tn * 64;
    const int64_t  bbase = c.coefoff + (int64_t)bb.rowoff * K + t.tn * 64;
    if ( bb.frownum < cf.fcolnum || bb.lrownum > cf.lcolnum ) {
        return;   // inconsistent structure: the host checks prevent this
    }

    if ( tid < 64 ) {
        int32_t r = t.tm * 64 + tid;
        if ( r < Mr ) {
            int32_t s  = bb.rowoff + r;
            int32_t bi = find_blok_by_row( bk, t.blok, lb, s );
            dblok   ib = bk[bi];
            int32_t g  = ib.frownum + ( s - ib.rowoff );
            sSrc[tid]   = c.coefoff + (int64_t)ib.rowoff * K + ( s - ib.rowoff );
            sSrcLd[tid] = ib.lrownum - ib.frownum + 1;
            int32_t fi  = find_blok_by_grow( bk, cf.fblok, flb, g );
            dblok   fb  = bk[fi];
            bool    ok  = ( g >= fb.frownum && g <= fb.lrownum );
            sDst[tid]   = ok ? cf.coefoff + (int64_t)fb.rowoff * Fn + ( g - fb.frownum ) : -1;
            sDstLd[tid] = fb.lrownum - fb.frownum + 1;
        }
        else {
            sSrc[tid] = -1;
            sDst[tid] = -1;
        }
    }
    __syncthreads();

    const int     lr     = tid & 63;
    const int64_t src    = sSrc[lr];
    const int32_t srcld  = sSrcLd[lr];
    const int32_t nvalid = min( 64, N - t.tn * 64 );

    acc_t acc[2][2];
#pragma unroll
    for ( int i = 0; i < 2; i++ )
#pragma unroll
        for ( int j = 0; j < 2; j++ ) acc[i][j] = M::zero();
    const int wr = 32 * ( w >> 1 ), wc = 32 * ( w & 1 );

    for ( int32_t k0 = 0; k0 < K; k0 += BK ) {
#pragma unroll
        for ( int kk = tid >> 6; kk < BK; kk += 4 ) {
            int32_t kg = k0 + kk;
            sA[kk][lr] = ( kg < K && src >= 0 ) ? L[src + (int64_t)kg * srcld] : (T)0;
            sB[kk][lr] = ( kg < K && lr < nvalid ) ? L[bbase + lr + (int64_t)kg * N] : (T)0;
        }
        __syncthreads();
#pragma unroll
        for ( int k4 = 0; k4 < BK; k4 += 4 ) {
            const int kr = k4 + ( lane >> 4 );
            T a0 = sA[kr][wr + ( lane & 15 )];
            T a1 = sA[kr][wr + 16 + ( lane & 15 )];
            T b0 = sB[kr][wc + ( lane & 15 )];
            T b1 = sB[kr][wc + 16 + ( lane & 15 )];
            acc[0][0] = M::mma( a0, b0, acc[0][0] );
            acc[0][1] = M::mma( a0, b1, acc[0][1] );
            acc[1][0] = M::mma( a1, b0, acc[1][0] );
            acc[1][1] = M::mma( a1, b1, acc[1][1] );
        }
        __syncthreads();
    }

    const bool atomic = ( t.flags & 1 );
#pragma unroll
    for ( int i = 0; i < 2; i++ ) {
#pragma unroll
        for ( int j = 0; j < 2; j++ ) {
            const int cc = wc + 16 * j + ( lane & 15 );
#pragma unroll
            for ( int v = 0; v < 4; v++ ) {
                const int r = wr + 16 * i + M::row( lane, v );
                if ( sDst[r] >= 0 && cc < nvalid ) {
                    T *dst = L + sDst[r] + (int64_t)( colF0 + cc ) * sDstLd[r];
                    if ( atomic ) {
                        px_atomic_add( dst, -acc[i][j][v] );
                    }
                    else {
                        *dst -= acc[i][j][v];
                    }
                }
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Diagonal solves with the explicit inverse: x_k <- Linv x_k (forward) or
// Linv^T x_k (backward) for a list of cblks, nrhs <= PXD_MAXRHS.
template <typename T>
__global__ void __launch_bounds__( PXD_THREADS )
k_solve_diag_inv( const int32_t *__restrict__ list, const dcblk *__restrict__ cb,
                  const T *__restrict__ Linv, T *__restrict__ X, int64_t ldx, int nrhs, int trans )
{
    extern __shared__ unsigned char smem_raw[];
    T *sb = reinterpret_cast<T *>( smem_raw );   // [n][nrhs]

    const int32_t k   = list[blockIdx.x];
    const dcblk   c   = cb[k];
    const int32_t n   = c.lcolnum - c.fcolnum + 1;
    const int     tid = threadIdx.x;
    const T      *LI  = Linv + c.dinvoff;
    T            *x   = X + c.fcolnum;

    for ( int idx = tid; idx < n * nrhs; idx += PXD_THREADS ) {
        int r = idx % n, q = idx / n;
        sb[r * nrhs + q] = x[r + q * ldx];
    }
    __syncthreads();
    if ( !trans ) {
        // x[r] = sum_{l<=r} Linv[r][l] b[l]; consecutive threads read consecutive rows
        for ( int32_t r = tid; r < n; r += PXD_THREADS ) {
            T acc[PXD_MAXRHS];
#pragma unroll
            for ( int q = 0; q < PXD_MAXRHS; q++ ) acc[q] = (T)0;
#pragma unroll 4
            for ( int32_t l = 0; l <= r; l++ ) {
                T a = LI[r + (int64_t)l * n];
#pragma unroll
                for ( int q = 0; q < PXD_MAXRHS; q++ ) {
                    if ( q < nrhs ) acc[q] += a * sb[l * nrhs + q];
                }
            }
            for ( int q = 0; q < nrhs; q++ ) x[r + q * ldx] = acc[q];
        }
    }
    else {
        // x[cc] = sum_{l>=cc} Linv[l][cc] b[l]; one wave per column, lanes along the column
        const int lane = tid & 63, w = tid >> 6;
        for ( int32_t cc = w; cc < n; cc += PXD_THREADS / 64 ) {
            T acc[PXD_MAXRHS];
#pragma unroll
            for ( int q = 0; q < PXD_MAXRHS; q++ ) acc[q] = (T)0;
            for ( int32_t l = cc + lane; l < n; l += 64 ) {
                T a = LI[l + (int64_t)cc * n];
#pragma unroll
                for ( int q = 0; q < PXD_MAXRHS; q++ ) {
                    if ( q < nrhs ) acc[q] += a * sb[l * nrhs + q];
                }
            }
            for ( int q = 0; q < nrhs; q++ ) {
                T s = wave_sum( acc[q] );
                if ( lane == 0 ) x[cc + q * ldx] = s;
            }
        }
    }
}

// Backward gather of a 64-row tile: x_k -= A(rows,:)^T X[grow].  The tile is
// staged through LDS by 64-column chunks (coalesced loads), four threads per
// column reduce 16 rows each, one atomic per column and RHS.
template <typename T>
__global__ void __launch_bounds__( PXD_THREADS )
k_solve_bwd_gather2( const row_task *__restrict__ tasks, const dcblk *__restrict__ cb,
                     const dblok *__restrict__ bk, const T *__restrict__ L, T *__restrict__ X,
                     int64_t ldx, int nrhs )
{
    __shared__ T       sA[64][64 + 1];           // sA[col][row]
    __shared__ T       sx[64][PXD_MAXRHS];
    __shared__ T       sred[4][64][PXD_MAXRHS + 1];
    __shared__ int64_t sBase[64];
    __shared__ int32_t sLd[64];

    const row_task t   = tasks[blockIdx.x];
    const dcblk    c   = cb[t.cblk];
    const int32_t  n   = c.lcolnum - c.fcolnum + 1;
    const int32_t  lb  = cb[t.cblk + 1].fblok;
    const int      tid = threadIdx.x;

    if ( tid < 64 ) {
        int32_t s = t.r0 + tid;
        if ( s < c.stride ) {
            int32_t b  = find_blok_by_row( bk, c.fblok, lb, s );
            dblok   bb = bk[b];
            sBase[tid] = c.coefoff + (int64_t)bb.rowoff * n + ( s - bb.rowoff );
            sLd[tid]   = bb.lrownum - bb.frownum + 1;
            int32_t g  = bb.frownum + ( s - bb.rowoff );
            for ( int q = 0; q < PXD_MAXRHS; q++ ) sx[tid][q] = ( q < nrhs ) ? X[g + q * ldx] : (T)0;
        }
        else {
            sBase[tid] = -1;
            sLd[tid]   = 0;
            for ( int q = 0; q < PXD_MAXRHS; q++ ) sx[tid][q] = (T)0;
        }
    }
    __syncthreads();
    const int col = tid & 63, part = tid >> 6;
    for ( int32_t c0 = 0; c0 < n; c0 += 64 ) {
        for ( int idx = tid; idx < 64 * 64; idx += PXD_THREADS ) {
            int r = idx & 63, cc = idx >> 6;
            sA[cc][r] = ( sBase[r] >= 0 && c0 + cc < n ) ? L[sBase[r] + (int64_t)( c0 + cc ) * sLd[r]] : (T)0;
        }
        __syncthreads();
        T acc[PXD_MAXRHS];
#pragma unroll
        for ( int q = 0; q < PXD_MAXRHS; q++ ) acc[q] = (T)0;
#pragma unroll
        for ( int r = 16 * part; r < 16 * part + 16; r++ ) {
            T a = sA[col][r];
#pragma unroll
            for ( int q = 0; q < PXD_MAXRHS; q++ ) acc[q] += a * sx[r][q];
        }
#pragma unroll
        for ( int q = 0; q < PXD_MAXRHS; q++ ) sred[part][col][q] = acc[q];
        __syncthreads();
        if ( part == 0 && c0 + col < n ) {
            for ( int q = 0; q < nrhs; q++ ) {
                T v = sred[0][col][q] + sred[1][col][q] + sred[2][col][q] + sred[3][col][q];
                px_atomic_add( X + c.fcolnum + c0 + col + q * ldx, -v );
            }
        }
        __syncthreads();
    }
}
