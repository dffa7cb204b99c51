// kernels_fact.hpp -- factorization kernels of the MI355X engine.
//
// Storage (solver matrix, 2D tile layout): cblk k with n = lcol-fcol+1
// columns owns the rows of its bloks stacked in order ("stacked rows", the
// diagonal blok first).  Blok b of cblk k is the column-major rows(b) x n tile
// at L + coefoff[k] + rowoff[b]*n, leading dimension rows(b).
//
// Numerics follow the reference kernels:
//  - panel: blocked right-looking Cholesky with static pivoting
//    (kernels/core_zpotrfsp.c:67-189: |a_kk| < tau -> a_kk = tau, counted);
//  - off-diagonal TRSM: L_off <- L_off * L_kk^{-T} (core_ztrsmsp.c:148-183);
//  - update: for source blok b and every blok i >= b of the same cblk,
//    C[f(i)] -= A_i * A_b^T into the facing cblk (core_zgemmsp.c:401-472).
// The 32x32 diagonal sub-blocks of every L_kk are also inverted (dinv) so the
// TRSM and the solve sweeps become products instead of dependent recurrences.
#pragma once
#include "px_device.hpp"

struct dcblk {
    int32_t fcolnum;
    int32_t lcolnum;
    int32_t fblok;
    int32_t stride;
    int64_t coefoff;
    int64_t dinvoff;   // offset of the 32x32 inverses of the diagonal sub-blocks
};

struct dblok {
    int32_t frownum;
    int32_t lrownum;
    int32_t fcblknm;
    int32_t rowoff;    // stacked row offset inside its cblk (coefind / n)
};

// update task: source cblk, source blok b, tile (tm, tn), flags (bit 0: atomic)
struct upd_task {
    int32_t cblk;
    int32_t blok;
    int32_t tm;
    int32_t tn;
    int32_t flags;
    int32_t pad;
};

// row task: cblk and first stacked row of a row tile
struct row_task {
    int32_t cblk;
    int32_t r0;
};

#define PXD_NB 32          // sub-block size of panel / inverses
#define PXD_TRSM_NT 22     // 16-column tiles of the widest cblk (352 columns)
#define PXD_THREADS 256

// Largest i in [0, n) with a[i] <= key (a ascending, a[0] <= key).
__device__ inline int32_t
lds_search_le( const int32_t *a, int32_t n, int32_t key )
{
    int32_t lo = 0, hi = n - 1;
    while ( lo < hi ) {
        const int32_t mid = ( lo + hi + 1 ) >> 1;
        if ( a[mid] <= key ) {
            lo = mid;
        }
        else {
            hi = mid - 1;
        }
    }
    return lo;
}

// Blok of cblk k (bloks [fb, lb)) that holds stacked row s.
__device__ inline int32_t
find_blok_by_row( const dblok *bk, int32_t fb, int32_t lb, int32_t s )
{
    int32_t lo = fb, hi = lb - 1;
    while ( lo < hi ) {
        int32_t mid = ( lo + hi + 1 ) >> 1;
        if ( bk[mid].rowoff <= s ) {
            lo = mid;
        }
        else {
            hi = mid - 1;
        }
    }
    return lo;
}

// Blok of cblk F (bloks [fb, lb)) whose row range holds global row g.
__device__ inline int32_t
find_blok_by_grow( const dblok *bk, int32_t fb, int32_t lb, int32_t g )
{
    int32_t lo = fb, hi = lb - 1;
    while ( lo < hi ) {
        int32_t mid = ( lo + hi + 1 ) >> 1;
        if ( bk[mid].frownum <= g ) {
            lo = mid;
        }
        else {
            hi = mid - 1;
        }
    }
    return lo;
}

// Row offsets of the bloks [fb, lb) of one cblk staged in LDS by one
// parallel load, for the row searches of a workgroup (lists longer than
// PXD_ROWKEY_CAP are searched in global memory).
#define PXD_ROWKEY_CAP 512
__device__ inline void
stage_row_keys( int32_t *keys, const dblok *__restrict__ bk, int32_t fb, int32_t lb )
{
    if ( lb - fb <= PXD_ROWKEY_CAP ) {
        for ( int32_t i = threadIdx.x; i < lb - fb; i += blockDim.x ) keys[i] = bk[fb + i].rowoff;
    }
}
__device__ inline int32_t
search_row_keys( const int32_t *keys, const dblok *__restrict__ bk, int32_t fb, int32_t lb, int32_t s )
{
    return ( lb - fb <= PXD_ROWKEY_CAP ) ? fb + lds_search_le( keys, lb - fb, s ) : find_blok_by_row( bk, fb, lb, s );
}

// ---------------------------------------------------------------------------
// Scatter of the permuted matrix into the zeroed coeftab.  One thread per
// column j of the CSC (colptr/rowtab/val); entries with row >= j go to the
// tile of column j's cblk.  Called with the transposed CSC (rows of A) to
// fill U^T with the strictly upper part for LU (strict = 1).
template <typename T, typename TI>
__global__ void
k_fill( int64_t n, const int64_t *__restrict__ colptr, const int64_t *__restrict__ rowtab,
        const TI *__restrict__ val, const int32_t *__restrict__ col2cblk,
        const dcblk *__restrict__ cb, const dblok *__restrict__ bk, T *__restrict__ L,
        int strict, double ovf_limit, int *__restrict__ ovf )
{
    int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if ( j >= n ) {
        return;
    }
    int32_t k    = col2cblk[j];
    dcblk   c    = cb[k];
    int32_t ncol = c.lcolnum - c.fcolnum + 1;
    int32_t jj   = (int32_t)j - c.fcolnum;
    int32_t b    = c.fblok;
    int32_t lb   = cb[k + 1].fblok;
    for ( int64_t e = colptr[j]; e < colptr[j + 1]; e++ ) {
        int64_t i = rowtab[e];
        if ( i < j || ( strict && i == j ) ) {
            continue;
        }
        while ( b < lb && bk[b].lrownum < i ) {
            b++;
        }
        if ( b >= lb || bk[b].frownum > i ) {
            continue;
        }
        dblok   bb   = bk[b];
        int32_t rows = bb.lrownum - bb.frownum + 1;
        int64_t off  = c.coefoff + (int64_t)bb.rowoff * ncol + ( i - bb.frownum ) + (int64_t)jj * rows;
        TI      v    = val[e];
        if ( px_abs( v ) > ovf_limit ) {
            atomicAdd( ovf, 1 );
        }
        L[off] = px_from( v, L );
    }
}

// ---------------------------------------------------------------------------
// Cholesky panel of the diagonal tiles of a list of cblks (one workgroup per
// cblk): blocked by PXD_NB columns, unblocked step with static pivoting, the
// inverse of every PXD_NB x PXD_NB diagonal sub-block written to dinv.
template <typename T>
__global__ void __launch_bounds__( PXD_THREADS )
k_potrf( const int32_t *__restrict__ list, const dcblk *__restrict__ cb, T *__restrict__ L,
         T *__restrict__ dinv, double thresh, int *__restrict__ npiv )
{
    __shared__ T sA[PXD_NB][PXD_NB + 1];   // sA[col][row]
    __shared__ T sI[PXD_NB][PXD_NB + 1];   // inverse, sI[col][row]
    __shared__ T sPi[PXD_NB][64 + 1];
    __shared__ T sPj[PXD_NB][64 + 1];

    const int32_t k   = list[blockIdx.x];
    const dcblk   c   = cb[k];
    const int32_t n   = c.lcolnum - c.fcolnum + 1;
    const int     tid = threadIdx.x;
    T *A  = L + c.coefoff;            // n x n, ld n
    T *DI = dinv + c.dinvoff;
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

        // unblocked Cholesky of the bs x bs block
        for ( int cc = 0; cc < bs; cc++ ) {
            if ( tid == 0 ) {
                T d = sA[cc][cc];
                if ( px_abs( d ) < tau ) {
                    d = tau;
                    piv++;
                }
                sA[cc][cc] = px_sqrt( d );
            }
            __syncthreads();
            if ( tid > cc && tid < bs ) {
                sA[cc][tid] = sA[cc][tid] / sA[cc][cc];
            }
            __syncthreads();
            const int m = bs - cc - 1;
            for ( int idx = tid; idx < m * m; idx += PXD_THREADS ) {
                int rr = cc + 1 + idx % m, c2 = cc + 1 + idx / m;
                if ( rr >= c2 ) {
                    sA[c2][rr] -= sA[cc][rr] * sA[cc][c2];
                }
            }
            __syncthreads();
        }

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

        for ( int idx = tid; idx < PXD_NB * PXD_NB; idx += PXD_THREADS ) {
            int r = idx % PXD_NB, cc = idx / PXD_NB;
            if ( r < bs && cc < bs && r >= cc ) {
                A[( kb + r ) + (int64_t)( kb + cc ) * n] = sA[cc][r];
            }
            DI[(int64_t)( kb / PXD_NB ) * PXD_NB * PXD_NB + r + cc * PXD_NB] =
                ( r < bs && cc < bs ) ? sI[cc][r] : ( ( r == cc ) ? (T)1 : (T)0 );
        }

        // rows below the block inside the tile: A21 <- A21 * L11^{-T}
        for ( int32_t r = kb + bs + tid; r < n; r += PXD_THREADS ) {
            T v[PXD_NB];
#pragma unroll
            for ( int l = 0; l < PXD_NB; l++ ) {
                v[l] = ( l < bs ) ? A[r + (int64_t)( kb + l ) * n] : (T)0;
            }
#pragma unroll
            for ( int cc = 0; cc < PXD_NB; cc++ ) {
                if ( cc < bs ) {
                    T s = (T)0;
                    for ( int l = 0; l <= cc; l++ ) {
                        s += v[l] * sI[l][cc];
                    }
                    A[r + (int64_t)( kb + cc ) * n] = s;
                }
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
// Off-diagonal TRSM of a list of RT-row tiles: X = B * L_kk^{-T}, done by
// PXD_NB column blocks with the inverted diagonal sub-blocks.  The tile rows
// (RT x n) live in dynamic LDS for the whole solve.
template <typename T, int RT>
__global__ void __launch_bounds__( PXD_THREADS )
k_trsm( const row_task *__restrict__ tasks, const dcblk *__restrict__ cb,
        const dblok *__restrict__ bk, T *__restrict__ L, const T *__restrict__ dinv )
{
    extern __shared__ unsigned char smem_raw[];
    T *sX = reinterpret_cast<T *>( smem_raw );          // [n][RT+1]
    __shared__ T       sT[PXD_NB][RT + 1];
    __shared__ int64_t sBase[RT];
    __shared__ int32_t sLd[RT];

    const row_task t   = tasks[blockIdx.x];
    const dcblk    c   = cb[t.cblk];
    const int32_t  n   = c.lcolnum - c.fcolnum + 1;
    const int      tid = threadIdx.x;
    const int32_t  lb  = cb[t.cblk + 1].fblok;
    const T       *Lkk = L + c.coefoff;
    const T       *DI  = dinv + c.dinvoff;
    const int      LD  = RT + 1;

    if ( tid < RT ) {
        int32_t s = t.r0 + tid;
        if ( s < c.stride ) {
            int32_t b = find_blok_by_row( bk, c.fblok, lb, s );
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
    for ( int idx = tid; idx < RT * n; idx += PXD_THREADS ) {
        int i = idx % RT, cc = idx / RT;
        sX[cc * LD + i] = ( sBase[i] >= 0 ) ? L[sBase[i] + (int64_t)cc * sLd[i]] : (T)0;
    }
    __syncthreads();

    for ( int32_t J = 0; J < n; J += PXD_NB ) {
        const int32_t bs = min( PXD_NB, n - J );
        for ( int idx = tid; idx < RT * bs; idx += PXD_THREADS ) {
            int i = idx % RT, cc = idx / RT;
            T s = sX[( J + cc ) * LD + i];
            for ( int32_t l = 0; l < J; l++ ) {
                s -= sX[l * LD + i] * Lkk[( J + cc ) + (int64_t)l * n];
            }
            sT[cc][i] = s;
        }
        __syncthreads();
        const T *Dj = DI + (int64_t)( J / PXD_NB ) * PXD_NB * PXD_NB;
        for ( int idx = tid; idx < RT * bs; idx += PXD_THREADS ) {
            int i = idx % RT, cc = idx / RT;
            T s = (T)0;
            for ( int l = 0; l <= cc; l++ ) {
                s += sT[l][i] * Dj[cc + l * PXD_NB];
            }
            sX[( J + cc ) * LD + i] = s;
        }
        __syncthreads();
    }
    for ( int idx = tid; idx < RT * n; idx += PXD_THREADS ) {
        int i = idx % RT, cc = idx / RT;
        if ( sBase[i] >= 0 ) {
            L[sBase[i] + (int64_t)cc * sLd[i]] = sX[cc * LD + i];
        }
    }
}

// ---------------------------------------------------------------------------
// Supernodal update C[f(i)] -= A_i * A_b^T for one 64x64 output tile, with
// the row remap into the facing cblk fused into the epilogue.  A rows are the
// stacked rows of cblk k from blok b down, B rows are the rows of blok b.
template <typename T>
__global__ void __launch_bounds__( PXD_THREADS )
k_update( const upd_task *__restrict__ tasks, const dcblk *__restrict__ cb,
          const dblok *__restrict__ bk, T *__restrict__ L )
{
    __shared__ T       sA[16][64 + 1];
    __shared__ T       sB[16][64 + 1];
    __shared__ int64_t sSrc[64];
    __shared__ int32_t sSrcLd[64];
    __shared__ int64_t sDst[64];
    __shared__ int32_t sDstLd[64];

    const upd_task t   = tasks[blockIdx.x];
    const int      tid = threadIdx.x;
    const dcblk    c   = cb[t.cblk];
    const int32_t  K   = c.lcolnum - c.fcolnum + 1;
    const int32_t  lb  = cb[t.cblk + 1].fblok;
    const dblok    bb  = bk[t.blok];
    const int32_t  N   = bb.lrownum - bb.frownum + 1;
    const int32_t  M   = c.stride - bb.rowoff;
    const int32_t  F   = bb.fcblknm;
    const dcblk    cf  = cb[F];
    const int32_t  Fn  = cf.lcolnum - cf.fcolnum + 1;
    const int32_t  flb = cb[F + 1].fblok;
    const int32_t  colF0 = bb.frownum - cf.fcolnum + t.tn * 64;
    const int64_t  bbase = c.coefoff + (int64_t)bb.rowoff * K + t.tn * 64;
    if ( bb.frownum < cf.fcolnum || bb.lrownum > cf.lcolnum ) {
        return;   // inconsistent structure: the host checks prevent this
    }

    if ( tid < 64 ) {
        int32_t r = t.tm * 64 + tid;
        if ( r < M ) {
            int32_t s  = bb.rowoff + r;
            int32_t bi = find_blok_by_row( bk, t.blok, lb, s );
            dblok   ib = bk[bi];
            int32_t g  = ib.frownum + ( s - ib.rowoff );
            sSrc[tid]   = c.coefoff + (int64_t)ib.rowoff * K + ( s - ib.rowoff );
            sSrcLd[tid] = ib.lrownum - ib.frownum + 1;
            int32_t fi  = find_blok_by_grow( bk, cf.fblok, flb, g );
            dblok   fb  = bk[fi];
            // a row outside the facing blok would mean an inconsistent structure
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

    const int tx = tid % 16, ty = tid / 16;
    T acc[4][4];
#pragma unroll
    for ( int a = 0; a < 4; a++ )
#pragma unroll
        for ( int b = 0; b < 4; b++ ) acc[a][b] = (T)0;

    const int32_t nvalid = min( 64, N - t.tn * 64 );
    for ( int32_t k0 = 0; k0 < K; k0 += 16 ) {
        for ( int idx = tid; idx < 64 * 16; idx += PXD_THREADS ) {
            int r = idx % 64, kk = idx / 64;
            int32_t kg = k0 + kk;
            sA[kk][r] = ( kg < K && sSrc[r] >= 0 ) ? L[sSrc[r] + (int64_t)kg * sSrcLd[r]] : (T)0;
            sB[kk][r] = ( kg < K && r < nvalid ) ? L[bbase + r + (int64_t)kg * N] : (T)0;
        }
        __syncthreads();
#pragma unroll
        for ( int kk = 0; kk < 16; kk++ ) {
            T a[4], b[4];
#pragma unroll
            for ( int q = 0; q < 4; q++ ) {
                a[q] = sA[kk][tx + 16 * q];
                b[q] = sB[kk][ty + 16 * q];
            }
#pragma unroll
            for ( int p = 0; p < 4; p++ )
#pragma unroll
                for ( int q = 0; q < 4; q++ ) acc[p][q] += a[p] * b[q];
        }
        __syncthreads();
    }

    const bool atomic = ( t.flags & 1 );
#pragma unroll
    for ( int p = 0; p < 4; p++ ) {
        int r = tx + 16 * p;
        if ( sDst[r] < 0 ) {
            continue;
        }
#pragma unroll
        for ( int q = 0; q < 4; q++ ) {
            int cc = ty + 16 * q;
            if ( cc < nvalid ) {
                T *dst = L + sDst[r] + (int64_t)( colF0 + cc ) * sDstLd[r];
                if ( atomic ) {
                    px_atomic_add( dst, -acc[p][q] );
                }
                else {
                    *dst -= acc[p][q];
                }
            }
        }
    }
}
