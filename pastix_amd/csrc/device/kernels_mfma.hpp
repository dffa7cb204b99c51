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
#include "kernels_solve.hpp"
#include "mfma.hpp"
#include <type_traits>

// tasks of the block-column inversion: cblk and first column of the block
struct inv_task {
    int32_t cblk;
    int32_t J;
};

// Waves of a kernel on the critical path of the factorization (panels,
// inverses, TRSM and the updates that the next level waits for) raise their
// instruction-issue priority, so that they keep their pace on CUs that they
// share with the bulk update kernels of the other stream.
__device__ inline void
wave_priority( int prio )
{
    if ( prio ) {
        __builtin_amdgcn_s_setprio( 3 );
    }
}

// Row i >= column j of entry idx of a row-major packed lower triangle.
__device__ inline void
tri_index( int idx, int &i, int &j )
{
    i = (int)( ( sqrtf( 8.f * (float)idx + 1.f ) - 1.f ) * 0.5f );
    if ( ( i + 1 ) * ( i + 2 ) / 2 <= idx ) i++;
    if ( i * ( i + 1 ) / 2 > idx ) i--;
    j = idx - i * ( i + 1 ) / 2;
}

// Unblocked eliminations of a 32 x 32 diagonal block in LDS (s[col][row],
// zero beyond bs rows and columns), one barrier per elimination step.  Each
// of the 256 threads keeps four entries of the block and of each inverse in
// registers: row r = tid & 31, columns g, g + 8, g + 16, g + 24 (g = tid >>
// 5).  The column and rows that a step reads go through a double-buffered
// exchange area x in LDS (4 rows of 32 for the symmetric forms, 8 for LU),
// written by their owners before the step's barrier.  On return the block and the inverses are back in LDS,
// sD holds the pivots, and the result is the number of replaced pivots
// (counted by thread 0 only; the other threads return 0).
static_assert( PXD_THREADS == 8 * PXD_NB, "the elimination maps 256 threads on a 32 x 32 block" );

template <typename T>
__device__ inline T
pick4( T v0, T v1, T v2, T v3, int u )
{
    return ( u == 0 ) ? v0 : ( u == 1 ) ? v1 : ( u == 2 ) ? v2 : v3;
}

// Complex double: each LDS read of a step as it is needed (the read-ahead
// form below would not fit the register file of the panel kernels).
template <typename T, bool LDLT>
__device__ inline int
elim_sym32_single( T ( *sA )[PXD_NB + 1], T ( *sI )[PXD_NB + 1], T *sD, T ( *x )[PXD_NB], int bs,
            typename px_real<T>::type tau, int tid )
{
    typedef typename px_real<T>::type R;
    const int r = tid & ( PXD_NB - 1 ), g = tid >> 5;
    T a[4], w[4];
#pragma unroll
    for ( int u = 0; u < 4; u++ ) {
        a[u] = sA[g + 8 * u][r];
        w[u] = sI[g + 8 * u][r];
    }
    int piv = 0;
    for ( int cc = 0; cc < bs; cc++ ) {
        T *colA = x[2 * ( cc & 1 )], *rowW = x[2 * ( cc & 1 ) + 1];
        if ( g == ( cc & 7 ) ) {
            colA[r] = pick4( a[0], a[1], a[2], a[3], cc >> 3 );   // column cc of A
        }
        if ( r == cc ) {
#pragma unroll
            for ( int u = 0; u < 4; u++ ) {
                rowW[g + 8 * u] = w[u];      // row cc of W
            }
        }
        __syncthreads();
        T d = colA[cc];
        if ( px_abs( d ) < tau ) {
            d = ( LDLT && px_re( d ) < (R)0 ) ? (T)-tau : (T)tau;
            piv += ( tid == 0 );
        }
        const T rinv = (T)1 / d;
        if ( r > cc && r < bs ) {
            const T ar = colA[r];
#pragma unroll
            for ( int u = 0; u < 4; u++ ) {
                const int c = g + 8 * u;
                if ( c > cc && c <= r ) {
                    a[u] -= ar * colA[c] * rinv;
                }
                else if ( c <= cc ) {
                    w[u] -= ar * rinv * rowW[c];
                }
            }
        }
        if ( tid == 0 ) {
            sD[cc] = d;
        }
    }
#pragma unroll
    for ( int u = 0; u < 4; u++ ) {
        sA[g + 8 * u][r] = a[u];
        sI[g + 8 * u][r] = w[u];
    }
    __syncthreads();
    return piv;
}

// Symmetric: A = U D U^T, U unit lower with multipliers u_rc = a_rc / d_c,
// on the lower triangle of sA; the same row operations on W = I (sI) leave
// W = U^{-1}.  A pivot below tau in modulus becomes tau (Cholesky,
// core_zpotf2sp) or sign(re d) tau (LDL^T, core_zsytf2sp).
template <typename T, bool LDLT>
__device__ inline int
elim_sym32_ahead( T ( *sA )[PXD_NB + 1], T ( *sI )[PXD_NB + 1], T *sD, T ( *x )[PXD_NB], int bs,
            typename px_real<T>::type tau, int tid )
{
    typedef typename px_real<T>::type R;
    const int r = tid & ( PXD_NB - 1 ), g = tid >> 5;
    T a[4], w[4];
#pragma unroll
    for ( int u = 0; u < 4; u++ ) {
        a[u] = sA[g + 8 * u][r];
        w[u] = sI[g + 8 * u][r];
    }
    int piv = 0;
    for ( int cc = 0; cc < bs; cc++ ) {
        T *colA = x[2 * ( cc & 1 )], *rowW = x[2 * ( cc & 1 ) + 1];
        if ( g == ( cc & 7 ) ) {
            colA[r] = pick4( a[0], a[1], a[2], a[3], cc >> 3 );   // column cc of A
        }
        if ( r == cc ) {
#pragma unroll
            for ( int u = 0; u < 4; u++ ) {
                rowW[g + 8 * u] = w[u];      // row cc of W
            }
        }
        __syncthreads();
        // every read of the step at once, then the updates as selects: one
        // LDS round trip per step
        T d = colA[cc];
        const T ar = colA[r];
        T ca[4], rw[4];
#pragma unroll
        for ( int u = 0; u < 4; u++ ) {
            ca[u] = colA[g + 8 * u];
            rw[u] = rowW[g + 8 * u];
        }
        if ( px_abs( d ) < tau ) {
            d = ( LDLT && px_re( d ) < (R)0 ) ? (T)-tau : (T)tau;
            piv += ( tid == 0 );
        }
        const T    rinv = (T)1 / d;
        const bool act  = r > cc && r < bs;
#pragma unroll
        for ( int u = 0; u < 4; u++ ) {
            const int c  = g + 8 * u;
            const T   na = a[u] - ar * ca[u] * rinv;
            const T   nw = w[u] - ar * rinv * rw[u];
            a[u] = ( act && c > cc && c <= r ) ? na : a[u];
            w[u] = ( act && c <= cc ) ? nw : w[u];
        }
        if ( tid == 0 ) {
            sD[cc] = d;
        }
    }
#pragma unroll
    for ( int u = 0; u < 4; u++ ) {
        sA[g + 8 * u][r] = a[u];
        sI[g + 8 * u][r] = w[u];
    }
    __syncthreads();
    return piv;
}

template <typename T, bool LDLT>
__device__ inline int
elim_sym32( T ( *sA )[PXD_NB + 1], T ( *sI )[PXD_NB + 1], T *sD, T ( *x )[PXD_NB], int bs,
            typename px_real<T>::type tau, int tid )
{
    if constexpr ( sizeof( T ) > 8 ) {
        return elim_sym32_single<T, LDLT>( sA, sI, sD, x, bs, tau, tid );
    }
    else {
        return elim_sym32_ahead<T, LDLT>( sA, sI, sD, x, bs, tau, tid );
    }
}

// LU without pivoting: multipliers a_r,cc / d left unscaled in sA; the row
// operations of L on W_L = I (sIL) leave L11^{-1}, those of U11^T = U~ D on
// W_U = I (sIU) leave U~^{-1}.  A pivot below tau becomes sign(re d) tau.
template <typename T>
__device__ inline int
elim_lu32_ahead( T ( *sA )[PXD_NB + 1], T ( *sIL )[PXD_NB + 1], T ( *sIU )[PXD_NB + 1], T *sD, T ( *x )[PXD_NB],
           int bs, typename px_real<T>::type tau, int tid )
{
    typedef typename px_real<T>::type R;
    const int r = tid & ( PXD_NB - 1 ), g = tid >> 5;
    T a[4], wl[4], wu[4];
#pragma unroll
    for ( int u = 0; u < 4; u++ ) {
        a[u]  = sA[g + 8 * u][r];
        wl[u] = sIL[g + 8 * u][r];
        wu[u] = sIU[g + 8 * u][r];
    }
    int piv = 0;
    for ( int cc = 0; cc < bs; cc++ ) {
        T *colA = x[4 * ( cc & 1 )], *rowA = x[4 * ( cc & 1 ) + 1];
        T *rowL = x[4 * ( cc & 1 ) + 2], *rowU = x[4 * ( cc & 1 ) + 3];
        if ( g == ( cc & 7 ) ) {
            colA[r] = pick4( a[0], a[1], a[2], a[3], cc >> 3 );
        }
        if ( r == cc ) {
#pragma unroll
            for ( int u = 0; u < 4; u++ ) {
                rowA[g + 8 * u] = a[u];
                rowL[g + 8 * u] = wl[u];
                rowU[g + 8 * u] = wu[u];
            }
        }
        __syncthreads();
        T d = colA[cc];
        const T cr = colA[r], rr = rowA[r];
        T ra[4], rl[4], ru[4];
#pragma unroll
        for ( int u = 0; u < 4; u++ ) {
            ra[u] = rowA[g + 8 * u];
            rl[u] = rowL[g + 8 * u];
            ru[u] = rowU[g + 8 * u];
        }
        if ( px_abs( d ) < tau ) {
            d = ( px_re( d ) < (R)0 ) ? (T)-tau : (T)tau;
            piv += ( tid == 0 );
        }
        const T    rinv = (T)1 / d;
        const bool act  = r > cc && r < bs;
#pragma unroll
        for ( int u = 0; u < 4; u++ ) {
            const int c  = g + 8 * u;
            const T   na = a[u] - cr * rinv * ra[u];
            const T   nl = wl[u] - cr * rinv * rl[u];
            const T   nu = wu[u] - rr * rinv * ru[u];
            a[u]  = ( act && c > cc && c < bs ) ? na : a[u];
            wl[u] = ( act && c <= cc ) ? nl : wl[u];
            wu[u] = ( act && c <= cc ) ? nu : wu[u];
        }
        if ( tid == 0 ) {
            sD[cc] = d;
        }
    }
#pragma unroll
    for ( int u = 0; u < 4; u++ ) {
        sA[g + 8 * u][r]  = a[u];
        sIL[g + 8 * u][r] = wl[u];
        sIU[g + 8 * u][r] = wu[u];
    }
    __syncthreads();
    return piv;
}

// complex double: each LDS read as it is needed (register file)
template <typename T>
__device__ inline int
elim_lu32_lean( T ( *sA )[PXD_NB + 1], T ( *sIL )[PXD_NB + 1], T ( *sIU )[PXD_NB + 1], T *sD, T ( *x )[PXD_NB],
           int bs, typename px_real<T>::type tau, int tid )
{
    typedef typename px_real<T>::type R;
    const int r = tid & ( PXD_NB - 1 ), g = tid >> 5;
    T a[4], wl[4], wu[4];
#pragma unroll
    for ( int u = 0; u < 4; u++ ) {
        a[u]  = sA[g + 8 * u][r];
        wl[u] = sIL[g + 8 * u][r];
        wu[u] = sIU[g + 8 * u][r];
    }
    int piv = 0;
    for ( int cc = 0; cc < bs; cc++ ) {
        T *colA = x[4 * ( cc & 1 )], *rowA = x[4 * ( cc & 1 ) + 1];
        T *rowL = x[4 * ( cc & 1 ) + 2], *rowU = x[4 * ( cc & 1 ) + 3];
        if ( g == ( cc & 7 ) ) {
            colA[r] = pick4( a[0], a[1], a[2], a[3], cc >> 3 );
        }
        if ( r == cc ) {
#pragma unroll
            for ( int u = 0; u < 4; u++ ) {
                rowA[g + 8 * u] = a[u];
                rowL[g + 8 * u] = wl[u];
                rowU[g + 8 * u] = wu[u];
            }
        }
        __syncthreads();
        T d = colA[cc];
        if ( px_abs( d ) < tau ) {
            d = ( px_re( d ) < (R)0 ) ? (T)-tau : (T)tau;
            piv += ( tid == 0 );
        }
        const T rinv = (T)1 / d;
        if ( r > cc && r < bs ) {
            const T cr = colA[r], rr = rowA[r];
#pragma unroll
            for ( int u = 0; u < 4; u++ ) {
                const int c = g + 8 * u;
                if ( c > cc && c < bs ) {
                    a[u] -= cr * rinv * rowA[c];
                }
                else if ( c <= cc ) {
                    wl[u] -= cr * rinv * rowL[c];
                    wu[u] -= rr * rinv * rowU[c];
                }
            }
        }
        if ( tid == 0 ) {
            sD[cc] = d;
        }
    }
#pragma unroll
    for ( int u = 0; u < 4; u++ ) {
        sA[g + 8 * u][r]  = a[u];
        sIL[g + 8 * u][r] = wl[u];
        sIU[g + 8 * u][r] = wu[u];
    }
    __syncthreads();
    return piv;
}

template <typename T>
__device__ inline int
elim_lu32( T ( *sA )[PXD_NB + 1], T ( *sIL )[PXD_NB + 1], T ( *sIU )[PXD_NB + 1], T *sD, T ( *x )[PXD_NB],
           int bs, typename px_real<T>::type tau, int tid )
{
    if constexpr ( sizeof( T ) > 8 ) {
        return elim_lu32_lean<T>( sA, sIL, sIU, sD, x, bs, tau, tid );
    }
    else {
        return elim_lu32_ahead<T>( sA, sIL, sIU, sD, x, bs, tau, tid );
    }
}

// Per-workgroup timing for the kernel timing log: when stamps is not null,
// task i of the launch writes its start and end times (constant-rate wall
// clock ticks) to stamps[2i] and stamps[2i+1].
__device__ inline uint64_t
stamp_begin( const uint64_t *stamps )
{
    return stamps ? (uint64_t)wall_clock64() : 0;
}
__device__ inline void
stamp_end( uint64_t *stamps, int64_t i, uint64_t t0 )
{
    if ( stamps ) {
        __syncthreads();
        if ( threadIdx.x == 0 ) {
            stamps[2 * i]     = t0;
            stamps[2 * i + 1] = (uint64_t)wall_clock64();
        }
    }
}

// Keys of the row searches of an update tile, staged in LDS with one parallel
// load: stacked row offsets of the source bloks [b, lb) and first rows of the
// facing cblk's bloks [ffb, flb).  Lists longer than PXD_SEARCH_CAP are
// searched in global memory instead.
#define PXD_SEARCH_CAP 512
struct search_keys {
    int32_t row[PXD_SEARCH_CAP];
    int32_t frow[PXD_SEARCH_CAP];
};

__device__ inline void
stage_search_keys( search_keys &sk, const dblok *__restrict__ bk, int32_t b, int32_t lb, int32_t ffb,
                   int32_t flb, int tid )
{
    const int32_t ns = min( lb - b, PXD_SEARCH_CAP ), nf = min( flb - ffb, PXD_SEARCH_CAP );
    for ( int32_t i = tid; i < ns; i += PXD_THREADS ) sk.row[i] = bk[b + i].rowoff;
    for ( int32_t i = tid; i < nf; i += PXD_THREADS ) sk.frow[i] = bk[ffb + i].frownum;
}

// Source blok (in [b, lb)) holding stacked row s, and facing blok (in
// [ffb, flb)) holding global row g, from the staged keys.
__device__ inline int32_t
search_row( const search_keys &sk, const dblok *__restrict__ bk, int32_t b, int32_t lb, int32_t s )
{
    return ( lb - b <= PXD_SEARCH_CAP ) ? b + lds_search_le( sk.row, lb - b, s ) : find_blok_by_row( bk, b, lb, s );
}
__device__ inline int32_t
search_grow( const search_keys &sk, const dblok *__restrict__ bk, int32_t ffb, int32_t flb, int32_t g )
{
    return ( flb - ffb <= PXD_SEARCH_CAP ) ? ffb + lds_search_le( sk.frow, flb - ffb, g )
                                           : find_blok_by_grow( bk, ffb, flb, g );
}


// ---------------------------------------------------------------------------
// Panel: Cholesky of the diagonal tile with static pivoting, right looking by
// PXD_NB columns; the unblocked step updates with unscaled columns so that it
// needs one barrier per column.  The inverses of the PXD_NB x PXD_NB diagonal
// sub-blocks go to the diagonal blocks of Linv.  The first pcap rows of the
// current block's L21 are kept in dynamic LDS; rows beyond pcap (wide cblks in
// complex double, whose strip would not fit the CU's LDS) are read back from
// the copy just written to A.
template <typename T, bool LDLT>
__global__ void __launch_bounds__( PXD_THREADS )
k_potrf2( const int32_t *__restrict__ list, const dcblk *__restrict__ cb, T *__restrict__ L,
          T *__restrict__ Linv, double thresh, int *__restrict__ npiv,
          uint64_t *__restrict__ stamps, int prio, int32_t pcap )
{
    __shared__ T sA[PXD_NB][PXD_NB + 1];   // sA[col][row]
    __shared__ T sI[PXD_NB][PXD_NB + 1];   // inverse, sI[col][row]
    __shared__ T sD[PXD_NB];
    __shared__ T sDi[PXD_NB];               // LDL^T: 1 / D of the block, 0 beyond it
    __shared__ T sX[4][PXD_NB];             // exchange area of the elimination
    // L21 of the current block, sP[row][k]: dynamic LDS for the widest cblk of the launch
    extern __shared__ unsigned char smem_raw[];
    T ( *sP )[PXD_NB + 1] = reinterpret_cast<T ( * )[PXD_NB + 1]>( smem_raw );

    const int32_t k   = list[blockIdx.x];
    const uint64_t t_stamp = stamp_begin( stamps );
    wave_priority( prio );
    const dcblk   c   = cb[k];
    const int32_t n   = c.lcolnum - c.fcolnum + 1;
    const int     tid = threadIdx.x;
    T *A  = L + c.coefoff;            // n x n, ld n
    T *LI = Linv + c.dinvoff;         // n x n, ld n
    int piv = 0;
    typedef typename px_real<T>::type R;
    const R tau = (R)thresh;

    for ( int32_t kb = 0; kb < n; kb += PXD_NB ) {
        const int32_t bs = min( PXD_NB, n - kb );

        {
            constexpr int NU = PXD_NB * PXD_NB / PXD_THREADS;
            T v[NU];
#pragma unroll
            for ( int u = 0; u < NU; u++ ) {
                const int idx = tid + u * PXD_THREADS, r = idx % PXD_NB, cc = idx / PXD_NB;
                v[u] = ( r < bs && cc < bs && r >= cc ) ? A[( kb + r ) + (int64_t)( kb + cc ) * n] : (T)0;
            }
#pragma unroll
            for ( int u = 0; u < NU; u++ ) {
                const int idx = tid + u * PXD_THREADS, r = idx % PXD_NB, cc = idx / PXD_NB;
                sA[cc][r] = v[u];
                sI[cc][r] = ( r == cc && r < bs ) ? (T)1 : (T)0;   // W = I
            }
        }
        __syncthreads();

        // unblocked elimination A = U D U^T, W = U^{-1} alongside
        piv += elim_sym32<T, LDLT>( sA, sI, sD, sX, bs, tau, tid );
        // Cholesky: L = U D^{1/2}, the columns scaled by 1/sqrt(pivot), and
        // L^{-1} = D^{-1/2} U^{-1}, the rows of W scaled likewise.  LDL^T:
        // L = U (column / d), L^{-1} = W, the diagonal holds D and the upper
        // part D L^T (the unscaled columns).
        for ( int idx = tid; idx < PXD_NB * PXD_NB; idx += PXD_THREADS ) {
            const int rr = idx % PXD_NB, cc = idx / PXD_NB;
            if ( rr >= bs || cc >= bs ) continue;
            if ( rr > cc ) {
                const T a = sA[cc][rr];
                if constexpr ( LDLT ) {
                    sA[rr][cc] = a;
                    sA[cc][rr] = a / sD[cc];
                }
                else {
                    sA[cc][rr] = a / px_sqrt( sD[cc] );
                }
            }
            if constexpr ( !LDLT ) {
                if ( rr >= cc ) {
                    sI[cc][rr] = sI[cc][rr] / px_sqrt( sD[rr] );
                }
            }
        }
        __syncthreads();
        if ( tid < bs ) {
            if constexpr ( LDLT ) {
                sA[tid][tid] = sD[tid];
            }
            else {
                sA[tid][tid] = px_sqrt( sD[tid] );
            }
        }
        if ( tid < PXD_NB ) {
            sDi[tid] = ( tid < bs ) ? (T)1 / sD[tid] : (T)0;
        }
        __syncthreads();

        for ( int idx = tid; idx < PXD_NB * PXD_NB; idx += PXD_THREADS ) {
            const int rr = idx % PXD_NB, cc = idx / PXD_NB;
            if ( rr >= bs || cc >= bs ) continue;
            if ( LDLT || rr >= cc ) {
                A[( kb + rr ) + (int64_t)( kb + cc ) * n] = sA[cc][rr];
            }
            LI[( kb + rr ) + (int64_t)( kb + cc ) * n] = sI[cc][rr];
        }

        // rows below the block inside the tile: A21 <- A21 * L11^{-T} on the
        // matrix cores, one wave per 16-row block (two 16 x 16 output tiles,
        // the inverse's transpose read from LDS); the result goes to A and to
        // sP for the trailing update
        typedef mfma16<T>         M;
        typedef typename M::acc_t acc_t;
        const int     lane = tid & 63, w = tid >> 6;
        const int32_t m0   = kb + bs;
        const int32_t m    = n - m0;
        for ( int32_t rb = 16 * w; rb < m; rb += 64 ) {
            T             a[PXD_NB / 4];
            const int32_t ra = m0 + rb + ( lane & 15 );
#pragma unroll
            for ( int s4 = 0; s4 < PXD_NB / 4; s4++ ) {
                const int l = 4 * s4 + ( lane >> 4 );
                a[s4] = ( ra < n && l < bs ) ? A[ra + (int64_t)( kb + l ) * n] : (T)0;
            }
            acc_t acc0 = M::zero(), acc1 = M::zero();
#pragma unroll
            for ( int s4 = 0; s4 < PXD_NB / 4; s4++ ) {
                const int l = 4 * s4 + ( lane >> 4 );
                acc0 = M::mma( a[s4], sI[l][( lane & 15 )], acc0 );
                acc1 = M::mma( a[s4], sI[l][16 + ( lane & 15 )], acc1 );
            }
#pragma unroll
            for ( int v = 0; v < 4; v++ ) {
                const int32_t rr = rb + M::row( lane, v );
                const int     c0 = lane & 15, c1 = 16 + ( lane & 15 );
                if ( rr < m ) {
                    // LDL^T: the product is L21 D; L21 = product / D goes below
                    // the block and the product itself, transposed, to the
                    // block's rows right of it (D L^T, core_zsytrfsp.c:176-186)
                    if ( rr < pcap ) {
                        sP[rr][c0] = M::get( acc0, v );   // zero beyond bs: sI is zero there
                        sP[rr][c1] = M::get( acc1, v );
                    }
                    if ( c0 < bs ) A[( m0 + rr ) + (int64_t)( kb + c0 ) * n] = LDLT ? M::get( acc0, v ) * sDi[c0] : M::get( acc0, v );
                    if ( c1 < bs ) A[( m0 + rr ) + (int64_t)( kb + c1 ) * n] = LDLT ? M::get( acc1, v ) * sDi[c1] : M::get( acc1, v );
                    if constexpr ( LDLT ) {
                        if ( c0 < bs ) A[( kb + c0 ) + (int64_t)( m0 + rr ) * n] = M::get( acc0, v );
                        if ( c1 < bs ) A[( kb + c1 ) + (int64_t)( m0 + rr ) * n] = M::get( acc1, v );
                    }
                }
            }
        }
        __syncthreads();

        // trailing update A22 -= L21 L21^T (lower part) on the matrix cores
        // from sP: 32 x 32 tiles, wave w takes every fourth tile, two at a
        // time, with the tiles' loads of A22 issued before their products.
        // Row ri of the product (L21, times D for LDL^T) beyond pcap comes
        // from A: the D L^T strip for LDL^T, the L21 column block otherwise.
        auto prow = [&]( int32_t ri, int l ) -> T {
            if ( ri < pcap ) {
                return sP[ri][l];
            }
            if ( l >= bs ) {
                return (T)0;
            }
            return LDLT ? A[( kb + l ) + (int64_t)( m0 + ri ) * n] : A[( m0 + ri ) + (int64_t)( kb + l ) * n];
        };
        {
            const int32_t nt = ( m + 31 ) / 32;
            const int32_t ntile = nt * ( nt + 1 ) / 2;
            for ( int32_t t0 = 2 * w; t0 < ntile; t0 += 8 ) {
                int32_t ti[2], tj[2];
                bool    tv[2];
#pragma unroll
                for ( int g = 0; g < 2; g++ ) {
                    // tile t0 + g in row-wise order of the lower triangle
                    int32_t tt = t0 + g, r = 0;
                    while ( ( r + 1 ) * ( r + 2 ) / 2 <= tt ) r++;
                    ti[g] = r;
                    tj[g] = tt - r * ( r + 1 ) / 2;
                    tv[g] = tt < ntile;
                }
                T old[2][2][2][4];
#pragma unroll
                for ( int g = 0; g < 2; g++ )
#pragma unroll
                    for ( int a = 0; a < 2; a++ )
#pragma unroll
                        for ( int b = 0; b < 2; b++ )
#pragma unroll
                            for ( int v = 0; v < 4; v++ ) {
                                const int32_t r  = 32 * ti[g] + 16 * a + M::row( lane, v );
                                const int32_t cc = 32 * tj[g] + 16 * b + ( lane & 15 );
                                const bool    ok = tv[g] && r < m && cc < m && r >= cc;
                                old[g][a][b][v] = ok ? A[( m0 + r ) + (int64_t)( m0 + cc ) * n] : (T)0;
                            }
                acc_t acc[2][2][2];
#pragma unroll
                for ( int g = 0; g < 2; g++ )
#pragma unroll
                    for ( int a = 0; a < 2; a++ )
#pragma unroll
                        for ( int b = 0; b < 2; b++ ) acc[g][a][b] = M::zero();
#pragma unroll
                for ( int s4 = 0; s4 < PXD_NB / 4; s4++ ) {
                    const int l = 4 * s4 + ( lane >> 4 );
#pragma unroll
                    for ( int g = 0; g < 2; g++ ) {
                        T af[2], bf[2];
#pragma unroll
                        for ( int a = 0; a < 2; a++ ) {
                            const int32_t ri = 32 * ti[g] + 16 * a + ( lane & 15 );
                            const int32_t rj = 32 * tj[g] + 16 * a + ( lane & 15 );
                            // LDL^T: A22 -= (L21 D) D^{-1} (L21 D)^T
                            af[a] = ( ri < m ) ? ( LDLT ? prow( ri, l ) * sDi[l] : prow( ri, l ) ) : (T)0;
                            bf[a] = ( rj < m ) ? prow( rj, l ) : (T)0;
                        }
#pragma unroll
                        for ( int a = 0; a < 2; a++ )
#pragma unroll
                            for ( int b = 0; b < 2; b++ ) acc[g][a][b] = M::mma( af[a], bf[b], acc[g][a][b] );
                    }
                }
#pragma unroll
                for ( int g = 0; g < 2; g++ )
#pragma unroll
                    for ( int a = 0; a < 2; a++ )
#pragma unroll
                        for ( int b = 0; b < 2; b++ )
#pragma unroll
                            for ( int v = 0; v < 4; v++ ) {
                                const int32_t r  = 32 * ti[g] + 16 * a + M::row( lane, v );
                                const int32_t cc = 32 * tj[g] + 16 * b + ( lane & 15 );
                                if ( tv[g] && r < m && cc < m && r >= cc ) {
                                    A[( m0 + r ) + (int64_t)( m0 + cc ) * n] = old[g][a][b][v] - M::get( acc[g][a][b], v );
                                }
                            }
            }
        }
        __syncthreads();
    }
    if ( tid == 0 && piv > 0 ) {
        atomicAdd( npiv, piv );
    }
    stamp_end( stamps, blockIdx.x, t_stamp );
}

// ---------------------------------------------------------------------------
// Multi-workgroup Cholesky of the diagonal tiles.  Task {cblk, p, flag0}:
// the workgroup owns the 32-row block p of the cblk's n x n diagonal tile,
// i.e. the 32 x 32 tiles T(p, 0..p), kept in LDS from the start.  For
// k = 0..p-1 it waits until block k has factored its diagonal tile
// (flag[k] > k), forms L(p,k) = T(p,k) L(k,k)^{-T} with the inverse that
// block k wrote to Linv, publishes it (flag[p] = k+1), and subtracts
// L(p,k) L(j,k)^T from its tiles T(p,j), j = k+1..p, each L(j,k) once
// block j has published it.  For k = p it factors T(p,p) like one block of
// k_potrf2 (static pivot, inverse built by the elimination) and publishes
// flag[p] = p+1.  The critical path is the chain of diagonal tiles instead
// of the n/32 sequential block steps of a single workgroup.
//
// Exchange between the workgroups: the L(p,k) tiles and the diagonal
// inverses that other blocks read are written and read with agent-scope
// relaxed atomic stores and loads (panel_st, panel_ld), which are coherent
// across the XCDs by themselves.  Publication: every thread waits until its
// own stores are complete, a barrier, then thread 0 stores the flag.  Waits:
// thread 0 polls the flag, then a barrier.  No agent-scope fence is needed,
// and none is issued: such a fence writes back (release) or invalidates
// (acquire) the XCD's whole L2, which the update kernels of the other
// stream keep filling (100^3: FP32 numfact 86.3 -> 85.4 ms).  Every wait is
// bounded by a wall-clock deadline; on expiry the kernel records it in *err
// and goes on, so every wave reaches the end.
struct pan_task {
    int32_t cblk;
    int32_t p;
    int32_t flag0;
    int32_t nblk;    // blocks of the tile: the last one has no consumer
};
// Bound of one wait in ticks of the constant clock (wall_clock64), set on
// every device from its clock rate when an engine context is created (2 s
// by default, engine option "wait_ms").
__constant__ uint64_t pxd_wait_ticks = 200000000ull;

__device__ inline void
panel_wait( const int *f, int target, int *err, int tid )
{
    if ( tid == 0 ) {
        // after one timeout anywhere every later wait falls through at once
        const uint64_t t0 = (uint64_t)wall_clock64();
        while ( __hip_atomic_load( f, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT ) < target &&
                __hip_atomic_load( err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT ) == 0 ) {
            if ( (uint64_t)wall_clock64() - t0 > pxd_wait_ticks ) {
                __hip_atomic_store( err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT );
                break;
            }
            __builtin_amdgcn_s_sleep( 1 );
        }
    }
    __syncthreads();
    __atomic_signal_fence( __ATOMIC_SEQ_CST );   // the block's loads stay after the wait
}

__device__ inline void
panel_publish( int *f, int value, int tid )
{
    // each thread's panel_st stores complete (vmcnt 0) before the barrier,
    // so the flag is stored after all of them
    __atomic_signal_fence( __ATOMIC_SEQ_CST );
    __builtin_amdgcn_s_waitcnt( 0 );
    __syncthreads();
    if ( tid == 0 ) {
        __hip_atomic_store( f, value, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT );
    }
}

// Loads and stores of the data that the panel's workgroups exchange.
template <typename T>
__device__ inline T
panel_ld( const T *p )
{
    return __hip_atomic_load( p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT );
}
template <typename T>
__device__ inline void
panel_st( T *p, T v )
{
    __hip_atomic_store( p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT );
}

// P[r][c] = sum_l X[r][l] Y[c][l] for 32 x 32 tiles held as s[col][row]:
// wave w computes the 16 x 16 quadrant (w >> 1, w & 1) on the matrix cores.
template <typename T>
__device__ inline typename mfma16<T>::acc_t
tile_xyt( const T ( *sX )[PXD_NB + 1], const T ( *sY )[PXD_NB + 1], int lane, int w )
{
    typedef mfma16<T> M;
    typename M::acc_t acc = M::zero();
    const int qr = w >> 1, qc = w & 1;
#pragma unroll
    for ( int s4 = 0; s4 < PXD_NB / 4; s4++ ) {
        const int l = 4 * s4 + ( lane >> 4 );
        acc = M::mma( sX[l][16 * qr + ( lane & 15 )], sY[l][16 * qc + ( lane & 15 )], acc );
    }
    return acc;
}

template <typename T>
__global__ void __launch_bounds__( PXD_THREADS )
k_potrf_mw( const pan_task *__restrict__ tasks, const dcblk *__restrict__ cb, T *__restrict__ L,
            T *__restrict__ Linv, double thresh, int *__restrict__ npiv, int *flags, int *err,
            uint64_t *__restrict__ stamps, int prio )
{
    typedef mfma16<T>         M;
    typedef typename M::acc_t acc_t;
    typedef T                 tile_t[PXD_NB][PXD_NB + 1];   // [col][row]
    __shared__ tile_t sI;                // inverse of a diagonal block
    __shared__ tile_t sO;                // L(j,k) of another block
    __shared__ T      sD[PXD_NB];
    extern __shared__ unsigned char smem_raw[];
    tile_t *sT = reinterpret_cast<tile_t *>( smem_raw );   // T(p, 0..p)

    const pan_task tk      = tasks[blockIdx.x];
    const uint64_t t_stamp = stamp_begin( stamps );
    wave_priority( prio );
    const dcblk   c    = cb[tk.cblk];
    const int32_t n    = c.lcolnum - c.fcolnum + 1;
    const int     p    = tk.p;
    const int32_t R0   = PXD_NB * p;
    const int32_t bs   = min( PXD_NB, n - R0 );
    const int     tid  = threadIdx.x;
    const int     lane = tid & 63, w = tid >> 6;
    T            *A    = L + c.coefoff;
    T            *LI   = Linv + c.dinvoff;
    int          *flag = flags + tk.flag0;
    const bool    last = ( p == tk.nblk - 1 );   // nobody waits for the last block
    constexpr int NU   = PXD_NB * PXD_NB / PXD_THREADS;
    int           piv  = 0;
    typedef typename px_real<T>::type R;
    const R tau = (R)thresh;

    // the block row: T(p, j) for j = 0..p (lower part only for j = p).
    // Every load is issued, the entries outside read a clamped in-range
    // address and are masked: a load under its own branch made the compiler
    // wait for the previous one.
    for ( int j0 = 0; j0 <= p; j0 += 4 ) {
        T v[4][NU];
#pragma unroll
        for ( int jj = 0; jj < 4; jj++ )
#pragma unroll
            for ( int u = 0; u < NU; u++ ) {
                const int idx = tid + u * PXD_THREADS, r = idx % PXD_NB, cc = idx / PXD_NB, j = j0 + jj;
                const bool    ok  = j <= p && r < bs && ( j < p || cc <= r );
                const int32_t col = min( PXD_NB * min( j, p ) + cc, n - 1 );
                const T       x   = A[( R0 + min( r, bs - 1 ) ) + (int64_t)col * n];
                v[jj][u] = ok ? x : (T)0;
            }
#pragma unroll
        for ( int jj = 0; jj < 4; jj++ )
#pragma unroll
            for ( int u = 0; u < NU; u++ ) {
                const int idx = tid + u * PXD_THREADS, r = idx % PXD_NB, cc = idx / PXD_NB;
                if ( j0 + jj <= p ) sT[j0 + jj][cc][r] = v[jj][u];
            }
    }

    for ( int k = 0; k < p; k++ ) {
        // L(p,k) = T(p,k) L(k,k)^{-T}
        panel_wait( flag + k, k + 1, err, tid );
        {
            T v[NU];
#pragma unroll
            for ( int u = 0; u < NU; u++ ) {
                const int idx = tid + u * PXD_THREADS, r = idx % PXD_NB, cc = idx / PXD_NB;
                v[u] = panel_ld( LI + ( PXD_NB * k + r ) + (int64_t)( PXD_NB * k + cc ) * n );
            }
#pragma unroll
            for ( int u = 0; u < NU; u++ ) {
                const int idx = tid + u * PXD_THREADS, r = idx % PXD_NB, cc = idx / PXD_NB;
                sI[cc][r] = v[u];
            }
        }
        __syncthreads();
        {
            acc_t acc = tile_xyt<T>( sT[k], sI, lane, w );
            __syncthreads();   // every wave has read T(p,k)
            const int qr = w >> 1, qc = w & 1;
#pragma unroll
            for ( int v = 0; v < 4; v++ ) {
                const int r = 16 * qr + M::row( lane, v ), cc = 16 * qc + ( lane & 15 );
                const T   x = ( r < bs ) ? M::get( acc, v ) : (T)0;
                sT[k][cc][r] = x;
                if ( r < bs ) panel_st( A + ( R0 + r ) + (int64_t)( PXD_NB * k + cc ) * n, x );
            }
        }
        if ( !last ) {
            panel_publish( flag + p, k + 1, tid );
        }
        __syncthreads();   // sT[k] complete for every wave
        // T(p,j) -= L(p,k) L(j,k)^T, j = k+1..p
        for ( int j = k + 1; j <= p; j++ ) {
            const tile_t *Y = &sT[k];
            if ( j < p ) {
                panel_wait( flag + j, k + 1, err, tid );
                T v[NU];
#pragma unroll
                for ( int u = 0; u < NU; u++ ) {
                    const int idx = tid + u * PXD_THREADS, r = idx % PXD_NB, cc = idx / PXD_NB;
                    v[u] = panel_ld( A + ( PXD_NB * j + r ) + (int64_t)( PXD_NB * k + cc ) * n );
                }
#pragma unroll
                for ( int u = 0; u < NU; u++ ) {
                    const int idx = tid + u * PXD_THREADS, r = idx % PXD_NB, cc = idx / PXD_NB;
                    sO[cc][r] = v[u];
                }
                __syncthreads();
                Y = &sO;
            }
            acc_t     acc = tile_xyt<T>( sT[k], *Y, lane, w );
            const int qr = w >> 1, qc = w & 1;
#pragma unroll
            for ( int v = 0; v < 4; v++ ) {
                const int r = 16 * qr + M::row( lane, v ), cc = 16 * qc + ( lane & 15 );
                sT[j][cc][r] -= M::get( acc, v );
            }
            __syncthreads();   // sO is reused by the next j
        }
    }

    // T(p,p): elimination with the inverse built alongside (as in k_potrf2);
    // sO is free now and serves as the exchange area
    T( *sA )[PXD_NB + 1] = sT[p];
    for ( int idx = tid; idx < PXD_NB * PXD_NB; idx += PXD_THREADS ) {
        const int r = idx % PXD_NB, cc = idx / PXD_NB;
        sI[cc][r] = ( r == cc && r < bs ) ? (T)1 : (T)0;
    }
    __syncthreads();
    piv += elim_sym32<T, false>( sA, sI, sD, reinterpret_cast<T ( * )[PXD_NB]>( &sO[0][0] ), bs, tau, tid );
    for ( int idx = tid; idx < PXD_NB * PXD_NB; idx += PXD_THREADS ) {
        const int rr = idx % PXD_NB, cc = idx / PXD_NB;
        if ( rr >= bs || cc >= bs ) continue;
        if ( rr > cc ) {
            sA[cc][rr] = sA[cc][rr] / px_sqrt( sD[cc] );
        }
        if ( rr >= cc ) {
            sI[cc][rr] = sI[cc][rr] / px_sqrt( sD[rr] );
        }
    }
    __syncthreads();
    if ( tid < bs ) {
        sA[tid][tid] = px_sqrt( sD[tid] );
    }
    __syncthreads();
    for ( int idx = tid; idx < PXD_NB * PXD_NB; idx += PXD_THREADS ) {
        const int rr = idx % PXD_NB, cc = idx / PXD_NB;
        if ( rr >= bs || cc >= bs ) continue;
        if ( rr >= cc ) {
            A[( R0 + rr ) + (int64_t)( R0 + cc ) * n] = sA[cc][rr];
        }
        panel_st( LI + ( R0 + rr ) + (int64_t)( R0 + cc ) * n, sI[cc][rr] );
    }
    if ( !last ) {
        panel_publish( flag + p, p + 1, tid );
    }
    if ( tid == 0 && piv > 0 ) {
        atomicAdd( npiv, piv );
    }
    stamp_end( stamps, blockIdx.x, t_stamp );
}

// ---------------------------------------------------------------------------
// X = B * S^T for a rows x 32 block B of the diagonal tile, in place, on the
// matrix cores: wave w takes 16-row blocks w, w+4, ...; row i of B is
// element (i, k) at base + i * rs + k * ks; S is a 32 x 32 matrix in LDS,
// S[row][col] = sS[col][row].  Only the first bs columns of B are read or
// written (S is zero outside its bs x bs corner).
template <typename T>
__device__ inline void
small_trsm_mfma( T *base, int64_t rs, int64_t ks, int32_t rows, int32_t bs, const T ( *sS )[PXD_NB + 1],
                 int tid )
{
    typedef mfma16<T>         M;
    typedef typename M::acc_t acc_t;
    const int lane = tid & 63, w = tid >> 6;
    for ( int32_t rb = 16 * w; rb < rows; rb += 64 ) {
        T             a[PXD_NB / 4];
        const int32_t ra = rb + ( lane & 15 );
#pragma unroll
        for ( int s4 = 0; s4 < PXD_NB / 4; s4++ ) {
            const int l = 4 * s4 + ( lane >> 4 );
            a[s4] = ( ra < rows && l < bs ) ? base[ra * rs + l * ks] : (T)0;
        }
        acc_t acc0 = M::zero(), acc1 = M::zero();
#pragma unroll
        for ( int s4 = 0; s4 < PXD_NB / 4; s4++ ) {
            const int l = 4 * s4 + ( lane >> 4 );
            acc0 = M::mma( a[s4], sS[l][( lane & 15 )], acc0 );
            acc1 = M::mma( a[s4], sS[l][16 + ( lane & 15 )], acc1 );
        }
#pragma unroll
        for ( int v = 0; v < 4; v++ ) {
            const int32_t rr = rb + M::row( lane, v );
            const int     c0 = lane & 15, c1 = 16 + ( lane & 15 );
            if ( rr < rows && c0 < bs ) base[rr * rs + c0 * ks] = M::get( acc0, v );
            if ( rr < rows && c1 < bs ) base[rr * rs + c1 * ks] = M::get( acc1, v );
        }
    }
}

// ---------------------------------------------------------------------------
// LU panel with static pivoting (core_zgetrfsp.c:64-272): the diagonal tile
// of L holds the lower part of A_kk and the U tile its strict upper part,
// transposed; they are summed (the reference's geadd), factorized without
// pivoting by PXD_NB-column blocks (|a_kk| < tau -> sign(a_kk) tau), and the
// packed LU is copied transposed into the U tile (the reference's getmo), so
// that the lower part of the U tile is U^T.  The inverses of the diagonal
// PXD_NB x PXD_NB blocks of L_kk (unit lower) and of U_kk^T go to the
// diagonal blocks of DL and DU.
template <typename T>
__global__ void __launch_bounds__( PXD_THREADS )
k_getrf2( const int32_t *__restrict__ list, const dcblk *__restrict__ cb, T *__restrict__ L, T *__restrict__ U,
          T *__restrict__ DL, T *__restrict__ DU, double thresh, int *__restrict__ npiv,
          uint64_t *__restrict__ stamps, int prio )
{
    __shared__ T sA[PXD_NB][PXD_NB + 1];    // sA[col][row]
    __shared__ T sIL[PXD_NB][PXD_NB + 1];   // L11^{-1}, [col][row]
    __shared__ T sIU[PXD_NB][PXD_NB + 1];   // (U11^T)^{-1}, [col][row]
    __shared__ T sD[PXD_NB];
    __shared__ T sX[8][PXD_NB];              // exchange area of the elimination
    __shared__ T sPi[PXD_NB][64 + 1];
    __shared__ T sPj[PXD_NB][64 + 1];

    const int32_t k   = list[blockIdx.x];
    const uint64_t t_stamp = stamp_begin( stamps );
    wave_priority( prio );
    const dcblk   c   = cb[k];
    const int32_t n   = c.lcolnum - c.fcolnum + 1;
    const int     tid = threadIdx.x;
    T *A  = L + c.coefoff;       // n x n, ld n
    T *AU = U + c.coefoff;       // n x n, ld n
    T *IL = DL + c.dinvoff;
    T *IU = DU + c.dinvoff;
    int     piv = 0;
    typedef typename px_real<T>::type R;
    const R tau = (R)thresh;
    const int64_t nn = (int64_t)n * n;

    // A_kk = L part + (U part)^T, 8 elements per thread in flight
    for ( int64_t i0 = 0; i0 < nn; i0 += 8 * PXD_THREADS ) {
        T va[8], vu[8];
#pragma unroll
        for ( int u = 0; u < 8; u++ ) {
            const int64_t idx = i0 + tid + u * PXD_THREADS;
            const int32_t r = (int32_t)( idx % n ), cc = (int32_t)( idx / n );
            const bool    up = idx < nn && r < cc;
            va[u] = up ? A[idx] : (T)0;
            vu[u] = up ? AU[cc + (int64_t)r * n] : (T)0;
        }
#pragma unroll
        for ( int u = 0; u < 8; u++ ) {
            const int64_t idx = i0 + tid + u * PXD_THREADS;
            const int32_t r = (int32_t)( idx % n ), cc = (int32_t)( idx / n );
            if ( idx < nn && r < cc ) A[idx] = va[u] + vu[u];
        }
    }
    __syncthreads();

    for ( int32_t kb = 0; kb < n; kb += PXD_NB ) {
        const int32_t bs = min( PXD_NB, n - kb );
        {
            constexpr int NU = PXD_NB * PXD_NB / PXD_THREADS;
            T v[NU];
#pragma unroll
            for ( int u = 0; u < NU; u++ ) {
                const int idx = tid + u * PXD_THREADS, r = idx % PXD_NB, cc = idx / PXD_NB;
                v[u] = ( r < bs && cc < bs ) ? A[( kb + r ) + (int64_t)( kb + cc ) * n] : (T)0;
            }
#pragma unroll
            for ( int u = 0; u < NU; u++ ) {
                const int idx = tid + u * PXD_THREADS, r = idx % PXD_NB, cc = idx / PXD_NB;
                sA[cc][r]  = v[u];
                sIL[cc][r] = ( r == cc && r < bs ) ? (T)1 : (T)0;   // W_L = I
                sIU[cc][r] = ( r == cc && r < bs ) ? (T)1 : (T)0;   // W_U = I
            }
        }
        __syncthreads();

        // unblocked right-looking LU, the multipliers left unscaled until the
        // end; W_L = L11^{-1} and W_U = U~^{-1} (U11^T = U~ D, so that
        // (U11^T)^{-1} = D^{-1} W_U) built alongside
        piv += elim_lu32<T>( sA, sIL, sIU, sD, sX, bs, tau, tid );
        for ( int idx = tid; idx < PXD_NB * PXD_NB; idx += PXD_THREADS ) {
            const int rr = idx % PXD_NB, cc = idx / PXD_NB;
            if ( rr >= bs || cc >= bs ) continue;
            if ( rr > cc ) {
                sA[cc][rr] = sA[cc][rr] / sD[cc];
            }
            else if ( rr == cc ) {
                sA[cc][cc] = sD[cc];
            }
        }
        __syncthreads();

        // (U11^T)^{-1} = D^{-1} W_U: rows scaled by 1/d (W_L is L11^{-1} as is)
        for ( int idx = tid; idx < PXD_NB * PXD_NB; idx += PXD_THREADS ) {
            const int rr = idx % PXD_NB, cc = idx / PXD_NB;
            if ( rr >= bs || cc >= bs ) continue;
            if ( rr >= cc ) {
                sIU[cc][rr] = sIU[cc][rr] / sD[rr];
            }
        }
        __syncthreads();

        for ( int idx = tid; idx < PXD_NB * PXD_NB; idx += PXD_THREADS ) {
            const int rr = idx % PXD_NB, cc = idx / PXD_NB;
            if ( rr >= bs || cc >= bs ) continue;
            A[( kb + rr ) + (int64_t)( kb + cc ) * n]  = sA[cc][rr];
            IL[( kb + rr ) + (int64_t)( kb + cc ) * n] = sIL[cc][rr];
            IU[( kb + rr ) + (int64_t)( kb + cc ) * n] = sIU[cc][rr];
        }

        // in the tile: L21 <- A21 U11^{-1} = A21 (sIU)^T and
        // U12^T <- A12^T L11^{-T} = A12^T (sIL)^T
        const int32_t m0 = kb + bs;
        const int32_t m  = n - m0;
        small_trsm_mfma<T>( A + m0 + (int64_t)kb * n, 1, n, m, bs, sIU, tid );
        small_trsm_mfma<T>( A + kb + (int64_t)m0 * n, n, 1, m, bs, sIL, tid );
        __syncthreads();

        // trailing update A22 -= L21 U12 (full square, 64 x 64 macro tiles)
        const int32_t nt = ( m + 63 ) / 64;
        const int     tx = tid % 16, ty = tid / 16;
        for ( int32_t ti = 0; ti < nt; ti++ ) {
            for ( int32_t tj = 0; tj < nt; tj++ ) {
                {
                    constexpr int NU = 64 * PXD_NB / PXD_THREADS;
                    T vi[NU], vj[NU];
#pragma unroll
                    for ( int u = 0; u < NU; u++ ) {
                        const int     idx = tid + u * PXD_THREADS, r = idx % 64, l = idx / 64;
                        const int32_t ri = ti * 64 + r, rj = tj * 64 + r;
                        vi[u] = ( ri < m && l < bs ) ? A[( m0 + ri ) + (int64_t)( kb + l ) * n] : (T)0;
                        vj[u] = ( rj < m && l < bs ) ? A[( kb + l ) + (int64_t)( m0 + rj ) * n] : (T)0;
                    }
#pragma unroll
                    for ( int u = 0; u < NU; u++ ) {
                        const int idx = tid + u * PXD_THREADS, r = idx % 64, l = idx / 64;
                        sPi[l][r] = vi[u];
                        sPj[l][r] = vj[u];
                    }
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
                T old[4][4];
#pragma unroll
                for ( int a = 0; a < 4; a++ )
#pragma unroll
                    for ( int b = 0; b < 4; b++ ) {
                        const int32_t r = ti * 64 + tx + 16 * a, cc = tj * 64 + ty + 16 * b;
                        old[a][b] = ( r < m && cc < m ) ? A[( m0 + r ) + (int64_t)( m0 + cc ) * n] : (T)0;
                    }
#pragma unroll
                for ( int a = 0; a < 4; a++ )
#pragma unroll
                    for ( int b = 0; b < 4; b++ ) {
                        const int32_t r = ti * 64 + tx + 16 * a, cc = tj * 64 + ty + 16 * b;
                        if ( r < m && cc < m ) A[( m0 + r ) + (int64_t)( m0 + cc ) * n] = old[a][b] - acc[a][b];
                    }
                __syncthreads();
            }
        }
        __syncthreads();
    }

    // U tile <- (packed LU)^T, so that its lower part is U^T
    for ( int64_t i0 = 0; i0 < nn; i0 += 8 * PXD_THREADS ) {
        T va[8];
#pragma unroll
        for ( int u = 0; u < 8; u++ ) {
            const int64_t idx = i0 + tid + u * PXD_THREADS;
            va[u] = ( idx < nn ) ? A[idx] : (T)0;
        }
#pragma unroll
        for ( int u = 0; u < 8; u++ ) {
            const int64_t idx = i0 + tid + u * PXD_THREADS;
            const int32_t r = (int32_t)( idx % n ), cc = (int32_t)( idx / n );
            if ( idx < nn ) AU[cc + (int64_t)r * n] = va[u];
        }
    }
    if ( tid == 0 && piv > 0 ) {
        atomicAdd( npiv, piv );
    }
    stamp_end( stamps, blockIdx.x, t_stamp );
}

// ---------------------------------------------------------------------------
// Inverse of the lower-triangular L_kk by block columns, one workgroup per
// (cblk, block column J).  The diagonal blocks D_K = L[K,K]^{-1} were written
// to Linv by the panel.  Column sweep: R = -L[J+1:, J] D_J, then for K = J+1,
// J+2, ...: X[K,J] = D_K R[K], R[K+1:] -= L[K+1:, K] X[K,J].  The first rcap
// rows of R stay in LDS, the rest (wide cblks in complex double) in place in
// Linv, where X[K,J] later overwrites them; both products run on the matrix
// cores.
template <typename T>
__global__ void __launch_bounds__( PXD_THREADS )
k_trtri( const inv_task *__restrict__ tasks, const dcblk *__restrict__ cb, const T *__restrict__ L,
         T *__restrict__ Linv, uint64_t *__restrict__ stamps, int prio, int32_t rcap )
{
    typedef mfma16<T>         M;
    typedef typename M::acc_t acc_t;
    __shared__ T sX[PXD_NB][PXD_NB + 1];   // the current X[K,J] (or D_J), sX[row][col]
    __shared__ T sD[PXD_NB][PXD_NB + 1];   // D_K, sD[row][col]
    extern __shared__ unsigned char smem_raw[];
    T ( *sR )[PXD_NB + 1] = reinterpret_cast<T ( * )[PXD_NB + 1]>( smem_raw );   // rows below block J

    const inv_task t    = tasks[blockIdx.x];
    const uint64_t t_stamp = stamp_begin( stamps );
    wave_priority( prio );
    const dcblk    c    = cb[t.cblk];
    const int32_t  n    = c.lcolnum - c.fcolnum + 1;
    const int      tid  = threadIdx.x;
    const int      lane = tid & 63, w = tid >> 6;
    const T       *A    = L + c.coefoff;
    T             *LI   = Linv + c.dinvoff;
    const int32_t  J    = t.J;                 // first column of the block column
    const int32_t  r0   = J + PXD_NB;          // first row of R
    const int32_t  nr   = n - r0;
    const int32_t  bj   = min( PXD_NB, n - J );
    // R(row, col): rows below rcap in LDS, the others at their place in Linv
    // (nr > 0 implies bj == PXD_NB, so every column is inside the tile)
    auto rget = [&]( int32_t rr, int cc ) -> T {
        return ( rr < rcap ) ? sR[rr][cc] : LI[( r0 + rr ) + (int64_t)( J + cc ) * n];
    };
    auto rset = [&]( int32_t rr, int cc, T v ) {
        if ( rr < rcap ) {
            sR[rr][cc] = v;
        }
        else {
            LI[( r0 + rr ) + (int64_t)( J + cc ) * n] = v;
        }
    };

    // load a 32 x 32 block of Linv (rows R0, cols C0) into dst[row][col]
    auto load_block = [&]( T ( *dst )[PXD_NB + 1], int32_t R0, int32_t C0, int32_t nrow, int32_t ncol ) {
        constexpr int NU = PXD_NB * PXD_NB / PXD_THREADS;
        T v[NU];
#pragma unroll
        for ( int u = 0; u < NU; u++ ) {
            const int idx = tid + u * PXD_THREADS, r = idx % PXD_NB, cc = idx / PXD_NB;
            v[u] = ( r < nrow && cc < ncol ) ? LI[( R0 + r ) + (int64_t)( C0 + cc ) * n] : (T)0;
        }
#pragma unroll
        for ( int u = 0; u < NU; u++ ) {
            const int idx = tid + u * PXD_THREADS, r = idx % PXD_NB, cc = idx / PXD_NB;
            dst[r][cc] = v[u];
        }
    };
    // R[rows >= rfirst] -= L[r0 + rows, K .. K+31] * sX  (16-row blocks per wave)
    auto sweep = [&]( int32_t rfirst, int32_t K, int32_t bk ) {
        for ( int32_t rb = rfirst + 16 * w; rb < nr; rb += 64 ) {
            T             a[PXD_NB / 4];
            const int32_t ra = rb + ( lane & 15 );
#pragma unroll
            for ( int s4 = 0; s4 < PXD_NB / 4; s4++ ) {
                const int l = 4 * s4 + ( lane >> 4 );
                a[s4] = ( ra < nr && l < bk ) ? -A[( r0 + ra ) + (int64_t)( K + l ) * n] : (T)0;
            }
            acc_t acc0, acc1;
#pragma unroll
            for ( int v = 0; v < 4; v++ ) {
                const int32_t rr = rb + M::row( lane, v );
                M::set( acc0, v, ( rr < nr ) ? rget( rr, lane & 15 ) : (T)0 );
                M::set( acc1, v, ( rr < nr ) ? rget( rr, 16 + ( lane & 15 ) ) : (T)0 );
            }
#pragma unroll
            for ( int s4 = 0; s4 < PXD_NB / 4; s4++ ) {
                const int l = 4 * s4 + ( lane >> 4 );
                acc0 = M::mma( a[s4], sX[l][lane & 15], acc0 );
                acc1 = M::mma( a[s4], sX[l][16 + ( lane & 15 )], acc1 );
            }
#pragma unroll
            for ( int v = 0; v < 4; v++ ) {
                const int32_t rr = rb + M::row( lane, v );
                if ( rr < nr ) {
                    rset( rr, lane & 15, M::get( acc0, v ) );
                    rset( rr, 16 + ( lane & 15 ), M::get( acc1, v ) );
                }
            }
        }
    };

    for ( int idx = tid; idx < min( nr, rcap ) * ( PXD_NB + 1 ); idx += PXD_THREADS ) {
        ( &sR[0][0] )[idx] = (T)0;
    }
    for ( int idx = tid; idx < ( nr - rcap ) * PXD_NB; idx += PXD_THREADS ) {
        rset( rcap + idx % ( nr - rcap ), idx / ( nr - rcap ), (T)0 );
    }
    load_block( sX, J, J, bj, bj );   // D_J
    __syncthreads();
    sweep( 0, J, bj );                // R = -L[r0:, J] D_J
    for ( int32_t K = r0; K < n; K += PXD_NB ) {
        const int32_t bk = min( PXD_NB, n - K );
        load_block( sD, K, K, bk, bk );
        __syncthreads();
        // X[K,J] = D_K R[K - r0 .. +31]: wave w computes the 16 x 16 block
        // (w >> 1, w & 1) of the 32 x 32 product
        {
            const int32_t bi = w >> 1, bc = w & 1;
            acc_t         acc = M::zero();
#pragma unroll
            for ( int s4 = 0; s4 < PXD_NB / 4; s4++ ) {
                const int l = 4 * s4 + ( lane >> 4 );
                const T   a = sD[16 * bi + ( lane & 15 )][l];
                const T   b = ( K - r0 + l < nr ) ? rget( K - r0 + l, 16 * bc + ( lane & 15 ) ) : (T)0;
                acc = M::mma( a, b, acc );
            }
            __syncthreads();   // every wave has read sX (previous step) and sR[K block]
#pragma unroll
            for ( int v = 0; v < 4; v++ ) {
                const int r = 16 * bi + M::row( lane, v ), cc = 16 * bc + ( lane & 15 );
                sX[r][cc] = M::get( acc, v );
                if ( r < bk && cc < bj ) LI[( K + r ) + (int64_t)( J + cc ) * n] = M::get( acc, v );
            }
        }
        __syncthreads();
        if ( K + PXD_NB < n ) {
            sweep( K + PXD_NB - r0, K, bk );
        }
    }
    stamp_end( stamps, blockIdx.x, t_stamp );
}

// ---------------------------------------------------------------------------
// Off-diagonal TRSM on the matrix cores: X = B * Linv^T for a tile of 64
// off-diagonal rows and all n columns of the cblk, in place (the workgroup
// owns complete rows, so every read of B precedes every write of X).
// Columns go in passes of NTP = 4 * CPW 16-column tiles, the highest pass
// first: Linv is lower triangular, so a pass reads only B columns up to its
// own last one and may overwrite its columns before the lower passes run.
// In a pass every wave covers all 64 rows against CPW column tiles (tiles
// w, w+4, ...), so each LDS fragment feeds 4 or CPW products; the k values
// of an MFMA step are permuted (lane group g takes k = KV*g + s) so that
// every fragment is one 16-byte read.  Stages whose k range lies above a
// column tile (Linv zero there) are skipped for that tile.
template <typename T> struct trsm_cpw;
template <> struct trsm_cpw<float>  { static constexpr int value = 2; };
template <> struct trsm_cpw<double> { static constexpr int value = 2; };
template <> struct trsm_cpw<pxcf>   { static constexpr int value = 2; };
template <> struct trsm_cpw<pxcd>   { static constexpr int value = 1; };

template <typename T>
__global__ void __launch_bounds__( PXD_THREADS, 2 )
k_trsm_mfma( const row_task *__restrict__ tasks, const dcblk *__restrict__ cb,
             const dblok *__restrict__ bk, T *__restrict__ L, const T *__restrict__ Linv, int ldlt,
             uint64_t *__restrict__ stamps, int prio )
{
    typedef mfma16<T>                 M;
    typedef typename M::acc_t         acc_t;
    typedef typename lds_vec<T>::type vec_t;
    constexpr int KV  = lds_vec<T>::n;
    constexpr int BK  = 4 * KV;                      // k values per stage
    constexpr int LDK = BK + KV;
    constexpr int CPW = trsm_cpw<T>::value;          // column tiles per wave
    constexpr int NTP = 4 * CPW;                     // column tiles per pass
    constexpr int CPT = NTP * 16;                    // columns per pass
    constexpr int TPC = PXD_THREADS / CPT;           // Linv loader threads per column
    constexpr int KPC = BK / TPC;                    // k values per Linv loader thread
    __shared__ __align__( 16 ) T sB[64 * LDK];       // B rows, k contiguous
    __shared__ __align__( 16 ) T sLi[CPT * LDK];     // Linv columns of the pass, k contiguous
    __shared__ int64_t sBase[64];
    __shared__ int32_t sLd[64];
    __shared__ T       sDi[PXD_TRSM_NT * 16];        // LDL^T: 1 / D of the cblk

    const int32_t  ti   = xcd_task( blockIdx.x, gridDim.x );
    const row_task t    = tasks[ti];
    const uint64_t t_stamp = stamp_begin( stamps );
    wave_priority( prio );
    const dcblk    c    = cb[t.cblk];
    const int32_t  n    = c.lcolnum - c.fcolnum + 1;
    const int32_t  lb   = cb[t.cblk + 1].fblok;
    const int      tid  = threadIdx.x;
    const int      lane = tid & 63, w = tid >> 6;
    const int      lr   = lane & 15, lg = lane >> 4;
    const T       *LI   = Linv + c.dinvoff;
    const int      nt   = ( n + 15 ) >> 4;

    if ( ldlt ) {
        for ( int32_t i = tid; i < n; i += PXD_THREADS ) sDi[i] = (T)1 / L[c.coefoff + (int64_t)i * ( n + 1 )];
    }
    __shared__ int32_t skeys[PXD_ROWKEY_CAP];
    stage_row_keys( skeys, bk, c.fblok, lb );
    __syncthreads();
    if ( tid < 64 ) {
        int32_t s = t.r0 + tid;
        if ( s < c.stride ) {
            int32_t b  = search_row_keys( skeys, bk, c.fblok, lb, s );
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

    // B loader: row br, k values [bq * KV, bq * KV + KV) of a stage
    const int     br    = tid & 63, bq = tid >> 6;
    const int64_t lbase = sBase[br];
    const int32_t lld   = sLd[br];
    // Linv loader: column lc of the pass, k values [lq * KPC, lq * KPC + KPC)
    const int     lc = tid % CPT, lq = tid / CPT;

    const int npass = ( nt + NTP - 1 ) / NTP;
    for ( int pass = npass - 1; pass >= 0; pass-- ) {
        const int32_t cb0  = pass * CPT;                       // first column of the pass
        const int32_t lend = min( n, cb0 + CPT );              // Linv[c][l] = 0 for l > c
        acc_t acc[4][CPW];
#pragma unroll
        for ( int i = 0; i < 4; i++ )
#pragma unroll
            for ( int j = 0; j < CPW; j++ ) acc[i][j] = M::zero();

        // the next stage's loads are issued before the current stage's products
        T    rb[KV], rl[KPC];
        auto load = [&]( int32_t l0 ) {
#pragma unroll
            for ( int q = 0; q < KV; q++ ) {
                const int32_t l = l0 + bq * KV + q;
                rb[q] = ( lbase >= 0 && l < lend ) ? L[lbase + (int64_t)l * lld] : (T)0;
            }
#pragma unroll
            for ( int q = 0; q < KPC; q++ ) {
                const int32_t cc = cb0 + lc, l = l0 + lq * KPC + q;
                rl[q] = ( cc < n && l <= cc ) ? LI[cc + (int64_t)l * n] : (T)0;
            }
        };
        load( 0 );
        for ( int32_t l0 = 0; l0 < lend; l0 += BK ) {
            {
                vec_t v;
#pragma unroll
                for ( int e = 0; e < KV; e++ ) v[e] = rb[e];
                *(vec_t *)&sB[br * LDK + bq * KV] = v;
            }
#pragma unroll
            for ( int q = 0; q < KPC; q += KV ) {
                vec_t v;
#pragma unroll
                for ( int e = 0; e < KV; e++ ) v[e] = rl[q + e];
                *(vec_t *)&sLi[lc * LDK + lq * KPC + q] = v;
            }
            __syncthreads();
            if ( l0 + BK < lend ) {
                load( l0 + BK );
            }
            vec_t av[4], bv[CPW];
#pragma unroll
            for ( int i = 0; i < 4; i++ ) av[i] = *(const vec_t *)&sB[( 16 * i + lr ) * LDK + KV * lg];
#pragma unroll
            for ( int j = 0; j < CPW; j++ ) bv[j] = *(const vec_t *)&sLi[( 16 * ( w + 4 * j ) + lr ) * LDK + KV * lg];
#pragma unroll
            for ( int j = 0; j < CPW; j++ ) {
                const int32_t c0 = cb0 + 16 * ( w + 4 * j );   // first column of the tile
                if ( c0 < n && c0 + 15 >= l0 ) {
#pragma unroll
                    for ( int e = 0; e < KV; e++ )
#pragma unroll
                        for ( int i = 0; i < 4; i++ ) acc[i][j] = M::mma( av[i][e], bv[j][e], acc[i][j] );
                }
            }
            __syncthreads();
        }
#pragma unroll
        for ( int j = 0; j < CPW; j++ ) {
            const int cc = cb0 + 16 * ( w + 4 * j ) + lr;
#pragma unroll
            for ( int i = 0; i < 4; i++ )
#pragma unroll
                for ( int v = 0; v < 4; v++ ) {
                    const int r = 16 * i + M::row( lane, v );
                    if ( sBase[r] >= 0 && cc < n ) {
                        // LDL^T: L_off = B L_kk^{-T} D^{-1}
                        L[sBase[r] + (int64_t)cc * sLd[r]] = ldlt ? M::get( acc[i][j], v ) * sDi[cc] : M::get( acc[i][j], v );
                    }
                }
        }
        __syncthreads();   // the pass's stores precede the next pass's loads
    }
    stamp_end( stamps, ti, t_stamp );
}


// ---------------------------------------------------------------------------
// Supernodal update on the matrix cores: C[f(i)] -= A_i * A_b^T for one
// 64 x 64 output tile (four waves of 32 x 32), fused row remap into the
// facing cblk in the epilogue (atomic adds when another source of the same
// level may hit the same columns).
template <typename T>
__global__ void __launch_bounds__( PXD_THREADS )
k_update_mfma( const upd_task *__restrict__ tasks, const dcblk *__restrict__ cb,
               const dblok *__restrict__ bk, T *L, T *U, int ldlt,
               uint64_t *__restrict__ stamps, int prio )
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
    __shared__ search_keys sk;
    __shared__ T           sDk[PXD_TRSM_NT * 16];   // LDL^T: D of the source cblk

    const int32_t  ti   = xcd_task( blockIdx.x, gridDim.x );
    const upd_task t    = tasks[ti];
    const uint64_t t_stamp = stamp_begin( stamps );
    wave_priority( prio );
    const int      tid  = threadIdx.x;
    const int      lane = tid & 63, w = tid >> 6;
    const dcblk    c    = cb[t.cblk];
    const int32_t  K    = c.lcolnum - c.fcolnum + 1;
    const int32_t  lb   = cb[t.cblk + 1].fblok;
    const dblok    bb   = bk[t.blok];
    const int32_t  N    = bb.lrownum - bb.frownum + 1;
    // U-side tile (LU): A = U^T rows after blok b, B = L rows of blok b, C = U
    const bool     uside = ( t.flags & 2 ) != 0;
    const T       *LA   = uside ? U : L;
    const T       *LB   = uside ? L : U;
    T             *LC   = uside ? U : L;
    const int32_t  r0   = bb.rowoff + ( uside ? N : 0 );
    const int32_t  Mr   = c.stride - r0;
    const int32_t  F    = bb.fcblknm;
    const dcblk    cf   = cb[F];
    const int32_t  Fn   = cf.lcolnum - cf.fcolnum + 1;
    const int32_t  flb  = cb[F + 1].fblok;
    const int32_t  colF0 = bb.frownum - cf.fcolnum + t.tn * 64;
    const int64_t  bbase = c.coefoff + (int64_t)bb.rowoff * K + t.tn * 64;
    if ( bb.frownum < cf.fcolnum || bb.lrownum > cf.lcolnum ) {
        return;   // inconsistent structure: the host checks prevent this
    }

    stage_search_keys( sk, bk, t.blok, lb, cf.fblok, flb, tid );
    if ( ldlt ) {
        // LDL^T: C -= A_i (A_b D)^T (the reference's DL^T workspace, formed on the fly)
        for ( int32_t i = tid; i < K; i += PXD_THREADS ) sDk[i] = L[c.coefoff + (int64_t)i * ( K + 1 )];
    }
    __syncthreads();
    if ( tid < 64 ) {
        int32_t r = t.tm * 64 + tid;
        if ( r < Mr ) {
            int32_t s  = r0 + r;
            int32_t bi = search_row( sk, bk, t.blok, lb, s );
            dblok   ib = bk[bi];
            int32_t g  = ib.frownum + ( s - ib.rowoff );
            sSrc[tid]   = c.coefoff + (int64_t)ib.rowoff * K + ( s - ib.rowoff );
            sSrcLd[tid] = ib.lrownum - ib.frownum + 1;
            int32_t fi  = search_grow( sk, bk, cf.fblok, flb, g );
            dblok   fb  = bk[fi];
            bool    ok  = ( g >= fb.frownum && g <= fb.lrownum );
            sDst[tid]   = ok ? cf.coefoff + (int64_t)fb.rowoff * Fn + ( g - fb.frownum ) : -1;
            sDstLd[tid] = fb.lrownum - fb.frownum + 1;
        }
        else {
            sSrc[tid]   = -1;
            sSrcLd[tid] = 0;
            sDst[tid]   = -1;
            sDstLd[tid] = 0;
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
        T va[BK / 4], vb[BK / 4];
#pragma unroll
        for ( int u = 0; u < BK / 4; u++ ) {
            const int32_t kg = k0 + ( tid >> 6 ) + 4 * u;
            va[u] = ( kg < K && src >= 0 ) ? LA[src + (int64_t)kg * srcld] : (T)0;
            vb[u] = ( kg < K && lr < nvalid ) ? LB[bbase + lr + (int64_t)kg * N] : (T)0;
        }
#pragma unroll
        for ( int u = 0; u < BK / 4; u++ ) {
            const int32_t kg = k0 + ( tid >> 6 ) + 4 * u;
            sA[( tid >> 6 ) + 4 * u][lr] = va[u];
            sB[( tid >> 6 ) + 4 * u][lr] = ( ldlt && kg < K ) ? vb[u] * sDk[kg] : vb[u];
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

    // scatter into F: all loads of the read-modify-writes first, then the stores
    const bool atomic = ( t.flags & 1 );
    T         *p[2][4];
    bool       pv[2][2][4];
#pragma unroll
    for ( int i = 0; i < 2; i++ ) {
#pragma unroll
        for ( int v = 0; v < 4; v++ ) {
            const int r = wr + 16 * i + M::row( lane, v );
            p[i][v]     = LC + ( sDst[r] >= 0 ? sDst[r] : 0 ) + (int64_t)colF0 * sDstLd[r];
#pragma unroll
            for ( int j = 0; j < 2; j++ ) pv[i][j][v] = sDst[r] >= 0 && wc + 16 * j + ( lane & 15 ) < nvalid;
        }
    }
    T old[2][2][4];
#pragma unroll
    for ( int i = 0; i < 2; i++ )
#pragma unroll
        for ( int j = 0; j < 2; j++ )
#pragma unroll
            for ( int v = 0; v < 4; v++ ) {
                const int r  = wr + 16 * i + M::row( lane, v );
                const int cc = wc + 16 * j + ( lane & 15 );
                T *dst = p[i][v] + (int64_t)cc * sDstLd[r];
                if ( atomic ) {
                    if ( pv[i][j][v] ) px_atomic_add( dst, -M::get( acc[i][j], v ) );
                }
                else {
                    old[i][j][v] = pv[i][j][v] ? *dst : (T)0;
                }
            }
    if ( !atomic ) {
#pragma unroll
        for ( int i = 0; i < 2; i++ )
#pragma unroll
            for ( int j = 0; j < 2; j++ )
#pragma unroll
                for ( int v = 0; v < 4; v++ ) {
                    const int r  = wr + 16 * i + M::row( lane, v );
                    const int cc = wc + 16 * j + ( lane & 15 );
                    if ( pv[i][j][v] ) p[i][v][(int64_t)cc * sDstLd[r]] = old[i][j][v] - M::get( acc[i][j], v );
                }
    }
    stamp_end( stamps, ti, t_stamp );
}

// ---------------------------------------------------------------------------
// Small supernodal updates, one wavefront per 64 stacked rows of a blok
// update (the lower tree has hundreds of thousands of them with a few rows
// and columns): task {cblk, blok, tm, flags} covers the stacked rows
// [64 tm, 64 tm + 64) of an update with at most PXD_SMALL_M stacked rows
// (Mr), PXD_SMALL_N facing columns (N) and PXD_SMALL_K source columns (K).
// Lane r owns stacked row 64 tm + r and keeps its K values of A in
// registers; the facing columns go by chunks of PXD_SMALL_NC: the chunk of
// B (the rows of blok b, D-scaled for LDL^T) is staged in the wavefront's
// LDS area, C[f(r), chunk] -= A_r B_chunk^T, plain read-modify-write or
// atomic adds when the task is flagged.  The four wavefronts of a workgroup
// are independent (no workgroup barrier).
#define PXD_SMALL_M  256
#define PXD_SMALL_N  64
#define PXD_SMALL_NC 16
#define PXD_SMALL_K  32
template <typename T>
__global__ void __launch_bounds__( PXD_THREADS )
k_update_small( const upd_task *__restrict__ tasks, int32_t ntasks, const dcblk *__restrict__ cb,
                const dblok *__restrict__ bk, T *L, T *U, int ldlt, uint64_t *__restrict__ stamps, int prio )
{
    constexpr int NC = PXD_SMALL_NC;
    __shared__ T sB[PXD_THREADS / 64][PXD_SMALL_K][NC + 1];
    const uint64_t t_stamp = stamp_begin( stamps );
    wave_priority( prio );
    const int     lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int32_t ti   = blockIdx.x * ( PXD_THREADS / 64 ) + w;
    if ( ti < ntasks ) {
        const upd_task t  = tasks[ti];
        const dcblk    c  = cb[t.cblk];
        const int32_t  K  = c.lcolnum - c.fcolnum + 1;
        const int32_t  lb = cb[t.cblk + 1].fblok;
        const dblok    bb = bk[t.blok];
        const int32_t  N  = bb.lrownum - bb.frownum + 1;
        const bool     uside = ( t.flags & 2 ) != 0;
        const T       *LA = uside ? U : L;
        const T       *LB = uside ? L : U;
        T             *LC = uside ? U : L;
        const int32_t  r0 = bb.rowoff + ( uside ? N : 0 );
        const int32_t  Mr = c.stride - r0;
        const int32_t  F  = bb.fcblknm;
        const dcblk    cf = cb[F];
        const int32_t  Fn = cf.lcolnum - cf.fcolnum + 1;
        const int32_t  flb = cb[F + 1].fblok;
        const int32_t  colF0 = bb.frownum - cf.fcolnum;
        const int64_t  bbase = c.coefoff + (int64_t)bb.rowoff * K;
        const int32_t  row   = 64 * t.tm + lane;   // stacked row of the lane, from r0
        // the lane's row: its K values of A and its place in F
        T       a[PXD_SMALL_K];
        T      *dst = nullptr;
        int32_t dld = 0;
        if ( row < Mr ) {
            const int32_t s   = r0 + row;
            const int32_t bi  = find_blok_by_row( bk, t.blok, lb, s );
            const dblok   ib  = bk[bi];
            const int64_t src = c.coefoff + (int64_t)ib.rowoff * K + ( s - ib.rowoff );
            const int32_t sld = ib.lrownum - ib.frownum + 1;
            const int32_t g   = ib.frownum + ( s - ib.rowoff );
#pragma unroll
            for ( int kk = 0; kk < PXD_SMALL_K; kk++ ) a[kk] = ( kk < K ) ? LA[src + (int64_t)kk * sld] : (T)0;
            const int32_t fi = find_blok_by_grow( bk, cf.fblok, flb, g );
            const dblok   fb = bk[fi];
            if ( g >= fb.frownum && g <= fb.lrownum ) {
                dld = fb.lrownum - fb.frownum + 1;
                dst = LC + cf.coefoff + (int64_t)fb.rowoff * Fn + ( g - fb.frownum ) + (int64_t)colF0 * dld;
            }
        }
        for ( int32_t j0 = 0; j0 < N; j0 += NC ) {
            const int32_t nc = min( NC, N - j0 );
            // B[j][k] = LB[bbase + j + k N] for the chunk's columns, times D_k for LDL^T
            for ( int idx = lane; idx < nc * K; idx += 64 ) {
                const int j = idx % nc, kk = idx / nc;
                const T   v = LB[bbase + j0 + j + (int64_t)kk * N];
                sB[w][kk][j] = ldlt ? v * L[c.coefoff + (int64_t)kk * ( K + 1 )] : v;
            }
            __builtin_amdgcn_fence( __ATOMIC_RELEASE, "wavefront" );
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence( __ATOMIC_ACQUIRE, "wavefront" );
            if ( dst != nullptr ) {
                T acc[NC];
#pragma unroll
                for ( int j = 0; j < NC; j++ ) acc[j] = (T)0;
#pragma unroll
                for ( int kk = 0; kk < PXD_SMALL_K; kk++ ) {
                    if ( kk < K ) {
#pragma unroll
                        for ( int j = 0; j < NC; j++ ) acc[j] += a[kk] * sB[w][kk][j];
                    }
                }
                T *d0 = dst + (int64_t)j0 * dld;
                if ( t.flags & 1 ) {
#pragma unroll
                    for ( int j = 0; j < NC; j++ ) {
                        if ( j < nc ) px_atomic_add( d0 + (int64_t)j * dld, -acc[j] );
                    }
                }
                else {
                    T old[NC];
#pragma unroll
                    for ( int j = 0; j < NC; j++ ) old[j] = ( j < nc ) ? d0[(int64_t)j * dld] : (T)0;
#pragma unroll
                    for ( int j = 0; j < NC; j++ ) {
                        if ( j < nc ) d0[(int64_t)j * dld] = old[j] - acc[j];
                    }
                }
            }
            // every lane has read the chunk before the next one is staged
            __builtin_amdgcn_fence( __ATOMIC_RELEASE, "wavefront" );
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence( __ATOMIC_ACQUIRE, "wavefront" );
        }
    }
    stamp_end( stamps, blockIdx.x, t_stamp );
}

// ---------------------------------------------------------------------------
// Supernodal update for tall tiles: C[f(i)] -= A_i * A_b^T for upd_rows<T>
// stacked rows x PXD_UPD_BN facing columns, wave w owning rows
// [w*BM/4, (w+1)*BM/4).  The next k-stage is loaded into registers while the
// current one is on the matrix cores.  LDS holds each row's k values
// contiguously, and the k values of one MFMA step are permuted (lane group g
// takes k = KV*g + s in step s) so that every fragment is one 16-byte read.
// C[rows, cols] -= acc for one 16 x 64 row block of an update tile: row r of
// the block goes to L + dst[r] + col * ld[r], columns colF0 + [0, nvalid).
template <typename T>
__device__ inline void
update_scatter( T *L, const typename mfma16<T>::acc_t ( &acc )[4], int r0, int lane,
                const int64_t *dst, const int32_t *ld, int32_t colF0, int32_t nvalid, bool atomic )
{
    typedef mfma16<T> M;
    T      *p[4];
    int32_t pl[4];
    bool    pv[4];
#pragma unroll
    for ( int v = 0; v < 4; v++ ) {
        const int r = r0 + M::row( lane, v );
        pv[v]       = dst[r] >= 0;
        pl[v]       = ld[r];
        p[v]        = L + ( pv[v] ? dst[r] : 0 ) + (int64_t)colF0 * pl[v];
    }
    if ( atomic ) {
#pragma unroll
        for ( int j = 0; j < 4; j++ ) {
            const int cc = 16 * j + ( lane & 15 );
#pragma unroll
            for ( int v = 0; v < 4; v++ ) {
                if ( pv[v] && cc < nvalid ) px_atomic_add( p[v] + (int64_t)cc * pl[v], -M::get( acc[j], v ) );
            }
        }
        return;
    }
    T old[4][4];
#pragma unroll
    for ( int j = 0; j < 4; j++ ) {
        const int cc = 16 * j + ( lane & 15 );
#pragma unroll
        for ( int v = 0; v < 4; v++ ) old[j][v] = ( pv[v] && cc < nvalid ) ? p[v][(int64_t)cc * pl[v]] : (T)0;
    }
#pragma unroll
    for ( int j = 0; j < 4; j++ ) {
        const int cc = 16 * j + ( lane & 15 );
#pragma unroll
        for ( int v = 0; v < 4; v++ ) {
            if ( pv[v] && cc < nvalid ) p[v][(int64_t)cc * pl[v]] = old[j][v] - M::get( acc[j], v );
        }
    }
}

#define PXD_UPD_BN 64
template <typename T> struct upd_rows;
template <> struct upd_rows<float>  { static constexpr int value = 256; };
template <> struct upd_rows<double> { static constexpr int value = 128; };
template <> struct upd_rows<pxcf>   { static constexpr int value = 128; };
template <> struct upd_rows<pxcd>   { static constexpr int value = 64; };

template <typename T, int NH>
__global__ void __launch_bounds__( PXD_THREADS, ( NH == 1 ) ? 3 : 2 )
k_update_big( const upd_task *__restrict__ tasks, int32_t ntasks, const dcblk *__restrict__ cb,
              const dblok *__restrict__ bk, T *L, T *U, int ldlt,
              uint64_t *__restrict__ stamps, int prio )
{
    typedef mfma16<T>          M;
    typedef typename M::acc_t  acc_t;
    typedef typename lds_vec<T>::type vec_t;
    constexpr int BM  = upd_rows<T>::value;
    constexpr int KV  = lds_vec<T>::n;    // k values per fragment
    constexpr int BK  = NH * 4 * KV;      // k values per stage (NH MFMA k-groups of 4*KV)
    constexpr int LDK = BK + KV;          // row stride of the staged tiles
    constexpr int TPR = PXD_THREADS / BM; // loader threads per row
    constexpr int KPT = BK / TPR;         // k values loaded per thread
    constexpr int RB  = BM / 64;          // 16-row blocks per wave
    constexpr uint32_t ES = sizeof( T );
    constexpr int WR  = 16 * RB;          // rows per wave
    constexpr int SE  = ( BM * LDK > 4 * WR * 17 ) ? BM * LDK : 4 * WR * 17;
    __shared__ __align__( 16 ) T sA[SE];  // staged A tile; after the k loop, the epilogue's transposes
    __shared__ __align__( 16 ) T sB[PXD_UPD_BN * LDK];
    __shared__ uint32_t sDst[BM];         // byte offset of the row in F (PXD_OOB: none)
    __shared__ uint32_t sDstLd[BM];       // byte stride between F's columns for the row
    __shared__ search_keys sk;
    __shared__ T           sDk[PXD_TRSM_NT * 16];   // LDL^T: D of the source cblk

    const int32_t  ti   = xcd_task( blockIdx.x, ntasks );
    const upd_task t    = tasks[ti];
    const uint64_t t_stamp = stamp_begin( stamps );
    wave_priority( prio );
    const int      tid  = threadIdx.x;
    const int      lane = tid & 63, w = tid >> 6;
    const dcblk    c    = cb[t.cblk];
    const int32_t  K    = c.lcolnum - c.fcolnum + 1;
    const int32_t  lb   = cb[t.cblk + 1].fblok;
    const dblok    bb   = bk[t.blok];
    const int32_t  N    = bb.lrownum - bb.frownum + 1;
    // U-side tile (LU): A = U^T rows after blok b, B = L rows of blok b, C = U
    const bool     uside = ( t.flags & 2 ) != 0;
    const T       *LA   = uside ? U : L;
    const T       *LB   = uside ? L : U;
    T             *LC   = uside ? U : L;
    const int32_t  r0   = bb.rowoff + ( uside ? N : 0 );
    const int32_t  Mr   = c.stride - r0;
    const int32_t  F    = bb.fcblknm;
    const dcblk    cf   = cb[F];
    const int32_t  Fn   = cf.lcolnum - cf.fcolnum + 1;
    const int32_t  flb  = cb[F + 1].fblok;
    const int32_t  colF0 = bb.frownum - cf.fcolnum + t.tn * PXD_UPD_BN;
    if ( bb.frownum < cf.fcolnum || bb.lrownum > cf.lcolnum ) {
        return;   // inconsistent structure: the host checks prevent this
    }
    const __amdgpu_buffer_rsrc_t rA = make_rsrc( LA + c.coefoff, (uint32_t)( (int64_t)c.stride * K * ES ) );
    const __amdgpu_buffer_rsrc_t rB = make_rsrc( LB + c.coefoff, (uint32_t)( (int64_t)c.stride * K * ES ) );
    const __amdgpu_buffer_rsrc_t rC = make_rsrc( LC + cf.coefoff, (uint32_t)( (int64_t)cf.stride * Fn * ES ) );

    stage_search_keys( sk, bk, t.blok, lb, cf.fblok, flb, tid );
    if ( ldlt ) {
        // LDL^T: C -= A_i (A_b D)^T (the reference's DL^T workspace, formed on the fly)
        for ( int32_t i = tid; i < K; i += PXD_THREADS ) sDk[i] = L[c.coefoff + (int64_t)i * ( K + 1 )];
    }
    __syncthreads();

    // loader row ar (k values [ap*KPT, ap*KPT+KPT) of each stage): byte
    // offset of its first element and byte stride between its columns; the
    // threads with ap == 0 also map the row into F
    const int ar = tid % BM, ap = tid / BM;
    uint32_t  abase = PXD_OOB, astep = 0;
    {
        const int32_t r   = t.tm * BM + ar;
        uint32_t      dst = PXD_OOB, dld = 0;
        if ( r < Mr ) {
            const int32_t s  = r0 + r;
            const dblok   ib = bk[search_row( sk, bk, t.blok, lb, s )];
            const int32_t g  = ib.frownum + ( s - ib.rowoff );
            abase = (uint32_t)( ib.rowoff * K + ( s - ib.rowoff ) ) * ES;
            astep = (uint32_t)( ib.lrownum - ib.frownum + 1 ) * ES;
            if ( ap == 0 ) {
                const dblok fb = bk[search_grow( sk, bk, cf.fblok, flb, g )];
                if ( g >= fb.frownum && g <= fb.lrownum ) {
                    dst = (uint32_t)( fb.rowoff * Fn + ( g - fb.frownum ) ) * ES;
                    dld = (uint32_t)( fb.lrownum - fb.frownum + 1 ) * ES;
                }
            }
        }
        if ( ap == 0 ) {
            sDst[ar]   = dst;
            sDstLd[ar] = dld;
        }
    }
    // B loader: facing column bc, k group bq
    const int      bc     = tid & ( PXD_UPD_BN - 1 ), bq = tid / PXD_UPD_BN;
    const int32_t  nvalid = min( PXD_UPD_BN, N - t.tn * PXD_UPD_BN );
    const uint32_t bbase  = ( bc < nvalid ) ? (uint32_t)( bb.rowoff * K + t.tn * PXD_UPD_BN + bc ) * ES : PXD_OOB;
    const uint32_t bstep  = ( bc < nvalid ) ? (uint32_t)N * ES : 0;

    T ra[KPT], rb[NH * KV];
    auto load = [&]( int32_t k0 ) {
#pragma unroll
        for ( int q = 0; q < KPT; q++ ) {
            const int32_t kg = k0 + ap * KPT + q;
            ra[q] = bufio<T>::load( rA, ( kg < K ) ? abase + (uint32_t)kg * astep : PXD_OOB );
        }
#pragma unroll
        for ( int q = 0; q < NH * KV; q++ ) {
            const int32_t kg = k0 + bq * NH * KV + q;
            rb[q] = bufio<T>::load( rB, ( kg < K ) ? bbase + (uint32_t)kg * bstep : PXD_OOB );
        }
    };

    acc_t acc[RB][4];
#pragma unroll
    for ( int i = 0; i < RB; i++ )
#pragma unroll
        for ( int j = 0; j < 4; j++ ) acc[i][j] = M::zero();

    const int lr = lane & 15, lg = lane >> 4;
    load( 0 );
    for ( int32_t k0 = 0; k0 < K; k0 += BK ) {
#pragma unroll
        for ( int q = 0; q < KPT; q += KV ) {
            vec_t v;
#pragma unroll
            for ( int e = 0; e < KV; e++ ) v[e] = ra[q + e];
            *(vec_t *)&sA[ar * LDK + ap * KPT + q] = v;
        }
#pragma unroll
        for ( int h = 0; h < NH; h++ ) {
            vec_t v;
#pragma unroll
            for ( int e = 0; e < KV; e++ ) {
                const int32_t kg = k0 + bq * NH * KV + h * KV + e;
                v[e] = ( ldlt && kg < K ) ? rb[h * KV + e] * sDk[kg] : rb[h * KV + e];
            }
            *(vec_t *)&sB[bc * LDK + bq * NH * KV + h * KV] = v;
        }
        __syncthreads();
        if ( k0 + BK < K ) {
            load( k0 + BK );
        }
#pragma unroll
        for ( int h = 0; h < NH; h++ ) {
            vec_t av[RB], bv[4];
#pragma unroll
            for ( int i = 0; i < RB; i++ )
                av[i] = *(const vec_t *)&sA[( 16 * ( RB * w + i ) + lr ) * LDK + 4 * KV * h + KV * lg];
#pragma unroll
            for ( int j = 0; j < 4; j++ ) bv[j] = *(const vec_t *)&sB[( 16 * j + lr ) * LDK + 4 * KV * h + KV * lg];
#pragma unroll
            for ( int s = 0; s < KV; s++ )
#pragma unroll
                for ( int i = 0; i < RB; i++ )
#pragma unroll
                    for ( int j = 0; j < 4; j++ ) acc[i][j] = M::mma( av[i][s], bv[j][s], acc[i][j] );
        }
        __syncthreads();
    }

    // scatter into F, one 16-column slab at a time: each wave transposes its
    // WR x 16 block through LDS so that a lane owns one row and every memory
    // instruction covers WR consecutive rows of one column (coalesced, as F is
    // column major).  Plain tiles issue all loads of a slab, then the stores
    // (out-of-range rows and columns read 0 and drop their stores); tiles
    // flagged atomic use atomic adds.
    const bool    atomic = ( t.flags & 1 );
    T            *sE     = sA + w * ( WR * 17 );
    constexpr int LPR    = 64 / WR;        // lanes per row
    constexpr int CPL    = 16 / LPR;       // columns per lane
    const int     erow   = lane % WR, epart = lane / WR;
    const uint32_t rowoff = sDst[WR * w + erow];
    const uint32_t rowld  = sDstLd[WR * w + erow];
#pragma unroll
    for ( int j = 0; j < 4; j++ ) {
#pragma unroll
        for ( int i = 0; i < RB; i++ )
#pragma unroll
            for ( int v = 0; v < 4; v++ ) sE[( 16 * i + M::row( lane, v ) ) * 17 + lr] = M::get( acc[i][j], v );
        __syncthreads();
        uint32_t off[CPL];
        T        val[CPL];
#pragma unroll
        for ( int c = 0; c < CPL; c++ ) {
            const int col = epart * CPL + c, cc = 16 * j + col;
            off[c] = ( rowoff != PXD_OOB && cc < nvalid ) ? rowoff + (uint32_t)( colF0 + cc ) * rowld : PXD_OOB;
            val[c] = sE[erow * 17 + col];
        }
        if ( atomic ) {
#pragma unroll
            for ( int c = 0; c < CPL; c++ ) {
                if ( off[c] != PXD_OOB ) px_atomic_add( LC + cf.coefoff + off[c] / ES, -val[c] );
            }
        }
        else {
            T old[CPL];
#pragma unroll
            for ( int c = 0; c < CPL; c++ ) old[c] = bufio<T>::load( rC, off[c] );
#pragma unroll
            for ( int c = 0; c < CPL; c++ ) bufio<T>::store( old[c] - val[c], rC, off[c] );
        }
        __syncthreads();
    }
    stamp_end( stamps, ti, t_stamp );
}


// ---------------------------------------------------------------------------
// Single-precision tall-tile update (the 256 x 64 tiles of k_update_big<float>,
// real LL^T and LU) with its operands moved from HBM straight into LDS by
// buffer loads that write LDS (no register staging), in a ring of NBUF
// buffers of BK k values: while a stage's products run, the loads of the
// next NBUF - 1 stages are in flight, and each stage ends with one barrier.  The staged tiles are k
// major (sA[k][256 rows], sB[k][64 columns]), so a load instruction moves one
// k value of 64 consecutive rows (one wave) and every lane writes LDS at its
// own consecutive word.  A 16-byte LDS read then gives a lane one k value of
// four rows (or columns): the wave's 64 x 64 block is cut into 16 x 16
// matrix-core tiles by interleaving, tile i (j) taking the rows (columns)
// 4 r + i of the block, so that the four values of one read feed the four
// row (column) tiles of one k step.  The scatter into the facing cblk is the
// one of k_update_big: a 16-column slab at a time, transposed through LDS so
// that a lane owns one row and each memory instruction covers 64 consecutive
// rows of one column.
//
// The row mapped by thread tid is row tid of the tile, for the loads and for
// the scatter alike, so its offsets stay in registers.  The LDS holds the
// ring and nothing else (NBUF x BK x 1280 bytes); the search keys live in the
// last buffer until its first load, and the scatter's transposes in the A
// buffers after the last stage.
#define PXD_DMA_BM 256
// measured on the 100^3 FP32 factorization: 8 k values in a ring of 3 (loads
// of the next two stages in flight, 30 KiB) ahead of 16 x 2 (40 KiB), 12 x 2,
// 12 x 3 and 8 x 4
#define PXD_DMA_BK   8
#define PXD_DMA_NBUF 3
template <int BK, int NBUF>
__global__ void __launch_bounds__( PXD_THREADS, 4 )
k_update_dma( const upd_task *__restrict__ tasks, int32_t ntasks, const dcblk *__restrict__ cb,
              const dblok *__restrict__ bk, float *L, float *U, uint64_t *__restrict__ stamps, int prio )
{
    typedef mfma16<float> M;
    typedef M::acc_t      acc_t;
    typedef float f4 __attribute__( ( ext_vector_type( 4 ) ) );
    constexpr int      BM = PXD_DMA_BM, BN = PXD_UPD_BN;
    constexpr uint32_t ES = sizeof( float );
    static_assert( BK % 4 == 0, "a stage is a whole number of 4-k matrix-core steps (one per wave for B)" );
    static_assert( NBUF >= 2 && NBUF <= 4, "ring of 2 to 4 stages" );
    static_assert( PXD_THREADS == BM, "thread tid maps row tid of the tile" );
    static_assert( upd_rows<float>::value == BM, "the tall-tile list of a single factor has 256-row tiles" );
    static_assert( sizeof( search_keys ) <= BK * BM * sizeof( float ), "search keys fit in one A buffer" );
    static_assert( 4 * 64 * 17 <= NBUF * BK * BM, "scatter transposes fit in the A buffers" );
    constexpr int LPS = BK + BK / 4;                    // loads per stage and wave
    constexpr int VMN = LPS * ( NBUF - 2 );             // loads issued after the next stage's
    static_assert( VMN < 64, "vmcnt holds 6 bits" );
    constexpr int WAIT_NEXT = ( VMN & 15 ) | ( ( VMN >> 4 ) << 14 ) | 0x0F70;   // s_waitcnt vmcnt(VMN)
    __shared__ __align__( 16 ) float smem[NBUF * BK * ( BM + BN )];
    search_keys &sk = *reinterpret_cast<search_keys *>( smem + ( NBUF - 1 ) * BK * BM );

    const int32_t  ti      = xcd_task( blockIdx.x, ntasks );
    const upd_task t       = tasks[ti];
    const uint64_t t_stamp = stamp_begin( stamps );
    wave_priority( prio );
    const int     tid = threadIdx.x;
    const int     lane = tid & 63, w = tid >> 6;
    const dcblk   c   = cb[t.cblk];
    const int32_t K   = c.lcolnum - c.fcolnum + 1;
    const int32_t lb  = cb[t.cblk + 1].fblok;
    const dblok   bb  = bk[t.blok];
    const int32_t N   = bb.lrownum - bb.frownum + 1;
    // U-side tile (LU): A = U^T rows after blok b, B = L rows of blok b, C = U
    const bool    uside = ( t.flags & 2 ) != 0;
    const float  *LA    = uside ? U : L;
    const float  *LB    = uside ? L : U;
    float        *LC    = uside ? U : L;
    const int32_t r0    = bb.rowoff + ( uside ? N : 0 );
    const int32_t Mr    = c.stride - r0;
    const int32_t F     = bb.fcblknm;
    const dcblk   cf    = cb[F];
    const int32_t Fn    = cf.lcolnum - cf.fcolnum + 1;
    const int32_t flb   = cb[F + 1].fblok;
    const int32_t colF0 = bb.frownum - cf.fcolnum + t.tn * BN;
    if ( bb.frownum < cf.fcolnum || bb.lrownum > cf.lcolnum ) {
        return;   // inconsistent structure: the host checks prevent this
    }
    const __amdgpu_buffer_rsrc_t rA = make_rsrc( LA + c.coefoff, (uint32_t)( (int64_t)c.stride * K * ES ) );
    const __amdgpu_buffer_rsrc_t rB = make_rsrc( LB + c.coefoff, (uint32_t)( (int64_t)c.stride * K * ES ) );
    const __amdgpu_buffer_rsrc_t rC = make_rsrc( LC + cf.coefoff, (uint32_t)( (int64_t)cf.stride * Fn * ES ) );

    stage_search_keys( sk, bk, t.blok, lb, cf.fblok, flb, tid );
    __syncthreads();

    // row tid of the tile: byte offset of its first element and byte stride
    // between its columns in the source, and its place in F
    uint32_t abase = PXD_OOB, astep = 0, dst = PXD_OOB, dld = 0;
    {
        const int32_t r = t.tm * BM + tid;
        if ( r < Mr ) {
            const int32_t s  = r0 + r;
            const dblok   ib = bk[search_row( sk, bk, t.blok, lb, s )];
            const int32_t g  = ib.frownum + ( s - ib.rowoff );
            abase = (uint32_t)( ib.rowoff * K + ( s - ib.rowoff ) ) * ES;
            astep = (uint32_t)( ib.lrownum - ib.frownum + 1 ) * ES;
            const dblok fb = bk[search_grow( sk, bk, cf.fblok, flb, g )];
            if ( g >= fb.frownum && g <= fb.lrownum ) {
                dst = (uint32_t)( fb.rowoff * Fn + ( g - fb.frownum ) ) * ES;
                dld = (uint32_t)( fb.lrownum - fb.frownum + 1 ) * ES;
            }
        }
    }
    // facing column `lane` of the tile
    const int32_t  nvalid = min( BN, N - t.tn * BN );
    const uint32_t bbase  = ( lane < nvalid ) ? (uint32_t)( bb.rowoff * K + t.tn * BN + lane ) * ES : PXD_OOB;
    const uint32_t bstep  = ( lane < nvalid ) ? (uint32_t)N * ES : 0;

    // stage k0 into buffer b: wave w loads the BK k values of rows
    // [64 w, 64 w + 64) and k values w, w + 4, ... of the 64 columns.  The
    // offsets advance by one column per load (ao, bo carried from stage to
    // stage), so that no per-k products stay live across the loop.
    const int wu = __builtin_amdgcn_readfirstlane( w );
    uint32_t  ao = abase, bo = ( bbase == PXD_OOB ) ? PXD_OOB : bbase + (uint32_t)wu * bstep;
    auto dma = [&]( auto bc, int32_t k0 ) {
        constexpr int b  = decltype( bc )::value;
        float        *dA = smem + b * ( BK * BM ) + 64 * wu;
        float        *dB = smem + NBUF * BK * BM + b * ( BK * BN ) + wu * BN;
        const bool    full = ( k0 + BK <= K );
#pragma unroll
        for ( int q = 0; q < BK; q++ ) {
            const uint32_t o = ( full || k0 + q < K ) ? ao : PXD_OOB;
            __builtin_amdgcn_raw_ptr_buffer_load_lds( rA, (__attribute__( ( address_space( 3 ) ) ) void *)( dA + q * BM ),
                                                      4, o, 0, 0, 0 );
            ao += astep;
        }
#pragma unroll
        for ( int q = 0; q < BK / 4; q++ ) {
            const uint32_t o = ( full || k0 + 4 * q + wu < K ) ? bo : PXD_OOB;
            __builtin_amdgcn_raw_ptr_buffer_load_lds(
                rB, (__attribute__( ( address_space( 3 ) ) ) void *)( dB + 4 * q * BN ), 4, o, 0, 0, 0 );
            bo += 4 * bstep;
        }
    };

    acc_t acc[4][4];
#pragma unroll
    for ( int i = 0; i < 4; i++ )
#pragma unroll
        for ( int j = 0; j < 4; j++ ) acc[i][j] = M::zero();

    const int     lr = lane & 15, lg = lane >> 4;
    const int32_t ns = ( K + BK - 1 ) / BK;
    // stage s from buffer b = s % NBUF, with the loads of stage s + NBUF - 1
    // into the buffer that stage s - 1 read (issued past the last stage too,
    // as out-of-range loads, so that every stage leaves the same count)
    auto stage = [&]( auto bc, int32_t s ) {
        constexpr int b  = decltype( bc )::value;
        const float  *SA = smem + b * ( BK * BM ) + 64 * w + 4 * lr;
        const float  *SB = smem + NBUF * BK * BM + b * ( BK * BN ) + 4 * lr;
        f4            av[BK / 4], bv[BK / 4];
#pragma unroll
        for ( int ks = 0; ks < BK / 4; ks++ ) {
            av[ks] = *(const f4 *)&SA[( 4 * ks + lg ) * BM];
            bv[ks] = *(const f4 *)&SB[( 4 * ks + lg ) * BN];
        }
        if ( NBUF > 2 || s + 1 < ns ) {
            dma( std::integral_constant<int, ( b + NBUF - 1 ) % NBUF>{}, ( s + NBUF - 1 ) * BK );
        }
#pragma unroll
        for ( int ks = 0; ks < BK / 4; ks++ )
#pragma unroll
            for ( int i = 0; i < 4; i++ )
#pragma unroll
                for ( int j = 0; j < 4; j++ ) acc[i][j] = M::mma( av[ks][i], bv[ks][j], acc[i][j] );
        // the products stay ahead of the wait for the next stage's loads
        __builtin_amdgcn_sched_barrier( 0 );
        __builtin_amdgcn_s_waitcnt( WAIT_NEXT );   // this wave's loads of stage s + 1 are in LDS
        __syncthreads();
    };
    dma( std::integral_constant<int, 0>{}, 0 );
    if constexpr ( NBUF > 2 ) dma( std::integral_constant<int, 1>{}, BK );
    if constexpr ( NBUF > 3 ) dma( std::integral_constant<int, 2>{}, 2 * BK );
    __builtin_amdgcn_s_waitcnt( WAIT_NEXT );
    __syncthreads();
    for ( int32_t s = 0; s < ns; s += NBUF ) {
        stage( std::integral_constant<int, 0>{}, s );
        if ( s + 1 < ns ) stage( std::integral_constant<int, 1>{}, s + 1 );
        if constexpr ( NBUF > 2 ) {
            if ( s + 2 < ns ) stage( std::integral_constant<int, 2>{}, s + 2 );
        }
        if constexpr ( NBUF > 3 ) {
            if ( s + 3 < ns ) stage( std::integral_constant<int, 3>{}, s + 3 );
        }
    }
    if constexpr ( NBUF > 2 ) {
        // the out-of-range loads past the last stage land before the scatter
        // reuses the buffers
        __builtin_amdgcn_s_waitcnt( 0x0F70 );
        __syncthreads();
    }

    // scatter: slab j holds the columns 4 r + j of the wave's block
    const bool atomic = ( t.flags & 1 );
    float     *sE     = smem + w * ( 64 * 17 );
#pragma unroll
    for ( int j = 0; j < 4; j++ ) {
#pragma unroll
        for ( int i = 0; i < 4; i++ )
#pragma unroll
            for ( int v = 0; v < 4; v++ ) sE[( 4 * M::row( lane, v ) + i ) * 17 + lr] = M::get( acc[i][j], v );
        __syncthreads();
        // two halves of 8 columns (the other slabs' accumulators are live)
#pragma unroll
        for ( int h = 0; h < 16; h += 8 ) {
            uint32_t off[8];
            float    val[8];
#pragma unroll
            for ( int q = 0; q < 8; q++ ) {
                const int cc = 4 * ( h + q ) + j;
                off[q] = ( dst != PXD_OOB && cc < nvalid ) ? dst + (uint32_t)( colF0 + cc ) * dld : PXD_OOB;
                val[q] = sE[lane * 17 + h + q];
            }
            if ( atomic ) {
#pragma unroll
                for ( int q = 0; q < 8; q++ ) {
                    if ( off[q] != PXD_OOB ) px_atomic_add( LC + cf.coefoff + off[q] / ES, -val[q] );
                }
            }
            else {
                float old[8];
#pragma unroll
                for ( int q = 0; q < 8; q++ ) old[q] = bufio<float>::load( rC, off[q] );
#pragma unroll
                for ( int q = 0; q < 8; q++ ) bufio<float>::store( old[q] - val[q], rC, off[q] );
            }
        }
        __syncthreads();
    }
    stamp_end( stamps, ti, t_stamp );
}
