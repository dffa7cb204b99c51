// kernels_sfsolve.hpp -- supernodal triangular solves in one launch per sweep
// (no level barriers), for the matrix-core engine.
//
// Forward sweep (reference kernels/solve_ztrsmsp.c:300-424): the task list
// follows the elimination order; cblk k contributes its diagonal chunks
// (16 rows of y_k = L_kk^{-1} x_k each, written to Y) and then its
// off-diagonal row tiles (64 stacked rows: x[rows] -= A_tile y_k, atomic
// adds into X).  A chunk starts when every tile that updates rows of k has
// finished (cnt[k] reaches need[k] per pass); a tile starts when all chunks
// of its cblk are done (rdy[k]).  Backward sweep (:449-625): the list runs
// the cblks in reverse; a tile gathers y_k -= A_tile^T x[rows] once the
// cblks its rows face have their final x (their chunks are done), and the
// chunks of x_k = L_kk^{-T} y_k wait for the tiles of k.  A small cblk (one
// chunk, at most one tile) is one task of both parts.
//
// Every wait is on a task earlier in the list and workgroups are dispatched
// in list order, so the lowest unfinished task can always run and the grid
// drains.  Counters grow by one pass's worth per pass (targets epoch *
// count, compared modulo 2^32), so they need no clearing between passes.
// Waits are bounded (2 s of the constant clock); a timeout sets a sticky
// error flag that makes every later wait fall through.
#pragma once
#include "kernels_mfma.hpp"

enum { SF_DIAG = 0, SF_TILE = 1, SF_LEAF = 2 };
struct sf_task {
    int32_t kind;
    int32_t cblk;
    int32_t r0;   // SF_DIAG: first row of the chunk; SF_TILE: first stacked off-diagonal row
    int32_t f0;   // SF_TILE, SF_LEAF: the cblks the tile's rows face, facing[f0 .. f1)
    int32_t f1;
    int32_t pad;
};
#define PXD_SF_ROWS 16   // rows of a diagonal chunk
#define PXD_SF_TILE 64   // stacked rows of an off-diagonal tile

__device__ inline float  sf_shfl_xor( float v, int m ) { return __shfl_xor( v, m, 64 ); }
__device__ inline double sf_shfl_xor( double v, int m ) { return __shfl_xor( v, m, 64 ); }
__device__ inline pxcf   sf_shfl_xor( pxcf v, int m ) { return pxcf( __shfl_xor( v.re, m, 64 ), __shfl_xor( v.im, m, 64 ) ); }
__device__ inline pxcd   sf_shfl_xor( pxcd v, int m ) { return pxcd( __shfl_xor( v.re, m, 64 ), __shfl_xor( v.im, m, 64 ) ); }

// thread 0: wait (bounded) until the counter reaches target.  Hundreds of
// workgroups can wait at once; each polls about every 0.5 us so that the
// polls do not crowd the counters' memory channel.
__device__ inline void
sf_spin( const uint32_t *c, uint32_t target, int *err )
{
    if ( (int32_t)( __hip_atomic_load( c, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT ) - target ) >= 0 ) {
        return;
    }
    const uint64_t t0 = (uint64_t)wall_clock64();
    do {
        __builtin_amdgcn_s_sleep( 20 );
        if ( __hip_atomic_load( err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT ) != 0 ) {
            break;
        }
        if ( (uint64_t)wall_clock64() - t0 > PXD_PANEL_WAIT_TICKS ) {
            __hip_atomic_store( err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT );
            break;
        }
    } while ( (int32_t)( __hip_atomic_load( c, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT ) - target ) < 0 );
}

// after thread 0's waits: the producers' writes become visible to the workgroup
__device__ inline void
sf_acquire( int tid )
{
    if ( tid == 0 ) {
        __builtin_amdgcn_fence( __ATOMIC_ACQUIRE, "agent" );
    }
    __syncthreads();
    __builtin_amdgcn_fence( __ATOMIC_ACQUIRE, "workgroup" );
}

// before thread 0 signals: every thread's stores and atomics are performed
__device__ inline void
sf_release( int tid )
{
    __builtin_amdgcn_fence( __ATOMIC_RELEASE, "workgroup" );
    __syncthreads();
    if ( tid == 0 ) {
        __builtin_amdgcn_fence( __ATOMIC_RELEASE, "agent" );
    }
}

__device__ inline void
sf_signal( uint32_t *c )
{
    __hip_atomic_fetch_add( c, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT );
}

__device__ inline int32_t
sf_nchunk( int32_t n )
{
    return ( n + PXD_SF_ROWS - 1 ) / PXD_SF_ROWS;
}
__device__ inline int32_t
sf_ntile( int32_t n, int32_t stride )
{
    return ( stride - n + PXD_SF_TILE - 1 ) / PXD_SF_TILE;
}

// Stage rows [j0, j1) of the NR columns of v (leading dimension ld) in LDS
// as s[j * NR + q]; columns q >= nrhs read 0.  When sub is given, s holds
// v - sub (sub[j * NR + q] in LDS).
template <typename T, int NR>
__device__ inline void
sf_stage( T *s, const T *v, int64_t ld, int32_t j0, int32_t j1, int nrhs, const T *sub, int tid )
{
    const int32_t len = j1 - j0;
    for ( int idx = tid; idx < len * NR; idx += PXD_THREADS ) {
        const int32_t j = j0 + idx % len;
        const int     q = idx / len;
        T             x = ( q < nrhs ) ? v[j + q * ld] : (T)0;
        if ( sub ) x -= sub[j * NR + q];
        s[j * NR + q] = x;
    }
}

// Rows R0 .. R0+15 of out = Linv in (forward, lower) or Linv^T in
// (backward, upper) for the staged in (sb[j * NR + q]); the n x n inverse
// LI of the diagonal block, ld n.  Thread (r, g) sums columns g, g + 16, ...
// of row R0 + r; the rows go to out (ld ldo) and, if keep is given, to
// keep[(R0 + r) * NR + q] in LDS.  sred: 4 x 16 x NR values of LDS.
template <typename T, typename V, int NR>
__device__ inline void
sf_diag_rows( const T *__restrict__ LI, int32_t n, int32_t R0, bool trans, const V *sb, V *sred, V *out,
              int64_t ldo, int nrhs, V *keep, int tid )
{
    constexpr int CR   = PXD_SF_ROWS;
    constexpr int NG   = PXD_THREADS / CR;
    constexpr int MAXU = ( PXD_TRSM_NT * 16 + NG - 1 ) / NG;
    const int     lane = tid & 63, w = tid >> 6;
    const int     r    = trans ? tid / NG : tid % CR;
    const int     g    = trans ? tid % NG : tid / CR;
    const int32_t i    = R0 + r;
    T             a[MAXU];
#pragma unroll
    for ( int u = 0; u < MAXU; u++ ) {
        const int32_t j  = g + NG * u;
        const bool    ok = i < n && j < n && ( trans ? j >= i : j <= i );
        a[u] = ok ? LI[trans ? j + (int64_t)i * n : i + (int64_t)j * n] : (T)0;
    }
    V acc[NR];
#pragma unroll
    for ( int q = 0; q < NR; q++ ) acc[q] = (V)0;
#pragma unroll
    for ( int u = 0; u < MAXU; u++ ) {
        const int32_t j = g + NG * u;
        if ( j < n && ( trans ? j >= i : j <= i ) ) {
            const V av = px_cvt<V>( a[u] );
#pragma unroll
            for ( int q = 0; q < NR; q++ ) acc[q] += av * sb[j * NR + q];
        }
    }
    if ( trans ) {
        // a wave holds rows 4w .. 4w+3, the 16 column phases of a row in 16 lanes
#pragma unroll
        for ( int q = 0; q < NR; q++ ) {
#pragma unroll
            for ( int m = 1; m < 16; m <<= 1 ) acc[q] += sf_shfl_xor( acc[q], m );
        }
        if ( ( lane & 15 ) == 0 && i < n ) {
#pragma unroll
            for ( int q = 0; q < NR; q++ ) {
                if ( q < nrhs ) out[i + q * ldo] = acc[q];
                if ( keep ) keep[i * NR + q] = acc[q];
            }
        }
    }
    else {
        // a wave holds the 16 rows for column phases 4w .. 4w+3
#pragma unroll
        for ( int q = 0; q < NR; q++ ) {
            acc[q] += sf_shfl_xor( acc[q], 16 );
            acc[q] += sf_shfl_xor( acc[q], 32 );
        }
        if ( lane < 16 ) {
#pragma unroll
            for ( int q = 0; q < NR; q++ ) sred[( w * CR + lane ) * NR + q] = acc[q];
        }
        __syncthreads();
        for ( int idx = tid; idx < CR * NR; idx += PXD_THREADS ) {
            const int r2 = idx % CR, q = idx / CR;
            const V   v  = sred[r2 * NR + q] + sred[( CR + r2 ) * NR + q] + sred[( 2 * CR + r2 ) * NR + q] +
                        sred[( 3 * CR + r2 ) * NR + q];
            if ( R0 + r2 < n ) {
                if ( q < nrhs ) out[R0 + r2 + q * ldo] = v;
                if ( keep ) keep[( R0 + r2 ) * NR + q] = v;
            }
        }
    }
}

// Row mapping of a tile: stacked row s = r0 + r of cblk c lies in one of
// its bloks; base = offset of the row's first column in the factor, ld =
// the blok's rows (column stride), g = the global row.  base < 0: no row.
template <typename T>
__device__ inline void
sf_row( const int32_t *keys, const dcblk &c, int32_t k, const dcblk *__restrict__ cb, const dblok *__restrict__ bk,
        int32_t n, int32_t s, int64_t &base, int32_t &ld, int32_t &g )
{
    base = -1;
    ld   = 0;
    g    = 0;
    if ( s < c.stride ) {
        const int32_t lb = cb[k + 1].fblok;
        const int32_t b  = search_row_keys( keys, bk, c.fblok, lb, s );
        const dblok   bb = bk[b];
        base = c.coefoff + (int64_t)bb.rowoff * n + ( s - bb.rowoff );
        ld   = bb.lrownum - bb.frownum + 1;
        g    = bb.frownum + ( s - bb.rowoff );
    }
}

// Forward tile: X[g(s)] -= A(s, :) y_k for the 64 stacked rows s = r0 ..
// r0 + 63 of cblk k; y_k staged in sx[j * NR + q].  Lane r of each wave
// takes row r0 + r; wave w the columns [n w / 4, n (w + 1) / 4), loaded B
// at a time.  red: 4 x 64 x NR values of LDS (may alias sx: it is written
// after a barrier that follows the last read of sx).
template <typename T, typename V, int NR, int B>
__device__ inline void
sf_fwd_tile( const int32_t *keys, const dcblk &c, int32_t k, const dcblk *__restrict__ cb,
             const dblok *__restrict__ bk, const T *__restrict__ L, int32_t r0, const V *sx, V *red, V *X,
             int64_t ldx, int nrhs, int tid )
{
    const int32_t n    = c.lcolnum - c.fcolnum + 1;
    const int     r    = tid & 63, part = tid >> 6;
    int64_t       base;
    int32_t       ld, g;
    sf_row<T>( keys, c, k, cb, bk, n, r0 + r, base, ld, g );
    V acc[NR];
#pragma unroll
    for ( int q = 0; q < NR; q++ ) acc[q] = (V)0;
    if ( base >= 0 ) {
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
                    for ( int q = 0; q < NR; q++ ) acc[q] += av * sx[( cb0 + u ) * NR + q];
                }
            }
        }
    }
    __syncthreads();   // every wave is done with sx (red may alias it)
#pragma unroll
    for ( int q = 0; q < NR; q++ ) red[( part * 64 + r ) * NR + q] = acc[q];
    __syncthreads();
    if ( part == 0 && base >= 0 ) {
#pragma unroll
        for ( int q = 0; q < NR; q++ ) {
            if ( q < nrhs ) {
                const V v = red[r * NR + q] + red[( 64 + r ) * NR + q] + red[( 128 + r ) * NR + q] +
                            red[( 192 + r ) * NR + q];
                px_atomic_add( X + g + q * ldx, -v );
            }
        }
    }
}

// Backward tile: y_k -= A(rows, :)^T X[g(rows)] for the 64 stacked rows
// r0 .. r0 + 63 of cblk k.  The tile goes through LDS by 64-column chunks
// (sA[col * 65 + row], coalesced loads along the rows); thread (col, part)
// reduces 16 rows; the four parts are summed through sA after the chunk's
// products.  The result is added atomically to Y (out == nullptr) or, for a
// one-task cblk, written to gk[j * NR + q] in LDS (n <= 64).
template <typename T, typename V, int NR>
__device__ inline void
sf_bwd_tile( const int32_t *keys, const dcblk &c, int32_t k, const dcblk *__restrict__ cb,
             const dblok *__restrict__ bk, const T *__restrict__ L, int32_t r0, const V *X, int64_t ldx, V *Y,
             int64_t ldy, int nrhs, V *sA, V *sx, int64_t *sBase, int32_t *sLd, V *gk, int tid )
{
    const int32_t n = c.lcolnum - c.fcolnum + 1;
    if ( tid < 64 ) {
        int64_t base;
        int32_t ld, g;
        sf_row<T>( keys, c, k, cb, bk, n, r0 + tid, base, ld, g );
        sBase[tid] = base;
        sLd[tid]   = ld;
#pragma unroll
        for ( int q = 0; q < NR; q++ ) sx[tid * NR + q] = ( base >= 0 && q < nrhs ) ? X[g + q * ldx] : (V)0;
    }
    __syncthreads();
    constexpr int NU   = 64 * 64 / PXD_THREADS;
    const int     col  = tid & 63, part = tid >> 6;
    const int64_t rb   = sBase[col];
    const int32_t rl   = sLd[col];
    for ( int32_t c0 = 0; c0 < n; c0 += 64 ) {
        T v[NU];
#pragma unroll
        for ( int u = 0; u < NU; u++ ) {
            const int32_t cc = c0 + part + 4 * u;
            v[u] = ( rb >= 0 && cc < n ) ? L[rb + (int64_t)cc * rl] : (T)0;
        }
#pragma unroll
        for ( int u = 0; u < NU; u++ ) sA[( part + 4 * u ) * 65 + col] = px_cvt<V>( v[u] );
        __syncthreads();
        V acc[NR];
#pragma unroll
        for ( int q = 0; q < NR; q++ ) acc[q] = (V)0;
#pragma unroll
        for ( int rr = 16 * part; rr < 16 * part + 16; rr++ ) {
            const V a = sA[col * 65 + rr];
#pragma unroll
            for ( int q = 0; q < NR; q++ ) acc[q] += a * sx[rr * NR + q];
        }
        __syncthreads();   // sA is reused for the partial sums
#pragma unroll
        for ( int q = 0; q < NR; q++ ) sA[( part * 64 + col ) * NR + q] = acc[q];
        __syncthreads();
        if ( part == 0 && c0 + col < n ) {
#pragma unroll
            for ( int q = 0; q < NR; q++ ) {
                const V s = sA[col * NR + q] + sA[( 64 + col ) * NR + q] + sA[( 128 + col ) * NR + q] +
                            sA[( 192 + col ) * NR + q];
                if ( gk ) {
                    gk[( c0 + col ) * NR + q] = s;
                }
                else if ( q < nrhs ) {
                    px_atomic_add( Y + c.fcolnum + c0 + col + q * ldy, -s );
                }
            }
        }
        __syncthreads();
    }
}

// Dynamic LDS of one sweep launch for cblks of at most nmax columns.
template <typename T, int NR>
constexpr size_t
sf_lds_bytes( int32_t nmax )
{
    const size_t a = (size_t)( nmax > 256 ? nmax : 256 ) * NR;   // sx / red, sb
    const size_t b = (size_t)64 * 65;                            // sA of the backward tile
    const size_t c = (size_t)4 * 64 * NR;                        // its partial sums
    size_t       m = a > b ? a : b;
    m              = m > c ? m : c;
    return m * sizeof( T );
}

// Forward sweep of one pass (NR right-hand sides at most): X holds b on
// entry and is consumed; Y receives y = L^{-1} b.
template <typename T, typename V, int NR, int B>
__global__ void __launch_bounds__( PXD_THREADS )
k_sf_forward( const sf_task *__restrict__ tasks, const int32_t *__restrict__ facing,
              const int32_t *__restrict__ need, const dcblk *__restrict__ cb, const dblok *__restrict__ bk,
              const T *__restrict__ L, const T *__restrict__ DI, V *X, int64_t ldx, V *Y, int64_t ldy, int nrhs,
              uint32_t epoch, uint32_t *cnt, uint32_t *rdy, int *err )
{
    extern __shared__ unsigned char smem_raw[];
    V *sbuf = reinterpret_cast<V *>( smem_raw );
    __shared__ int32_t keys[PXD_ROWKEY_CAP];
    __shared__ V       sred[4 * PXD_SF_ROWS * NR];
    __shared__ V       sy[PXD_SF_ROWS * NR];

    const sf_task t   = tasks[blockIdx.x];
    const int     tid = threadIdx.x;
    const int32_t k   = t.cblk;
    const dcblk   c   = cb[k];
    const int32_t n   = c.lcolnum - c.fcolnum + 1;

    if ( t.kind == SF_DIAG || t.kind == SF_LEAF ) {
        if ( tid == 0 && need[k] > 0 ) {
            sf_spin( cnt + k, epoch * (uint32_t)need[k], err );
        }
        sf_acquire( tid );
        const int32_t R0 = ( t.kind == SF_DIAG ) ? t.r0 : 0;
        sf_stage<V, NR>( sbuf, X + c.fcolnum, ldx, 0, min( n, R0 + PXD_SF_ROWS ), nrhs, nullptr, tid );
        __syncthreads();
        sf_diag_rows<T, V, NR>( DI + c.dinvoff, n, R0, false, sbuf, sred, Y + c.fcolnum, ldy, nrhs,
                                ( t.kind == SF_LEAF ) ? sy : nullptr, tid );
        if ( t.kind == SF_DIAG ) {
            sf_release( tid );
            if ( tid == 0 ) {
                sf_signal( rdy + k );
            }
            return;
        }
        // one-task cblk: its tile (if any) from the rows just solved
        if ( t.f1 > t.f0 ) {
            stage_row_keys( keys, bk, c.fblok, cb[k + 1].fblok );
            __syncthreads();
            sf_fwd_tile<T, V, NR, B>( keys, c, k, cb, bk, L, n, sy, sbuf, X, ldx, nrhs, tid );
        }
    }
    else {
        if ( tid == 0 ) {
            sf_spin( rdy + k, epoch * (uint32_t)sf_nchunk( n ), err );
        }
        sf_acquire( tid );
        sf_stage<V, NR>( sbuf, Y + c.fcolnum, ldy, 0, n, nrhs, nullptr, tid );
        stage_row_keys( keys, bk, c.fblok, cb[k + 1].fblok );
        __syncthreads();
        sf_fwd_tile<T, V, NR, B>( keys, c, k, cb, bk, L, t.r0, sbuf, sbuf, X, ldx, nrhs, tid );
    }
    sf_release( tid );
    if ( tid == 0 ) {
        for ( int32_t i = t.f0; i < t.f1; i++ ) sf_signal( cnt + facing[i] );
    }
}

// Backward sweep of one pass: Y holds y on entry (and is consumed), X
// receives x = L^{-T} y.  LB, DB: the tiles and diagonal inverses of the
// backward factor (L for Cholesky and LDL^T, U^T for LU).
template <typename T, typename V, int NR>
__global__ void __launch_bounds__( PXD_THREADS )
k_sf_backward( const sf_task *__restrict__ tasks, const int32_t *__restrict__ facing, const dcblk *__restrict__ cb,
               const dblok *__restrict__ bk, const T *__restrict__ LB, const T *__restrict__ DB, V *X, int64_t ldx,
               V *Y, int64_t ldy, int nrhs, uint32_t epoch, uint32_t *cnt, uint32_t *rdy, int *err )
{
    extern __shared__ unsigned char smem_raw[];
    V *sbuf = reinterpret_cast<V *>( smem_raw );
    __shared__ int32_t keys[PXD_ROWKEY_CAP];
    __shared__ V       sred[4 * PXD_SF_ROWS * NR];
    __shared__ V       sx[64 * NR];
    __shared__ V       gk[PXD_SF_ROWS * NR];
    __shared__ int64_t sBase[64];
    __shared__ int32_t sLd[64];

    const sf_task t   = tasks[blockIdx.x];
    const int     tid = threadIdx.x;
    const int32_t k   = t.cblk;
    const dcblk   c   = cb[k];
    const int32_t n   = c.lcolnum - c.fcolnum + 1;

    if ( t.kind == SF_TILE || t.kind == SF_LEAF ) {
        if ( tid == 0 ) {
            for ( int32_t i = t.f0; i < t.f1; i++ ) {
                const int32_t f  = facing[i];
                const dcblk   cf = cb[f];
                sf_spin( rdy + f, epoch * (uint32_t)sf_nchunk( cf.lcolnum - cf.fcolnum + 1 ), err );
            }
        }
        sf_acquire( tid );
        if ( t.f1 > t.f0 ) {
            stage_row_keys( keys, bk, c.fblok, cb[k + 1].fblok );
            __syncthreads();
            sf_bwd_tile<T, V, NR>( keys, c, k, cb, bk, LB, ( t.kind == SF_LEAF ) ? n : t.r0, X, ldx, Y, ldy, nrhs,
                                   sbuf, sx, sBase, sLd, ( t.kind == SF_LEAF ) ? gk : nullptr, tid );
        }
        else if ( t.kind == SF_LEAF ) {
            for ( int idx = tid; idx < PXD_SF_ROWS * NR; idx += PXD_THREADS ) gk[idx] = (V)0;
        }
        if ( t.kind == SF_TILE ) {
            sf_release( tid );
            if ( tid == 0 ) {
                sf_signal( cnt + k );
            }
            return;
        }
        __syncthreads();
        sf_stage<V, NR>( sbuf, Y + c.fcolnum, ldy, 0, n, nrhs, gk, tid );
    }
    else {
        if ( tid == 0 && c.stride > n ) {
            sf_spin( cnt + k, epoch * (uint32_t)sf_ntile( n, c.stride ), err );
        }
        sf_acquire( tid );
        sf_stage<V, NR>( sbuf, Y + c.fcolnum, ldy, t.r0, n, nrhs, nullptr, tid );
    }
    __syncthreads();
    sf_diag_rows<T, V, NR>( DB + c.dinvoff, n, ( t.kind == SF_DIAG ) ? t.r0 : 0, true, sbuf, sred, X + c.fcolnum,
                            ldx, nrhs, nullptr, tid );
    sf_release( tid );
    if ( tid == 0 ) {
        sf_signal( rdy + k );
    }
}
