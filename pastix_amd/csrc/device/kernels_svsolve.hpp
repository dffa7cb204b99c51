// kernels_svsolve.hpp -- level-by-level supernodal solve kernels with few
// dependent memory round trips (reference kernels/solve_ztrsmsp.c:300-424
// forward, :449-625 backward).
//
// A workgroup of the older kernels looked its rows up at run time: it read
// the task, then the cblk, then staged the cblk's blok keys and searched
// them, then read the blok, and only then loaded the factor, about ten
// dependent trips to memory (~20 us per workgroup measured, which bounded
// every level of the upper tree).  Here the tasks carry the cblk's sizes,
// each off-diagonal row has a precomputed record (factor offset, column
// stride, global row), and the loads that do not depend on each other are
// issued together:
//   trip 1  the task;
//   trip 2  the row records and the right-hand-side block (staged to LDS);
//   trip 3  the factor entries (all columns of a batch in flight);
// then the products, one reduction through LDS and the atomics.
#pragma once
#include "kernels_mfma.hpp"

// A 64-row tile of the stacked off-diagonal rows of a cblk: rows r0 .. r0+63
// (below the diagonal block, r0 >= n) have their records at row0 ...
struct sv_tile {
    int32_t n;         // columns of the cblk
    int32_t fcolnum;   // its first column
    int32_t r0;        // first stacked row of the tile
    int32_t stride;    // stacked rows of the cblk
    int64_t row0;      // record of stacked row r0
    int32_t soff;      // the cblk's columns in the backward partial sums, or -1
    int32_t pad;
};
// ... one record per off-diagonal stacked row
struct sv_row {
    int64_t base;   // offset of the row's first column in the factor
    int32_t ld;     // column stride of its blok
    int32_t g;      // its global row (column of the unknowns)
};
// A PXD_DIAG_ROWS-row chunk of a diagonal block's inverse
#define PXD_DIAG_ROWS 16
struct sv_chunk {
    int32_t n;
    int32_t fcolnum;
    int32_t R0;
    int32_t soff;      // the cblk's columns in the backward partial sums, or -1
    int64_t dinvoff;
};

// Backward partial sums.  The backward tiles of one cblk all subtract into
// the same n x nrhs entries of Y, so on the levels of one or two cblks (150
// to 230 tiles per cblk at 100^3) each entry took that many atomics in a
// row.  A cblk with at least PXD_SV_SPREAD_MIN tiles has instead its own
// columns soff .. soff + n - 1 in PXD_SV_SPREAD copies of a partial-sum
// array (copy c at S + c * nrhs * lds, right-hand side q at + q * lds); tile
// t adds into copy t % PXD_SV_SPREAD, and the cblk's diagonal solve adds the
// copies to Y as it loads it.  The copies are cleared before each backward
// sweep.
// Only where the copies' loads fit in registers: NR * sizeof(V) <= 32 bytes
// (4 right-hand sides up to complex float, 8 in float); elsewhere the tiles
// add into Y directly.
#define PXD_SV_SPREAD 8
#define PXD_SV_SPREAD_MIN 16
template <typename V, int NR>
struct sv_spread_ok {
    static constexpr bool value = NR * sizeof( V ) <= 32;
};

#define PXD_SV_MAXN ( PXD_TRSM_NT * 16 )   // widest cblk (352 columns)

// Stage columns [0, n) of the NR right-hand sides of v (rows fcolnum..,
// leading dimension ld) as s[j * NR + q] (q >= nrhs read 0).  Every load is
// issued before the first LDS store.
template <typename V, int NR>
__device__ inline void
sv_stage( V *s, const V *__restrict__ v, int64_t ld, int32_t n, int nrhs, int tid )
{
    constexpr int MAXE = ( PXD_SV_MAXN * NR + PXD_THREADS - 1 ) / PXD_THREADS;
    V             t[MAXE];
#pragma unroll
    for ( int e = 0; e < MAXE; e++ ) {
        const int idx = tid + PXD_THREADS * e;
        const int j = idx % n, q = idx / n;
        t[e] = ( idx < n * NR && q < nrhs ) ? v[j + q * ld] : (V)0;
    }
#pragma unroll
    for ( int e = 0; e < MAXE; e++ ) {
        const int idx = tid + PXD_THREADS * e;
        if ( idx < n * NR ) s[( idx % n ) * NR + idx / n] = t[e];
    }
}

// sv_stage of v plus the PXD_SV_SPREAD partial-sum copies of S (copy c at
// S + c * cst, right-hand side q at + q * lds).  The entries past n * NR
// redo the last one (same loads, same sums, so the same bits to the same LDS
// slot): no load or store sits under a branch, which had the compiler sink
// the copies' loads into the store's branch and wait for each in turn.  One
// 32-bit offset per entry against the copies' uniform bases lets the loads
// share their address registers.
template <typename V, int NR>
__device__ inline void
sv_stage_sum( V *s, const V *__restrict__ v, int64_t ld, const V *__restrict__ S, int64_t cst, int64_t lds,
              int32_t n, int nrhs, int tid )
{
    constexpr int MAXE = ( PXD_SV_MAXN * NR + PXD_THREADS - 1 ) / PXD_THREADS;
    V             t[MAXE];
    V             p[MAXE][PXD_SV_SPREAD];
    int32_t       off[MAXE];
#pragma unroll
    for ( int e = 0; e < MAXE; e++ ) {
        const int idx = min( tid + PXD_THREADS * e, n * NR - 1 );
        const int j = idx % n, q = min( idx / n, nrhs - 1 );
        off[e] = j + q * (int32_t)lds;
        t[e]   = v[j + q * ld];
    }
#pragma unroll
    for ( int c = 0; c < PXD_SV_SPREAD; c++ ) {
        const V *__restrict__ Sc = S + c * cst;
#pragma unroll
        for ( int e = 0; e < MAXE; e++ ) p[e][c] = Sc[off[e]];
    }
#pragma unroll
    for ( int e = 0; e < MAXE; e++ ) {
        const int idx = min( tid + PXD_THREADS * e, n * NR - 1 );
        V         a   = t[e];
#pragma unroll
        for ( int c = 0; c < PXD_SV_SPREAD; c++ ) a += p[e][c];
        s[( idx % n ) * NR + idx / n] = ( idx / n < nrhs ) ? a : (V)0;
    }
}

// Forward tile: X[g(s)] -= A(s, :) y_k for its 64 rows, y_k = Y[fcolnum ..].
// `split` workgroups share a tile (launches with few tiles): workgroup part
// ps takes the columns [n ps / split, n (ps + 1) / split) and adds its
// partial sums.  Lane r of each wave takes row r0 + r; wave w a quarter of
// the workgroup's columns, B at a time.  Dynamic LDS: n * NR values of V.
template <typename T, typename V, int NR, int B>
__global__ void __launch_bounds__( PXD_THREADS )
k_sv_fwd_tile( const sv_tile *__restrict__ tiles, const sv_row *__restrict__ rows, const T *__restrict__ L,
               V *X, int64_t ldx, const V *__restrict__ Y, int64_t ldy, int nrhs, int split )
{
    extern __shared__ unsigned char smem_raw[];
    V *sy = reinterpret_cast<V *>( smem_raw );
    __shared__ V red[4][64][NR];

    const int     idx = xcd_task( blockIdx.x, gridDim.x );
    const sv_tile t   = tiles[idx / split];
    const int     ps  = idx % split;
    const int     tid = threadIdx.x;
    const int     r = tid & 63, part = tid >> 6;
    const bool    valid = t.r0 + r < t.stride;
    sv_row        ri{ 0, 0, 0 };
    if ( valid ) {
        ri = rows[t.row0 + r];
    }
    sv_stage<V, NR>( sy, Y + t.fcolnum, ldy, t.n, nrhs, tid );
    V acc[NR];
#pragma unroll
    for ( int q = 0; q < NR; q++ ) acc[q] = (V)0;
    const int32_t cs0 = ( t.n * ps ) / split, cw = ( t.n * ( ps + 1 ) ) / split - cs0;
    const int32_t c0 = cs0 + ( cw * part ) / 4, c1 = cs0 + ( cw * ( part + 1 ) ) / 4;
    __syncthreads();
    if ( valid ) {
        for ( int32_t cb0 = c0; cb0 < c1; cb0 += B ) {
            // every load of the batch is issued, the columns past c1 reading
            // column c1 - 1 again: a load under its own branch made the
            // compiler wait for the previous one (one round trip per column)
            T a[B];
#pragma unroll
            for ( int u = 0; u < B; u++ ) a[u] = L[ri.base + (int64_t)min( cb0 + u, c1 - 1 ) * ri.ld];
#pragma unroll
            for ( int u = 0; u < B; u++ ) {
                if ( cb0 + u < c1 ) {
                    const V av = px_cvt<V>( a[u] );
#pragma unroll
                    for ( int q = 0; q < NR; q++ ) acc[q] += av * sy[( cb0 + u ) * NR + q];
                }
            }
        }
    }
#pragma unroll
    for ( int q = 0; q < NR; q++ ) red[part][r][q] = acc[q];
    __syncthreads();
    if ( part == 0 && valid ) {
#pragma unroll
        for ( int q = 0; q < NR; q++ ) {
            if ( q < nrhs ) {
                const V v = red[0][r][q] + red[1][r][q] + red[2][r][q] + red[3][r][q];
                px_atomic_add( X + ri.g + q * ldx, -v );
            }
        }
    }
}

// Backward tile: Y[fcolnum + j] -= sum over its rows s of A(s, j) X[g(s)].
// The tile passes through LDS in 64-column chunks (coalesced loads along the
// rows); four threads per column reduce 16 rows each; NCH chunks are loaded
// at once (all of them for launches with few workgroups).  `split`
// workgroups share a tile: workgroup part ps takes the chunks ps, ps +
// split, ...
template <typename T, typename V, int NR, int NCH>
__global__ void __launch_bounds__( PXD_THREADS )
k_sv_bwd_tile( const sv_tile *__restrict__ tiles, const sv_row *__restrict__ rows, const T *__restrict__ L,
               const V *__restrict__ X, int64_t ldx, V *Y, int64_t ldy, int nrhs, int split, V *S, int64_t lds )
{
    __shared__ T sA[64][64 + 1];   // sA[col][row]
    __shared__ V sx[64][NR];
    __shared__ V red[4][64][NR + 1];

    const int     idx = xcd_task( blockIdx.x, gridDim.x );
    const sv_tile t   = tiles[idx / split];
    const int     ps  = idx % split;
    const int     tid = threadIdx.x;
    const int     rr = tid & 63, part = tid >> 6;
    const bool    valid = t.r0 + rr < t.stride;
    sv_row        ri{ 0, 0, 0 };
    if ( valid ) {
        ri = rows[t.row0 + rr];
    }
    // the sums go to Y, or to the tile's copy of the cblk's partial sums
    const bool spread = sv_spread_ok<V, NR>::value && S != nullptr && t.soff >= 0;
    V *const   dst    = spread ? S + (int64_t)( ( idx / split ) % PXD_SV_SPREAD ) * nrhs * lds + t.soff
                               : Y + t.fcolnum;
    const int64_t ldd = spread ? lds : ldy;
    // X of the tile's rows (wave 0) and the first NCH chunks of the tile
    if ( part == 0 ) {
        V xv[NR];
#pragma unroll
        for ( int q = 0; q < NR; q++ ) xv[q] = ( valid && q < nrhs ) ? X[ri.g + q * ldx] : (V)0;
#pragma unroll
        for ( int q = 0; q < NR; q++ ) sx[rr][q] = xv[q];
    }
    constexpr int NU = 64 * 64 / PXD_THREADS;
    T             v[NCH][NU];
    int kk = 0;
    for ( int32_t c0 = 64 * ps; c0 < t.n; c0 += 64 * split, kk++ ) {
        const int ch = kk % NCH;
        if ( ch == 0 ) {
#pragma unroll
            for ( int h = 0; h < NCH; h++ )
#pragma unroll
                for ( int u = 0; u < NU; u++ ) {
                    const int32_t cc = c0 + 64 * split * h + part + 4 * u;
                    v[h][u] = ( valid && cc < t.n ) ? L[ri.base + (int64_t)cc * ri.ld] : (T)0;
                }
        }
#pragma unroll
        for ( int h = 0; h < NCH; h++ ) {
            if ( h == ch ) {
#pragma unroll
                for ( int u = 0; u < NU; u++ ) sA[part + 4 * u][rr] = v[h][u];
            }
        }
        __syncthreads();
        const int col = rr;
        V         acc[NR];
#pragma unroll
        for ( int q = 0; q < NR; q++ ) acc[q] = (V)0;
#pragma unroll
        for ( int r = 16 * part; r < 16 * part + 16; r++ ) {
            const V a = px_cvt<V>( sA[col][r] );
#pragma unroll
            for ( int q = 0; q < NR; q++ ) acc[q] += a * sx[r][q];
        }
#pragma unroll
        for ( int q = 0; q < NR; q++ ) red[part][col][q] = acc[q];
        __syncthreads();
        if ( part == 0 && c0 + col < t.n ) {
#pragma unroll
            for ( int q = 0; q < NR; q++ ) {
                if ( q < nrhs ) {
                    const V s = red[0][col][q] + red[1][col][q] + red[2][col][q] + red[3][col][q];
                    px_atomic_add( dst + c0 + col + q * ldd, -s );
                }
            }
        }
        __syncthreads();
    }
}

// Rows R0 .. R0 + 15 of out = Linv in (forward, lower) or Linv^T in
// (backward, upper), in taken from Xin (rows fcolnum ..).  Thread (r, g) sums
// columns g, g + 16, ... of row R0 + r.  The inverse's entries and the
// right-hand sides are loaded together.  Dynamic LDS: n * NR values of V.
template <typename T, typename V, int NR>
__global__ void __launch_bounds__( PXD_THREADS )
k_sv_diag( const sv_chunk *__restrict__ chunks, const T *__restrict__ Linv, const V *__restrict__ Xin,
           int64_t ldx, V *__restrict__ Yout, int64_t ldy, int nrhs, int trans, const V *__restrict__ S,
           int64_t lds )
{
    constexpr int CR   = PXD_DIAG_ROWS;
    constexpr int NG   = PXD_THREADS / CR;   // column phases
    constexpr int MAXU = ( PXD_SV_MAXN + NG - 1 ) / NG;
    extern __shared__ unsigned char smem_raw[];
    V *sb = reinterpret_cast<V *>( smem_raw );
    __shared__ V sred[NG][CR][NR + 1];

    const sv_chunk ch  = chunks[xcd_task( blockIdx.x, gridDim.x )];
    const int32_t  n   = ch.n;
    const int      tid = threadIdx.x;
    const T       *LI  = Linv + ch.dinvoff;
    const int      r   = trans ? tid / NG : tid % CR;
    const int      g   = trans ? tid % NG : tid / CR;
    const int32_t  i   = ch.R0 + r;

    T a[MAXU];
#pragma unroll
    for ( int u = 0; u < MAXU; u++ ) {
        const int32_t j  = g + NG * u;
        const bool    ok = i < n && j < n && ( trans ? j >= i : j <= i );
        a[u] = ok ? LI[trans ? j + (int64_t)i * n : i + (int64_t)j * n] : (T)0;
    }
    bool summed = false;
    if constexpr ( sv_spread_ok<V, NR>::value ) {
        if ( S != nullptr && ch.soff >= 0 ) {
            // backward: Y plus the cblk's partial sums
            sv_stage_sum<V, NR>( sb, Xin + ch.fcolnum, ldx, S + ch.soff, (int64_t)nrhs * lds, lds, n, nrhs, tid );
            summed = true;
        }
    }
    if ( !summed ) {
        sv_stage<V, NR>( sb, Xin + ch.fcolnum, ldx, n, nrhs, tid );
    }
    __syncthreads();

    V acc[NR];
#pragma unroll
    for ( int q = 0; q < NR; q++ ) acc[q] = (V)0;
#pragma unroll
    for ( int u = 0; u < MAXU; u++ ) {
        const int32_t j = g + NG * u;
        if ( j < n ) {
            const V av = px_cvt<V>( a[u] );
#pragma unroll
            for ( int q = 0; q < NR; q++ ) acc[q] += av * sb[j * NR + q];
        }
    }
#pragma unroll
    for ( int q = 0; q < NR; q++ ) sred[g][r][q] = acc[q];
    __syncthreads();
    if ( tid < CR * NR ) {
        const int r2 = tid % CR, q = tid / CR;
        V         v  = (V)0;
#pragma unroll
        for ( int p2 = 0; p2 < NG; p2++ ) v += sred[p2][r2][q];
        if ( ch.R0 + r2 < n && q < nrhs ) Yout[ch.fcolnum + ch.R0 + r2 + q * ldy] = v;
    }
}

// ---------------------------------------------------------------------------
// Levels whose cblks all fit one diagonal chunk (at most PXD_DIAG_ROWS
// columns: the leaves of the nested dissection, ~60k cblks of 1-16 columns
// for 100^3) run the diagonal solve and the off-diagonal rows of a cblk in
// one wavefront, four cblks per workgroup, in one launch per level.  The
// wavefront exchanges its right-hand-side block through its own LDS area;
// its LDS accesses complete in program order, so no workgroup barrier is
// needed.  NR = 4 right-hand sides at most: lane (i, q) = (l & 15, l >> 4).
struct sv_small {
    int32_t n;         // columns of the cblk (<= PXD_DIAG_ROWS)
    int32_t fcolnum;
    int32_t nrows;     // off-diagonal stacked rows
    int32_t pad;
    int64_t dinvoff;   // its diagonal block's inverse
    int64_t row0;      // record of its first off-diagonal row
};

__device__ inline void
sv_wave_sync()
{
    __builtin_amdgcn_s_waitcnt( 0xC07F );   // lgkmcnt(0)
    __builtin_amdgcn_fence( __ATOMIC_ACQ_REL, "wavefront" );
}

// Forward: y_k = Linv_k X[k rows] into Y, then X[g(s)] -= A(s, :) y_k for
// the cblk's off-diagonal rows s (atomics: other cblks add to the same rows).
template <typename T, typename V>
__global__ void __launch_bounds__( PXD_THREADS )
k_sv_small_fwd( const sv_small *__restrict__ tasks, int64_t ntasks, const sv_row *__restrict__ rows,
                const T *__restrict__ L, const T *__restrict__ Linv, V *X, int64_t ldx, V *__restrict__ Y,
                int64_t ldy, int nrhs )
{
    constexpr int CR = PXD_DIAG_ROWS;
    static_assert( CR * 4 == 64, "lane (i, q) covers 16 rows and 4 right-hand sides" );
    __shared__ V sb[4][CR][4];
    const int     tid = threadIdx.x, w = tid >> 6, lane = tid & 63;
    const int64_t ti  = (int64_t)xcd_task( blockIdx.x, gridDim.x ) * 4 + w;
    if ( ti >= ntasks ) {
        return;   // no workgroup barrier follows
    }
    const sv_small t = tasks[ti];
    const int      i = lane & ( CR - 1 ), q = lane >> 4;
    const T       *LI = Linv + t.dinvoff;
    // the record of the lane's first off-diagonal row travels with the
    // inverse and the right-hand side, and that row's factor entries are in
    // flight while the diagonal block is solved: three dependent trips to
    // memory instead of four
    const bool has0 = lane < t.nrows;
    sv_row     ri0{ 0, 0, 0 };
    if ( has0 ) {
        ri0 = rows[t.row0 + lane];
    }
    T a[CR];
#pragma unroll
    for ( int j = 0; j < CR; j++ ) a[j] = ( i < t.n && j <= i ) ? LI[i + (int64_t)j * t.n] : (T)0;
    const V b = ( i < t.n && q < nrhs ) ? X[t.fcolnum + i + q * ldx] : (V)0;
    T       av0[CR];
#pragma unroll
    for ( int j = 0; j < CR; j++ ) av0[j] = ( has0 && j < t.n ) ? L[ri0.base + (int64_t)j * ri0.ld] : (T)0;
    sb[w][i][q] = b;
    sv_wave_sync();
    V y = (V)0;
#pragma unroll
    for ( int j = 0; j < CR; j++ ) y += px_cvt<V>( a[j] ) * sb[w][j][q];
    sv_wave_sync();
    sb[w][i][q] = y;
    if ( i < t.n && q < nrhs ) Y[t.fcolnum + i + q * ldy] = y;
    sv_wave_sync();
    for ( int32_t s = lane; s < t.nrows; s += 64 ) {
        sv_row ri = ri0;
        T      av[CR];
#pragma unroll
        for ( int j = 0; j < CR; j++ ) av[j] = av0[j];
        if ( s != lane ) {
            ri = rows[t.row0 + s];
#pragma unroll
            for ( int j = 0; j < CR; j++ ) av[j] = ( j < t.n ) ? L[ri.base + (int64_t)j * ri.ld] : (T)0;
        }
#pragma unroll
        for ( int qq = 0; qq < 4; qq++ ) {
            if ( qq < nrhs ) {
                V acc = (V)0;
#pragma unroll
                for ( int j = 0; j < CR; j++ ) acc += px_cvt<V>( av[j] ) * sb[w][j][qq];
                px_atomic_add( X + ri.g + qq * ldx, -acc );
            }
        }
    }
}

// Backward: y_k = Y[k rows] - sum over the off-diagonal rows s of A(s, :)^T
// X[g(s)], then X[k rows] = Linv_k^T y_k.  The gather's partial sums of the
// wavefront's lanes (one row each) meet through LDS, one right-hand side at
// a time.
template <typename T, typename V>
__global__ void __launch_bounds__( PXD_THREADS )
k_sv_small_bwd( const sv_small *__restrict__ tasks, int64_t ntasks, const sv_row *__restrict__ rows,
                const T *__restrict__ L, const T *__restrict__ Linv, V *X, int64_t ldx, const V *__restrict__ Y,
                int64_t ldy, int nrhs )
{
    constexpr int CR = PXD_DIAG_ROWS;
    __shared__ V sb[4][CR][4];
    __shared__ V sred[4][64][CR + 1];
    const int     tid = threadIdx.x, w = tid >> 6, lane = tid & 63;
    const int64_t ti  = (int64_t)xcd_task( blockIdx.x, gridDim.x ) * 4 + w;
    if ( ti >= ntasks ) {
        return;   // no workgroup barrier follows
    }
    const sv_small t = tasks[ti];
    const int      i = lane & ( CR - 1 ), q = lane >> 4;
    // Linv^T row i (column i of Linv) and y_k, issued with the gather's loads
    const T *LI = Linv + t.dinvoff;
    T        a[CR];
#pragma unroll
    for ( int j = 0; j < CR; j++ ) a[j] = ( i < t.n && j >= i && j < t.n ) ? LI[j + (int64_t)i * t.n] : (T)0;
    const V yk = ( i < t.n && q < nrhs ) ? Y[t.fcolnum + i + q * ldy] : (V)0;
    V       p[4][CR];
#pragma unroll
    for ( int qq = 0; qq < 4; qq++ )
#pragma unroll
        for ( int j = 0; j < CR; j++ ) p[qq][j] = (V)0;
    for ( int32_t s = lane; s < t.nrows; s += 64 ) {
        const sv_row ri = rows[t.row0 + s];
        T            av[CR];
#pragma unroll
        for ( int j = 0; j < CR; j++ ) av[j] = ( j < t.n ) ? L[ri.base + (int64_t)j * ri.ld] : (T)0;
        V xs[4];
#pragma unroll
        for ( int qq = 0; qq < 4; qq++ ) xs[qq] = ( qq < nrhs ) ? X[ri.g + qq * ldx] : (V)0;
#pragma unroll
        for ( int qq = 0; qq < 4; qq++ )
#pragma unroll
            for ( int j = 0; j < CR; j++ ) p[qq][j] += px_cvt<V>( av[j] ) * xs[qq];
    }
    // gather sum of (j, qq): lane (j, part) adds the rows part * 16 .. + 15,
    // then the four parts meet in LDS
    V g = (V)0;
#pragma unroll
    for ( int qq = 0; qq < 4; qq++ ) {
#pragma unroll
        for ( int j = 0; j < CR; j++ ) sred[w][lane][j] = p[qq][j];
        sv_wave_sync();
        V part = (V)0;
#pragma unroll
        for ( int r = 0; r < 16; r++ ) part += sred[w][16 * q + r][i];
        sv_wave_sync();
        sred[w][lane][0] = part;
        sv_wave_sync();
        const V tot = sred[w][i][0] + sred[w][16 + i][0] + sred[w][32 + i][0] + sred[w][48 + i][0];
        if ( q == qq ) g = tot;
        sv_wave_sync();
    }
    sb[w][i][q] = yk - g;
    sv_wave_sync();
    V x = (V)0;
#pragma unroll
    for ( int j = 0; j < CR; j++ ) x += px_cvt<V>( a[j] ) * sb[w][j][q];
    if ( i < t.n && q < nrhs ) X[t.fcolnum + i + q * ldx] = x;
}
