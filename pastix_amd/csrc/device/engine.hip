// engine.hip -- the single-GPU numerical engine behind pastix_amd_device.h.
//
// Scheduling.  The reference runs the cblk tasks of blend's static list on
// CPU threads that spin on dependency counters (sopalin/sequential_zpotrf.c:
// 102-147, kernels/cpucblk_zmpi_coeftab.c:541-622).  On the GPU the same task
// graph is walked by dependency levels: level(k) = 1 + max level of the cblks
// that update k (browtab), so every cblk of a level is ready once the previous
// levels are done.  A level is three batched launches: diagonal panels, the
// off-diagonal TRSM tiles, and the update tiles of all its cblks.  Updates of
// one level that can touch the same facing-cblk columns (two sources, one
// fcblk, overlapping column ranges) use atomic adds; all other updates write
// with plain read-modify-write, which is what the reference's per-cblk lock
// serializes (kernels/core_zgemmsp.c:462,470).
#include <algorithm>
#include <dlfcn.h>
#include <cfloat>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <limits>
#include <mutex>
#include <type_traits>
#include <vector>
#include "kernels_solve.hpp"
#include "kernels_mfma.hpp"
#include "kernels_svsolve.hpp"
#include "pastix/kernels_enums.h"

namespace {

enum { FT_DOUBLE = 3, FT_FLOAT = 2, FT_CPLX64 = 5, FT_CPLX32 = 4 };
enum { FACT_POTRF = 0, FACT_SYTRF = 1, FACT_GETRF = 2 };

struct level_t {
    int64_t c0, c1;     // cblk list
    int64_t t0, t1;     // trsm tasks (32-row tiles)
    int64_t u0, uc, u1; // update tasks (64 x 64 tiles); [u0, uc) face the next level
    int64_t b0, bc, b1; // update tasks (tall tiles); [b0, bc) face the next level
    int64_t s0, sc, s1; // small updates (one wavefront each); [s0, sc) face the next level
    int64_t f0, fs, f1; // solve row tasks (64-row tiles of off-diagonal rows); [fs, f1): Schur cblks
    int64_t d0, ds, d1; // solve diagonal tasks (PXD_DIAG_ROWS-row chunks of the diagonal blocks); [ds, d1): Schur
    int64_t m0, m1;     // one-wavefront solve tasks (levels whose cblks all fit one chunk, no Schur part)
    int64_t i0, i1;     // block-column inversion tasks
    int64_t g0, g1;     // blok updates (upd_group)
    int64_t q0, q1;     // multi-workgroup panel tasks (32-row blocks of the diagonal tiles)
    int32_t maxn;
};

// The tiles of one blok update (source cblk, facing blok b, side), in one of
// the two tile lists, and its GEMM sizes as the reference logs them.
struct upd_group {
    int32_t cblk, blok;
    int32_t M, N, K;
    int32_t big, side;
    int64_t t0, t1;
};

}  // namespace

// solve launches with fewer workgroups than this load every column at once
#define PXD_SOLVE_WIDE_MAX 128

// A captured solve pass (option solve_graph): the launch sequence of both
// sweeps on the instance's own vectors for nq columns of one factor type,
// replayed as one graph.  It stays valid while everything the capture baked
// in is unchanged (sig).
// What a captured factorization baked in: it is replayed only while these
// are unchanged.
struct fact_graph_sig {
    int         tag;
    uint64_t    gen;
    const void *L, *U, *DI, *DU;
    double      thresh;
    int         prio, panel_mw, ldlt;
    bool operator==( const fact_graph_sig &o ) const
    {
        return tag == o.tag && gen == o.gen && L == o.L && U == o.U && DI == o.DI && DU == o.DU &&
               thresh == o.thresh && prio == o.prio && panel_mw == o.panel_mw && ldlt == o.ldlt;
    }
};

struct solve_graph {
    int            nq, tag;
    const void    *sig[7];   // vectors, Y, L, U, the two inverse sets and the partial sums
    uint64_t       gen;      // task-list generation
    int64_t        wide;     // solve_wide_max at capture
    hipGraphExec_t exec;
};

struct pxd_ctx_s {
    int         device = 0;
    hipStream_t stream = nullptr;
    hipStream_t stream2 = nullptr;    // panels of the next level (lookahead)

    hipEvent_t  ev_ready = nullptr, ev_panel = nullptr;
    size_t      lds_max = 65536;

    // structure
    int64_t n = 0, cblknbr = 0, bloknbr = 0, coefnbr = 0, dinvnbr = 0;
    int32_t colmax = 0;
    std::vector<dcblk> hcb;
    std::vector<dblok> hbk;
    dcblk   *dcb = nullptr;
    dblok   *dbk = nullptr;
    int32_t *dcol2cblk = nullptr;

    // schedule
    std::vector<int32_t> level;     // dependency level of each cblk
    int32_t              nlev = 0;
    int                  tasks_dirty = 0;   // options changed the task lists
    std::vector<level_t> levels;
    int32_t  *dlist = nullptr;
    pan_task *dpan  = nullptr;
    int      *dflags = nullptr;     // progress counters of the panel tasks
    int64_t   npan  = 0;
    row_task *dtrsm = nullptr;
    upd_task *dupd  = nullptr;
    upd_task *dupdb = nullptr;
    upd_task *dupds = nullptr;
    row_task *dsolv = nullptr;
    sv_tile  *dsvt  = nullptr;   // the solv tasks with their cblk sizes
    sv_row   *dsvr  = nullptr;   // records of the off-diagonal stacked rows
    sv_chunk *dsvc  = nullptr;   // the diag tasks with their cblk sizes
    sv_small *dsvs  = nullptr;   // the one-wavefront solve tasks
    // backward partial sums of the cblks with many tiles (PXD_SV_SPREAD
    // copies of spread_cols columns x PXD_MAXRHS, factor precision)
    void     *dspread       = nullptr;
    size_t    dspread_bytes = 0;
    int64_t   spread_cols   = 0;
    int       solve_spread  = PXD_SV_SPREAD_MIN;   // fewest tiles of a cblk that spreads; 0: none
    // 1: levels of cblks of at most PXD_DIAG_ROWS columns solve with one
    // wavefront per cblk (diagonal and off-diagonal rows in one launch)
    int       solve_small = 1;
    row_task *ddiag = nullptr;
    inv_task *dinvt = nullptr;
    int64_t   nupd = 0, ntrsm = 0, nsolv = 0;

    // matrix
    int      mflt = 0, mtx = 0;
    int64_t  nnz = 0;
    int64_t *dcolptr = nullptr, *drowtab = nullptr, *dtcolptr = nullptr, *dtrowtab = nullptr;
    void    *dval = nullptr, *dtval = nullptr;
    std::vector<int64_t> hcolptr, hrowtab;   // kept for the fill of the U side
    double   norm_inf = 0.;

    // factor
    int    fflt = 0, facto = -1;
    void  *dL = nullptr, *dU = nullptr, *ddinv = nullptr, *dudinv = nullptr;
    size_t coef_bytes = 0;
    int   *dscratch = nullptr;     // counters (pivots, overflow)
    void  *dvecs = nullptr;        // solve scratch in the factor precision
    std::vector<solve_graph> sgraphs;
    uint64_t sched_gen   = 0;      // bumped when the task lists are rebuilt
    int      solve_graph = 1;
    // the factorization's level loop as a captured graph (option fact_graph)
    int            fact_graph = 1;
    hipGraphExec_t fgraph     = nullptr;
    fact_graph_sig fgraph_sig{};
    int64_t        fgraph_launches = 0;
    int64_t        stats_graph_captures = 0;
    std::string tuning_file;       // the tuning file applied at creation, if any
    int         clock_khz = 0;     // rate of the device's constant clock
    size_t dvecs_bytes = 0;
    void  *dsolvy = nullptr;       // the sweeps' second vector (n x PXD_MAXRHS, factor precision)
    size_t dsolvy_bytes = 0;
    void  *dred = nullptr;         // reduction scratch

    // rows per TRSM tile: 64 for the matrix-core kernels; for the scalar
    // kernels 32, or 16 when the LDS cannot hold 32 x colmax
    int          trsm_rt = 64;
    // 1: matrix-core kernels with the explicit diagonal inverse; 0: scalar kernels
    int          kernels = 1;
    // blok updates with at least this many stacked rows use the tall tiles
    int          upd_big_rows = 192;
    std::vector<int8_t> upd_cls;   // per (blok, side): kernel class chosen by the caller, -1: the rule
    // 1: updates of at most 64 rows, PXD_SMALL_N columns and PXD_SMALL_K
    // source columns take one wavefront each (k_update_small)
    int          upd_small = 1;
    // the cut-off rule gives the one-wavefront kernel the updates of at most
    // upd_small_m stacked rows and upd_small_n facing columns (and at most
    // PXD_SMALL_K source columns); a caller's choice may use the kernel's
    // whole range (PXD_SMALL_M, PXD_SMALL_N)
    int          upd_small_m = 64;
    int          upd_small_n = 16;
    // stacked rows per tall tile (upd_rows<T>, set with the factor precision)
    int          upd_bm = 256;
    // 1: single-precision LL^T / LU tall tiles on the kernel that loads its
    // operands straight into LDS (k_update_dma); 0: on k_update_big
    int          upd_dma = 1;
    // level-solve launches with fewer than solve_wide_max tiles share each
    // tile among up to solve_split_max workgroups, solve_split_wg in all
    int          solve_split_max = 4;
    int64_t      solve_split_wg  = 512;
    // dynamic LDS bytes added to each k_update_dma workgroup: 4 KiB makes it
    // 34 KiB, four workgroups per CU, and leaves 24 KiB of the CU's LDS to the
    // panel and TRSM workgroups of the lookahead stream (without it, or with
    // less free, the next level's panels wait for a CU)
    unsigned     upd_dma_pad = 4096;

    // host copies of the panel, TRSM and inversion lists, and the blok
    // updates in launch order (kernel timing log)
    std::vector<int32_t>   hlist;
    std::vector<pan_task>  hpan;
    std::vector<row_task>  htrsm;
    std::vector<inv_task>  hinvt;
    std::vector<upd_group> groups;

    // level solve launches with at least this many tiles load 16 columns
    // per batch (higher occupancy); smaller ones load every column at once
    int64_t   solve_wide_max = PXD_SOLVE_WIDE_MAX;

    // Schur complement: the cblks from column schurfcol on (the last ones)
    // hold the Schur unknowns.  schur_fact_local: they receive the updates of
    // the other cblks but are not factorized (PastixFactModeLocal).
    // schur_solv_mode: 0 local (x11 = A11^{-1} b11, rows of the Schur part
    // untouched), 1 interface (plus b22 -= L21 x11 forward, b11 -= L21^T x22
    // backward), 2 the whole matrix (reference pastix_solv_mode_t)
    int64_t      schurfcol = 0;
    int          schur_fact_local = 1;
    int          schur_solv_mode = 0;

    // options / stats
    int          timing = 0;
    // 1: panels, inverses, TRSM and the updates into the next level raise
    // their waves' issue priority (s_setprio) over the bulk updates
    int          prio = 1;
    // 1: Cholesky diagonal tiles factored by one workgroup per 32-row block
    // (k_potrf_mw); 0: one workgroup per cblk (k_potrf2)
    int          panel_mw = 1;
    // 1: the factorization runs one stream, launches every blok update on
    // its own and records per-workgroup times into model_entries
    int          model_log = 0;
    std::vector<pxd_model_entry_t> model_entries;
    pxd_stats_t  stats{};
};

namespace {

template <typename T> int flt_of();
template <> int flt_of<double>() { return FT_DOUBLE; }
template <> int flt_of<float>() { return FT_FLOAT; }

size_t
elt_size( int flt )
{
    switch ( flt ) {
    case FT_FLOAT: return 4;
    case FT_DOUBLE: return 8;
    case FT_CPLX32: return 8;
    case FT_CPLX64: return 16;
    }
    return 0;
}

template <typename T>
int
upload( T **d, const std::vector<T> &h )
{
    if ( *d ) {
        PXD_CHECK( hipFree( *d ) );
        *d = nullptr;
    }
    size_t bytes = std::max<size_t>( 1, h.size() ) * sizeof( T );
    PXD_CHECK( hipMalloc( (void **)d, bytes ) );
    if ( !h.empty() ) {
        PXD_CHECK( hipMemcpy( *d, h.data(), h.size() * sizeof( T ), hipMemcpyHostToDevice ) );
    }
    return 0;
}

inline unsigned
nblocks( int64_t n, int bs )
{
    return (unsigned)( ( n + bs - 1 ) / bs );
}

// Dynamic LDS that a launch of kernel fn may take on a CU with lds_max bytes
// of LDS: what the kernel's static LDS leaves.
size_t
lds_room( const void *fn, size_t lds_max )
{
    hipFuncAttributes a;
    if ( hipFuncGetAttributes( &a, fn ) != hipSuccess || a.sharedSizeBytes >= lds_max ) {
        return 0;
    }
    return lds_max - a.sharedSizeBytes;
}

// LDS bytes of the scalar TRSM tile: rt rows x n columns
template <typename T>
size_t
trsm_lds( int32_t n, int rt )
{
    return (size_t)n * ( rt + 1 ) * sizeof( T );
}

// Dependency level of every cblk: 1 + the highest level of its sources.
void
compute_levels( pxd_ctx_t *ctx, const pxd_cblk_t *cblks, const pxd_blok_t *bloks, const int64_t *browtab )
{
    const int64_t         nc    = ctx->cblknbr;
    std::vector<int32_t> &level = ctx->level;
    level.assign( nc, 0 );
    ctx->nlev = 0;
    for ( int64_t k = 0; k < nc; k++ ) {
        int32_t l = 0;
        for ( int64_t e = cblks[k].brownum; e < cblks[k + 1].brownum; e++ ) {
            int64_t src = bloks[browtab[e]].lcblknm;
            l = std::max( l, level[src] + 1 );
        }
        level[k]  = l;
        ctx->nlev = std::max( ctx->nlev, l + 1 );
    }
}


void drop_solve_graphs( pxd_ctx_t *ctx );

// a number for each factor type (keys of the captured graphs)
template <typename T> struct type_tag;
template <> struct type_tag<double> { static constexpr int value = 0; };
template <> struct type_tag<float> { static constexpr int value = 1; };
template <> struct type_tag<pxcd> { static constexpr int value = 2; };
template <> struct type_tag<pxcf> { static constexpr int value = 3; };

// Build the task lists of the factorization and the solve, level by level.
int
build_schedule( pxd_ctx_t *ctx )
{
    const int64_t               nc    = ctx->cblknbr;
    const std::vector<int32_t> &level = ctx->level;
    const int32_t               nlev  = ctx->nlev;

    // the trsm tile keeps RT rows x n columns in LDS (static LDS of the kernel ~9 KB)
    if ( ctx->kernels ) {
        ctx->trsm_rt = 64;
        if ( ctx->colmax > PXD_TRSM_NT * 16 ) {
            fprintf( stderr, "pastix_amd device: cblk width %d above the %d columns of the TRSM tile\n",
                     ctx->colmax, PXD_TRSM_NT * 16 );
            return 3;
        }
    }
    else {
        ctx->trsm_rt = ( trsm_lds<double>( ctx->colmax, 32 ) + 12288 <= ctx->lds_max ) ? 32 : 16;
    }
    if ( !ctx->kernels && trsm_lds<double>( ctx->colmax, 16 ) + 12288 > ctx->lds_max ) {
        fprintf( stderr, "pastix_amd device: cblk width %d exceeds the LDS budget of the TRSM tile\n",
                 ctx->colmax );
        return 3;
    }

    // within a level, the cblks of the Schur part come after the others
    // (their solve tasks can then be left out as a suffix of the level)
    std::vector<std::vector<int32_t>> bylev( nlev );
    for ( int pass = 0; pass < 2; pass++ ) {
        for ( int64_t k = 0; k < nc; k++ ) {
            if ( ( ctx->hcb[k].fcolnum >= ctx->schurfcol ) == ( pass == 1 ) ) {
                bylev[level[k]].push_back( (int32_t)k );
            }
        }
    }
    // cblks of the Schur part that the factorization leaves alone
    auto unfactored = [&]( int32_t k ) { return ctx->schur_fact_local && ctx->hcb[k].fcolnum >= ctx->schurfcol; };

    std::vector<int32_t>  list;
    std::vector<pan_task> pan;
    std::vector<row_task> trsm, solv, diag;
    std::vector<inv_task> invt;
    std::vector<upd_task> upd, updb, upds;
    ctx->levels.assign( nlev, level_t{} );
    const bool big_ok = ctx->kernels != 0;

    struct iv { int32_t side, F, f, l; std::vector<upd_task> *list; int64_t first, last; };
    const bool lu = ( ctx->facto == FACT_GETRF );
    std::vector<iv> ivs;
    std::vector<int32_t> smallk;
    ctx->groups.clear();

    for ( int32_t L = 0; L < nlev; L++ ) {
        level_t &lv = ctx->levels[L];
        lv.c0   = list.size();
        lv.q0   = pan.size();
        lv.t0   = trsm.size();
        lv.u0   = upd.size();
        lv.b0   = updb.size();
        lv.s0   = upds.size();
        lv.f0   = solv.size();
        lv.d0   = diag.size();
        lv.i0   = invt.size();
        lv.g0   = ctx->groups.size();
        lv.maxn = 0;
        ivs.clear();
        lv.ds = -1;
        lv.fs = -1;
        int32_t maxn_all = 0;
        for ( int32_t k : bylev[L] ) {
            const dcblk &c  = ctx->hcb[k];
            int32_t      nk = c.lcolnum - c.fcolnum + 1;
            maxn_all        = std::max( maxn_all, nk );
            if ( lv.ds < 0 && c.fcolnum >= ctx->schurfcol ) {
                lv.ds = diag.size();
                lv.fs = solv.size();
            }
            for ( int32_t r0 = nk; r0 < c.stride; r0 += 64 ) {
                solv.push_back( row_task{ k, r0 } );
            }
            for ( int32_t r0 = 0; r0 < nk; r0 += PXD_DIAG_ROWS ) {
                diag.push_back( row_task{ k, r0 } );
            }
            if ( unfactored( k ) ) {
                continue;
            }
            lv.maxn = std::max( lv.maxn, nk );
            list.push_back( k );
            {
                const int32_t f0 = (int32_t)pan.size(), nblk = ( nk + PXD_NB - 1 ) / PXD_NB;
                for ( int32_t pb = 0; pb < nblk; pb++ ) {
                    pan.push_back( pan_task{ k, pb, f0, nblk } );
                }
            }
            for ( int32_t r0 = nk; r0 < c.stride; r0 += ctx->trsm_rt ) {
                trsm.push_back( row_task{ k, r0 } );
            }
            for ( int32_t J = 0; J + PXD_NB < nk; J += PXD_NB ) {
                invt.push_back( inv_task{ k, J } );
            }
        }
        if ( lv.ds < 0 ) {
            lv.ds = diag.size();
            lv.fs = solv.size();
        }
        lv.m0 = smallk.size();
        if ( lv.ds == (int64_t)diag.size() && maxn_all <= PXD_DIAG_ROWS ) {
            for ( int32_t k : bylev[L] ) smallk.push_back( k );
        }
        lv.m1 = smallk.size();
        // update tiles: first those into cblks of the next level (the
        // critical part that the next level's panels wait for), then the rest
        for ( int pass = 0; pass < 2; pass++ ) {
            if ( pass == 1 ) {
                lv.uc = upd.size();
                lv.bc = updb.size();
                lv.sc = upds.size();
            }
            for ( int32_t k : bylev[L] ) {
                const dcblk &c  = ctx->hcb[k];
                int32_t      lb = ctx->hcb[k + 1].fblok;
                if ( unfactored( k ) ) {
                    continue;
                }
                for ( int32_t b = c.fblok + 1; b < lb; b++ ) {
                    const dblok &bb = ctx->hbk[b];
                    if ( ( level[bb.fcblknm] == L + 1 ) != ( pass == 0 ) ) {
                        continue;
                    }
                    const int32_t N = bb.lrownum - bb.frownum + 1;
                    // L side: rows of bloks b.. ; U side (LU): rows of bloks b+1..
                    for ( int side = 0; side < ( lu ? 2 : 1 ); side++ ) {
                        const int32_t M = c.stride - bb.rowoff - ( side ? N : 0 );
                        if ( M <= 0 ) {
                            continue;
                        }
                        const int32_t fl = side ? 2 : 0;
                        const int32_t K  = c.lcolnum - c.fcolnum + 1;
                        // kernel class: the caller's choice when it fits the
                        // update, else the cut-offs
                        const bool small_fits = big_ok && M <= PXD_SMALL_M && N <= PXD_SMALL_N && K <= PXD_SMALL_K;
                        const bool small_cut  = small_fits && M <= ctx->upd_small_m && N <= ctx->upd_small_n;
                        int        cls        = ctx->upd_cls.empty() ? -1 : ctx->upd_cls[2 * (size_t)b + side];
                        if ( ( cls == 2 && !small_fits ) || ( cls == 1 && !big_ok ) || cls < 0 || cls > 2 ) {
                            cls = ( small_cut && ctx->upd_small ) ? 2 : ( big_ok && M >= ctx->upd_big_rows ) ? 1 : 0;
                        }
                        if ( cls == 2 ) {
                            // one wavefront per 64 stacked rows
                            int64_t first = upds.size();
                            for ( int32_t tm = 0; tm * 64 < M; tm++ ) {
                                upds.push_back( upd_task{ k, b, tm, 0, fl, 0 } );
                            }
                            ivs.push_back( iv{ side, bb.fcblknm, bb.frownum, bb.lrownum, &upds, first,
                                               (int64_t)upds.size() } );
                            ctx->groups.push_back( upd_group{ k, b, M, N, K, 2, side, first, (int64_t)upds.size() } );
                        }
                        else if ( cls == 1 ) {
                            int64_t first = updb.size();
                            // column tiles fastest: the tiles that share a row
                            // panel run together and share its L2 lines
                            for ( int32_t tm = 0; tm * ctx->upd_bm < M; tm++ ) {
                                for ( int32_t tn = 0; tn * PXD_UPD_BN < N; tn++ ) {
                                    updb.push_back( upd_task{ k, b, tm, tn, fl, 0 } );
                                }
                            }
                            ivs.push_back( iv{ side, bb.fcblknm, bb.frownum, bb.lrownum, &updb, first,
                                               (int64_t)updb.size() } );
                            ctx->groups.push_back( upd_group{ k, b, M, N, c.lcolnum - c.fcolnum + 1, 1, side,
                                                              first, (int64_t)updb.size() } );
                        }
                        else {
                            int64_t first = upd.size();
                            for ( int32_t tn = 0; tn * 64 < N; tn++ ) {
                                for ( int32_t tm = 0; tm * 64 < M; tm++ ) {
                                    upd.push_back( upd_task{ k, b, tm, tn, fl, 0 } );
                                }
                            }
                            ivs.push_back( iv{ side, bb.fcblknm, bb.frownum, bb.lrownum, &upd, first,
                                               (int64_t)upd.size() } );
                            ctx->groups.push_back( upd_group{ k, b, M, N, c.lcolnum - c.fcolnum + 1, 0, side,
                                                              first, (int64_t)upd.size() } );
                        }
                    }
                }
            }
        }
        // atomic flags for updates whose facing columns overlap another source's
        std::sort( ivs.begin(), ivs.end(), []( const iv &a, const iv &b ) {
            if ( a.side != b.side ) return a.side < b.side;
            return ( a.F != b.F ) ? ( a.F < b.F ) : ( a.f < b.f );
        } );
        for ( size_t i = 0; i < ivs.size(); ) {
            size_t j = i;
            while ( j < ivs.size() && ivs[j].F == ivs[i].F && ivs[j].side == ivs[i].side ) {
                j++;
            }
            int32_t maxl = -1;
            for ( size_t p = i; p < j; p++ ) {
                bool conflict = ( maxl >= ivs[p].f ) || ( p + 1 < j && ivs[p + 1].f <= ivs[p].l );
                maxl = std::max( maxl, ivs[p].l );
                if ( conflict ) {
                    for ( int64_t u = ivs[p].first; u < ivs[p].last; u++ ) {
                        ( *ivs[p].list )[u].flags |= 1;
                    }
                }
            }
            i = j;
        }
        lv.c1 = list.size();
        lv.q1 = pan.size();
        lv.t1 = trsm.size();
        lv.u1 = upd.size();
        lv.b1 = updb.size();
        lv.s1 = upds.size();
        lv.f1 = solv.size();
        lv.d1 = diag.size();
        lv.i1 = invt.size();
        lv.g1 = ctx->groups.size();
    }
    ctx->nupd  = upd.size() + updb.size() + upds.size();
    ctx->ntrsm = trsm.size();
    ctx->nsolv = solv.size();
    ctx->stats.levels       = nlev;
    ctx->stats.update_tiles = upd.size() + updb.size() + upds.size();
    ctx->hlist = list;
    ctx->hpan  = pan;
    ctx->npan  = (int64_t)pan.size();
    if ( ctx->dflags ) {
        PXD_CHECK( hipFree( ctx->dflags ) );
        ctx->dflags = nullptr;
    }
    PXD_CHECK( hipMalloc( (void **)&ctx->dflags, sizeof( int ) * std::max<int64_t>( 1, ctx->npan ) ) );
    ctx->htrsm = trsm;
    ctx->hinvt = invt;
    if ( upload( &ctx->dlist, list ) || upload( &ctx->dpan, pan ) || upload( &ctx->dtrsm, trsm ) || upload( &ctx->dupd, upd ) ||
         upload( &ctx->dupdb, updb ) || upload( &ctx->dupds, upds ) ||
         upload( &ctx->dsolv, solv ) || upload( &ctx->dinvt, invt ) || upload( &ctx->ddiag, diag ) ) {
        return 6;
    }
    {
        // the solve's row records and tasks with the sizes of their cblk
        std::vector<int64_t> rowstart( nc + 1, 0 );
        for ( int64_t k = 0; k < nc; k++ ) {
            const dcblk &c = ctx->hcb[k];
            rowstart[k + 1] = rowstart[k] + ( c.stride - ( c.lcolnum - c.fcolnum + 1 ) );
        }
        std::vector<sv_row> rows( rowstart[nc] );
        for ( int64_t k = 0; k < nc; k++ ) {
            const dcblk  &c  = ctx->hcb[k];
            const int32_t n  = c.lcolnum - c.fcolnum + 1;
            const int32_t lb = ctx->hcb[k + 1].fblok;
            for ( int32_t b = c.fblok + 1; b < lb; b++ ) {
                const dblok  &bb = ctx->hbk[b];
                const int32_t nr = bb.lrownum - bb.frownum + 1;
                for ( int32_t i = 0; i < nr; i++ ) {
                    rows[rowstart[k] + bb.rowoff + i - n] =
                        sv_row{ c.coefoff + (int64_t)bb.rowoff * n + i, nr, bb.frownum + i };
                }
            }
        }
        // columns in the backward partial sums for the cblks with at least
        // solve_spread off-diagonal tiles (PXD_SV_SPREAD_MIN by default)
        std::vector<int32_t> soff( nc, -1 );
        ctx->spread_cols = 0;
        for ( int64_t k = 0; k < nc; k++ ) {
            const dcblk  &c = ctx->hcb[k];
            const int32_t n = c.lcolnum - c.fcolnum + 1;
            if ( ctx->solve_spread > 0 && ( c.stride - n + 63 ) / 64 >= ctx->solve_spread ) {
                soff[k] = (int32_t)ctx->spread_cols;
                ctx->spread_cols += n;
            }
        }
        std::vector<sv_tile>  svt( solv.size() );
        std::vector<sv_chunk> svc( diag.size() );
        for ( size_t i = 0; i < solv.size(); i++ ) {
            const dcblk  &c = ctx->hcb[solv[i].cblk];
            const int32_t n = c.lcolnum - c.fcolnum + 1;
            svt[i] = sv_tile{ n, c.fcolnum, solv[i].r0, c.stride, rowstart[solv[i].cblk] + solv[i].r0 - n,
                              soff[solv[i].cblk], 0 };
        }
        for ( size_t i = 0; i < diag.size(); i++ ) {
            const dcblk &c = ctx->hcb[diag[i].cblk];
            svc[i] = sv_chunk{ c.lcolnum - c.fcolnum + 1, c.fcolnum, diag[i].r0, soff[diag[i].cblk], c.dinvoff };
        }
        std::vector<sv_small> svs( smallk.size() );
        for ( size_t i = 0; i < smallk.size(); i++ ) {
            const dcblk  &c = ctx->hcb[smallk[i]];
            const int32_t n = c.lcolnum - c.fcolnum + 1;
            svs[i] = sv_small{ n, c.fcolnum, c.stride - n, 0, c.dinvoff, rowstart[smallk[i]] };
        }
        if ( upload( &ctx->dsvt, svt ) || upload( &ctx->dsvr, rows ) || upload( &ctx->dsvc, svc ) ||
             upload( &ctx->dsvs, svs ) ) {
            return 6;
        }
    }
    ctx->sched_gen++;
    drop_solve_graphs( ctx );
    return 0;
}

// Elapsed milliseconds between two events (the stream must be synchronized).
float
ev_ms( hipEvent_t a, hipEvent_t b )
{
    float ms = 0.f;
    (void)hipEventElapsedTime( &ms, a, b );
    return ms;
}

// Records of the kernel timing log from the workgroup stamps of one
// factorization (layout of the stamp buffer: panels, inversions L and U
// sides, TRSM tiles L and U sides, 64 x 64 update tiles, tall update tiles).
// A panel is logged as (n, 0, 0) like cpucblk_zpotrfsp1d
// (kernels/core_zpotrfsp.c:258); the TRSM of a cblk as (stride - n, n, 0) like
// cpucblk_ztrsmsp (kernels/core_ztrsmsp.c:366-397), with its block-column
// inversions counted in since they replace the reference's triangular solve;
// a blok update as (M, N, K) like the GEMMCblk2d2d record
// (kernels/gpu_zgemmsp.c:242-251).  A task's time is the span from its first
// workgroup's start to its last workgroup's end.
void
model_log_entries( pxd_ctx_t *ctx, const std::vector<uint64_t> &hs, int64_t n_list, int64_t n_inv,
                   int64_t n_trsm, int64_t n_upd, int64_t n_updb, double sec_per_tick, bool mw )
{
    const bool      lu    = ( ctx->facto == FACT_GETRF );
    const int       nside = lu ? 2 : 1;
    const uint64_t *sp    = hs.data();
    const uint64_t *si    = sp + 2 * n_list;
    const uint64_t *sm    = si + 4 * n_inv;
    const uint64_t *su    = sm + 4 * n_trsm;
    const uint64_t *sb    = su + 2 * n_upd;
    const uint64_t *ss    = sb + 2 * n_updb;   // small updates: one task per launch in this mode
    // span of the stamped tasks [a, b) of one buffer; false when none ran
    auto span = [&]( const uint64_t *base, int64_t a, int64_t b, double &t ) {
        uint64_t lo = UINT64_MAX, hi = 0;
        for ( int64_t i = a; i < b; i++ ) {
            if ( base[2 * i + 1] == 0 ) {
                continue;
            }
            lo = std::min( lo, base[2 * i] );
            hi = std::max( hi, base[2 * i + 1] );
        }
        if ( hi == 0 ) {
            return false;
        }
        t += (double)( hi - lo ) * sec_per_tick;
        return true;
    };
    // tasks of each cblk in the inversion and TRSM lists (contiguous runs)
    const int64_t        nc = ctx->cblknbr;
    std::vector<int64_t> ib( nc + 1, 0 ), ie( nc + 1, 0 ), tb( nc + 1, 0 ), te( nc + 1, 0 );
    for ( int64_t i = (int64_t)ctx->hinvt.size() - 1; i >= 0; i-- ) {
        const int32_t k = ctx->hinvt[i].cblk;
        if ( ie[k] == 0 ) ie[k] = i + 1;
        ib[k] = i;
    }
    for ( int64_t i = (int64_t)ctx->htrsm.size() - 1; i >= 0; i-- ) {
        const int32_t k = ctx->htrsm[i].cblk;
        if ( te[k] == 0 ) te[k] = i + 1;
        tb[k] = i;
    }
    const int kpanel = lu ? PastixKernelGETRF : ( ctx->facto == FACT_SYTRF ) ? PastixKernelSYTRF : PastixKernelPOTRF;
    std::vector<pxd_model_entry_t> &out = ctx->model_entries;
    out.clear();
    // multi-workgroup panels: the tasks of cblk k in the panel-task list
    std::vector<int64_t> qb( nc + 1, 0 ), qe( nc + 1, 0 );
    if ( mw ) {
        for ( int64_t i = (int64_t)ctx->hpan.size() - 1; i >= 0; i-- ) {
            const int32_t k = ctx->hpan[i].cblk;
            if ( qe[k] == 0 ) qe[k] = i + 1;
            qb[k] = i;
        }
    }
    for ( const level_t &lv : ctx->levels ) {
        for ( int64_t i = lv.c0; i < lv.c1; i++ ) {
            const int32_t k = ctx->hlist[i];
            const dcblk  &c = ctx->hcb[k];
            const int32_t n = c.lcolnum - c.fcolnum + 1;
            double        t = 0.;
            if ( mw ? span( sp, qb[k], qe[k], t ) : span( sp, i, i + 1, t ) ) {
                out.push_back( pxd_model_entry_t{ kpanel, n, 0, 0, t, -1, 0 } );
            }
            if ( c.stride == n ) {
                continue;
            }
            for ( int side = 0; side < nside; side++ ) {
                double tt = 0.;
                span( si + 2 * side * n_inv, ib[k], ie[k], tt );
                if ( span( sm + 2 * side * n_trsm, tb[k], te[k], tt ) ) {
                    out.push_back( pxd_model_entry_t{ PastixKernelTRSMCblk2d, c.stride - n, n, 0, tt, -1, 0 } );
                }
            }
        }
        for ( int64_t g = lv.g0; g < lv.g1; g++ ) {
            const upd_group &gr = ctx->groups[g];
            double           t  = 0.;
            const bool ok = ( gr.big == 2 ) ? span( ss, gr.t0, gr.t1, t ) : span( gr.big ? sb : su, gr.t0, gr.t1, t );
            if ( ok ) {
                // variant: 0 64 x 64 tiles, 1 tall tiles, 2 one wavefront
                const int32_t var = ( gr.big == 2 ) ? 2 : gr.big;
                out.push_back( pxd_model_entry_t{ PastixKernelGEMMCblk2d2d, gr.M, gr.N, gr.K, t, var, 0 } );
            }
        }
    }
}

// ---------------------------------------------------------------------------
template <typename T>
int
factorize( pxd_ctx_t *ctx, double thresh, int64_t *nbpivots )
{
    const bool lu   = ( ctx->facto == FACT_GETRF );
    const int  ldlt = ( ctx->facto == FACT_SYTRF );
    T *L  = (T *)ctx->dL;
    T *DI = (T *)ctx->ddinv;                        // Cholesky: L_kk^{-1}; LU: (unit L_kk)^{-1}
    T *U  = lu ? (T *)ctx->dU : L;                  // LU: U^T tiles
    T *DU = lu ? (T *)ctx->dudinv : DI;             // LU: (U_kk^T)^{-1}
    hipStream_t s = ctx->stream;
    hipEvent_t  e0, e1, ea, eb;
    PXD_CHECK( hipEventCreate( &e0 ) );
    PXD_CHECK( hipEventCreate( &e1 ) );
    PXD_CHECK( hipEventCreate( &ea ) );
    PXD_CHECK( hipEventCreate( &eb ) );
    double t_panel = 0., t_trsm = 0., t_upd = 0.;
    int64_t launches = 0;

    PXD_CHECK( hipMemsetAsync( ctx->dscratch, 0, 4 * sizeof( int ), s ) );
    PXD_CHECK( hipMemsetAsync( ctx->dflags, 0, sizeof( int ) * std::max<int64_t>( 1, ctx->npan ), s ) );
    PXD_CHECK( hipEventRecord( e0, s ) );

    // fill: zero + scatter of the lower triangle of P A P^T (LU: and of the
    // strict upper triangle, transposed, into the U tiles)
    PXD_CHECK( hipMemsetAsync( L, 0, ctx->coefnbr * sizeof( T ), s ) );
    typedef typename px_real<T>::type R;
    typedef typename std::conditional<std::is_same<T, R>::value, double, pxcd>::type TI;   // matrix values
    const double ovf = (double)std::numeric_limits<R>::max();
    hipLaunchKernelGGL( ( k_fill<T, TI> ), dim3( nblocks( ctx->n, 256 ) ), dim3( 256 ), 0, s,
                        ctx->n, ctx->dcolptr, ctx->drowtab, (const TI *)ctx->dval, ctx->dcol2cblk,
                        ctx->dcb, ctx->dbk, L, 0, ovf, ctx->dscratch + 1 );
    launches += 2;
    if ( lu ) {
        PXD_CHECK( hipMemsetAsync( U, 0, ctx->coefnbr * sizeof( T ), s ) );
        // the U side scatters A^T (a symmetric matrix is its own transpose)
        const bool     tr = ( ctx->dtval != nullptr );
        const int64_t *tc = tr ? ctx->dtcolptr : ctx->dcolptr;
        const int64_t *tt = tr ? ctx->dtrowtab : ctx->drowtab;
        const TI      *tv = (const TI *)( tr ? ctx->dtval : ctx->dval );
        hipLaunchKernelGGL( ( k_fill<T, TI> ), dim3( nblocks( ctx->n, 256 ) ), dim3( 256 ), 0, s,
                            ctx->n, tc, tt, tv, ctx->dcol2cblk, ctx->dcb, ctx->dbk, U, 1, ovf, ctx->dscratch + 1 );
        launches += 2;
    }
    if ( ctx->timing ) {
        PXD_CHECK( hipEventRecord( ea, s ) );
        PXD_CHECK( hipEventSynchronize( ea ) );
        ctx->stats.fill_seconds = 1e-3 * ev_ms( e0, ea );
    }

    // kernel timing log: stamps of every panel, inversion, TRSM and update
    // workgroup (LU: inversions and TRSM of both sides)
    const bool mlog = ctx->model_log && ctx->kernels;
    uint64_t *stamps_all = nullptr, *st_panel = nullptr, *st_inv = nullptr, *st_trsm = nullptr;
    uint64_t *st_upd = nullptr, *st_updb = nullptr, *st_upds = nullptr;
    const bool    use_mw = ctx->panel_mw && ctx->kernels && !lu && !ldlt;
    const int64_t n_list = use_mw ? (int64_t)ctx->hpan.size() : (int64_t)ctx->hlist.size();
    const int64_t n_inv = ctx->hinvt.size(), n_trsm = ctx->htrsm.size();
    int64_t n_upd = 0, n_updb = 0, n_upds = 0;
    for ( const level_t &lv : ctx->levels ) {
        n_upd  = std::max( n_upd, lv.u1 );
        n_updb = std::max( n_updb, lv.b1 );
        n_upds = std::max( n_upds, lv.s1 );
    }
    const int64_t n_st = n_list + 2 * n_inv + 2 * n_trsm + n_upd + n_updb + n_upds;
    if ( mlog ) {
        PXD_CHECK( hipMalloc( (void **)&stamps_all, 2 * sizeof( uint64_t ) * std::max<int64_t>( 1, n_st ) ) );
        PXD_CHECK( hipMemsetAsync( stamps_all, 0, 2 * sizeof( uint64_t ) * std::max<int64_t>( 1, n_st ), s ) );
        st_panel = stamps_all;
        st_inv   = st_panel + 2 * n_list;
        st_trsm  = st_inv + 4 * n_inv;
        st_upd   = st_trsm + 4 * n_trsm;
        st_updb  = st_upd + 2 * n_upd;
        st_upds  = st_updb + 2 * n_updb;
    }
    // rows of the panel's L21 strip and of the inverse's R block kept in
    // LDS (PXD_NB + 1 values each); the kernels keep the rest in HBM
    const size_t  lds_row = ( PXD_NB + 1 ) * sizeof( T );
    const void   *potrf2  = (const void *)k_potrf2<T, true>;
    if constexpr ( std::is_same<T, typename px_real<T>::type>::value ) {
        if ( !ldlt ) {
            potrf2 = (const void *)k_potrf2<T, false>;
        }
    }
    const int32_t pcap = (int32_t)( lds_room( potrf2, ctx->lds_max ) / lds_row );
    const int32_t rcap = (int32_t)( lds_room( (const void *)k_trtri<T>, ctx->lds_max ) / lds_row );
    if ( ctx->kernels && ( rcap < 1 || ( !lu && pcap < 1 ) ) ) {
        if ( stamps_all ) {
            (void)hipFree( stamps_all );
        }
        return 3;
    }
    // stamps of the tasks [i0, ...) of a launch (side 1: the U side's copy)
    auto stp = [&]( uint64_t *base, int64_t i0, int side, int64_t nside ) -> uint64_t * {
        return base ? base + 2 * ( i0 + side * nside ) : nullptr;
    };

    if ( ctx->kernels ) {
        PXD_CHECK( hipMemsetAsync( DI, 0, ctx->dinvnbr * sizeof( T ), s ) );
        launches++;
        if ( lu ) {
            PXD_CHECK( hipMemsetAsync( DU, 0, ctx->dinvnbr * sizeof( T ), s ) );
            launches++;
        }
    }
    // launchers of one level's phases on a stream
    auto panel = [&]( const level_t &lv, hipStream_t st ) {
        unsigned np = (unsigned)( lv.c1 - lv.c0 );
        unsigned ni = (unsigned)( lv.i1 - lv.i0 );
        if ( np == 0 ) {
            return;   // a level of unfactored Schur cblks only
        }
        // rows below a block column of the widest cblk of the level
        const size_t trtri_lds = (size_t)std::max( 1, std::min( lv.maxn - PXD_NB, rcap ) ) * lds_row;
        if ( lu ) {
            hipLaunchKernelGGL( ( k_getrf2<T> ), dim3( np ), dim3( PXD_THREADS ), 0, st, ctx->dlist + lv.c0,
                                ctx->dcb, L, U, DI, DU, thresh, ctx->dscratch, stp( st_panel, lv.c0, 0, 0 ), ctx->prio );
            launches++;
            if ( ni ) {
                hipLaunchKernelGGL( ( k_trtri<T> ), dim3( ni ), dim3( PXD_THREADS ), trtri_lds, st,
                                    ctx->dinvt + lv.i0, ctx->dcb, (const T *)L, DI, stp( st_inv, lv.i0, 0, n_inv ), ctx->prio, rcap );
                hipLaunchKernelGGL( ( k_trtri<T> ), dim3( ni ), dim3( PXD_THREADS ), trtri_lds, st,
                                    ctx->dinvt + lv.i0, ctx->dcb, (const T *)U, DU, stp( st_inv, lv.i0, 1, n_inv ), ctx->prio, rcap );
                launches += 2;
            }
        }
        else if ( ctx->kernels ) {
            const size_t sh = (size_t)std::max( 1, std::min( lv.maxn, pcap ) ) * lds_row;
            if ( ldlt ) {
                hipLaunchKernelGGL( ( k_potrf2<T, true> ), dim3( np ), dim3( PXD_THREADS ), sh, st, ctx->dlist + lv.c0,
                                    ctx->dcb, L, DI, thresh, ctx->dscratch, stp( st_panel, lv.c0, 0, 0 ), ctx->prio, pcap );
            }
            else if constexpr ( std::is_same<T, R>::value ) {
                if ( ctx->panel_mw ) {
                    // one workgroup per 32-row block of each diagonal tile
                    const unsigned nq  = (unsigned)( lv.q1 - lv.q0 );
                    const size_t   shq = (size_t)( ( std::max( 1, lv.maxn ) + PXD_NB - 1 ) / PXD_NB ) * PXD_NB *
                                       ( PXD_NB + 1 ) * sizeof( T );
                    hipLaunchKernelGGL( ( k_potrf_mw<T> ), dim3( nq ), dim3( PXD_THREADS ), shq, st,
                                        ctx->dpan + lv.q0, ctx->dcb, L, DI, thresh, ctx->dscratch, ctx->dflags,
                                        ctx->dscratch + 2, stp( st_panel, lv.q0, 0, 0 ), ctx->prio );
                }
                else {
                    hipLaunchKernelGGL( ( k_potrf2<T, false> ), dim3( np ), dim3( PXD_THREADS ), sh, st,
                                        ctx->dlist + lv.c0, ctx->dcb, L, DI, thresh, ctx->dscratch,
                                        stp( st_panel, lv.c0, 0, 0 ), ctx->prio, pcap );
                }
            }
            launches++;
            if ( ni ) {
                hipLaunchKernelGGL( ( k_trtri<T> ), dim3( ni ), dim3( PXD_THREADS ), trtri_lds, st,
                                    ctx->dinvt + lv.i0, ctx->dcb, (const T *)L, DI, stp( st_inv, lv.i0, 0, n_inv ), ctx->prio, rcap );
                launches++;
            }
        }
        else if constexpr ( std::is_same<T, typename px_real<T>::type>::value ) {
            hipLaunchKernelGGL( ( k_potrf<T> ), dim3( np ), dim3( PXD_THREADS ), 0, st, ctx->dlist + lv.c0,
                                ctx->dcb, L, DI, thresh, ctx->dscratch );
            launches++;
        }
    };
    auto trsm = [&]( const level_t &lv, hipStream_t st ) {
        unsigned nt = (unsigned)( lv.t1 - lv.t0 );
        if ( !nt ) {
            return;
        }
        if ( lu ) {
            // L_off <- L_off U_kk^{-1};  U^T_off <- U^T_off L_kk^{-T}
            hipLaunchKernelGGL( ( k_trsm_mfma<T> ), dim3( nt ), dim3( PXD_THREADS ), 0, st, ctx->dtrsm + lv.t0,
                                ctx->dcb, ctx->dbk, L, (const T *)DU, 0, stp( st_trsm, lv.t0, 0, n_trsm ), ctx->prio );
            hipLaunchKernelGGL( ( k_trsm_mfma<T> ), dim3( nt ), dim3( PXD_THREADS ), 0, st, ctx->dtrsm + lv.t0,
                                ctx->dcb, ctx->dbk, U, (const T *)DI, 0, stp( st_trsm, lv.t0, 1, n_trsm ), ctx->prio );
            launches++;
        }
        else if ( ctx->kernels ) {
            hipLaunchKernelGGL( ( k_trsm_mfma<T> ), dim3( nt ), dim3( PXD_THREADS ), 0, st, ctx->dtrsm + lv.t0,
                                ctx->dcb, ctx->dbk, L, (const T *)DI, ldlt, stp( st_trsm, lv.t0, 0, n_trsm ), ctx->prio );
        }
        else if constexpr ( std::is_same<T, typename px_real<T>::type>::value ) {
            if ( ctx->trsm_rt == 32 ) {
                hipLaunchKernelGGL( ( k_trsm<T, 32> ), dim3( nt ), dim3( PXD_THREADS ), trsm_lds<T>( lv.maxn, 32 ),
                                    st, ctx->dtrsm + lv.t0, ctx->dcb, ctx->dbk, L, (const T *)DI );
            }
            else {
                hipLaunchKernelGGL( ( k_trsm<T, 16> ), dim3( nt ), dim3( PXD_THREADS ), trsm_lds<T>( lv.maxn, 16 ),
                                    st, ctx->dtrsm + lv.t0, ctx->dcb, ctx->dbk, L, (const T *)DI );
            }
        }
        launches++;
    };
    // update tiles [u0, u1) of the 64 x 64 list and [b0, b1) of the tall list
    auto update = [&]( int64_t u0, int64_t u1, int64_t b0, int64_t b1, int64_t q0, int64_t q1, hipStream_t st,
                       int prio ) {
        unsigned nu = (unsigned)( u1 - u0 );
        unsigned nb = (unsigned)( b1 - b0 );
        unsigned nq = (unsigned)( q1 - q0 );
        if ( nq ) {
            constexpr unsigned per = PXD_THREADS / 64;
            hipLaunchKernelGGL( ( k_update_small<T> ), dim3( ( nq + per - 1 ) / per ), dim3( PXD_THREADS ), 0, st,
                                ctx->dupds + q0, (int32_t)nq, ctx->dcb, ctx->dbk, L, U, ldlt,
                                stp( st_upds, q0, 0, 0 ), prio );
            launches++;
        }
        if ( nb ) {
            bool done = false;
            if constexpr ( std::is_same<T, float>::value ) {
                if ( ctx->upd_dma && !ldlt ) {
                    hipLaunchKernelGGL( ( k_update_dma<PXD_DMA_BK, PXD_DMA_NBUF> ), dim3( nb ), dim3( PXD_THREADS ),
                                        ctx->upd_dma_pad, st, ctx->dupdb + b0, (int32_t)nb, ctx->dcb, ctx->dbk, L, U,
                                        stp( st_updb, b0, 0, 0 ), prio );
                    done = true;
                }
            }
            if ( !done ) {
                hipLaunchKernelGGL( ( k_update_big<T, 1> ), dim3( nb ), dim3( PXD_THREADS ), 0, st, ctx->dupdb + b0,
                                    (int32_t)nb, ctx->dcb, ctx->dbk, L, U, ldlt, stp( st_updb, b0, 0, 0 ), prio );
            }
            launches++;
        }
        if ( nu && ctx->kernels ) {
            hipLaunchKernelGGL( ( k_update_mfma<T> ), dim3( nu ), dim3( PXD_THREADS ), 0, st, ctx->dupd + u0,
                                ctx->dcb, ctx->dbk, L, U, ldlt, stp( st_upd, u0, 0, 0 ), prio );
            launches++;
        }
        else if ( nu ) {
            if constexpr ( std::is_same<T, typename px_real<T>::type>::value ) {
                hipLaunchKernelGGL( ( k_update<T> ), dim3( nu ), dim3( PXD_THREADS ), 0, st, ctx->dupd + u0,
                                    ctx->dcb, ctx->dbk, L );
                launches++;
            }
        }
    };

    if ( mlog ) {
        // one stream; every blok update is a launch of its own, so that its
        // workgroups' times are those of the update alone on the device
        for ( const level_t &lv : ctx->levels ) {
            panel( lv, s );
            trsm( lv, s );
            for ( int64_t g = lv.g0; g < lv.g1; g++ ) {
                const upd_group &gr = ctx->groups[g];
                if ( gr.big == 2 ) {
                    update( 0, 0, 0, 0, gr.t0, gr.t1, s, 0 );
                }
                else if ( gr.big ) {
                    update( 0, 0, gr.t0, gr.t1, 0, 0, s, 0 );
                }
                else {
                    update( gr.t0, gr.t1, 0, 0, 0, 0, s, 0 );
                }
            }
        }
    }
    else if ( ctx->timing ) {
        // one stream, every phase timed on its own
        for ( const level_t &lv : ctx->levels ) {
            PXD_CHECK( hipEventRecord( ea, s ) );
            panel( lv, s );
            PXD_CHECK( hipEventRecord( eb, s ) );
            PXD_CHECK( hipEventSynchronize( eb ) );
            t_panel += 1e-3 * ev_ms( ea, eb );
            PXD_CHECK( hipEventRecord( ea, s ) );
            trsm( lv, s );
            PXD_CHECK( hipEventRecord( eb, s ) );
            PXD_CHECK( hipEventSynchronize( eb ) );
            t_trsm += 1e-3 * ev_ms( ea, eb );
            PXD_CHECK( hipEventRecord( ea, s ) );
            update( lv.u0, lv.u1, lv.b0, lv.b1, lv.s0, lv.s1, s, 0 );
            PXD_CHECK( hipEventRecord( eb, s ) );
            PXD_CHECK( hipEventSynchronize( eb ) );
            t_upd += 1e-3 * ev_ms( ea, eb );
        }
    }
    else {
        // Lookahead: the panels, inverses and TRSM of level L+1 (stream s2)
        // start as soon as the updates into level L+1 are done, while the
        // remaining updates of level L go on in stream s.  Those only write
        // cblks of later levels, so the two streams touch disjoint data.
        hipStream_t s2 = ctx->stream2;
        hipStream_t su = s;
        hipEvent_t  ready = ctx->ev_ready, paneldone = ctx->ev_panel;
        auto        issue = [&]() -> int {
            PXD_CHECK( hipEventRecord( ready, su ) );
            for ( const level_t &lv : ctx->levels ) {
                PXD_CHECK( hipStreamWaitEvent( s2, ready, 0 ) );
                panel( lv, s2 );
                trsm( lv, s2 );
                PXD_CHECK( hipEventRecord( paneldone, s2 ) );
                PXD_CHECK( hipStreamWaitEvent( su, paneldone, 0 ) );
                update( lv.u0, lv.uc, lv.b0, lv.bc, lv.s0, lv.sc, su, ctx->prio );
                PXD_CHECK( hipEventRecord( ready, su ) );
                update( lv.uc, lv.u1, lv.bc, lv.b1, lv.sc, lv.s1, su, 0 );
            }
            // the last inversions ran on s2 after the last handoff
            PXD_CHECK( hipEventRecord( paneldone, s2 ) );
            PXD_CHECK( hipStreamWaitEvent( su, paneldone, 0 ) );
            return 0;
        };
        if ( ctx->fact_graph ) {
            // the level loop (both streams) as one graph, captured once per
            // structure and factor buffers; the clearing and the fill above
            // stay outside it
            const fact_graph_sig sig{ type_tag<T>::value, ctx->sched_gen, L, U, DI, DU, thresh, ctx->prio,
                                      ctx->panel_mw, ldlt };
            if ( ctx->fgraph == nullptr || !( ctx->fgraph_sig == sig ) ) {
                if ( ctx->fgraph ) {
                    (void)hipGraphExecDestroy( ctx->fgraph );
                    ctx->fgraph = nullptr;
                }
                const int64_t l0 = launches;
                hipGraph_t    g  = nullptr;
                PXD_CHECK( hipStreamBeginCapture( s, hipStreamCaptureModeThreadLocal ) );
                const int  rc = issue();
                hipError_t e  = hipStreamEndCapture( s, &g );
                if ( rc != 0 || e != hipSuccess || g == nullptr ||
                     hipGraphInstantiate( &ctx->fgraph, g, nullptr, nullptr, 0 ) != hipSuccess ) {
                    if ( g ) (void)hipGraphDestroy( g );
                    (void)hipGetLastError();
                    ctx->fgraph = nullptr;
                    return rc ? rc : 6;
                }
                (void)hipGraphDestroy( g );
                ctx->fgraph_sig      = sig;
                ctx->fgraph_launches = launches - l0;
                ctx->stats_graph_captures++;
            }
            else {
                launches += ctx->fgraph_launches;
            }
            PXD_CHECK( hipGraphLaunch( ctx->fgraph, s ) );
        }
        else if ( int rc = issue() ) {
            return rc;
        }
    }
    PXD_CHECK( hipGetLastError() );
    PXD_CHECK( hipEventRecord( e1, s ) );
    PXD_CHECK( hipEventSynchronize( e1 ) );
    ctx->stats.fact_seconds   = 1e-3 * ev_ms( e0, e1 );
    ctx->stats.panel_seconds  = t_panel;
    ctx->stats.trsm_seconds   = t_trsm;
    ctx->stats.update_seconds = t_upd;
    ctx->stats.launches       = launches;
    int h[4];
    PXD_CHECK( hipMemcpy( h, ctx->dscratch, 4 * sizeof( int ), hipMemcpyDeviceToHost ) );
    *nbpivots = h[0];
    if ( mlog ) {
        std::vector<uint64_t> hs( 2 * n_st );
        PXD_CHECK( hipMemcpy( hs.data(), stamps_all, hs.size() * sizeof( uint64_t ), hipMemcpyDeviceToHost ) );
        PXD_CHECK( hipFree( stamps_all ) );
        model_log_entries( ctx, hs, n_list, n_inv, n_trsm, n_upd, n_updb, 1. / ( 1e3 * std::max( 1, ctx->clock_khz ) ),
                           use_mw );
    }
    (void)hipEventDestroy( e0 );
    (void)hipEventDestroy( e1 );
    (void)hipEventDestroy( ea );
    (void)hipEventDestroy( eb );
    if ( h[1] > 0 ) {
        fprintf( stderr, "pastix_amd device: %d matrix entries overflow the factor precision\n", h[1] );
        return 7; /* PASTIX_ERR_BADPARAMETER */
    }
    if ( h[2] != 0 ) {
        fprintf( stderr, "pastix_amd device: a panel workgroup timed out waiting for its neighbour\n" );
        return 6; /* PASTIX_ERR_INTERNAL */
    }
    return 0;
}


// Right-hand sides per solve pass.
template <typename V>
int
solve_pass_rhs( const pxd_ctx_t * )
{
    return PXD_MAXRHS;
}


// Steps of a solve (bit mask): the forward sweep, the diagonal D^{-1} of an
// LDL^T factor, the backward sweep.
enum { SOLVE_FWD = 1, SOLVE_DIAG = 2, SOLVE_BWD = 4, SOLVE_ALL = 7 };

// The sweeps of `what` for nrhs columns of X (leading dimension ldx) with
// the factor of type T, vectors of type V (= T).  swap exchanges the two sides of an LU factor: the forward
// sweep runs on U^T and the backward sweep on L^T (the solve with A^T,
// reference sopalin/pastix_task_solve.c:597-601).
// One level of the level-by-level sweeps with the sv kernels, NR columns
// per launch (nrhs <= NR): forward, the diagonal chunks (X -> Y) then the
// tiles (Y -> atomics into X); backward, the tiles (X -> atomics into Y)
// then the chunks (Y -> X).  nd chunks from lv.d0, nf tiles from lv.f0.
template <typename T, typename V, int NR>
void
sv_level( pxd_ctx_t *ctx, bool forward, const level_t &lv, unsigned nd, unsigned nf, const T *L, const T *DI,
          V *X, int64_t ldx, V *Y, int64_t ldy, int nrhs, V *S )
{
    const int64_t lds = ctx->spread_cols;   // the backward partial sums S (or none)
    constexpr int WB = solve_wide<T>::value;   // column batch of launches with few workgroups
    hipStream_t   s  = ctx->stream;
    const size_t  sh = (size_t)lv.maxn * NR * sizeof( V );
    const bool    many = (int64_t)nf >= ctx->solve_wide_max;
    // launches with few tiles share each tile among `split` workgroups (the
    // columns, or the 64-column chunks backward), up to solve_split_wg
    // workgroups in all
    int split = 1;
    if ( !many && nf ) {
        split = std::max( 1, std::min( ctx->solve_split_max, (int)( ctx->solve_split_wg / (int64_t)nf ) ) );
    }
    const int split_b = std::max( 1, std::min( split, ( lv.maxn + 63 ) / 64 ) );
    const int64_t nm  = lv.m1 - lv.m0;
    if constexpr ( NR == 4 && sizeof( V ) <= 8 ) {   // complex double keeps the chunk and tile launches
        if ( ctx->solve_small && nm > 0 && nrhs <= 4 ) {
            const unsigned nb = (unsigned)( ( nm + 3 ) / 4 );
            if ( forward ) {
                hipLaunchKernelGGL( ( k_sv_small_fwd<T, V> ), dim3( nb ), dim3( PXD_THREADS ), 0, s, ctx->dsvs + lv.m0,
                                    nm, ctx->dsvr, L, DI, X, ldx, Y, ldy, nrhs );
            }
            else {
                hipLaunchKernelGGL( ( k_sv_small_bwd<T, V> ), dim3( nb ), dim3( PXD_THREADS ), 0, s, ctx->dsvs + lv.m0,
                                    nm, ctx->dsvr, L, DI, X, ldx, (const V *)Y, ldy, nrhs );
            }
            return;
        }
    }
    if ( forward ) {
        if ( nd ) {
            hipLaunchKernelGGL( ( k_sv_diag<T, V, NR> ), dim3( nd ), dim3( PXD_THREADS ), sh, s, ctx->dsvc + lv.d0,
                                DI, (const V *)X, ldx, Y, ldy, nrhs, 0, (const V *)nullptr, (int64_t)0 );
        }
        if ( nf && many ) {
            hipLaunchKernelGGL( ( k_sv_fwd_tile<T, V, NR, 16> ), dim3( nf ), dim3( PXD_THREADS ), sh, s,
                                ctx->dsvt + lv.f0, ctx->dsvr, L, X, ldx, (const V *)Y, ldy, nrhs, 1 );
        }
        else if ( nf ) {
            hipLaunchKernelGGL( ( k_sv_fwd_tile<T, V, NR, WB> ), dim3( nf * split ), dim3( PXD_THREADS ), sh, s,
                                ctx->dsvt + lv.f0, ctx->dsvr, L, X, ldx, (const V *)Y, ldy, nrhs, split );
        }
    }
    else {
        if ( nf && many ) {
            hipLaunchKernelGGL( ( k_sv_bwd_tile<T, V, NR, 1> ), dim3( nf ), dim3( PXD_THREADS ), 0, s,
                                ctx->dsvt + lv.f0, ctx->dsvr, L, (const V *)X, ldx, Y, ldy, nrhs, 1, S, lds );
        }
        else if ( nf ) {
            hipLaunchKernelGGL( ( k_sv_bwd_tile<T, V, NR, WB / 16> ), dim3( nf * split_b ), dim3( PXD_THREADS ), 0,
                                s, ctx->dsvt + lv.f0, ctx->dsvr, L, (const V *)X, ldx, Y, ldy, nrhs, split_b, S,
                                lds );
        }
        if ( nd ) {
            hipLaunchKernelGGL( ( k_sv_diag<T, V, NR> ), dim3( nd ), dim3( PXD_THREADS ), sh, s, ctx->dsvc + lv.d0,
                                DI, (const V *)Y, ldy, X, ldx, nrhs, 1, (const V *)S, lds );
        }
    }
}

template <typename T, typename V>
int
solve_sweeps( pxd_ctx_t *ctx, V *X, int64_t ldx, int nrhs, int what = SOLVE_ALL, bool swap = false,
              V *Yscr = nullptr )
{
    // forward sweep: L (unit for LU and LDL^T); with swap, U^T
    const T    *L  = (const T *)( swap ? ctx->dU : ctx->dL );
    const T    *DI = (const T *)( swap ? ctx->dudinv : ctx->ddinv );
    // backward sweep: Cholesky L^T; LU: U, read from the U^T tiles with
    // the inverses of the U_kk^T; with swap, L^T
    const bool  lu = ( ctx->facto == FACT_GETRF );
    const T    *LB = (const T *)( lu && !swap ? ctx->dU : ctx->dL );
    const T    *DB = (const T *)( lu && !swap ? ctx->dudinv : ctx->ddinv );
    hipStream_t s  = ctx->stream;
    const int   nl = (int)ctx->levels.size();
    // Schur complement solved apart (solve modes local and interface): the
    // rows [sf0, n) belong to cblks whose diagonal and tiles are left out
    // (reference kernels/solve_ztrsmsp.c:358-376,513-557)
    const bool    schur = ctx->schurfcol < ctx->n && ctx->schur_solv_mode != 2;
    const bool    iface = ctx->schur_solv_mode == 1;
    const int64_t sf0   = schur ? ctx->schurfcol : ctx->n;
    const int64_t ns    = ctx->n - sf0;
    // rows of the Schur part: copied between the vectors, or cleared
    auto schur_copy = [&]( const V *from, int64_t ldf, V *to, int64_t ldt ) {
        if ( ns > 0 ) {
            hipLaunchKernelGGL( ( k_copy_cols<V> ), dim3( nblocks( ns, 256 ), nrhs ), dim3( 256 ), 0, s, ns,
                                from + sf0, ldf, to + sf0, ldt );
        }
    };
    auto schur_zero = [&]( V *to, int64_t ldt ) {
        if ( ns > 0 ) {
            hipLaunchKernelGGL( ( k_zero_cols<V> ), dim3( nblocks( ns, 256 ), nrhs ), dim3( 256 ), 0, s, ns,
                                to + sf0, ldt );
        }
    };

    if ( what == SOLVE_DIAG ) {
        // x <- D^{-1} x in place, D the diagonal of the L-side diagonal tiles
        // (solve_cblk_zdiag divides by it whatever the factorization), the
        // Schur part aside
        if ( sf0 > 0 ) {
            hipLaunchKernelGGL( ( k_diag_scale<T, V> ), dim3( nblocks( sf0, 256 ) ), dim3( 256 ), 0, s, sf0,
                                ctx->dcol2cblk, ctx->dcb, (const T *)ctx->dL, X, ldx, nrhs );
        }
        PXD_CHECK( hipGetLastError() );
        return 0;
    }
    if ( !ctx->kernels && ( what != SOLVE_ALL || ctx->schurfcol < ctx->n ) ) {
        return 3;   // the scalar kernels solve in place, both sweeps at once, without a Schur part
    }

    if ( ctx->kernels ) {
        // matrix-core path, diagonal solves chunked and out of place:
        // forward  -- the diagonal reads X and writes Y; the off-diagonal
        //             update reads x_k from Y and subtracts into X;
        // backward -- the gather reads the solved rows from X and subtracts
        //             into Y; the diagonal reads Y and writes X.
        // Y: the instance's buffer (allocated on first use), or the scratch
        // the caller hands over (n x nrhs entries)
        if ( Yscr == nullptr ) {
            size_t need = (size_t)ctx->n * PXD_MAXRHS * sizeof( V );
            if ( ctx->dsolvy_bytes < need ) {
                if ( ctx->dsolvy ) PXD_CHECK( hipFree( ctx->dsolvy ) );
                PXD_CHECK( hipMalloc( &ctx->dsolvy, need ) );
                ctx->dsolvy_bytes = need;
            }
        }
        V            *Y   = Yscr ? Yscr : (V *)ctx->dsolvy;
        const int64_t ldy = ctx->n;
        // the backward partial sums: not with the caller's scratch (nothing
        // is allocated then), nor where the kernels keep the direct atomics
        V         *S     = nullptr;
        const bool sp_ok = nrhs <= 4 ? sv_spread_ok<V, 4>::value : sv_spread_ok<V, PXD_MAXRHS>::value;
        if ( Yscr == nullptr && ctx->solve_spread && ctx->spread_cols > 0 && sp_ok ) {
            const size_t sneed = (size_t)PXD_SV_SPREAD * PXD_MAXRHS * ctx->spread_cols * sizeof( V );
            if ( ctx->dspread_bytes < sneed ) {
                if ( ctx->dspread ) PXD_CHECK( hipFree( ctx->dspread ) );
                PXD_CHECK( hipMalloc( &ctx->dspread, sneed ) );
                ctx->dspread_bytes = sneed;
            }
            S = (V *)ctx->dspread;
        }
        const int     nlv = nl;
        // the widest cblk's right-hand-side block must fit next to each
        // kernel's static LDS
        const size_t shmax = (size_t)ctx->colmax * PXD_MAXRHS * sizeof( V );
        if ( nlv > 0 && ( nrhs > PXD_MAXRHS || ctx->colmax > PXD_SV_MAXN ||
                          shmax > lds_room( (const void *)k_sv_diag<T, V, PXD_MAXRHS>, ctx->lds_max ) ||
                          shmax > lds_room( (const void *)k_sv_fwd_tile<T, V, PXD_MAXRHS, 16>, ctx->lds_max ) ||
                          shmax > lds_room( (const void *)k_sv_fwd_tile<T, V, PXD_MAXRHS, solve_wide<T>::value>,
                                            ctx->lds_max ) ) ) {
            return 3;
        }
        if ( !( what & SOLVE_FWD ) ) {
            // the backward sweep reads its input from Y
            hipLaunchKernelGGL( ( k_copy_cols<V> ), dim3( nblocks( ctx->n, 256 ), nrhs ), dim3( 256 ), 0, s, ctx->n,
                                (const V *)X, ldx, Y, ldy );
        }
        if ( ( what & SOLVE_FWD ) && schur && !iface ) {
            schur_copy( X, ldx, Y, ldy );   // local: b22 is left as it is
        }
        for ( int li = 0; li < nlv && ( what & SOLVE_FWD ); li++ ) {
            const level_t &lv = ctx->levels[li];
            unsigned       nd = (unsigned)( ( schur ? lv.ds : lv.d1 ) - lv.d0 );
            unsigned       nf = (unsigned)( ( schur ? lv.fs : lv.f1 ) - lv.f0 );
            if ( nrhs <= 4 ) {
                sv_level<T, V, 4>( ctx, true, lv, nd, nf, L, DI, X, ldx, Y, ldy, nrhs, S );
            }
            else {
                sv_level<T, V, PXD_MAXRHS>( ctx, true, lv, nd, nf, L, DI, X, ldx, Y, ldy, nrhs, S );
            }
        }
        if ( ( what & SOLVE_FWD ) && schur && iface ) {
            schur_copy( X, ldx, Y, ldy );   // interface: b22 - L21 x11
        }
        if ( ( what & SOLVE_DIAG ) && ctx->facto == FACT_SYTRF && sf0 > 0 ) {
            hipLaunchKernelGGL( ( k_diag_scale<T, V> ), dim3( nblocks( sf0, 256 ) ), dim3( 256 ), 0, s, sf0,
                                ctx->dcol2cblk, ctx->dcb, (const T *)ctx->dL, Y, ldy, nrhs );
        }
        if ( !( what & SOLVE_BWD ) ) {
            // the forward sweep leaves its result in Y
            hipLaunchKernelGGL( ( k_copy_cols<V> ), dim3( nblocks( ctx->n, 256 ), nrhs ), dim3( 256 ), 0, s, ctx->n,
                                (const V *)Y, ldy, X, ldx );
            PXD_CHECK( hipGetLastError() );
            return 0;
        }
        if ( schur ) {
            // the gathers read x22 (interface) or nothing of it (local)
            if ( iface ) {
                schur_copy( Y, ldy, X, ldx );
            }
            else {
                schur_zero( X, ldx );
            }
        }
        if ( S ) {
            // cleared by a kernel (the sweeps may be captured in a graph)
            hipLaunchKernelGGL( ( k_zero_cols<V> ), dim3( nblocks( ctx->spread_cols, 256 ), PXD_SV_SPREAD * nrhs ),
                                dim3( 256 ), 0, s, ctx->spread_cols, S, ctx->spread_cols );
        }
        for ( int li = nlv - 1; li >= 0; li-- ) {
            const level_t &lv = ctx->levels[li];
            unsigned       nd = (unsigned)( ( schur ? lv.ds : lv.d1 ) - lv.d0 );
            unsigned       nf = (unsigned)( ( schur ? lv.fs : lv.f1 ) - lv.f0 );
            if ( nrhs <= 4 ) {
                sv_level<T, V, 4>( ctx, false, lv, nd, nf, LB, DB, X, ldx, Y, ldy, nrhs, S );
            }
            else {
                sv_level<T, V, PXD_MAXRHS>( ctx, false, lv, nd, nf, LB, DB, X, ldx, Y, ldy, nrhs, S );
            }
        }
        if ( schur && !iface ) {
            schur_copy( Y, ldy, X, ldx );   // local: the rows of the Schur part are untouched
        }
        PXD_CHECK( hipGetLastError() );
        return 0;
    }

    // scalar kernels (real factors, vectors in the factor precision), in place
    if constexpr ( !std::is_same<T, typename px_real<T>::type>::value || !std::is_same<T, V>::value ) {
        return 3;
    }
    else {
    for ( int li = 0; li < nl; li++ ) {
        const level_t &lv = ctx->levels[li];
        unsigned       np = (unsigned)( lv.c1 - lv.c0 );
        unsigned       nf = (unsigned)( lv.f1 - lv.f0 );
        size_t         sh = (size_t)lv.maxn * nrhs * sizeof( T );
        hipLaunchKernelGGL( ( k_solve_diag<T> ), dim3( np ), dim3( PXD_THREADS ), sh, s,
                            ctx->dlist + lv.c0, ctx->dcb, L, DI, X, ldx, nrhs, 0, 0 );
        if ( nf ) {
            hipLaunchKernelGGL( ( k_solve_fwd_upd<T, T, 16> ), dim3( nf ), dim3( PXD_THREADS ), sh, s,
                                ctx->dsolv + lv.f0, ctx->dcb, ctx->dbk, L, X, ldx, nrhs, (const T *)X, ldx );
        }
    }
    for ( int li = nl - 1; li >= 0; li-- ) {
        const level_t &lv = ctx->levels[li];
        unsigned       np = (unsigned)( lv.c1 - lv.c0 );
        unsigned       nf = (unsigned)( lv.f1 - lv.f0 );
        size_t         sh = (size_t)lv.maxn * nrhs * sizeof( T );
        if ( nf ) {
            hipLaunchKernelGGL( ( k_solve_bwd_gather<T> ), dim3( nf ), dim3( PXD_THREADS ), 0, s,
                                ctx->dsolv + lv.f0, ctx->dcb, ctx->dbk, L, X, ldx, nrhs );
        }
        hipLaunchKernelGGL( ( k_solve_diag<T> ), dim3( np ), dim3( PXD_THREADS ), sh, s,
                            ctx->dlist + lv.c0, ctx->dcb, L, DI, X, ldx, nrhs, 1, 0 );
    }
    }
    PXD_CHECK( hipGetLastError() );
    return 0;
}

// dynamic-LDS limits: each kernel may take what its static LDS leaves of
// the CU's lds bytes
inline void
set_lds_room( const void *fn, size_t lds )
{
    (void)hipFuncSetAttribute( fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_room( fn, lds ) );
}

// the solve kernels for factor entries T and vectors V
template <typename T, typename V>
void
set_lds_attributes_solve( size_t lds )
{
    set_lds_room( (const void *)k_sv_diag<T, V, 4>, lds );
    set_lds_room( (const void *)k_sv_diag<T, V, PXD_MAXRHS>, lds );
    set_lds_room( (const void *)k_sv_fwd_tile<T, V, 4, 16>, lds );
    set_lds_room( (const void *)k_sv_fwd_tile<T, V, PXD_MAXRHS, 16>, lds );
    set_lds_room( (const void *)k_sv_fwd_tile<T, V, 4, solve_wide<T>::value>, lds );
    set_lds_room( (const void *)k_sv_fwd_tile<T, V, PXD_MAXRHS, solve_wide<T>::value>, lds );
}

// the matrix-core kernels (all four arithmetics)
template <typename T>
void
set_lds_attributes_mfma( size_t lds )
{
    if constexpr ( std::is_same<T, typename px_real<T>::type>::value ) {
        set_lds_room( (const void *)k_potrf2<T, false>, lds );
        set_lds_room( (const void *)k_potrf_mw<T>, lds );
    }
    set_lds_room( (const void *)k_potrf2<T, true>, lds );
    set_lds_room( (const void *)k_trtri<T>, lds );
    set_lds_attributes_solve<T, T>( lds );
}

// and the scalar kernels (real arithmetics)
template <typename T>
void
set_lds_attributes( size_t lds )
{
    set_lds_attributes_mfma<T>( lds );
    set_lds_room( (const void *)k_trsm<T, 32>, lds );
    set_lds_room( (const void *)k_trsm<T, 16>, lds );
    set_lds_room( (const void *)k_solve_diag<T>, lds );
    set_lds_room( (const void *)k_solve_fwd_upd<T, T, 16>, lds );
}

// Solve of nrhs vectors of type TV (double or complex double) with a factor
// of type T: directly when the types agree, otherwise through vectors
// converted to the factor precision and back, as the reference's mixed
// solve does.

void
drop_solve_graphs( pxd_ctx_t *ctx )
{
    for ( solve_graph &g : ctx->sgraphs ) {
        (void)hipGraphExecDestroy( g.exec );
    }
    ctx->sgraphs.clear();
}

// Whether the sweeps of a pass run as a captured graph: the full solve of
// the matrix-core path without the one-launch sweeps (their pass counters
// change from call to call) and without a Schur part.
bool
graph_ok( const pxd_ctx_t *ctx, int what, bool swap )
{
    return ctx->solve_graph && ctx->kernels && what == SOLVE_ALL && !swap &&
           ctx->schurfcol >= ctx->n;
}

// Instance buffers of a graph pass: the vectors (n x PXD_MAXRHS of T, ld n)
// and the sweeps' Y, allocated before any capture.
template <typename T>
int
graph_buffers( pxd_ctx_t *ctx )
{
    const size_t need = (size_t)ctx->n * PXD_MAXRHS * sizeof( T );
    if ( ctx->dvecs_bytes < need ) {
        if ( ctx->dvecs ) PXD_CHECK( hipFree( ctx->dvecs ) );
        PXD_CHECK( hipMalloc( &ctx->dvecs, need ) );
        ctx->dvecs_bytes = need;
    }
    if ( ctx->dsolvy_bytes < need ) {
        if ( ctx->dsolvy ) PXD_CHECK( hipFree( ctx->dsolvy ) );
        PXD_CHECK( hipMalloc( &ctx->dsolvy, need ) );
        ctx->dsolvy_bytes = need;
    }
    // the backward partial sums, so that no allocation happens under capture
    const size_t sneed = (size_t)PXD_SV_SPREAD * PXD_MAXRHS * ctx->spread_cols * sizeof( T );
    if ( sv_spread_ok<T, 4>::value && ctx->solve_spread && ctx->dspread_bytes < sneed ) {
        if ( ctx->dspread ) PXD_CHECK( hipFree( ctx->dspread ) );
        PXD_CHECK( hipMalloc( &ctx->dspread, sneed ) );
        ctx->dspread_bytes = sneed;
    }
    return 0;
}

// Both sweeps of nq columns held in ctx->dvecs, replayed from a graph that
// is captured on first use (and again when what it baked in has changed).
template <typename T>
int
graph_pass( pxd_ctx_t *ctx, int nq )
{
    const void *sig[7] = { ctx->dvecs, ctx->dsolvy, ctx->dL, ctx->dU, ctx->ddinv, ctx->dudinv,
                           ctx->solve_spread ? ctx->dspread : nullptr };
    solve_graph *hit  = nullptr;
    for ( size_t i = 0; i < ctx->sgraphs.size(); i++ ) {
        solve_graph &g = ctx->sgraphs[i];
        if ( g.nq == nq && g.tag == type_tag<T>::value ) {
            if ( g.gen == ctx->sched_gen && g.wide == ctx->solve_wide_max && !memcmp( g.sig, sig, sizeof( sig ) ) ) {
                hit = &g;
            }
            else {
                (void)hipGraphExecDestroy( g.exec );
                ctx->sgraphs.erase( ctx->sgraphs.begin() + i );
            }
            break;
        }
    }
    if ( hit == nullptr ) {
        hipGraph_t     g  = nullptr;
        hipGraphExec_t ex = nullptr;
        PXD_CHECK( hipStreamBeginCapture( ctx->stream, hipStreamCaptureModeThreadLocal ) );
        const int  rc = solve_sweeps<T, T>( ctx, (T *)ctx->dvecs, ctx->n, nq, SOLVE_ALL, false, nullptr );
        hipError_t e  = hipStreamEndCapture( ctx->stream, &g );
        if ( rc != 0 || e != hipSuccess || g == nullptr ) {
            if ( g ) (void)hipGraphDestroy( g );
            (void)hipGetLastError();
            return rc ? rc : 6;
        }
        e = hipGraphInstantiate( &ex, g, nullptr, nullptr, 0 );
        (void)hipGraphDestroy( g );
        if ( e != hipSuccess ) {
            (void)hipGetLastError();
            return 6;
        }
        solve_graph ng;
        ng.nq   = nq;
        ng.tag  = type_tag<T>::value;
        memcpy( ng.sig, sig, sizeof( sig ) );
        ng.gen  = ctx->sched_gen;
        ng.wide = ctx->solve_wide_max;
        ng.exec = ex;
        ctx->sgraphs.push_back( ng );
        hit = &ctx->sgraphs.back();
    }
    PXD_CHECK( hipGraphLaunch( hit->exec, ctx->stream ) );
    return 0;
}

template <typename TV, typename T>
int
solve_vectors( pxd_ctx_t *ctx, int64_t nrhs, void *dx, int64_t ldx, int what = SOLVE_ALL, bool swap = false,
               void *scr = nullptr )
{
    int rc = 0;
    const bool graph = scr == nullptr && graph_ok( ctx, what, swap );
    if ( graph && ( rc = graph_buffers<T>( ctx ) ) != 0 ) {
        return rc;
    }
    if constexpr ( std::is_same<TV, T>::value ) {
        const int pass = solve_pass_rhs<T>( ctx );
        for ( int64_t q0 = 0; q0 < nrhs && !rc; q0 += pass ) {
            int nq = (int)std::min<int64_t>( pass, nrhs - q0 );
            if ( graph ) {
                // through the instance's vectors, which the graph reads
                T *xs = (T *)ctx->dvecs;
                hipLaunchKernelGGL( ( k_copy_cols<T> ), dim3( nblocks( ctx->n, 256 ), nq ), dim3( 256 ), 0,
                                    ctx->stream, ctx->n, (const T *)dx + q0 * ldx, ldx, xs, ctx->n );
                rc = graph_pass<T>( ctx, nq );
                hipLaunchKernelGGL( ( k_copy_cols<T> ), dim3( nblocks( ctx->n, 256 ), nq ), dim3( 256 ), 0,
                                    ctx->stream, ctx->n, (const T *)xs, ctx->n, (T *)dx + q0 * ldx, ldx );
            }
            else {
                rc = solve_sweeps<T, T>( ctx, (T *)dx + q0 * ldx, ldx, nq, what, swap, (T *)scr );
            }
        }
        return rc;
    }
    else {
        const int pass = solve_pass_rhs<T>( ctx );
        size_t need = (size_t)ctx->n * std::min<int64_t>( nrhs, pass ) * sizeof( T );
        // the factor-precision copies: the instance's buffer, or the start of
        // the caller's scratch with the sweeps' Y after them
        T *xs = (T *)scr;
        T *ys = nullptr;
        if ( scr == nullptr ) {
            if ( ctx->dvecs_bytes < need ) {
                if ( ctx->dvecs ) PXD_CHECK( hipFree( ctx->dvecs ) );
                PXD_CHECK( hipMalloc( &ctx->dvecs, need ) );
                ctx->dvecs_bytes = need;
            }
            xs = (T *)ctx->dvecs;
        }
        else {
            ys = (T *)( (char *)scr + ( ( need + 255 ) & ~(size_t)255 ) );
        }
        for ( int64_t q0 = 0; q0 < nrhs && !rc; q0 += pass ) {
            int nq = (int)std::min<int64_t>( pass, nrhs - q0 );
            for ( int q = 0; q < nq; q++ ) {
                hipLaunchKernelGGL( ( k_convert<T, TV> ), dim3( nblocks( ctx->n, 256 ) ), dim3( 256 ), 0, ctx->stream,
                                    ctx->n, (const TV *)dx + ( q0 + q ) * ldx, xs + q * ctx->n );
            }
            rc = graph ? graph_pass<T>( ctx, nq ) : solve_sweeps<T, T>( ctx, xs, ctx->n, nq, what, swap, ys );
            for ( int q = 0; q < nq; q++ ) {
                hipLaunchKernelGGL( ( k_convert<TV, T> ), dim3( nblocks( ctx->n, 256 ) ), dim3( 256 ), 0, ctx->stream,
                                    ctx->n, xs + q * ctx->n, (TV *)dx + ( q0 + q ) * ldx );
            }
        }
    }
    return rc;
}

// The solve steps `what` (swap: the sides of an LU factor exchanged) on
// nrhs device vectors in the double (complex double) precision.
int
solve_steps( pxd_ctx_t *ctx, int what, bool swap, int64_t nrhs, void *dx, int64_t ldx )
{
    PXD_CHECK( hipSetDevice( ctx->device ) );
    if ( ctx->dL == nullptr || nrhs < 0 || ( nrhs > 0 && ( dx == nullptr || ldx < ctx->n ) ) ) {
        return 7;
    }
    if ( nrhs == 0 || ctx->n == 0 ) {
        return 0;
    }
    if ( ctx->schurfcol < ctx->n && ctx->schur_solv_mode == 2 && ctx->schur_fact_local ) {
        return 7;   // the Schur part was not factorized
    }
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int        rc = 0;
    if ( hipEventCreate( &e0 ) != hipSuccess || hipEventCreate( &e1 ) != hipSuccess ) {
        if ( e0 ) (void)hipEventDestroy( e0 );
        return 3;
    }
    if ( hipEventRecord( e0, ctx->stream ) != hipSuccess ) {
        rc = 3;
    }
    if ( rc == 0 ) {
        switch ( ctx->fflt ) {
        case FT_DOUBLE: rc = solve_vectors<double, double>( ctx, nrhs, dx, ldx, what, swap ); break;
        case FT_FLOAT: rc = solve_vectors<double, float>( ctx, nrhs, dx, ldx, what, swap ); break;
        case FT_CPLX64: rc = solve_vectors<pxcd, pxcd>( ctx, nrhs, dx, ldx, what, swap ); break;
        default: rc = solve_vectors<pxcd, pxcf>( ctx, nrhs, dx, ldx, what, swap ); break;
        }
    }
    if ( hipEventRecord( e1, ctx->stream ) != hipSuccess || hipEventSynchronize( e1 ) != hipSuccess ) {
        rc = rc ? rc : 3;
    }
    else {
        ctx->stats.solve_seconds = 1e-3 * ev_ms( e0, e1 );
    }
    (void)hipEventDestroy( e0 );
    (void)hipEventDestroy( e1 );
    return rc;
}

// ---------------------------------------------------------------------------
// Scratch of the allocation-free solve (pxd_solve_noalloc): the sweeps'
// out-of-place vectors and, for a single-precision factor, the copies of the
// right-hand sides in the factor precision.  One copy of the array per GPU,
// held until the process ends; all calls on one GPU share it.  Host code
// never passes the array's name to a kernel (on the host it only stands for
// the symbol): it uses the device address that hipGetSymbolAddress returns
// for each device, kept in scratch_ptr.
}  // namespace

static __device__ __align__( 256 ) unsigned char scratch[PXD_NOALLOC_BYTES];

namespace {

std::mutex          scratch_mu;
std::vector<void *> scratch_ptr;   // per device ordinal

// Once per device, with that device current (from pxd_create).
int
scratch_setup( int device )
{
    std::lock_guard<std::mutex> lk( scratch_mu );
    if ( (int)scratch_ptr.size() <= device ) {
        scratch_ptr.resize( device + 1, nullptr );
    }
    if ( scratch_ptr[device] != nullptr ) {
        return 0;
    }
    void      *ptr = nullptr;
    hipError_t e   = hipGetSymbolAddress( (void **)&ptr, HIP_SYMBOL( scratch ) );
    if ( e != hipSuccess || ptr == nullptr ) {
        (void)hipGetLastError();
        return 6;
    }
    scratch_ptr[device] = ptr;
    return 0;
}

void *
scratch_address( int device )
{
    std::lock_guard<std::mutex> lk( scratch_mu );
    return ( device >= 0 && device < (int)scratch_ptr.size() ) ? scratch_ptr[device] : nullptr;
}

// Bytes of scratch a solve of nrhs columns takes: Y of one pass (at most
// PXD_MAXRHS columns) in the factor precision, plus for a single factor the
// factor-precision copies of the double vectors (256-byte aligned).
size_t
noalloc_need( const pxd_ctx_t *ctx, int64_t nrhs )
{
    const size_t cols = (size_t)std::min<int64_t>( nrhs, PXD_MAXRHS );
    const size_t n    = (size_t)ctx->n;
    auto         al   = []( size_t b ) { return ( b + 255 ) & ~(size_t)255; };
    switch ( ctx->fflt ) {
    case FT_DOUBLE: return n * cols * sizeof( double );
    case FT_CPLX64: return n * cols * sizeof( pxcd );
    case FT_FLOAT: return al( n * cols * sizeof( float ) ) + n * cols * sizeof( float );
    default: return al( n * cols * sizeof( pxcf ) ) + n * cols * sizeof( pxcf );
    }
}

}  // namespace

// ===========================================================================
extern "C" {

int
pxd_device_count( void )
{
    int n = 0;
    if ( hipGetDeviceCount( &n ) != hipSuccess ) {
        return 0;
    }
    return n;
}

// The bound of the panel's waits on the context's device, in ms of the
// constant clock (whose rate the device reports).
static int
set_wait_ms( pxd_ctx_t *ctx, double ms )
{
    const uint64_t ticks = (uint64_t)( std::max( 0., ms ) * std::max( 1, ctx->clock_khz ) );
    PXD_CHECK( hipSetDevice( ctx->device ) );
    PXD_CHECK( hipMemcpyToSymbol( HIP_SYMBOL( pxd_wait_ticks ), &ticks, sizeof( ticks ) ) );
    return 0;
}

// Engine options chosen by the autotuner (pastix_amd/autotune.py): a text
// file of "key value" lines ('#' starts a comment), read at every
// pxd_create from $PXD_TUNING_FILE, or else from tuning_mi355x.txt next to
// this library.  Each line is applied with pxd_set_option; a missing file
// leaves the built-in defaults, an unknown key is reported and skipped.
static void
load_tuning( pxd_ctx_t *ctx )
{
    std::string path;
    if ( const char *env = getenv( "PXD_TUNING_FILE" ) ) {
        path = env;
    }
    else {
        Dl_info info;
        if ( dladdr( (void *)&load_tuning, &info ) && info.dli_fname ) {
            path = info.dli_fname;
            const size_t slash = path.rfind( '/' );
            path = ( slash == std::string::npos ? std::string( "." ) : path.substr( 0, slash ) ) +
                   "/tuning_mi355x.txt";
        }
    }
    FILE *f = path.empty() ? nullptr : fopen( path.c_str(), "r" );
    if ( f == nullptr ) {
        return;
    }
    char line[256];
    while ( fgets( line, sizeof( line ), f ) ) {
        char  *hash = strchr( line, '#' );
        char   key[128];
        double value;
        if ( hash ) {
            *hash = '\0';
        }
        if ( sscanf( line, "%127s %lf", key, &value ) == 2 && pxd_set_option( ctx, key, value ) != 0 ) {
            fprintf( stderr, "pastix_amd device: %s: unknown option %s\n", path.c_str(), key );
        }
    }
    fclose( f );
    ctx->tuning_file = path;
}

int
pxd_create( pxd_ctx_t **pctx, int device )
{
    int ndev = pxd_device_count();
    if ( ndev <= 0 ) {
        fprintf( stderr, "pastix_amd device: no GPU visible; the numerical steps run only on the GPU\n" );
        return 3; /* PASTIX_ERR_NOTIMPLEMENTED */
    }
    if ( device < 0 || device >= ndev ) {
        return 7;
    }
    pxd_ctx_t *ctx = new pxd_ctx_t();
    ctx->device = device;
    PXD_CHECK( hipSetDevice( device ) );
    PXD_CHECK( hipStreamCreateWithFlags( &ctx->stream, hipStreamNonBlocking ) );
    PXD_CHECK( hipStreamCreateWithFlags( &ctx->stream2, hipStreamNonBlocking ) );
    PXD_CHECK( hipEventCreateWithFlags( &ctx->ev_ready, hipEventDisableTiming ) );
    PXD_CHECK( hipEventCreateWithFlags( &ctx->ev_panel, hipEventDisableTiming ) );
    hipDeviceProp_t prop;
    PXD_CHECK( hipGetDeviceProperties( &prop, device ) );
    ctx->lds_max = prop.sharedMemPerBlock;
    PXD_CHECK( hipMalloc( (void **)&ctx->dscratch, 64 * sizeof( int ) ) );
    PXD_CHECK( hipMalloc( &ctx->dred, 4096 ) );
    if ( scratch_setup( device ) != 0 ) {
        // every other entry point works without it
        fprintf( stderr, "pastix_amd device: no address for the allocation-free solve's scratch on device %d; "
                         "pxd_solve_noalloc will refuse\n", device );
    }
    set_lds_attributes<double>( ctx->lds_max );
    set_lds_attributes<float>( ctx->lds_max );
    set_lds_attributes_mfma<pxcd>( ctx->lds_max );
    set_lds_attributes_mfma<pxcf>( ctx->lds_max );
    (void)hipGetLastError();   // an attribute refused above must not be reported later
    PXD_CHECK( hipDeviceGetAttribute( &ctx->clock_khz, hipDeviceAttributeWallClockRate, device ) );
    if ( set_wait_ms( ctx, 2000. ) != 0 ) {
        return 6;
    }
    load_tuning( ctx );
    if ( getenv( "PXD_VERBOSE" ) ) {
        fprintf( stderr, "pastix_amd device: %s, %d CUs, LDS/block %zu, L2 %d, tuning %s\n", prop.gcnArchName,
                 prop.multiProcessorCount, ctx->lds_max, prop.l2CacheSize,
                 ctx->tuning_file.empty() ? "(defaults)" : ctx->tuning_file.c_str() );
    }
    *pctx = ctx;
    return 0;
}

void
pxd_destroy( pxd_ctx_t *ctx )
{
    if ( ctx == nullptr ) {
        return;
    }
    (void)hipSetDevice( ctx->device );
    (void)hipStreamSynchronize( ctx->stream );
    drop_solve_graphs( ctx );
    if ( ctx->fgraph ) {
        (void)hipGraphExecDestroy( ctx->fgraph );
    }
    void *ptrs[] = { ctx->dcb, ctx->dbk, ctx->dcol2cblk, ctx->dlist, ctx->dpan, ctx->dflags, ctx->dtrsm, ctx->dupd, ctx->dupdb, ctx->dupds, ctx->dsolv, ctx->dsvt, ctx->dsvr, ctx->dsvc, ctx->dsvs, ctx->dspread, ctx->ddiag, ctx->dsolvy, ctx->dinvt,
                     ctx->dcolptr, ctx->drowtab, ctx->dtcolptr, ctx->dtrowtab, ctx->dval, ctx->dtval,
                     ctx->dL, ctx->dU, ctx->ddinv, ctx->dudinv, ctx->dscratch, ctx->dvecs, ctx->dred,
                     };
    for ( void *p : ptrs ) {
        if ( p ) {
            (void)hipFree( p );
        }
    }
    (void)hipStreamSynchronize( ctx->stream2 );
    (void)hipEventDestroy( ctx->ev_ready );
    (void)hipEventDestroy( ctx->ev_panel );
    (void)hipStreamDestroy( ctx->stream2 );

    (void)hipStreamDestroy( ctx->stream );
    delete ctx;
}

int
pxd_sync( pxd_ctx_t *ctx )
{
    PXD_CHECK( hipStreamSynchronize( ctx->stream ) );
    return 0;
}

int
pxd_get_stats( const pxd_ctx_t *ctx, pxd_stats_t *stats )
{
    *stats = ctx->stats;
    return 0;
}

int
pxd_model_log( const pxd_ctx_t *ctx, pxd_model_entry_t *entries, int64_t max, int64_t *count )
{
    const int64_t n = (int64_t)ctx->model_entries.size();
    if ( count ) {
        *count = n;
    }
    if ( entries && max > 0 ) {
        std::memcpy( entries, ctx->model_entries.data(), sizeof( pxd_model_entry_t ) * std::min( n, max ) );
    }
    return 0;
}

int
pxd_set_option( pxd_ctx_t *ctx, const char *key, double value )
{
    std::string k( key );
    if ( k == "upd_small" ) {
        ctx->upd_small   = ( value != 0. );
        ctx->tasks_dirty = 1;
        return 0;
    }
    if ( k == "solve_spread" ) {
        if ( value < 0. || value > 1e6 ) {
            return 7;
        }
        ctx->solve_spread = (int)value;
        ctx->tasks_dirty  = 1;   // the records take it at the next factorization
        return 0;
    }
    if ( k == "solve_small" ) {
        ctx->solve_small = ( value != 0. );
        drop_solve_graphs( ctx );
        return 0;
    }
    if ( k == "solve_split_max" || k == "solve_split_wg" ) {
        if ( value < 1. || value > 1e6 ) {
            return 7;
        }
        if ( k == "solve_split_max" ) ctx->solve_split_max = (int)value;
        else ctx->solve_split_wg = (int64_t)value;
        drop_solve_graphs( ctx );
        return 0;
    }
    if ( k == "upd_dma_pad" ) {
        if ( value < 0. || value > 65536. ) {
            return 7;
        }
        ctx->upd_dma_pad = (unsigned)value;
        ctx->tasks_dirty = 1;
        return 0;
    }
    if ( k == "upd_dma" ) {
        ctx->upd_dma     = ( value != 0. );
        ctx->tasks_dirty = 1;
        return 0;
    }
    if ( k == "upd_small_m" || k == "upd_small_n" ) {
        const int v = (int)std::max( 0., std::min( value, 1e6 ) );
        if ( k == "upd_small_m" ) ctx->upd_small_m = std::min( v, PXD_SMALL_M );
        else ctx->upd_small_n = std::min( v, PXD_SMALL_N );
        ctx->tasks_dirty = 1;
        return 0;
    }
    if ( k == "wait_ms" ) {
        // for every context on this device: the bound lives in device memory
        return set_wait_ms( ctx, value );
    }
    if ( k == "fact_graph" ) {
        ctx->fact_graph = ( value != 0. );
        return 0;
    }
    if ( k == "solve_graph" ) {
        ctx->solve_graph = ( value != 0. );
        return 0;
    }
    if ( k == "solve_wide_max" ) {
        ctx->solve_wide_max = (int64_t)std::max( 0., std::min( value, 1e15 ) );
        return 0;
    }
    if ( k == "panel_mw" ) {
        ctx->panel_mw = ( value != 0. );
        return 0;
    }
    if ( k == "prio" ) {
        ctx->prio = ( value != 0. );
        return 0;
    }
    if ( k == "model_log" ) {
        ctx->model_log = ( value != 0. );
        return 0;
    }
    if ( k == "timing" ) {
        ctx->timing = ( value != 0. );
        return 0;
    }
    if ( k == "schur_solv_mode" ) {
        if ( value < 0. || value > 2. ) {
            return 7;
        }
        ctx->schur_solv_mode = (int)value;
        return 0;
    }
    if ( k == "kernels" ) {
        // takes effect at the next factorization
        ctx->kernels     = ( value != 0. );
        ctx->tasks_dirty = 1;
        return 0;
    }
    if ( k == "upd_big_rows" ) {
        // takes effect at the next factorization
        ctx->upd_big_rows = (int)std::max( 1., std::min( value, 2e9 ) );
        ctx->tasks_dirty  = 1;
        return 0;
    }
    return 7;
}

int
pxd_malloc( pxd_ctx_t *ctx, size_t bytes, void **dptr )
{
    (void)ctx;
    PXD_CHECK( hipMalloc( dptr, std::max<size_t>( bytes, 1 ) ) );
    return 0;
}

int
pxd_free( pxd_ctx_t *ctx, void *dptr )
{
    (void)ctx;
    if ( dptr ) {
        PXD_CHECK( hipFree( dptr ) );
    }
    return 0;
}

int
pxd_memset0( pxd_ctx_t *ctx, void *dptr, size_t bytes )
{
    PXD_CHECK( hipMemsetAsync( dptr, 0, bytes, ctx->stream ) );
    return 0;
}

int
pxd_h2d( pxd_ctx_t *ctx, void *dst, const void *src, size_t bytes )
{
    PXD_CHECK( hipMemcpyAsync( dst, src, bytes, hipMemcpyHostToDevice, ctx->stream ) );
    PXD_CHECK( hipStreamSynchronize( ctx->stream ) );
    return 0;
}

int
pxd_d2h( pxd_ctx_t *ctx, void *dst, const void *src, size_t bytes )
{
    PXD_CHECK( hipMemcpyAsync( dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream ) );
    PXD_CHECK( hipStreamSynchronize( ctx->stream ) );
    return 0;
}

int
pxd_set_update_classes( pxd_ctx_t *ctx, const int8_t *cls, int64_t count )
{
    if ( count != 0 && ( cls == nullptr || count != 2 * ctx->bloknbr ) ) {
        return 7;
    }
    ctx->upd_cls.assign( cls, cls + count );
    ctx->tasks_dirty = 1;
    return 0;
}

int
pxd_set_schur( pxd_ctx_t *ctx, int64_t schurfcol, int fact_local )
{
    if ( schurfcol < 0 || schurfcol > ctx->n ) {
        return 7;
    }
    for ( int64_t k = 0; k < ctx->cblknbr; k++ ) {
        // the Schur part must start at a cblk boundary
        if ( ctx->hcb[k].fcolnum < schurfcol && ctx->hcb[k].lcolnum >= schurfcol ) {
            return 7;
        }
    }
    if ( ctx->schurfcol != schurfcol || ctx->schur_fact_local != ( fact_local != 0 ) ) {
        ctx->schurfcol        = schurfcol;
        ctx->schur_fact_local = fact_local != 0;
        ctx->tasks_dirty      = 1;
    }
    return 0;
}

int
pxd_set_structure( pxd_ctx_t *ctx, int64_t n, int64_t cblknbr, int64_t bloknbr,
                   const pxd_cblk_t *cblks, const pxd_blok_t *bloks, const int64_t *browtab,
                   int64_t coefnbr )
{
    if ( n >= INT32_MAX || cblknbr >= INT32_MAX || bloknbr >= INT32_MAX ) {
        return 3;
    }
    PXD_CHECK( hipSetDevice( ctx->device ) );
    ctx->n       = n;
    ctx->schurfcol = n;   // no Schur part until pxd_set_schur
    ctx->cblknbr = cblknbr;
    ctx->bloknbr = bloknbr;
    ctx->coefnbr = coefnbr;
    ctx->hcb.resize( cblknbr + 1 );
    ctx->hbk.resize( bloknbr + 1 );
    ctx->colmax = 0;
    int64_t dinv = 0;
    std::vector<int32_t> col2cblk( std::max<int64_t>( n, 1 ) );
    for ( int64_t k = 0; k <= cblknbr; k++ ) {
        dcblk &c = ctx->hcb[k];
        c.fcolnum = (int32_t)cblks[k].fcolnum;
        c.lcolnum = (int32_t)cblks[k].lcolnum;
        c.fblok   = (int32_t)cblks[k].fblok;
        c.stride  = (int32_t)cblks[k].stride;
        c.coefoff = cblks[k].coefoff;
        c.dinvoff = dinv;
        if ( k < cblknbr ) {
            int32_t nk = c.lcolnum - c.fcolnum + 1;
            ctx->colmax = std::max( ctx->colmax, nk );
            // n x n explicit inverse (matrix-core path) or the inverted
            // PXD_NB x PXD_NB diagonal sub-blocks (scalar path)
            dinv += std::max<int64_t>( (int64_t)nk * nk,
                                       (int64_t)( ( nk + PXD_NB - 1 ) / PXD_NB ) * PXD_NB * PXD_NB );
            for ( int32_t j = c.fcolnum; j <= c.lcolnum; j++ ) {
                col2cblk[j] = (int32_t)k;
            }
        }
    }
    ctx->dinvnbr = dinv;
    // the update kernels address a cblk's coefficients with 32-bit byte
    // offsets (buffer loads): every cblk must stay below 2 GiB in double
    for ( int64_t k = 0; k < cblknbr; k++ ) {
        const dcblk &c = ctx->hcb[k];
        if ( (double)c.stride * ( c.lcolnum - c.fcolnum + 1 ) * 8. >= 2147483648. ) {
            fprintf( stderr, "pastix_amd device: cblk %lld holds more than 2 GiB of coefficients\n", (long long)k );
            return 3;
        }
    }
    for ( int64_t b = 0; b <= bloknbr; b++ ) {
        dblok &bb  = ctx->hbk[b];
        bb.frownum = (int32_t)bloks[b].frownum;
        bb.lrownum = (int32_t)bloks[b].lrownum;
        bb.fcblknm = (int32_t)bloks[b].fcblknm;
        if ( b < bloknbr ) {
            int64_t lc = bloks[b].lcblknm;
            int64_t nk = cblks[lc].lcolnum - cblks[lc].fcolnum + 1;
            bb.rowoff = (int32_t)( bloks[b].coefind / nk );
        }
        else {
            bb.rowoff = 0;
        }
    }
    if ( upload( &ctx->dcb, ctx->hcb ) || upload( &ctx->dbk, ctx->hbk ) || upload( &ctx->dcol2cblk, col2cblk ) ) {
        return 6;
    }
    compute_levels( ctx, cblks, bloks, browtab );
    ctx->tasks_dirty = 0;
    return build_schedule( ctx );
}

int
pxd_set_matrix( pxd_ctx_t *ctx, int flttype, int mtxtype, int64_t n, int64_t nnz,
                const int64_t *colptr, const int64_t *rowtab, const void *values,
                const int64_t *tcolptr, const int64_t *trowtab, const void *tvalues )
{
    PXD_CHECK( hipSetDevice( ctx->device ) );
    if ( flttype != FT_DOUBLE && flttype != FT_CPLX64 ) {
        return 7;
    }
    if ( n != ctx->n ) {
        return 7;
    }
    ctx->mflt = flttype;
    ctx->mtx  = mtxtype;
    ctx->nnz  = nnz;
    size_t es = elt_size( flttype );
    void *ptrs[] = { ctx->dcolptr, ctx->drowtab, ctx->dval, ctx->dtcolptr, ctx->dtrowtab, ctx->dtval };
    for ( void *p : ptrs ) {
        if ( p ) PXD_CHECK( hipFree( p ) );
    }
    ctx->dcolptr = ctx->drowtab = ctx->dtcolptr = ctx->dtrowtab = nullptr;
    ctx->dval = ctx->dtval = nullptr;
    PXD_CHECK( hipMalloc( (void **)&ctx->dcolptr, ( n + 1 ) * sizeof( int64_t ) ) );
    PXD_CHECK( hipMalloc( (void **)&ctx->drowtab, std::max<int64_t>( nnz, 1 ) * sizeof( int64_t ) ) );
    PXD_CHECK( hipMalloc( &ctx->dval, std::max<int64_t>( nnz, 1 ) * es ) );
    PXD_CHECK( hipMemcpy( ctx->dcolptr, colptr, ( n + 1 ) * sizeof( int64_t ), hipMemcpyHostToDevice ) );
    PXD_CHECK( hipMemcpy( ctx->drowtab, rowtab, nnz * sizeof( int64_t ), hipMemcpyHostToDevice ) );
    PXD_CHECK( hipMemcpy( ctx->dval, values, nnz * es, hipMemcpyHostToDevice ) );
    if ( tcolptr ) {
        PXD_CHECK( hipMalloc( (void **)&ctx->dtcolptr, ( n + 1 ) * sizeof( int64_t ) ) );
        PXD_CHECK( hipMalloc( (void **)&ctx->dtrowtab, std::max<int64_t>( nnz, 1 ) * sizeof( int64_t ) ) );
        PXD_CHECK( hipMalloc( &ctx->dtval, std::max<int64_t>( nnz, 1 ) * es ) );
        PXD_CHECK( hipMemcpy( ctx->dtcolptr, tcolptr, ( n + 1 ) * sizeof( int64_t ), hipMemcpyHostToDevice ) );
        PXD_CHECK( hipMemcpy( ctx->dtrowtab, trowtab, nnz * sizeof( int64_t ), hipMemcpyHostToDevice ) );
        PXD_CHECK( hipMemcpy( ctx->dtval, tvalues, nnz * es, hipMemcpyHostToDevice ) );
    }
    // ||A||_inf on the host copy (rows of P A P^T = columns of the transposed storage)
    {
        std::vector<double> rs( n, 0. );
        for ( int64_t e = 0; e < nnz; e++ ) {
            double a;
            if ( flttype == FT_DOUBLE ) {
                a = std::fabs( ( (const double *)values )[e] );
            }
            else {
                const double *z = (const double *)values + 2 * e;
                a = std::hypot( z[0], z[1] );
            }
            rs[rowtab[e]] += a;
        }
        ctx->norm_inf = 0.;
        for ( double v : rs ) ctx->norm_inf = std::max( ctx->norm_inf, v );
    }
    return 0;
}

int
pxd_matrix_norm_inf( pxd_ctx_t *ctx, double *norm )
{
    *norm = ctx->norm_inf;
    return 0;
}

int
pxd_factorize( pxd_ctx_t *ctx, int factotype, int factor_flttype, double threshold, int64_t *nbpivots )
{
    PXD_CHECK( hipSetDevice( ctx->device ) );
    if ( ctx->dval == nullptr || ctx->dcb == nullptr ) {
        return 7;
    }
    if ( factotype != FACT_POTRF && factotype != FACT_GETRF && factotype != FACT_SYTRF ) {
        return 3;
    }
    if ( factotype == FACT_SYTRF && !ctx->kernels ) {
        return 7;   // LDL^T runs on the matrix-core kernels only
    }
    const bool cplx = ( factor_flttype == FT_CPLX64 || factor_flttype == FT_CPLX32 );
    if ( factor_flttype != FT_DOUBLE && factor_flttype != FT_FLOAT && !cplx ) {
        return 3;
    }
    if ( cplx && ( factotype == FACT_POTRF || !ctx->kernels ) ) {
        return 3;   // complex: symmetric LDL^T and LU on the matrix-core kernels
    }
    if ( factotype == FACT_GETRF && ( ( ctx->dtval == nullptr && ctx->mtx == 111 ) || !ctx->kernels ) ) {
        // LU of a general matrix needs its transpose for the U side, and the
        // matrix-core kernels
        return 7;
    }
    const bool lu    = ( factotype == FACT_GETRF );
    size_t     es    = elt_size( factor_flttype );
    size_t     bytes = std::max<int64_t>( ctx->coefnbr, 1 ) * es;
    size_t     dinvb = std::max<int64_t>( ctx->dinvnbr, 1 ) * es;
    if ( ctx->dL == nullptr || ctx->coef_bytes != bytes || ctx->fflt != factor_flttype ) {
        void *old[] = { ctx->dL, ctx->ddinv, ctx->dU, ctx->dudinv };
        for ( void *q : old ) {
            if ( q ) PXD_CHECK( hipFree( q ) );
        }
        ctx->dL = ctx->ddinv = ctx->dU = ctx->dudinv = nullptr;
        PXD_CHECK( hipMalloc( &ctx->dL, bytes ) );
        PXD_CHECK( hipMalloc( &ctx->ddinv, dinvb ) );
        ctx->coef_bytes = bytes;
    }
    if ( lu && ctx->dU == nullptr ) {
        PXD_CHECK( hipMalloc( &ctx->dU, bytes ) );
        PXD_CHECK( hipMalloc( &ctx->dudinv, dinvb ) );
    }
    if ( !lu && ctx->dU != nullptr ) {
        PXD_CHECK( hipFree( ctx->dU ) );
        PXD_CHECK( hipFree( ctx->dudinv ) );
        ctx->dU = ctx->dudinv = nullptr;
    }
    ctx->fflt  = factor_flttype;
    ctx->stats.coef_bytes = bytes * ( lu ? 2 : 1 );
    const int bm = ( factor_flttype == FT_DOUBLE ) ? upd_rows<double>::value
                 : ( factor_flttype == FT_FLOAT ) ? upd_rows<float>::value
                 : ( factor_flttype == FT_CPLX64 ) ? upd_rows<pxcd>::value : upd_rows<pxcf>::value;
    if ( ctx->tasks_dirty || ctx->upd_bm != bm || ctx->facto != factotype ) {
        ctx->upd_bm      = bm;
        ctx->facto       = factotype;
        ctx->tasks_dirty = 0;
        int rc = build_schedule( ctx );
        if ( rc ) {
            return rc;
        }
    }
    switch ( factor_flttype ) {
    case FT_DOUBLE: return factorize<double>( ctx, threshold, nbpivots );
    case FT_FLOAT: return factorize<float>( ctx, threshold, nbpivots );
    case FT_CPLX64: return factorize<pxcd>( ctx, threshold, nbpivots );
    default: return factorize<pxcf>( ctx, threshold, nbpivots );
    }
}

int
pxd_get_coeftab( pxd_ctx_t *ctx, int side, void *host, size_t bytes )
{
    void *src = ( side == 0 ) ? ctx->dL : ctx->dU;
    if ( src == nullptr || bytes > ctx->coef_bytes ) {
        return 7;
    }
    PXD_CHECK( hipStreamSynchronize( ctx->stream ) );
    PXD_CHECK( hipMemcpy( host, src, bytes, hipMemcpyDeviceToHost ) );
    return 0;
}

int
pxd_get_coeftab_range( pxd_ctx_t *ctx, int side, int64_t offset, int64_t count, void *host )
{
    const void  *src = ( side == 0 ) ? ctx->dL : ctx->dU;
    const size_t es  = elt_size( ctx->fflt );
    if ( src == nullptr || offset < 0 || count < 0 || (size_t)( offset + count ) * es > ctx->coef_bytes ) {
        return 7;
    }
    PXD_CHECK( hipStreamSynchronize( ctx->stream ) );
    PXD_CHECK( hipMemcpy( host, (const char *)src + (size_t)offset * es, (size_t)count * es, hipMemcpyDeviceToHost ) );
    return 0;
}

int
pxd_solve( pxd_ctx_t *ctx, int64_t nrhs, void *dx, int64_t ldx )
{
    return solve_steps( ctx, SOLVE_ALL, false, nrhs, dx, ldx );
}

int
pxd_solve_noalloc( pxd_ctx_t *ctx, int64_t nrhs, void *dx, int64_t ldx )
{
    PXD_CHECK( hipSetDevice( ctx->device ) );
    if ( ctx->dL == nullptr || nrhs < 0 || ( nrhs > 0 && ( dx == nullptr || ldx < ctx->n ) ) ) {
        return 7;
    }
    if ( nrhs == 0 || ctx->n == 0 ) {
        return 0;
    }
    if ( ctx->schurfcol < ctx->n && ctx->schur_solv_mode == 2 && ctx->schur_fact_local ) {
        return 7;   // the Schur part was not factorized
    }
    void *scr = scratch_address( ctx->device );
    if ( scr == nullptr ) {
        fprintf( stderr, "pxd_solve_noalloc: the scratch of device %d has no address\n", ctx->device );
        return 6;
    }
    const size_t need = noalloc_need( ctx, nrhs );
    if ( need > (size_t)PXD_NOALLOC_BYTES ) {
        fprintf( stderr,
                 "pxd_solve_noalloc: a solve of %lld unknowns x %lld right-hand sides needs %zu bytes of scratch, "
                 "more than its %zu (use pxd_solve)\n",
                 (long long)ctx->n, (long long)nrhs, need, (size_t)PXD_NOALLOC_BYTES );
        return 7;
    }
    int rc = 0;
    switch ( ctx->fflt ) {
    case FT_DOUBLE: rc = solve_vectors<double, double>( ctx, nrhs, dx, ldx, SOLVE_ALL, false, scr ); break;
    case FT_FLOAT: rc = solve_vectors<double, float>( ctx, nrhs, dx, ldx, SOLVE_ALL, false, scr ); break;
    case FT_CPLX64: rc = solve_vectors<pxcd, pxcd>( ctx, nrhs, dx, ldx, SOLVE_ALL, false, scr ); break;
    default: rc = solve_vectors<pxcd, pxcf>( ctx, nrhs, dx, ldx, SOLVE_ALL, false, scr ); break;
    }
    if ( hipStreamSynchronize( ctx->stream ) != hipSuccess ) {
        rc = rc ? rc : 6;
    }
    return rc;
}

int
pxd_solve_trans( pxd_ctx_t *ctx, int trans, int64_t nrhs, void *dx, int64_t ldx )
{
    // (L U)^T = U^T L^T; Cholesky and LDL^T factors are symmetric (complex
    // symmetric for LDL^T), so A^T is solved as A
    return solve_steps( ctx, SOLVE_ALL, trans != 0 && ctx->facto == FACT_GETRF, nrhs, dx, ldx );
}

int
pxd_solve_step( pxd_ctx_t *ctx, int step, int64_t nrhs, void *dx, int64_t ldx )
{
    const bool lu = ( ctx->facto == FACT_GETRF );
    switch ( step ) {
    case PXD_STEP_L_FWD: return solve_steps( ctx, SOLVE_FWD, false, nrhs, dx, ldx );
    case PXD_STEP_LT_BWD: return solve_steps( ctx, SOLVE_BWD, lu, nrhs, dx, ldx );
    case PXD_STEP_U_BWD: return lu ? solve_steps( ctx, SOLVE_BWD, false, nrhs, dx, ldx ) : 7;
    case PXD_STEP_UT_FWD: return lu ? solve_steps( ctx, SOLVE_FWD, true, nrhs, dx, ldx ) : 7;
    case PXD_STEP_DIAG: return solve_steps( ctx, SOLVE_DIAG, false, nrhs, dx, ldx );
    default: return 7;
    }
}

int
pxd_get_diag( pxd_ctx_t *ctx, void *dd )
{
    PXD_CHECK( hipSetDevice( ctx->device ) );
    if ( ctx->dL == nullptr || dd == nullptr ) {
        return 7;
    }
    const unsigned nb = nblocks( ctx->n, 256 );
    switch ( ctx->fflt ) {
    case FT_DOUBLE:
        hipLaunchKernelGGL( ( k_get_diag<double, double> ), dim3( nb ), dim3( 256 ), 0, ctx->stream, ctx->n,
                            ctx->dcol2cblk, ctx->dcb, (const double *)ctx->dL, (double *)dd );
        break;
    case FT_FLOAT:
        hipLaunchKernelGGL( ( k_get_diag<float, double> ), dim3( nb ), dim3( 256 ), 0, ctx->stream, ctx->n,
                            ctx->dcol2cblk, ctx->dcb, (const float *)ctx->dL, (double *)dd );
        break;
    case FT_CPLX64:
        hipLaunchKernelGGL( ( k_get_diag<pxcd, pxcd> ), dim3( nb ), dim3( 256 ), 0, ctx->stream, ctx->n,
                            ctx->dcol2cblk, ctx->dcb, (const pxcd *)ctx->dL, (pxcd *)dd );
        break;
    default:
        hipLaunchKernelGGL( ( k_get_diag<pxcf, pxcd> ), dim3( nb ), dim3( 256 ), 0, ctx->stream, ctx->n,
                            ctx->dcol2cblk, ctx->dcb, (const pxcf *)ctx->dL, (pxcd *)dd );
        break;
    }
    PXD_CHECK( hipGetLastError() );
    PXD_CHECK( hipStreamSynchronize( ctx->stream ) );
    return 0;
}

extern "C++" {
template <typename T>
static void
spmv_launch( int G, dim3 grid, hipStream_t st, int64_t n, const int64_t *cp, const int64_t *rt, const T *v,
             const T *x, T *y, int64_t ld, int nrhs, double alpha, double beta )
{
    switch ( G ) {
    case 4: hipLaunchKernelGGL( ( k_spmv_grp<T, T, 4> ), grid, dim3( 256 ), 0, st, n, cp, rt, v, x, y, ld, nrhs, alpha, beta ); break;
    case 8: hipLaunchKernelGGL( ( k_spmv_grp<T, T, 8> ), grid, dim3( 256 ), 0, st, n, cp, rt, v, x, y, ld, nrhs, alpha, beta ); break;
    case 16: hipLaunchKernelGGL( ( k_spmv_grp<T, T, 16> ), grid, dim3( 256 ), 0, st, n, cp, rt, v, x, y, ld, nrhs, alpha, beta ); break;
    case 32: hipLaunchKernelGGL( ( k_spmv_grp<T, T, 32> ), grid, dim3( 256 ), 0, st, n, cp, rt, v, x, y, ld, nrhs, alpha, beta ); break;
    default: hipLaunchKernelGGL( ( k_spmv_grp<T, T, 64> ), grid, dim3( 256 ), 0, st, n, cp, rt, v, x, y, ld, nrhs, alpha, beta ); break;
    }
}
}

int
pxd_spmv( pxd_ctx_t *ctx, double alpha, const void *dx, double beta, void *dy, int64_t nrhs, int64_t ld )
{
    const int64_t *cp = ctx->dcolptr, *rt = ctx->drowtab;
    const void    *vv = ctx->dval;
    if ( ctx->dtval != nullptr ) {
        // general or Hermitian: rows of A are the columns of the transpose
        cp = ctx->dtcolptr;
        rt = ctx->dtrowtab;
        vv = ctx->dtval;
    }
    // lanes per row: the power of two at or above the mean entries per row
    const double avg = ctx->n > 0 ? (double)ctx->nnz / (double)ctx->n : 1.;
    const int    G   = avg <= 4. ? 4 : avg <= 8. ? 8 : avg <= 16. ? 16 : avg <= 32. ? 32 : 64;
    const dim3   grid( (unsigned)( ( ctx->n * G + 255 ) / 256 ) );
    if ( ctx->mflt == FT_DOUBLE ) {
        spmv_launch<double>( G, grid, ctx->stream, ctx->n, cp, rt, (const double *)vv, (const double *)dx,
                             (double *)dy, ld, (int)nrhs, alpha, beta );
    }
    else {
        spmv_launch<pxcd>( G, grid, ctx->stream, ctx->n, cp, rt, (const pxcd *)vv, (const pxcd *)dx, (pxcd *)dy,
                           ld, (int)nrhs, alpha, beta );
    }
    PXD_CHECK( hipGetLastError() );
    return 0;
}

static unsigned
red_blocks( int64_t n )
{
    // each workgroup ends with one atomic on its column's result, and the
    // atomics on one address serialize: few workgroups, many values each
    return (unsigned)std::min<int64_t>( 64, std::max<int64_t>( 1, ( n + 4095 ) / 4096 ) );
}

int
pxd_vec_dot( pxd_ctx_t *ctx, int64_t n, int64_t nrhs, const void *x, const void *y, void *dres )
{
    if ( ctx->mflt == FT_CPLX64 ) {
        hipLaunchKernelGGL( ( k_zero<pxcd> ), dim3( 1 ), dim3( 64 ), 0, ctx->stream, (pxcd *)dres, nrhs );
        hipLaunchKernelGGL( ( k_dot<pxcd> ), dim3( red_blocks( n ), nrhs ), dim3( 256 ), 0, ctx->stream, n,
                            (const pxcd *)x, (const pxcd *)y, (pxcd *)dres );
    }
    else {
        hipLaunchKernelGGL( ( k_zero<double> ), dim3( 1 ), dim3( 64 ), 0, ctx->stream, (double *)dres, nrhs );
        hipLaunchKernelGGL( ( k_dot<double> ), dim3( red_blocks( n ), nrhs ), dim3( 256 ), 0, ctx->stream, n,
                            (const double *)x, (const double *)y, (double *)dres );
    }
    PXD_CHECK( hipGetLastError() );
    return 0;
}

int
pxd_vec_nrm2( pxd_ctx_t *ctx, int64_t n, int64_t nrhs, const void *x, double *dres )
{
    hipLaunchKernelGGL( ( k_zero<double> ), dim3( 1 ), dim3( 64 ), 0, ctx->stream, dres, nrhs );
    if ( ctx->mflt == FT_CPLX64 ) {
        hipLaunchKernelGGL( ( k_nrm2sq<pxcd> ), dim3( red_blocks( n ), nrhs ), dim3( 256 ), 0, ctx->stream, n,
                            (const pxcd *)x, dres );
    }
    else {
        hipLaunchKernelGGL( ( k_nrm2sq<double> ), dim3( red_blocks( n ), nrhs ), dim3( 256 ), 0, ctx->stream, n,
                            (const double *)x, dres );
    }
    hipLaunchKernelGGL( k_sqrt_inplace, dim3( 1 ), dim3( 64 ), 0, ctx->stream, dres, (int)nrhs );
    PXD_CHECK( hipGetLastError() );
    return 0;
}

int
pxd_vec_maxabs( pxd_ctx_t *ctx, int64_t n, int64_t nrhs, const void *x, double *dres )
{
    hipLaunchKernelGGL( ( k_zero<double> ), dim3( 1 ), dim3( 64 ), 0, ctx->stream, dres, nrhs );
    if ( ctx->mflt == FT_CPLX64 ) {
        hipLaunchKernelGGL( ( k_maxabs<pxcd> ), dim3( red_blocks( n ), nrhs ), dim3( 256 ), 0, ctx->stream, n,
                            (const pxcd *)x, (unsigned long long *)dres );
    }
    else {
        hipLaunchKernelGGL( ( k_maxabs<double> ), dim3( red_blocks( n ), nrhs ), dim3( 256 ), 0, ctx->stream, n,
                            (const double *)x, (unsigned long long *)dres );
    }
    PXD_CHECK( hipGetLastError() );
    return 0;
}

int
pxd_vec_axpy( pxd_ctx_t *ctx, int64_t n, int64_t nrhs, double sign, const void *dalpha, const void *x, void *y )
{
    if ( ctx->mflt == FT_CPLX64 ) {
        hipLaunchKernelGGL( ( k_axpy<pxcd> ), dim3( nblocks( n, 256 ) ), dim3( 256 ), 0, ctx->stream, n, (int)nrhs,
                            sign, (const pxcd *)dalpha, (const pxcd *)x, (pxcd *)y );
    }
    else {
        hipLaunchKernelGGL( ( k_axpy<double> ), dim3( nblocks( n, 256 ) ), dim3( 256 ), 0, ctx->stream, n, (int)nrhs,
                            sign, (const double *)dalpha, (const double *)x, (double *)y );
    }
    PXD_CHECK( hipGetLastError() );
    return 0;
}

int
pxd_vec_scal( pxd_ctx_t *ctx, int64_t n, int64_t nrhs, const void *dalpha, int inv, void *x )
{
    if ( ctx->mflt == FT_CPLX64 ) {
        hipLaunchKernelGGL( ( k_scal<pxcd> ), dim3( nblocks( n, 256 ) ), dim3( 256 ), 0, ctx->stream, n, (int)nrhs,
                            (const pxcd *)dalpha, inv, (pxcd *)x );
    }
    else {
        hipLaunchKernelGGL( ( k_scal<double> ), dim3( nblocks( n, 256 ) ), dim3( 256 ), 0, ctx->stream, n, (int)nrhs,
                            (const double *)dalpha, inv, (double *)x );
    }
    PXD_CHECK( hipGetLastError() );
    return 0;
}

int
pxd_vec_copy( pxd_ctx_t *ctx, int64_t n, int64_t nrhs, const void *x, void *y )
{
    size_t es = ( ctx->mflt == FT_CPLX64 ) ? 16 : 8;
    PXD_CHECK( hipMemcpyAsync( y, x, n * nrhs * es, hipMemcpyDeviceToDevice, ctx->stream ) );
    return 0;
}

int
pxd_vec_gemv( pxd_ctx_t *ctx, int64_t n, int64_t nrhs, int64_t m, const void *V, const void *dy, void *x )
{
    if ( ctx->mflt == FT_CPLX64 ) {
        hipLaunchKernelGGL( ( k_gemv<pxcd> ), dim3( nblocks( n, 256 ) ), dim3( 256 ), 0, ctx->stream, n, (int)nrhs,
                            (int)m, (const pxcd *)V, (const pxcd *)dy, (pxcd *)x );
    }
    else {
        hipLaunchKernelGGL( ( k_gemv<double> ), dim3( nblocks( n, 256 ) ), dim3( 256 ), 0, ctx->stream, n, (int)nrhs,
                            (int)m, (const double *)V, (const double *)dy, (double *)x );
    }
    PXD_CHECK( hipGetLastError() );
    return 0;
}

}  // extern "C"
