/*
 * options.c -- pastixGetOptions and the iparm / dparm / enum name tables it
 * parses (reference common/get_options.c:149-330, common/parse_options.c):
 * the command-line options of the reference's examples and tests, so that a
 * program built on them reads its matrix and parameters the same way.
 */
#include <getopt.h>
#include <strings.h>
#include "px_internal.h"

typedef struct { const char *name; long value; } px_name_t;
#define PX_E( x ) { #x, (long)( x ) }

static const px_name_t px_iparm_names[] = {
    PX_E( IPARM_VERBOSE ), PX_E( IPARM_IO_STRATEGY ), PX_E( IPARM_NNZEROS ),
    PX_E( IPARM_NNZEROS_BLOCK_LOCAL ), PX_E( IPARM_ALLOCATED_TERMS ), PX_E( IPARM_PRODUCE_STATS ),
    PX_E( IPARM_TRACE ), PX_E( IPARM_MC64 ), PX_E( IPARM_ORDERING ),
    PX_E( IPARM_ORDERING_DEFAULT ), PX_E( IPARM_SCOTCH_MT ), PX_E( IPARM_SCOTCH_SWITCH_LEVEL ),
    PX_E( IPARM_SCOTCH_CMIN ), PX_E( IPARM_SCOTCH_CMAX ), PX_E( IPARM_SCOTCH_FRAT ),
    PX_E( IPARM_METIS_CTYPE ), PX_E( IPARM_METIS_RTYPE ), PX_E( IPARM_METIS_NO2HOP ),
    PX_E( IPARM_METIS_NSEPS ), PX_E( IPARM_METIS_NITER ), PX_E( IPARM_METIS_UFACTOR ),
    PX_E( IPARM_METIS_COMPRESS ), PX_E( IPARM_METIS_CCORDER ), PX_E( IPARM_METIS_PFACTOR ),
    PX_E( IPARM_METIS_SEED ), PX_E( IPARM_METIS_DBGLVL ), PX_E( IPARM_AMALGAMATION_LVLBLAS ),
    PX_E( IPARM_AMALGAMATION_LVLCBLK ), PX_E( IPARM_REORDERING_SPLIT ), PX_E( IPARM_REORDERING_STOP ),
    PX_E( IPARM_SPLITTING_STRATEGY ), PX_E( IPARM_SPLITTING_LEVELS_PROJECTIONS ), PX_E( IPARM_SPLITTING_LEVELS_KWAY ),
    PX_E( IPARM_SPLITTING_PROJECTIONS_DEPTH ), PX_E( IPARM_SPLITTING_PROJECTIONS_DISTANCE ), PX_E( IPARM_SPLITTING_PROJECTIONS_WIDTH ),
    PX_E( IPARM_MIN_BLOCKSIZE ), PX_E( IPARM_MAX_BLOCKSIZE ), PX_E( IPARM_TASKS2D_LEVEL ),
    PX_E( IPARM_TASKS2D_WIDTH ), PX_E( IPARM_ALLCAND ), PX_E( IPARM_INCOMPLETE ),
    PX_E( IPARM_LEVEL_OF_FILL ), PX_E( IPARM_FACTORIZATION ), PX_E( IPARM_FACTO_LOOK_SIDE ),
    PX_E( IPARM_STATIC_PIVOTING ), PX_E( IPARM_FREE_CSCUSER ), PX_E( IPARM_SCHUR_FACT_MODE ),
    PX_E( IPARM_TRANSPOSE_SOLVE ), PX_E( IPARM_SCHUR_SOLV_MODE ), PX_E( IPARM_APPLYPERM_WS ),
    PX_E( IPARM_REFINEMENT ), PX_E( IPARM_NBITER ), PX_E( IPARM_ITERMAX ),
    PX_E( IPARM_GMRES_IM ), PX_E( IPARM_SCHEDULER ), PX_E( IPARM_THREAD_NBR ),
    PX_E( IPARM_SOCKET_NBR ), PX_E( IPARM_AUTOSPLIT_COMM ), PX_E( IPARM_GPU_NBR ),
    PX_E( IPARM_GPU_MEMORY_PERCENTAGE ), PX_E( IPARM_GPU_MEMORY_BLOCK_SIZE ), PX_E( IPARM_GLOBAL_ALLOCATION ),
    PX_E( IPARM_COMPRESS_MIN_WIDTH ), PX_E( IPARM_COMPRESS_MIN_HEIGHT ), PX_E( IPARM_COMPRESS_WHEN ),
    PX_E( IPARM_COMPRESS_METHOD ), PX_E( IPARM_COMPRESS_ORTHO ), PX_E( IPARM_COMPRESS_RELTOL ),
    PX_E( IPARM_COMPRESS_PRESELECT ), PX_E( IPARM_COMPRESS_ILUK ), PX_E( IPARM_MIXED ),
    PX_E( IPARM_FTZ ), PX_E( IPARM_MPI_THREAD_LEVEL ), PX_E( IPARM_MODIFY_PARAMETER ),
    PX_E( IPARM_START_TASK ), PX_E( IPARM_END_TASK ), PX_E( IPARM_FLOAT ),
    PX_E( IPARM_MTX_TYPE ), PX_E( IPARM_DOF_NBR ),
};

static const px_name_t px_dparm_names[] = {
    PX_E( DPARM_FILL_IN ), PX_E( DPARM_EPSILON_REFINEMENT ), PX_E( DPARM_RELATIVE_ERROR ),
    PX_E( DPARM_EPSILON_MAGN_CTRL ), PX_E( DPARM_ORDER_TIME ), PX_E( DPARM_SYMBFACT_TIME ),
    PX_E( DPARM_REORDER_TIME ), PX_E( DPARM_BLEND_TIME ), PX_E( DPARM_ANALYZE_TIME ),
    PX_E( DPARM_PRED_FACT_TIME ), PX_E( DPARM_FACT_TIME ), PX_E( DPARM_FACT_FLOPS ),
    PX_E( DPARM_FACT_THFLOPS ), PX_E( DPARM_FACT_RLFLOPS ), PX_E( DPARM_FACT_ENERGY ),
    PX_E( DPARM_MEM_FR ), PX_E( DPARM_MEM_LR ), PX_E( DPARM_SOLV_TIME ),
    PX_E( DPARM_SOLV_FLOPS ), PX_E( DPARM_SOLV_THFLOPS ), PX_E( DPARM_SOLV_RLFLOPS ),
    PX_E( DPARM_SOLV_ENERGY ), PX_E( DPARM_REFINE_TIME ), PX_E( DPARM_A_NORM ),
    PX_E( DPARM_COMPRESS_TOLERANCE ), PX_E( DPARM_COMPRESS_MIN_RATIO ),
};

/* enum values that an -i option may name (also the spm arithmetic aliases) */
static const px_name_t px_enum_names[] = {
    PX_E( PastixColMajor ), PX_E( PastixCompressMethodPQRCP ), PX_E( PastixCompressMethodRQRCP ),
    PX_E( PastixCompressMethodRQRRT ), PX_E( PastixCompressMethodSVD ), PX_E( PastixCompressMethodTQRCP ),
    PX_E( PastixCompressNever ), PX_E( PastixCompressOrthoCGS ), PX_E( PastixCompressOrthoPartialQR ),
    PX_E( PastixCompressOrthoQR ), PX_E( PastixCompressWhenBegin ), PX_E( PastixCompressWhenDuring ),
    PX_E( PastixCompressWhenEnd ), PX_E( PastixConjTrans ), PX_E( PastixDirBackward ),
    PX_E( PastixDirForward ), PX_E( PastixFactGETRF ), PX_E( PastixFactHETRF ),
    PX_E( PastixFactLDLH ), PX_E( PastixFactLDLT ), PX_E( PastixFactLLH ),
    PX_E( PastixFactLLT ), PX_E( PastixFactLU ), PX_E( PastixFactLeftLooking ),
    PX_E( PastixFactModeBoth ), PX_E( PastixFactModeLocal ), PX_E( PastixFactModeSchur ),
    PX_E( PastixFactPOTRF ), PX_E( PastixFactPXTRF ), PX_E( PastixFactRightLooking ),
    PX_E( PastixFactSYTRF ), PX_E( PastixFrobeniusNorm ), PX_E( PastixIOLoad ),
    PX_E( PastixIOLoadCSC ), PX_E( PastixIOLoadGraph ), PX_E( PastixIONo ),
    PX_E( PastixIOSave ), PX_E( PastixIOSaveCSC ), PX_E( PastixIOSaveGraph ),
    PX_E( PastixInfNorm ), PX_E( PastixLCoef ), PX_E( PastixLUCoef ),
    PX_E( PastixLeft ), PX_E( PastixLower ), PX_E( PastixMaxNorm ),
    PX_E( PastixMpiNone ), PX_E( PastixMpiThreadFunneled ), PX_E( PastixMpiThreadMultiple ),
    PX_E( PastixMpiThreadSerialized ), PX_E( PastixMpiThreadSingle ), PX_E( PastixNoTrans ),
    PX_E( PastixNonUnit ), PX_E( PastixOneNorm ), PX_E( PastixOrderMetis ),
    PX_E( PastixOrderParMetis ), PX_E( PastixOrderPersonal ), PX_E( PastixOrderPtScotch ),
    PX_E( PastixOrderScotch ), PX_E( PastixRefineBiCGSTAB ), PX_E( PastixRefineCG ),
    PX_E( PastixRefineGMRES ), PX_E( PastixRefineSR ), PX_E( PastixRight ),
    PX_E( PastixRowMajor ), PX_E( PastixSchedDynamic ), PX_E( PastixSchedParsec ),
    PX_E( PastixSchedSequential ), PX_E( PastixSchedStarPU ), PX_E( PastixSchedStatic ),
    PX_E( PastixSolvModeInterface ), PX_E( PastixSolvModeLocal ), PX_E( PastixSolvModeSchur ),
    PX_E( PastixSplitKway ), PX_E( PastixSplitKwayProjections ), PX_E( PastixSplitNot ),
    PX_E( PastixTaskAnalyze ), PX_E( PastixTaskClean ), PX_E( PastixTaskInit ),
    PX_E( PastixTaskNumfact ), PX_E( PastixTaskOrdering ), PX_E( PastixTaskRefine ),
    PX_E( PastixTaskSolve ), PX_E( PastixTaskSymbfact ), PX_E( PastixTraceNumfact ),
    PX_E( PastixTraceSolve ), PX_E( PastixTrans ), PX_E( PastixUCoef ),
    PX_E( PastixUnit ), PX_E( PastixUpper ), PX_E( PastixUpperLower ),
    PX_E( PastixVerboseNo ), PX_E( PastixVerboseNot ), PX_E( PastixVerboseYes ),
    { "PastixPattern", SpmPattern }, { "PastixFloat", SpmFloat }, { "PastixDouble", SpmDouble },
    { "PastixComplex32", SpmComplex32 }, { "PastixComplex64", SpmComplex64 },
    { "PastixGeneral", SpmGeneral }, { "PastixSymmetric", SpmSymmetric }, { "PastixHermitian", SpmHermitian },
};

static long
px_lookup( const px_name_t *t, size_t n, const char *s )
{
    for ( size_t i = 0; i < n; i++ ) {
        if ( strcasecmp( t[i].name, s ) == 0 ) {
            return t[i].value;
        }
    }
    return -1;
}

/* index of an iparm given by name ("iparm_verbose", any case), -1 if none */
int
pastix_amd_parse_iparm( const char *s )
{
    return (int)px_lookup( px_iparm_names, sizeof( px_iparm_names ) / sizeof( px_iparm_names[0] ), s );
}

int
pastix_amd_parse_dparm( const char *s )
{
    return (int)px_lookup( px_dparm_names, sizeof( px_dparm_names ) / sizeof( px_dparm_names[0] ), s );
}

/* value of an enum given by name ("pastixrefinecg") or as an integer; -1 if neither */
int
pastix_amd_parse_enum( const char *s )
{
    long v = px_lookup( px_enum_names, sizeof( px_enum_names ) / sizeof( px_enum_names[0] ), s );
    int  iv;
    if ( v != -1 ) {
        return (int)v;
    }
    if ( sscanf( s, "%d", &iv ) != 1 ) {
        return -1;
    }
    return iv;
}

static void
px_usage( void )
{
    fprintf( stderr,
             "Matrix input (mandatory):\n"
             " -0 --rsa          : RSA/Matrix Market Fortran driver (read as Harwell-Boeing)\n"
             " -1 --hb           : Harwell Boeing C driver\n"
             " -2 --ijv          : IJV coordinate C driver\n"
             " -3 --mm           : Matrix Market C driver\n"
             " -4 --spm          : SPM Matrix driver\n"
             " -9 --lap          : Generate a Laplacian (7-points stencil)\n"
             " -x --xlap         : Generate an extended Laplacian\n"
             " -G --graph        : SCOTCH Graph file\n"
             "\n"
             "Architecture arguments:\n"
             " -a --scatter      : Scatter the spm (default: 0)\n"
             " -t --threads      : Number of threads per node (default: -1)\n"
             " -g --gpus         : Number of gpus per node (default: 0)\n"
             " -s --sched        : Set the default scheduler (default: 4)\n"
             "                     0: Sequential, 1: Static, 2: PaRSEC, 3: StarPU, 4: Dynamic\n"
             "\n"
             "Optional arguments:\n"
             " -f --fact                     : Factorization (default: LU)\n"
             "                                 0: Cholesky, 1: LDL^[th], 2: LU, 3:LL^t, 4:LDL^t\n"
             " -c --check                    : Level of check (default: 1)\n"
             " -o --ord                      : Ordering library (scotch, ptscotch, metis, parmetis, personal)\n"
             " -i --iparm <IPARM_ID> <value> : set any given integer parameter\n"
             " -d --dparm <DPARM_ID> <value> : set any given floating parameter\n"
             "\n"
             " -v --verbose[=lvl] : extra verbose output\n"
             " -h --help          : this message\n"
             "\n" );
}

#define PX_GETOPT_STRING "0:1:2:3:4:9:x:G:a:t:g:s:o:f:c:i:d:v::h"

static struct option px_long_options[] = {
    { "rsa", required_argument, 0, '0' },     { "hb", required_argument, 0, '1' },
    { "ijv", required_argument, 0, '2' },     { "mm", required_argument, 0, '3' },
    { "spm", required_argument, 0, '4' },     { "lap", required_argument, 0, '9' },
    { "xlap", required_argument, 0, 'x' },    { "graph", required_argument, 0, 'G' },
    { "scatter", required_argument, 0, 'a' }, { "threads", required_argument, 0, 't' },
    { "gpus", required_argument, 0, 'g' },    { "sched", required_argument, 0, 's' },
    { "ord", required_argument, 0, 'o' },     { "fact", required_argument, 0, 'f' },
    { "check", required_argument, 0, 'c' },   { "iparm", required_argument, 0, 'i' },
    { "dparm", required_argument, 0, 'd' },   { "verbose", optional_argument, 0, 'v' },
    { "help", no_argument, 0, 'h' },          { 0, 0, 0, 0 } };

/*
 * Options of the reference's examples (common/get_options.c:149-330): the
 * matrix driver and file, iparm / dparm settings by name, factorization,
 * scheduler, threads, GPUs, check level.  Without arguments, or with -h or an
 * unknown option, it prints the usage and exits as the reference does.
 */
void
pastixGetOptions( int argc, char **argv, pastix_int_t *iparm, double *dparm, int *check, int *scatter,
                  spm_driver_t *driver, char **filename )
{
    int c;
    if ( argc == 1 ) {
        px_usage();
        exit( 0 );
    }
    *driver = (spm_driver_t)-1;
    optind  = 1;
    for ( ;; ) {
        c = getopt_long( argc, argv, PX_GETOPT_STRING, px_long_options, NULL );
        if ( c == -1 ) {
            break;
        }
        switch ( c ) {
        case '0':
            fprintf( stderr, "RSA driver is no longer supported and is replaced by the HB driver\n" );
            /* fall through */
        case '1': *driver = SpmDriverHB; free( *filename ); *filename = strdup( optarg ); break;
        case '2': *driver = SpmDriverIJV; free( *filename ); *filename = strdup( optarg ); break;
        case '3': *driver = SpmDriverMM; free( *filename ); *filename = strdup( optarg ); break;
        case '4': *driver = SpmDriverSPM; free( *filename ); *filename = strdup( optarg ); break;
        case '9': *driver = SpmDriverLaplacian; free( *filename ); *filename = strdup( optarg ); break;
        case 'x': *driver = SpmDriverXLaplacian; free( *filename ); *filename = strdup( optarg ); break;
        case 'G': *driver = SpmDriverGraph; free( *filename ); *filename = strdup( optarg ); break;
        case 'a': {
            int v = atoi( optarg );
            if ( v < 0 || v >= 6 ) {
                fprintf( stderr, "\nInvalid value for scatter option: %s\n\n", optarg );
                goto unknown;
            }
            if ( scatter ) *scatter = v;
        } break;
        case 't': iparm[IPARM_THREAD_NBR] = atoi( optarg ); break;
        case 'g': iparm[IPARM_GPU_NBR] = atoi( optarg ); break;
        case 'o':
            if ( strncasecmp( optarg, "scotch", 6 ) == 0 ) iparm[IPARM_ORDERING] = PastixOrderScotch;
            else if ( strncasecmp( optarg, "metis", 5 ) == 0 ) iparm[IPARM_ORDERING] = PastixOrderMetis;
            else if ( strncasecmp( optarg, "ptscotch", 8 ) == 0 ) iparm[IPARM_ORDERING] = PastixOrderPtScotch;
            else if ( strncasecmp( optarg, "parmetis", 8 ) == 0 ) iparm[IPARM_ORDERING] = PastixOrderParMetis;
            else if ( strncasecmp( optarg, "personal", 8 ) == 0 ) iparm[IPARM_ORDERING] = PastixOrderPersonal;
            else {
                fprintf( stderr, "\nInvalid value for ordering option: %s\n\n", optarg );
                goto unknown;
            }
            break;
        case 'f': {
            int v = atoi( optarg );
            if ( v < 0 || v > 4 ) {
                fprintf( stderr, "\nInvalid value for factorization option: %s\n\n", optarg );
                goto unknown;
            }
            iparm[IPARM_FACTORIZATION] = v;
        } break;
        case 'c': {
            int v = atoi( optarg );
            if ( v < 0 || v >= 6 ) {
                fprintf( stderr, "\nInvalid value for check option: %s\n\n", optarg );
                goto unknown;
            }
            if ( check ) *check = v;
        } break;
        case 's': {
            int v = atoi( optarg );
            if ( v < 0 || v > 4 ) {
                fprintf( stderr, "\nInvalid value for scheduler option: %s\n\n", optarg );
                goto unknown;
            }
            iparm[IPARM_SCHEDULER] = v;
        } break;
        case 'i': {
            int idx = pastix_amd_parse_iparm( optarg );
            int val;
            if ( idx < 0 ) {
                fprintf( stderr, "\n%s is not a correct iparm parameter\n\n", optarg );
                goto unknown;
            }
            if ( optind >= argc || ( val = pastix_amd_parse_enum( argv[optind] ) ) == -1 ) {
                fprintf( stderr, "\n%s is not a correct value for the iparm parameters\n\n",
                         optind < argc ? argv[optind] : "(none)" );
                goto unknown;
            }
            iparm[idx] = val;
            optind++;
        } break;
        case 'd': {
            int idx = pastix_amd_parse_dparm( optarg );
            if ( idx < 0 ) {
                fprintf( stderr, "\n%s is not a correct dparm parameter\n\n", optarg );
                goto unknown;
            }
            if ( optind >= argc ) {
                goto unknown;
            }
            dparm[idx] = atof( argv[optind] );
            optind++;
        } break;
        case 'v':
            iparm[IPARM_VERBOSE] = optarg ? atoi( optarg ) : PastixVerboseYes;
            break;
        case 'h':
            px_usage();
            exit( 0 );
        default:
            goto unknown;
        }
    }
    return;

unknown:
    px_usage();
    exit( 0 );
}
