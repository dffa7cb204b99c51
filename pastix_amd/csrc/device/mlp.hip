// mlp.hip -- learned kernel cost model: a multilayer perceptron
// (PXD_MLP_F inputs, two tanh layers of PXD_MLP_H units, one linear output)
// trained on the device by full-batch Adam on a mean squared error.
//
// The reference has no learned model: its kernel costs are the polynomials
// of common/models.h:38-83.  This model is fitted to the same timing log
// (kernels/kernels_trace.c:222-249 records) and compared with them.
//
// Parameters, one float array: W1[F][H], b1[H], W2[H][H], b2[H], w3[H], b3.
// One workgroup (one wavefront) takes 64 rows; the forward pass, the
// backward pass and the workgroup's share of the parameter gradients are one
// kernel, whose per-workgroup partial sums a second kernel adds up in a fixed
// order (the gradient of a given data set does not depend on scheduling).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <vector>

#include "pastix_amd_device.h"
#include "px_device.hpp"

#define PXD_MLP_F    10
#define PXD_MLP_H    32
#define PXD_MLP_ROWS 64

namespace {

constexpr int F = PXD_MLP_F, H = PXD_MLP_H, RB = PXD_MLP_ROWS;
constexpr int OFF_W1 = 0, OFF_B1 = F * H, OFF_W2 = OFF_B1 + H, OFF_B2 = OFF_W2 + H * H;
constexpr int OFF_W3 = OFF_B2 + H, OFF_B3 = OFF_W3 + H, NPARAM = OFF_B3 + 1;

// Forward pass of one row: h1, h2 and the output.
__device__ inline float
mlp_row_forward( const float *P, const float *x, float *h1, float *h2 )
{
    for ( int j = 0; j < H; j++ ) {
        float z = P[OFF_B1 + j];
        for ( int i = 0; i < F; i++ ) z += x[i] * P[OFF_W1 + i * H + j];
        h1[j] = tanhf( z );
    }
    for ( int j = 0; j < H; j++ ) {
        float z = P[OFF_B2 + j];
        for ( int i = 0; i < H; i++ ) z += h1[i] * P[OFF_W2 + i * H + j];
        h2[j] = tanhf( z );
    }
    float y = P[OFF_B3];
    for ( int j = 0; j < H; j++ ) y += h2[j] * P[OFF_W3 + j];
    return y;
}

__global__ void __launch_bounds__( RB )
k_mlp_forward( const float *__restrict__ X, int64_t n, const float *__restrict__ params, float *__restrict__ y )
{
    __shared__ float sP[NPARAM];
    for ( int p = threadIdx.x; p < NPARAM; p += RB ) sP[p] = params[p];
    __syncthreads();
    const int64_t r = (int64_t)blockIdx.x * RB + threadIdx.x;
    if ( r >= n ) {
        return;
    }
    float x[F], h1[H], h2[H];
    for ( int i = 0; i < F; i++ ) x[i] = X[r * F + i];
    y[r] = mlp_row_forward( sP, x, h1, h2 );
}

// Loss sum_r (y_r - t_r)^2 / n and its gradient: the workgroup's partial sums
// go to part[blockIdx.x * (NPARAM + 1) + p] (p = NPARAM: the loss).
__global__ void __launch_bounds__( RB )
k_mlp_grad( const float *__restrict__ X, const float *__restrict__ t, int64_t n,
            const float *__restrict__ params, float *__restrict__ part )
{
    __shared__ float sP[NPARAM];
    __shared__ float sX[RB][F + 1];
    __shared__ float sH1[RB][H + 1], sG1[RB][H + 1];
    __shared__ float sH2[RB][H + 1], sG2[RB][H + 1];
    __shared__ float sDy[RB], sL[RB];
    const int tid = threadIdx.x;
    for ( int p = tid; p < NPARAM; p += RB ) sP[p] = params[p];
    __syncthreads();

    const int64_t r     = (int64_t)blockIdx.x * RB + tid;
    const bool    valid = r < n;
    float x[F], h1[H], h2[H];
    for ( int i = 0; i < F; i++ ) x[i] = valid ? X[r * F + i] : 0.f;
    const float y  = mlp_row_forward( sP, x, h1, h2 );
    const float e  = valid ? y - t[r] : 0.f;
    const float dy = 2.f * e / (float)n;
    // backward: g2 = dy w3 (1 - h2^2), g1 = (W2 g2) (1 - h1^2)
    float g2[H];
    for ( int j = 0; j < H; j++ ) g2[j] = dy * sP[OFF_W3 + j] * ( 1.f - h2[j] * h2[j] );
    for ( int i = 0; i < H; i++ ) {
        float s = 0.f;
        for ( int j = 0; j < H; j++ ) s += sP[OFF_W2 + i * H + j] * g2[j];
        sG1[tid][i] = s * ( 1.f - h1[i] * h1[i] );
        sH1[tid][i] = h1[i];
        sH2[tid][i] = h2[i];
        sG2[tid][i] = g2[i];
    }
    for ( int i = 0; i < F; i++ ) sX[tid][i] = x[i];
    sDy[tid] = dy;
    sL[tid]  = e * e / (float)n;
    __syncthreads();

    // parameter gradients: sums over the workgroup's rows
    float *out = part + (int64_t)blockIdx.x * ( NPARAM + 1 );
    for ( int p = tid; p <= NPARAM; p += RB ) {
        float s = 0.f;
        if ( p < OFF_B1 ) {
            const int i = p / H, j = p % H;
            for ( int q = 0; q < RB; q++ ) s += sX[q][i] * sG1[q][j];
        }
        else if ( p < OFF_W2 ) {
            const int j = p - OFF_B1;
            for ( int q = 0; q < RB; q++ ) s += sG1[q][j];
        }
        else if ( p < OFF_B2 ) {
            const int i = ( p - OFF_W2 ) / H, j = ( p - OFF_W2 ) % H;
            for ( int q = 0; q < RB; q++ ) s += sH1[q][i] * sG2[q][j];
        }
        else if ( p < OFF_W3 ) {
            const int j = p - OFF_B2;
            for ( int q = 0; q < RB; q++ ) s += sG2[q][j];
        }
        else if ( p < OFF_B3 ) {
            const int j = p - OFF_W3;
            for ( int q = 0; q < RB; q++ ) s += sH2[q][j] * sDy[q];
        }
        else if ( p == OFF_B3 ) {
            for ( int q = 0; q < RB; q++ ) s += sDy[q];
        }
        else {
            for ( int q = 0; q < RB; q++ ) s += sL[q];
        }
        out[p] = s;
    }
}

// grad[p] = sum over workgroups of part[b][p], b in increasing order
__global__ void
k_mlp_reduce( const float *__restrict__ part, int64_t nblocks, float *__restrict__ grad )
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if ( p > NPARAM ) {
        return;
    }
    double s = 0.;
    for ( int64_t b = 0; b < nblocks; b++ ) s += part[b * ( NPARAM + 1 ) + p];
    grad[p] = (float)s;
}

// Adam (Kingma and Ba, 2015) with bias correction; step counts from 1.
__global__ void
k_mlp_adam( float *__restrict__ params, const float *__restrict__ grad, float *__restrict__ m,
            float *__restrict__ v, float lr, float beta1, float beta2, float eps, int step )
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if ( p >= NPARAM ) {
        return;
    }
    const float g  = grad[p];
    const float mp = beta1 * m[p] + ( 1.f - beta1 ) * g;
    const float vp = beta2 * v[p] + ( 1.f - beta2 ) * g * g;
    m[p] = mp;
    v[p] = vp;
    const float mh = mp / ( 1.f - powf( beta1, (float)step ) );
    const float vh = vp / ( 1.f - powf( beta2, (float)step ) );
    params[p] -= lr * mh / ( sqrtf( vh ) + eps );
}

}  // namespace

struct pxd_mlp_s {
    int         device = 0;
    hipStream_t stream = nullptr;
    float      *dparams = nullptr, *dm = nullptr, *dv = nullptr, *dgrad = nullptr;
    float      *dX = nullptr, *dt = nullptr, *dpart = nullptr;
    int64_t     n = 0;
    int         step = 0;
};

namespace {

int
mlp_free_data( pxd_mlp_t *m )
{
    PXD_CHECK( hipFree( m->dX ) );
    PXD_CHECK( hipFree( m->dt ) );
    PXD_CHECK( hipFree( m->dpart ) );
    m->dX = m->dt = m->dpart = nullptr;
    m->n = 0;
    return 0;
}

int64_t
mlp_blocks( int64_t n )
{
    return ( n + RB - 1 ) / RB;
}

// loss and gradient of the uploaded data into m->dgrad (NPARAM + 1 floats)
int
mlp_grad( pxd_mlp_t *m )
{
    const int64_t nb = mlp_blocks( m->n );
    hipLaunchKernelGGL( k_mlp_grad, dim3( (unsigned)nb ), dim3( RB ), 0, m->stream, m->dX, m->dt, m->n,
                        m->dparams, m->dpart );
    hipLaunchKernelGGL( k_mlp_reduce, dim3( ( NPARAM + 1 + 255 ) / 256 ), dim3( 256 ), 0, m->stream, m->dpart,
                        nb, m->dgrad );
    PXD_CHECK( hipGetLastError() );
    return 0;
}

}  // namespace

extern "C" {

int
pxd_mlp_nparams( void )
{
    return NPARAM;
}

int
pxd_mlp_create( pxd_mlp_t **pm, int device, uint64_t seed )
{
    int ndev = 0;
    if ( hipGetDeviceCount( &ndev ) != hipSuccess || ndev <= 0 ) {
        fprintf( stderr, "pastix_amd device: no GPU visible; the learned model trains on the GPU\n" );
        return 3;
    }
    if ( device < 0 || device >= ndev ) {
        return 7;
    }
    pxd_mlp_t *m = new pxd_mlp_t();
    *pm          = m;
    m->device    = device;
    PXD_CHECK( hipSetDevice( device ) );
    PXD_CHECK( hipStreamCreateWithFlags( &m->stream, hipStreamNonBlocking ) );
    PXD_CHECK( hipMalloc( (void **)&m->dparams, NPARAM * sizeof( float ) ) );
    PXD_CHECK( hipMalloc( (void **)&m->dm, NPARAM * sizeof( float ) ) );
    PXD_CHECK( hipMalloc( (void **)&m->dv, NPARAM * sizeof( float ) ) );
    PXD_CHECK( hipMalloc( (void **)&m->dgrad, ( NPARAM + 1 ) * sizeof( float ) ) );
    // Glorot-uniform weights from a splitmix64 stream, zero biases
    std::vector<float> p( NPARAM, 0.f );
    uint64_t           s    = seed;
    auto               unif = [&]() {
        uint64_t z = ( s += 0x9E3779B97F4A7C15ull );
        z          = ( z ^ ( z >> 30 ) ) * 0xBF58476D1CE4E5B9ull;
        z          = ( z ^ ( z >> 27 ) ) * 0x94D049BB133111EBull;
        z ^= z >> 31;
        return (double)( z >> 11 ) * ( 1.0 / 9007199254740992.0 ) * 2. - 1.;
    };
    auto fill = [&]( int off, int fan_in, int fan_out, int count ) {
        const double a = std::sqrt( 6. / ( fan_in + fan_out ) );
        for ( int i = 0; i < count; i++ ) p[off + i] = (float)( a * unif() );
    };
    fill( OFF_W1, F, H, F * H );
    fill( OFF_W2, H, H, H * H );
    fill( OFF_W3, H, 1, H );
    PXD_CHECK( hipMemcpy( m->dparams, p.data(), NPARAM * sizeof( float ), hipMemcpyHostToDevice ) );
    PXD_CHECK( hipMemset( m->dm, 0, NPARAM * sizeof( float ) ) );
    PXD_CHECK( hipMemset( m->dv, 0, NPARAM * sizeof( float ) ) );
    return 0;
}

void
pxd_mlp_destroy( pxd_mlp_t *m )
{
    if ( m == nullptr ) {
        return;
    }
    (void)hipSetDevice( m->device );
    (void)hipStreamSynchronize( m->stream );
    (void)mlp_free_data( m );
    (void)hipFree( m->dparams );
    (void)hipFree( m->dm );
    (void)hipFree( m->dv );
    (void)hipFree( m->dgrad );
    (void)hipStreamDestroy( m->stream );
    delete m;
}

int
pxd_mlp_get_params( pxd_mlp_t *m, float *params )
{
    PXD_CHECK( hipSetDevice( m->device ) );
    PXD_CHECK( hipStreamSynchronize( m->stream ) );
    PXD_CHECK( hipMemcpy( params, m->dparams, NPARAM * sizeof( float ), hipMemcpyDeviceToHost ) );
    return 0;
}

int
pxd_mlp_set_params( pxd_mlp_t *m, const float *params )
{
    PXD_CHECK( hipSetDevice( m->device ) );
    PXD_CHECK( hipStreamSynchronize( m->stream ) );
    PXD_CHECK( hipMemcpy( m->dparams, params, NPARAM * sizeof( float ), hipMemcpyHostToDevice ) );
    PXD_CHECK( hipMemset( m->dm, 0, NPARAM * sizeof( float ) ) );
    PXD_CHECK( hipMemset( m->dv, 0, NPARAM * sizeof( float ) ) );
    m->step = 0;
    return 0;
}

int
pxd_mlp_set_data( pxd_mlp_t *m, const float *X, const float *t, int64_t n )
{
    if ( n <= 0 || X == nullptr || t == nullptr ) {
        return 7;
    }
    PXD_CHECK( hipSetDevice( m->device ) );
    PXD_CHECK( hipStreamSynchronize( m->stream ) );
    if ( mlp_free_data( m ) ) {
        return 6;
    }
    PXD_CHECK( hipMalloc( (void **)&m->dX, n * F * sizeof( float ) ) );
    PXD_CHECK( hipMalloc( (void **)&m->dt, n * sizeof( float ) ) );
    PXD_CHECK( hipMalloc( (void **)&m->dpart, mlp_blocks( n ) * ( NPARAM + 1 ) * sizeof( float ) ) );
    PXD_CHECK( hipMemcpy( m->dX, X, n * F * sizeof( float ), hipMemcpyHostToDevice ) );
    PXD_CHECK( hipMemcpy( m->dt, t, n * sizeof( float ), hipMemcpyHostToDevice ) );
    m->n = n;
    return 0;
}

int
pxd_mlp_loss_grad( pxd_mlp_t *m, double *loss, float *grad )
{
    if ( m->n <= 0 ) {
        return 7;
    }
    PXD_CHECK( hipSetDevice( m->device ) );
    if ( mlp_grad( m ) ) {
        return 6;
    }
    std::vector<float> g( NPARAM + 1 );
    PXD_CHECK( hipMemcpyAsync( g.data(), m->dgrad, ( NPARAM + 1 ) * sizeof( float ), hipMemcpyDeviceToHost,
                               m->stream ) );
    PXD_CHECK( hipStreamSynchronize( m->stream ) );
    if ( loss ) {
        *loss = g[NPARAM];
    }
    if ( grad ) {
        std::copy( g.begin(), g.begin() + NPARAM, grad );
    }
    return 0;
}

int
pxd_mlp_train( pxd_mlp_t *m, int steps, double lr, double *loss )
{
    if ( m->n <= 0 || steps < 0 ) {
        return 7;
    }
    PXD_CHECK( hipSetDevice( m->device ) );
    for ( int s = 0; s < steps; s++ ) {
        if ( mlp_grad( m ) ) {
            return 6;
        }
        m->step++;
        hipLaunchKernelGGL( k_mlp_adam, dim3( ( NPARAM + 255 ) / 256 ), dim3( 256 ), 0, m->stream, m->dparams,
                            m->dgrad, m->dm, m->dv, (float)lr, 0.9f, 0.999f, 1e-8f, m->step );
    }
    PXD_CHECK( hipGetLastError() );
    if ( loss ) {
        // loss of the final parameters
        return pxd_mlp_loss_grad( m, loss, nullptr );
    }
    PXD_CHECK( hipStreamSynchronize( m->stream ) );
    return 0;
}

int
pxd_mlp_predict( pxd_mlp_t *m, const float *X, int64_t n, float *y )
{
    if ( n <= 0 ) {
        return 0;
    }
    PXD_CHECK( hipSetDevice( m->device ) );
    float *dX = nullptr, *dy = nullptr;
    PXD_CHECK( hipMalloc( (void **)&dX, n * F * sizeof( float ) ) );
    PXD_CHECK( hipMalloc( (void **)&dy, n * sizeof( float ) ) );
    PXD_CHECK( hipMemcpyAsync( dX, X, n * F * sizeof( float ), hipMemcpyHostToDevice, m->stream ) );
    hipLaunchKernelGGL( k_mlp_forward, dim3( (unsigned)mlp_blocks( n ) ), dim3( RB ), 0, m->stream, dX, n,
                        m->dparams, dy );
    PXD_CHECK( hipGetLastError() );
    PXD_CHECK( hipMemcpyAsync( y, dy, n * sizeof( float ), hipMemcpyDeviceToHost, m->stream ) );
    PXD_CHECK( hipStreamSynchronize( m->stream ) );
    PXD_CHECK( hipFree( dX ) );
    PXD_CHECK( hipFree( dy ) );
    return 0;
}

}  // extern "C"
