// px_device.hpp -- shared definitions of the HIP engine (device side).
//
// Scalar types: double / float for real factors, pxcd / pxcf (complex double
// / complex float, stored as two consecutive reals like C99 _Complex) for the
// complex-symmetric LDL^T path.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include "pastix_amd_device.h"

#define PXD_CHECK( call )                                                              \
    do {                                                                               \
        hipError_t e_ = ( call );                                                      \
        if ( e_ != hipSuccess ) {                                                      \
            fprintf( stderr, "pastix_amd device: %s failed: %s (%s:%d)\n", #call,       \
                     hipGetErrorString( e_ ), __FILE__, __LINE__ );                    \
            return 6; /* PASTIX_ERR_INTERNAL */                                        \
        }                                                                              \
    } while ( 0 )

struct pxcd {
    double re, im;
    pxcd() = default;
    __host__ __device__ constexpr pxcd( double r, double i = 0. ) : re( r ), im( i ) {}
};
struct pxcf {
    float re, im;
    pxcf() = default;
    __host__ __device__ constexpr pxcf( float r, float i = 0.f ) : re( r ), im( i ) {}
};

// ---------------------------------------------------------------------------
// Scalar helpers usable for the four arithmetics.
template <typename T> struct px_real;
template <> struct px_real<double> { typedef double type; };
template <> struct px_real<float>  { typedef float  type; };
template <> struct px_real<pxcd>   { typedef double type; };
template <> struct px_real<pxcf>   { typedef float  type; };

__host__ __device__ inline double px_abs( double a ) { return fabs( a ); }
__host__ __device__ inline float  px_abs( float a ) { return fabsf( a ); }
__host__ __device__ inline double px_abs( pxcd a ) { return hypot( a.re, a.im ); }
__host__ __device__ inline float  px_abs( pxcf a ) { return hypotf( a.re, a.im ); }

__host__ __device__ inline double px_zero( double ) { return 0.; }
__host__ __device__ inline float  px_zero( float ) { return 0.f; }
__host__ __device__ inline pxcd   px_zero( pxcd ) { return pxcd{ 0., 0. }; }
__host__ __device__ inline pxcf   px_zero( pxcf ) { return pxcf{ 0.f, 0.f }; }

__host__ __device__ inline pxcd operator+( pxcd a, pxcd b ) { return pxcd{ a.re + b.re, a.im + b.im }; }
__host__ __device__ inline pxcd operator-( pxcd a, pxcd b ) { return pxcd{ a.re - b.re, a.im - b.im }; }
__host__ __device__ inline pxcd operator*( pxcd a, pxcd b ) { return pxcd{ a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re }; }
__host__ __device__ inline pxcd operator-( pxcd a ) { return pxcd{ -a.re, -a.im }; }
__host__ __device__ inline pxcf operator+( pxcf a, pxcf b ) { return pxcf{ a.re + b.re, a.im + b.im }; }
__host__ __device__ inline pxcf operator-( pxcf a, pxcf b ) { return pxcf{ a.re - b.re, a.im - b.im }; }
__host__ __device__ inline pxcf operator*( pxcf a, pxcf b ) { return pxcf{ a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re }; }
__host__ __device__ inline pxcf operator-( pxcf a ) { return pxcf{ -a.re, -a.im }; }

__host__ __device__ inline pxcd operator*( pxcd a, double b ) { return pxcd{ a.re * b, a.im * b }; }
__host__ __device__ inline pxcf operator*( pxcf a, float b ) { return pxcf{ a.re * b, a.im * b }; }
__host__ __device__ inline pxcd &operator+=( pxcd &a, pxcd b ) { a = a + b; return a; }
__host__ __device__ inline pxcd &operator-=( pxcd &a, pxcd b ) { a = a - b; return a; }
__host__ __device__ inline pxcf &operator+=( pxcf &a, pxcf b ) { a = a + b; return a; }
__host__ __device__ inline pxcf &operator-=( pxcf &a, pxcf b ) { a = a - b; return a; }

// real part (the sign of a static pivot follows it, core_zsytf2sp)
__host__ __device__ inline double px_re( double a ) { return a; }
__host__ __device__ inline float  px_re( float a ) { return a; }
__host__ __device__ inline double px_re( pxcd a ) { return a.re; }
__host__ __device__ inline float  px_re( pxcf a ) { return a.re; }

__host__ __device__ inline pxcd px_div( pxcd a, pxcd b )
{
    // Smith's algorithm
    if ( fabs( b.re ) >= fabs( b.im ) ) {
        double r = b.im / b.re, d = b.re + b.im * r;
        return pxcd{ ( a.re + a.im * r ) / d, ( a.im - a.re * r ) / d };
    }
    double r = b.re / b.im, d = b.im + b.re * r;
    return pxcd{ ( a.re * r + a.im ) / d, ( a.im * r - a.re ) / d };
}
__host__ __device__ inline pxcf px_div( pxcf a, pxcf b )
{
    if ( fabsf( b.re ) >= fabsf( b.im ) ) {
        float r = b.im / b.re, d = b.re + b.im * r;
        return pxcf{ ( a.re + a.im * r ) / d, ( a.im - a.re * r ) / d };
    }
    float r = b.re / b.im, d = b.im + b.re * r;
    return pxcf{ ( a.re * r + a.im ) / d, ( a.im * r - a.re ) / d };
}
__host__ __device__ inline pxcd operator/( pxcd a, pxcd b ) { return px_div( a, b ); }
__host__ __device__ inline pxcf operator/( pxcf a, pxcf b ) { return px_div( a, b ); }
__host__ __device__ inline double px_div( double a, double b ) { return a / b; }
__host__ __device__ inline float  px_div( float a, float b ) { return a / b; }

// fused multiply-add c + a*b
// Widening of a factor entry to the vector precision of a solve (single
// factor, double vectors) or identity.
template <typename V, typename T> struct px_cvt_s;
template <typename T> struct px_cvt_s<T, T> {
    __device__ static inline T f( T x ) { return x; }
};
template <> struct px_cvt_s<double, float> {
    __device__ static inline double f( float x ) { return (double)x; }
};
template <> struct px_cvt_s<pxcd, pxcf> {
    __device__ static inline pxcd f( pxcf x ) { return pxcd( (double)x.re, (double)x.im ); }
};
template <typename V, typename T>
__device__ inline V
px_cvt( T x )
{
    return px_cvt_s<V, T>::f( x );
}

__device__ inline double px_fma( double a, double b, double c ) { return fma( a, b, c ); }
__device__ inline float  px_fma( float a, float b, float c ) { return fmaf( a, b, c ); }
__device__ inline pxcd   px_fma( pxcd a, pxcd b, pxcd c ) { return c + a * b; }
__device__ inline pxcf   px_fma( pxcf a, pxcf b, pxcf c ) { return c + a * b; }

// conversions from the input (double / complex double) to the factor type
__device__ inline double px_from( double v, const double * ) { return v; }
__device__ inline float  px_from( double v, const float * ) { return (float)v; }
__device__ inline pxcd   px_from( pxcd v, const pxcd * ) { return v; }
__device__ inline pxcf   px_from( pxcd v, const pxcf * ) { return pxcf{ (float)v.re, (float)v.im }; }

// atomic add on global memory (hardware float atomics on gfx950)
__device__ inline void px_atomic_add( double *p, double v ) { unsafeAtomicAdd( p, v ); }
__device__ inline void px_atomic_add( float *p, float v ) { unsafeAtomicAdd( p, v ); }
__device__ inline void px_atomic_add( pxcd *p, pxcd v )
{
    unsafeAtomicAdd( &p->re, v.re );
    unsafeAtomicAdd( &p->im, v.im );
}
__device__ inline void px_atomic_add( pxcf *p, pxcf v )
{
    unsafeAtomicAdd( &p->re, v.re );
    unsafeAtomicAdd( &p->im, v.im );
}

// sqrt used by the Cholesky panel (real symmetric positive definite only)
__device__ inline double px_sqrt( double a ) { return sqrt( a ); }
__device__ inline float  px_sqrt( float a ) { return sqrtf( a ); }

// generic conversions used by the vector <-> factor precision kernels
template <typename TO, typename TI> __device__ inline TO px_from_any( TI v );
template <> __device__ inline double px_from_any<double, double>( double v ) { return v; }
template <> __device__ inline float  px_from_any<float, double>( double v ) { return (float)v; }
template <> __device__ inline double px_from_any<double, float>( float v ) { return (double)v; }
template <> __device__ inline float  px_from_any<float, float>( float v ) { return v; }
template <> __device__ inline pxcd   px_from_any<pxcd, pxcd>( pxcd v ) { return v; }
template <> __device__ inline pxcf   px_from_any<pxcf, pxcd>( pxcd v ) { return pxcf{ (float)v.re, (float)v.im }; }
template <> __device__ inline pxcd   px_from_any<pxcd, pxcf>( pxcf v ) { return pxcd{ (double)v.re, (double)v.im }; }
template <> __device__ inline pxcf   px_from_any<pxcf, pxcf>( pxcf v ) { return v; }
