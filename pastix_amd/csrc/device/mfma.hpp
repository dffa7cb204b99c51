// mfma.hpp -- CDNA4 matrix-core helpers (v_mfma_*_16x16x4 for f32 and f64).
//
// Lane layout of the 16x16x4 instructions (checked on gfx950): lane l holds
// A[l%16][l/16] and B[l/16][l%16]; the accumulator holds 4 values of D with
// column l%16 and row 4*(l/16)+v for f32, row (l/16)+4*v for f64.
#pragma once
#include "px_device.hpp"

typedef float  pxf4 __attribute__( ( ext_vector_type( 4 ) ) );
typedef double pxd4 __attribute__( ( ext_vector_type( 4 ) ) );

template <typename T> struct mfma16;

template <> struct mfma16<float> {
    typedef pxf4 acc_t;
    __device__ static inline acc_t zero() { return acc_t{ 0.f, 0.f, 0.f, 0.f }; }
    __device__ static inline acc_t mma( float a, float b, acc_t c )
    {
        return __builtin_amdgcn_mfma_f32_16x16x4f32( a, b, c, 0, 0, 0 );
    }
    __device__ static inline int row( int lane, int v ) { return 4 * ( lane >> 4 ) + v; }
    __device__ static inline float get( const acc_t &c, int v ) { return c[v]; }
    __device__ static inline void  set( acc_t &c, int v, float x ) { c[v] = x; }
};

template <> struct mfma16<double> {
    typedef pxd4 acc_t;
    __device__ static inline acc_t zero() { return acc_t{ 0., 0., 0., 0. }; }
    __device__ static inline acc_t mma( double a, double b, acc_t c )
    {
        return __builtin_amdgcn_mfma_f64_16x16x4f64( a, b, c, 0, 0, 0 );
    }
    __device__ static inline int row( int lane, int v ) { return ( lane >> 4 ) + 4 * v; }
    __device__ static inline double get( const acc_t &c, int v ) { return c[v]; }
    __device__ static inline void   set( acc_t &c, int v, double x ) { c[v] = x; }
};

// Complex 16x16x4 products as four real ones: (ar + i ai)(br + i bi)
// accumulated into separate real and imaginary accumulators (no conjugation:
// the complex factorizations here are symmetric, not Hermitian).
template <typename C, typename R> struct mfma16_cplx {
    typedef typename mfma16<R>::acc_t racc;
    struct acc_t {
        racc re, im;
    };
    __device__ static inline acc_t zero() { return acc_t{ mfma16<R>::zero(), mfma16<R>::zero() }; }
    __device__ static inline acc_t mma( C a, C b, acc_t c )
    {
        c.re = mfma16<R>::mma( a.re, b.re, c.re );
        c.re = mfma16<R>::mma( -a.im, b.im, c.re );
        c.im = mfma16<R>::mma( a.re, b.im, c.im );
        c.im = mfma16<R>::mma( a.im, b.re, c.im );
        return c;
    }
    __device__ static inline int row( int lane, int v ) { return mfma16<R>::row( lane, v ); }
    __device__ static inline C    get( const acc_t &c, int v ) { return C( c.re[v], c.im[v] ); }
    __device__ static inline void set( acc_t &c, int v, C x )
    {
        c.re[v] = x.re;
        c.im[v] = x.im;
    }
};
template <> struct mfma16<pxcf> : mfma16_cplx<pxcf, float> {};
template <> struct mfma16<pxcd> : mfma16_cplx<pxcd, double> {};

// 16-byte LDS fragments: KV consecutive k values of one row
template <typename T> struct lds_vec;
template <> struct lds_vec<float> {
    typedef float type __attribute__( ( ext_vector_type( 4 ) ) );
    static constexpr int n = 4;
};
template <> struct lds_vec<double> {
    typedef double type __attribute__( ( ext_vector_type( 2 ) ) );
    static constexpr int n = 2;
};
template <typename C, int N> struct cvec {
    C e[N];
    __device__ inline C       &operator[]( int i ) { return e[i]; }
    __device__ inline const C &operator[]( int i ) const { return e[i]; }
} __attribute__( ( aligned( 16 ) ) );
template <> struct lds_vec<pxcf> {
    typedef cvec<pxcf, 2> type;
    static constexpr int n = 2;
};
template <> struct lds_vec<pxcd> {
    typedef cvec<pxcd, 1> type;
    static constexpr int n = 1;
};

// Task of workgroup b among nb: the workgroups that the dispatcher sends to
// one XCD (b % 8) take a contiguous range of the task list, so that tiles
// that share source rows share that XCD's L2.
__device__ inline int32_t
xcd_task( int32_t b, int32_t nb )
{
    const int32_t q = nb >> 3, r = nb & 7, x = b & 7, i = b >> 3;
    return x * q + min( x, r ) + i;
}

// Buffer loads and stores relative to one cblk's coefficients: a byte offset
// at or beyond the range of the resource (PXD_OOB) reads 0 and drops the
// store, which replaces per-element bounds branches.
#define PXD_OOB 0x80000000u
__device__ inline __amdgpu_buffer_rsrc_t
make_rsrc( const void *p, uint32_t bytes )
{
    return __builtin_amdgcn_make_buffer_rsrc( const_cast<void *>( p ), 0, (int)bytes, 0x00020000 );
}
template <typename T> struct bufio;
template <> struct bufio<float> {
    __device__ static inline float load( __amdgpu_buffer_rsrc_t r, uint32_t off )
    {
        return __builtin_bit_cast( float, __builtin_amdgcn_raw_buffer_load_b32( r, off, 0, 0 ) );
    }
    __device__ static inline void store( float v, __amdgpu_buffer_rsrc_t r, uint32_t off )
    {
        __builtin_amdgcn_raw_buffer_store_b32( __builtin_bit_cast( unsigned, v ), r, off, 0, 0 );
    }
};
// The 64- and 128-bit loads are bit-cast whole: taking 32-bit elements of the
// builtin's vector result compiles (ROCm 7.2 clang, gfx950) to a single
// dword load whose value is copied into every element.
template <> struct bufio<pxcf> {
    typedef unsigned u2 __attribute__( ( ext_vector_type( 2 ) ) );
    __device__ static inline pxcf load( __amdgpu_buffer_rsrc_t r, uint32_t off )
    {
        return __builtin_bit_cast( pxcf, __builtin_amdgcn_raw_buffer_load_b64( r, off, 0, 0 ) );
    }
    __device__ static inline void store( pxcf v, __amdgpu_buffer_rsrc_t r, uint32_t off )
    {
        u2 w;
        w.x = __builtin_bit_cast( unsigned, v.re );
        w.y = __builtin_bit_cast( unsigned, v.im );
        __builtin_amdgcn_raw_buffer_store_b64( w, r, off, 0, 0 );
    }
};
template <> struct bufio<pxcd> {
    typedef unsigned u4 __attribute__( ( ext_vector_type( 4 ) ) );
    typedef unsigned u2 __attribute__( ( ext_vector_type( 2 ) ) );
    __device__ static inline pxcd load( __amdgpu_buffer_rsrc_t r, uint32_t off )
    {
        return __builtin_bit_cast( pxcd, __builtin_amdgcn_raw_buffer_load_b128( r, off, 0, 0 ) );
    }
    __device__ static inline void store( pxcd v, __amdgpu_buffer_rsrc_t r, uint32_t off )
    {
        u2 a = __builtin_bit_cast( u2, v.re ), b = __builtin_bit_cast( u2, v.im );
        u4 w = { a.x, a.y, b.x, b.y };
        __builtin_amdgcn_raw_buffer_store_b128( w, r, off, 0, 0 );
    }
};
template <> struct bufio<double> {
    typedef unsigned u2 __attribute__( ( ext_vector_type( 2 ) ) );
    __device__ static inline double load( __amdgpu_buffer_rsrc_t r, uint32_t off )
    {
        return __builtin_bit_cast( double, __builtin_amdgcn_raw_buffer_load_b64( r, off, 0, 0 ) );
    }
    __device__ static inline void store( double v, __amdgpu_buffer_rsrc_t r, uint32_t off )
    {
        __builtin_amdgcn_raw_buffer_store_b64( __builtin_bit_cast( u2, v ), r, off, 0, 0 );
    }
};
