// Forward-mode dual numbers for the adjoint and tangent-linear kernels.
//
// Dual<N> carries a double value and N directional derivatives.  The solver's device building blocks
// (hydro_math.hpp) are templates on their arithmetic type, so instantiating them with Dual<N> evaluates the
// very same operation sequence on the value part while propagating derivatives.  Every comparison, min/max
// and select acts on the value part, so derivatives follow the branch the plain FP64 evaluation takes.
// The derivative parts use explicit fused multiply-adds: the file that instantiates them is compiled without
// contraction.
#pragma once
#include "hydro_math.hpp"

namespace gamer_amd {

template <int N>
struct Dual {
   double v;
   double d[N];

   __device__ __forceinline__ Dual() {}
   __device__ __forceinline__ Dual( double x ) : v( x )
   {
#     pragma unroll
      for (int k=0; k<N; k++) d[k] = 0.0;
   }

   friend __device__ __forceinline__ Dual operator+( const Dual &a, const Dual &b )
   {
      Dual r;  r.v = a.v + b.v;
#     pragma unroll
      for (int k=0; k<N; k++) r.d[k] = a.d[k] + b.d[k];
      return r;
   }
   friend __device__ __forceinline__ Dual operator-( const Dual &a, const Dual &b )
   {
      Dual r;  r.v = a.v - b.v;
#     pragma unroll
      for (int k=0; k<N; k++) r.d[k] = a.d[k] - b.d[k];
      return r;
   }
   friend __device__ __forceinline__ Dual operator-( const Dual &a )
   {
      Dual r;  r.v = -a.v;
#     pragma unroll
      for (int k=0; k<N; k++) r.d[k] = -a.d[k];
      return r;
   }
   friend __device__ __forceinline__ Dual operator*( const Dual &a, const Dual &b )
   {
      Dual r;  r.v = a.v*b.v;
#     pragma unroll
      for (int k=0; k<N; k++) r.d[k] = __builtin_fma( a.d[k], b.v, a.v*b.d[k] );
      return r;
   }
   friend __device__ __forceinline__ Dual operator/( const Dual &a, const Dual &b )
   {
      Dual r;  r.v = a.v/b.v;
      const double ib = 1.0/b.v;
#     pragma unroll
      for (int k=0; k<N; k++) r.d[k] = __builtin_fma( -r.v, b.d[k], a.d[k] )*ib;
      return r;
   }
   __device__ __forceinline__ Dual &operator+=( const Dual &b ) { *this = *this + b;  return *this; }
   __device__ __forceinline__ Dual &operator-=( const Dual &b ) { *this = *this - b;  return *this; }
   __device__ __forceinline__ Dual &operator*=( const Dual &b ) { *this = *this * b;  return *this; }

   friend __device__ __forceinline__ bool operator< ( const Dual &a, const Dual &b ) { return a.v <  b.v; }
   friend __device__ __forceinline__ bool operator> ( const Dual &a, const Dual &b ) { return a.v >  b.v; }
   friend __device__ __forceinline__ bool operator<=( const Dual &a, const Dual &b ) { return a.v <= b.v; }
   friend __device__ __forceinline__ bool operator>=( const Dual &a, const Dual &b ) { return a.v >= b.v; }
   friend __device__ __forceinline__ bool operator==( const Dual &a, const Dual &b ) { return a.v == b.v; }
   friend __device__ __forceinline__ bool operator!=( const Dual &a, const Dual &b ) { return a.v != b.v; }
};

// a*b + c with the value rounded once, as the plain-type fmad
template <int N> __device__ __forceinline__ Dual<N> fmad( const Dual<N> &a, const Dual<N> &b, const Dual<N> &c )
{
   Dual<N> r;  r.v = fmad( a.v, b.v, c.v );
#  pragma unroll
   for (int k=0; k<N; k++) r.d[k] = __builtin_fma( a.d[k], b.v, __builtin_fma( a.v, b.d[k], c.d[k] ) );
   return r;
}
// d(1/x) = -dx/x^2 ; the value is the solver's own reciprocal
template <int N> __device__ __forceinline__ Dual<N> rcp( const Dual<N> &a )
{
   Dual<N> r;  r.v = rcp( a.v );
   const double s = -r.v*r.v;
#  pragma unroll
   for (int k=0; k<N; k++) r.d[k] = s*a.d[k];
   return r;
}
template <int N> __device__ __forceinline__ Dual<N> vsqrt( const Dual<N> &a )
{
   Dual<N> r;  r.v = vsqrt( a.v );
   const double s = 0.5/r.v;
#  pragma unroll
   for (int k=0; k<N; k++) r.d[k] = s*a.d[k];
   return r;
}
// fmax/fmin semantics: a NaN operand loses
template <int N> __device__ __forceinline__ Dual<N> vmax( const Dual<N> &a, const Dual<N> &b )
{ return ( b.v != b.v || a.v >= b.v ) ? a : b; }
template <int N> __device__ __forceinline__ Dual<N> vmin( const Dual<N> &a, const Dual<N> &b )
{ return ( b.v != b.v || a.v <= b.v ) ? a : b; }
template <int N> __device__ __forceinline__ Dual<N> vabs( const Dual<N> &a ) { return __builtin_signbit( a.v ) ? -a : a; }
template <int N> __device__ __forceinline__ Dual<N> vcopysign( const Dual<N> &m, const Dual<N> &s )
{
   const Dual<N> am = vabs( m );
   return __builtin_signbit( s.v ) ? -am : am;
}
template <int N> __device__ __forceinline__ Dual<N> clamp_between( const Dual<N> &x, const Dual<N> &a, const Dual<N> &b )
{ return vmin( vmax( x, vmin( a, b ) ), vmax( a, b ) ); }

template <typename T> struct ValueOf            { __device__ static double get( const T &x ) { return x; } };
template <int N>      struct ValueOf< Dual<N> > { __device__ static double get( const Dual<N> &x ) { return x.v; } };

// solver constants in another arithmetic type (derivative-free constants)
template <typename T>
__device__ __forceinline__ Prm<T> prm_as( const Prm<double> &P )
{
   Prm<T> Q;
   Q.gamma = P.gamma;  Q.gm1 = P.gm1;  Q.inv_gm1 = P.inv_gm1;
   Q.min_dens = P.min_dens;  Q.min_pres = P.min_pres;  Q.min_eint = P.min_eint;  Q.coeff = P.coeff;
   Q.dt_dh2 = P.dt_dh2;  Q.dt_dh = P.dt_dh;  Q.gm1_d = P.gm1_d;
   Q.min_dens_d = P.min_dens_d;  Q.min_pres_d = P.min_pres_d;  Q.min_eint_d = P.min_eint_d;
   Q.limiter = P.limiter;
   return Q;
}

} // namespace gamer_amd
