// Device-side building blocks of the MHM/MHM_RP patch-group solver: ideal-gas conversions, the HLLC / HLLE /
// Roe interface solvers and the PLM slope limiters.  Written for gfx950 wave64 VALU; templated on the bulk
// arithmetic type R (float or double).  Pressure recovery E - |m|^2/(2 rho) goes through type A (double).
//
// Operation order follows the reference where it matters for parity:
//   Riemann solvers   src/Model_Hydro/CPU_Hydro/CPU_Shared_RiemannSolver_{HLLC:391-655, HLLE:301-698, Roe:87-711}.cpp
//   limiter + clamp   src/Model_Hydro/CPU_Hydro/CPU_Shared_DataReconstruction.cpp:403-430, 1917-2032
//   conversions       src/Model_Hydro/CPU_Hydro/CPU_Shared_FluUtility.cpp:209-551, 681-805
// Interface states arrive already rotated to the face normal with their pressure, 1/rho and sound speed
// evaluated once (per cell for the half step, per face for the full step) instead of inside every solve.
#pragma once
#include <hip/hip_runtime.h>
#include <type_traits>

namespace gamer_amd {

constexpr int NV = 5;

// ---------------- scalar helpers ----------------
__device__ __forceinline__ float  rcp( float x )  { return __builtin_amdgcn_rcpf( x ); }
__device__ __forceinline__ double rcp( double x )
{
   // hardware estimate + two Newton steps: within 1 ulp of 1/x for normal x
   double r = __builtin_amdgcn_rcp( x );
   double e = __builtin_fma( -x, r, 1.0 );
   r = __builtin_fma( r, e, r );
   e = __builtin_fma( -x, r, 1.0 );
   return __builtin_fma( r, e, r );
}
__device__ __forceinline__ float  vsqrt( float x )  { return __builtin_amdgcn_sqrtf( x ); }
__device__ __forceinline__ double vsqrt( double x ) { return __builtin_sqrt( x ); }
// fused multiply-add written out where the rounding of an expression must not depend on how the compiler
// contracts it (the same face state is built by different code paths)
__device__ __forceinline__ float  fmad( float a, float b, float c )    { return __builtin_fmaf( a, b, c ); }
__device__ __forceinline__ double fmad( double a, double b, double c ) { return __builtin_fma( a, b, c ); }
template <typename R> __device__ __forceinline__ R vmax( R a, R b ) { return __builtin_fmax( a, b ); }
template <typename R> __device__ __forceinline__ R vmin( R a, R b ) { return __builtin_fmin( a, b ); }
template <typename R> __device__ __forceinline__ R vabs( R a )      { return __builtin_fabs( a ); }
template <typename R> __device__ __forceinline__ R vcopysign( R m, R s ) { return __builtin_copysign( m, s ); }

// reciprocal for the FP64 pressure recovery of the FP32-bulk kernels: hardware estimate + one Newton step, a
// relative error near 1e-14, far below what the float result keeps (two steps, as in rcp(double), for FP64)
__device__ __forceinline__ double rcp_recover( double x )
{
   const double r = __builtin_amdgcn_rcp( x );
   return __builtin_fma( r, __builtin_fma( -x, r, 1.0 ), r );
}

template <typename R> __device__ __forceinline__ double rcp_pres( double x ) { return sizeof( R ) == 4 ? rcp_recover( x ) : rcp( x ); }

// sound speed (gamma P / rho)^1/2 from gp = gamma P, rho and ir = 1/rho.  Generic form sqrt(gp ir).  FP64:
// gp (gp rho)^-1/2 from the hardware reciprocal square root and two Newton steps, about 2 ulp, one transcendental
// and 9 instructions where the reciprocal's Newton steps and the IEEE square root take about 20; gp == 0 gives 0
// and NaN passes.  The derivative kernels define HYDRO_IEEE_SOUND_SPEED before including this header and keep the
// generic form, which their dual-number twins evaluate bit for bit.
template <typename R> __device__ __forceinline__ R sound_speed( R gp, R rho, R ir ) { (void)rho; return vsqrt( gp*ir ); }
#ifndef HYDRO_IEEE_SOUND_SPEED
template <> __device__ __forceinline__ double sound_speed<double>( double gp, double rho, double ir )
{
   (void)ir;
   const double a = gp*rho, h = 0.5*a;
   double y = __builtin_amdgcn_rsq( a );
   double r = __builtin_fma( -h*y, y, 0.5 );        // y (1.5 - a y^2 / 2) = y + y (0.5 - h y^2)
   y = __builtin_fma( y, r, y );
   r = __builtin_fma( -h*y, y, 0.5 );
   y = __builtin_fma( y, r, y );
   return ( gp == 0.0 ) ? 0.0 : gp*y;
}
#endif

// pressure floor that lets NaN through (FluUtility.cpp:681-690): a NaN compares false and is returned as is.
// One compare + select; max() would first have to quiet operands read from memory (IEEE mode).
template <typename R> __device__ __forceinline__ R floor_pres( R p, R pmin ) { return ( p < pmin ) ? pmin : p; }

// ---------------- raw buffer access ----------------
// Global loads and stores of the z-marching kernels go through a buffer descriptor per patch group: the 64-bit
// base stays in SGPRs and every access takes a lane-constant 32-bit byte offset plus a wave-uniform byte offset,
// so the march does no 64-bit address arithmetic per access.  The hardware range-checks each access against
// `bytes`: a read past it returns 0 and a write past it is dropped.  Word 3 = 0x00020000 is the gfx9-family
// raw-buffer format (32-bit data format, no swizzle).
__device__ __forceinline__ __amdgpu_buffer_rsrc_t make_buf( const void *base, unsigned bytes )
{
   return __builtin_amdgcn_make_buffer_rsrc( const_cast<void *>( base ), (short)0, (int)bytes, 0x00020000 );
}
__device__ __forceinline__ void buf_load( __amdgpu_buffer_rsrc_t r, int voff, int soff, float &x )
{
   x = __builtin_bit_cast( float, __builtin_amdgcn_raw_buffer_load_b32( r, voff, soff, 0 ) );
}
__device__ __forceinline__ void buf_load( __amdgpu_buffer_rsrc_t r, int voff, int soff, double &x )
{
   x = __builtin_bit_cast( double, __builtin_amdgcn_raw_buffer_load_b64( r, voff, soff, 0 ) );
}
__device__ __forceinline__ void buf_store( float x, __amdgpu_buffer_rsrc_t r, int voff, int soff )
{
   __builtin_amdgcn_raw_buffer_store_b32( __builtin_bit_cast( unsigned, x ), r, voff, soff, 0 );
}
__device__ __forceinline__ void buf_store( double x, __amdgpu_buffer_rsrc_t r, int voff, int soff )
{
   using U2 = decltype( __builtin_amdgcn_raw_buffer_load_b64( __amdgpu_buffer_rsrc_t(), 0, 0, 0 ) );
   __builtin_amdgcn_raw_buffer_store_b64( __builtin_bit_cast( U2, x ), r, voff, soff, 0 );
}

// ---------------- solver constants ----------------
template <typename R>
struct Prm {
   R gamma, gm1, inv_gm1, min_dens, min_pres, min_eint, coeff;
   double dt_dh2, dt_dh, gm1_d, min_dens_d, min_pres_d, min_eint_d;
   int limiter;
};

// Multi-level batch description: optional per-PG (dt, dh) pairs in device memory (uneven AMR batches, local
// time steps) and an optional global work-queue counter for the persistent mode (each workgroup pulls tickets of
// `claim` consecutive patch groups until the queue is exhausted).  pg_dt_dh == nullptr: every PG uses Prm's dt/dh;
// counter == nullptr: one PG per workgroup.  `waves` selects the MHM_RP kernel's workgroup shape.
struct Batch {
   const double *pg_dt_dh;
   unsigned     *counter;
   double        safety;
   int           claim;
   int           waves;         // launch shape: 4 or 8 waves per workgroup, 0 = the kernel's default (host side only)
};

// an interface state rotated to the face normal: conserved fields + derived quantities
template <typename R>
struct St { R d, mn, m1, m2, e, p, ir, cs, u; };

// ---------------- Riemann solvers (Davis wave speeds), flux returned in the rotated frame ----------------
template <typename R>
__device__ __forceinline__ void xflux_minus( const St<R> &s, R sp, R f[NV] )
{
   // physical x-flux minus sp*U (HLLC.cpp:600-612, HLLE.cpp:640-651), with (u - sp) factored out: one
   // instruction per transverse component instead of two
   const R us = s.u - sp;
   f[0] = s.mn - sp*s.d;
   f[1] = us*s.mn + s.p;
   f[2] = us*s.m1;
   f[3] = us*s.m2;
   f[4] = us*s.e + s.u*s.p;
}

template <typename R>
__device__ __forceinline__ void xflux( const St<R> &s, R f[NV] )
{
   f[0] = s.mn;
   f[1] = s.u*s.mn + s.p;
   f[2] = s.u*s.m1;
   f[3] = s.u*s.m2;
   f[4] = s.u*( s.e + s.p );
}

template <typename R>
__device__ __forceinline__ void rs_hlle( const St<R> &L, const St<R> &Q, R f[NV] )
{
   const R Z = (R)0;
   R sl = vmin( L.u - L.cs, Q.u - Q.cs );
   R sr = vmax( L.u + L.cs, Q.u + Q.cs );
   sl = vmin( sl, Z );
   sr = vmax( sr, Z );
   R fl[NV], fr[NV];
   xflux_minus( L, sl, fl );
   xflux_minus( Q, sr, fr );
   const bool both0 = ( sl == Z ) && ( sr == Z );
   const R inv = rcp( sr - sl );
#  pragma unroll
   for (int v=0; v<NV; v++) f[v] = both0 ? fl[v] : inv*( sr*fl[v] - sl*fr[v] );
}

template <typename R>
__device__ __forceinline__ void rs_hllc( const St<R> &L, const St<R> &Q, R pmin, R f[NV] )
{
   const R Z = (R)0, ONE = (R)1;
   const R wl1 = L.u - L.cs, wl2 = Q.u - Q.cs, wr1 = L.u + L.cs, wr2 = Q.u + Q.cs;
   const R wl = vmin( wl1, wl2 ), wr = vmax( wr1, wr2 );
   const R du = L.u - Q.u;
   const R tl =  L.d*( ( wl1 < wl2 ) ? L.cs : du + Q.cs );
   const R tr = -Q.d*( ( wr2 > wr1 ) ? Q.cs : du + L.cs );
   const R t2 = rcp( tl - tr );
   const R vs = t2*( L.p - Q.p + tl*L.u - tr*Q.u );
   R ps = t2*( tl*( Q.p + tr*Q.u ) - tr*( L.p + tl*L.u ) );
   ps = floor_pres( ps, pmin );

   const bool left = ( vs >= Z );
   St<R> s;
   s.d  = left ? L.d  : Q.d;   s.mn = left ? L.mn : Q.mn;  s.m1 = left ? L.m1 : Q.m1;
   s.m2 = left ? L.m2 : Q.m2;  s.e  = left ? L.e  : Q.e;   s.p  = left ? L.p  : Q.p;
   s.u  = left ? L.u  : Q.u;
   const R sp = left ? vmin( wl, Z ) : vmax( wr, Z );
   R F[NV];
   xflux_minus( s, sp, F );
   const bool deg = left && ( vs == Z ) && ( sp == Z );
   const R t4 = rcp( vs - sp );
   const R cf = deg ? ONE : t4*vs;
   const R cs = deg ? Z   : -t4*sp*ps;
   f[0] = cf*F[0];
   f[1] = cf*F[1] + cs;
   f[2] = cf*F[2];
   f[3] = cf*F[3];
   f[4] = cf*F[4] + cs*vs;
}

// Roe solver with the intermediate-state positivity check falling back to HLLE (Roe.cpp:87-711).  GAMER compiles
// that check only into builds whose RSOLVER is ROE (include/CUFLU.h:178-186, CHECK_INTERMEDIATE); CHECK = false is
// the Roe solve of the other builds, which CorrectUnphysical's first-order correction can still select.
template <typename R, bool CHECK = true>
__device__ __forceinline__ void rs_roe( const St<R> &L, const St<R> &Q, const Prm<R> &P, R f[NV] )
{
   const R Z = (R)0, HALF = (R)0.5;
   const R hl = L.ir*( L.e + L.p ), hr = Q.ir*( Q.e + Q.p );
   const R sql = vsqrt( L.d ), sqr = vsqrt( Q.d );
   const R rho = sql*sqr, irho = rcp( rho );
   const R isql = rcp( sql ), isqr = rcp( sqr ), isum = rcp( sql + sqr );
   const R u = isum*( isql*L.mn + isqr*Q.mn );
   const R v = isum*( isql*L.m1 + isqr*Q.m1 );
   const R w = isum*( isql*L.m2 + isqr*Q.m2 );
   const R V2 = u*u + v*v + w*w;
   const R H  = isum*( sql*hl + sqr*hr );
   R gpr = P.gm1*( H - HALF*V2 );
   gpr = P.gamma*irho*floor_pres( gpr*rho*rcp( P.gamma ), P.min_pres );
   const R a2 = gpr, a = vsqrt( a2 );
   const R e0 = u - a, e4 = u + a;

   R fl[NV], fr[NV];
   xflux( L, fl );
   xflux( Q, fr );
   if ( e0 >= Z ) { for (int k=0; k<NV; k++) f[k] = fl[k];  return; }
   if ( e4 <= Z ) { for (int k=0; k<NV; k++) f[k] = fr[k];  return; }

   const R j0 = Q.d - L.d, j1 = Q.mn - L.mn, j2 = Q.m1 - L.m1, j3 = Q.m2 - L.m2, j4 = Q.e - L.e;
   const R a2_ = j2 - v*j0;
   const R a3_ = j3 - w*j0;
   const R a1_ = P.gm1*rcp( a2 )*( j0*( H - u*u ) + u*j1 - j4 + v*a2_ + w*a3_ );
   const R a0_ = HALF*rcp( a )*( j0*( u + a ) - j1 - a*a1_ );
   const R a4_ = j0 - a0_ - a1_;

   // intermediate states after the 1st and 4th waves (the degenerate middle waves are skipped)
   if constexpr ( CHECK )
   {
      const R d1 = L.d + a0_, m1 = L.mn + a0_*e0, t1 = L.m1 + a0_*v, t2 = L.m2 + a0_*w, en1 = L.e + a0_*( H - u*a );
      const R ie1 = en1 - HALF*( m1*m1 + t1*t1 + t2*t2 )*rcp( d1 );
      bool bad = ( u > e0 ) && ( d1 <= Z || ie1*P.gm1 <= Z );
      const R d4 = d1 + a1_, m4 = m1 + a1_*u, s4 = t1 + a1_*v + a2_, r4 = t2 + a1_*w + a3_,
              en4 = en1 + a1_*( HALF*V2 ) + a2_*v + a3_*w;
      const R ie4 = en4 - HALF*( m4*m4 + s4*s4 + r4*r4 )*rcp( d4 );
      bad = bad || ( ( e4 > u ) && ( d4 <= Z || ie4*P.gm1 <= Z ) );
      if ( bad ) { rs_hlle( L, Q, f );  return; }
   }

   const R A0 = a0_*vabs( e0 ), A1 = a1_*vabs( u ), A2 = a2_*vabs( u ), A3 = a3_*vabs( u ), A4 = a4_*vabs( e4 );
   f[0] = HALF*( fl[0] + fr[0] - A0 - A1 - A4 );
   f[1] = HALF*( fl[1] + fr[1] - A0*e0 - A1*u - A4*e4 );
   f[2] = HALF*( fl[2] + fr[2] - A0*v - A1*v - A2 - A4*v );
   f[3] = HALF*( fl[3] + fr[3] - A0*w - A1*w - A3 - A4*w );
   f[4] = HALF*( fl[4] + fr[4] - A0*( H - u*a ) - A1*( HALF*V2 ) - A2*v - A3*w - A4*( H + u*a ) );
}

// ---------------- PLM reconstruction of one cell along one direction ----------------
// Limited slope (DataReconstruction.cpp:1917-2032), faces C -/+ slope/2, then clamp each face between the
// cell and its neighbour and mirror the other face about the cell value (:415-430).  LIM is the limiter id
// (include/Typedef.h:188-193) when known at compile time, or 0 to dispatch on `lim` at run time.
template <int LIM, typename R>
__device__ __forceinline__ R plm_slope( R sl, R sr, R lim_rt, int lim, R coeff )
{
   const R Z = (R)0;
   const R sc = (R)0.5*( sl + sr ), lr = sl*sr;
   const int L = ( LIM != 0 ) ? LIM : lim;
   R s;
   if ( L == 6 )      s = sc;                                                   // central
   else if ( L == 1 ) s = (R)2.0*lr*rcp( sl + sr );                             // van Leer
   else if ( L == 3 ) s = lr*( sl + sr )*rcp( sl*sl + sr*sr );                  // van Albada
   else
   {
      R m = vmin( vabs( sc ), vmin( vabs( sl*coeff ), vabs( sr*coeff ) ) );     // generalised MinMod
      if ( L == 4 ) m = vmin( vabs( (R)2.0*sl*sr*rcp( sl + sr ) ), m );          // + van Leer
      s = ( sc < Z ) ? -m : m;
   }
   (void)lim_rt;
   return ( lr > Z ) ? s : Z;
}

// clamp x into [min(a,b), max(a,b)] = median of (x, a, b)
__device__ __forceinline__ float  clamp_between( float x, float a, float b )   { return __builtin_amdgcn_fmed3f( x, a, b ); }
__device__ __forceinline__ double clamp_between( double x, double a, double b ) { return vmin( vmax( x, vmin( a, b ) ), vmax( a, b ) ); }

// van Leer + generalised MinMod slope without branches:
//   sign(sc) * min( theta|sl|, theta|sr|, |sc|, |2 sl sr/(sl+sr)| )  if sl*sr > 0, else 0.
// h = 2 sl sr / |sl+sr| equals the harmonic-mean magnitude when sl*sr > 0 and is <= 0 otherwise, or NaN when
// sl*sr underflows to zero while 1/|sl+sr| overflows (both slopes below about 1e-38 in FP32); max(h, 0) maps
// both cases to 0, which realises the sign test.  |sc| >= |harmonic mean| is implied and dropped.
template <typename R>
__device__ __forceinline__ R gminmod_mag( R h, R x ) { return vmin( x, vmax( h, (R)0 ) ); }
// FP32, x >= 0: med3(h, 0, x) is 0, h or x for h < 0, 0 <= h <= x or h > x, and 0 for a NaN h (with a NaN
// operand v_med3_f32 returns the minimum of the other two)
__device__ __forceinline__ float gminmod_mag( float h, float x ) { return __builtin_amdgcn_fmed3f( h, 0.0f, x ); }

template <typename R>
__device__ __forceinline__ R slope_vl_gminmod( R sl, R sr, R coeff )
{
   const R t = sl + sr;
   const R h = (R)2.0*( sl*sr )*rcp( vabs( t ) );
   return vcopysign( gminmod_mag( h, vmin( vabs( sl )*coeff, vabs( sr )*coeff ) ), t );
}

// Coefficients up to this value leave the PLM face clamps inactive under VL_GMINMOD (plm_face).
constexpr double VLGM_NOCLAMP_COEFF = 1.9999;

// Limiter id to instantiate a kernel with: 4 (VL_GMINMOD, resolved at compile time, face clamps left out) only
// for a MinMod coefficient <= VLGM_NOCLAMP_COEFF; otherwise 0, the run-time dispatch that keeps the clamps.
// The MinMod retries only lower the coefficient.
inline int kernel_limiter( int limiter, double minmod_coeff )
{
   return ( limiter == 4 && minmod_coeff <= VLGM_NOCLAMP_COEFF ) ? 4 : 0;
}

// Faces C -/+ s/2, the left one clamped between C and L, the right one 2C - (left face) clamped between C and R
// (DataReconstruction.cpp:415-430).  With LIM == 4 and plain FP32/FP64 arithmetic the clamps never act, and are
// left out: the slope is at most theta*|C-L|*(1+u)^2 <= 2|C-L| for theta <= 1.9999, so the exact C - s/2 lies
// in [min(C,L), max(C,L)], whose ends are representable, and rounding keeps it there; likewise s/2 <= |R-C|
// puts 2C - (left face) between C and R before its rounding.  Same bits as the clamped form; dual numbers keep
// the clamps for their derivative tie rules.
template <int LIM, typename R>
__device__ __forceinline__ void plm_face( R L, R C, R Rr, int lim, R coeff, R &fL, R &fR )
{
   const R s = ( LIM == 4 ) ? slope_vl_gminmod( C - L, Rr - C, coeff ) : plm_slope<LIM>( C - L, Rr - C, (R)0, lim, coeff );
   R fl, fr;
   if constexpr ( LIM == 4 && std::is_floating_point<R>::value )
   {
      fl = C - (R)0.5*s;
      fr = (R)2.0*C - fl;
   }
   else
   {
      fl = clamp_between( C - (R)0.5*s, C, L );
      fr = clamp_between( (R)2.0*C - fl, C, Rr );
   }
   fL = (R)2.0*C - fr;
   fR = fr;
}

// primitive face state (rho, vx, vy, vz, P) -> interface state rotated to direction D
// (Pri2Con, FluUtility.cpp:343-420; the solver takes the face pressure and velocity directly instead of
// recovering them from the conserved face state)
template <int D, typename R>
__device__ __forceinline__ St<R> face_state( const R w[NV], const Prm<R> &P )
{
   St<R> s;
   const R vn = w[1 + D], v1 = w[1 + ( D + 1 ) % 3], v2 = w[1 + ( D + 2 ) % 3];
   s.d  = w[0];
   s.mn = w[0]*vn;
   s.m1 = w[0]*v1;
   s.m2 = w[0]*v2;
   s.u  = vn;
   s.ir = rcp( w[0] );
   s.p  = floor_pres( w[4], P.min_pres );
   s.e  = fmad( (R)0.5*w[0], fmad( w[3], w[3], fmad( w[2], w[2], w[1]*w[1] ) ), w[4]*P.inv_gm1 );
   s.cs = sound_speed( P.gamma*s.p, s.d, s.ir );
   return s;
}

// face_state<0> or face_state<1> chosen per lane (dy = y-normal face): the same arithmetic, the rotation by select
template <typename R>
__device__ __forceinline__ St<R> face_state_xy( const R w[NV], bool dy, const Prm<R> &P )
{
   St<R> s;
   const R vn = dy ? w[2] : w[1], v1 = dy ? w[3] : w[2], v2 = dy ? w[1] : w[3];
   s.d  = w[0];
   s.mn = w[0]*vn;
   s.m1 = w[0]*v1;
   s.m2 = w[0]*v2;
   s.u  = vn;
   s.ir = rcp( w[0] );
   s.p  = floor_pres( w[4], P.min_pres );
   s.e  = fmad( (R)0.5*w[0], fmad( w[3], w[3], fmad( w[2], w[2], w[1]*w[1] ) ), w[4]*P.inv_gm1 );
   s.cs = sound_speed( P.gamma*s.p, s.d, s.ir );
   return s;
}

// conserved cell (with its pressure, 1/rho and sound speed) -> interface state rotated to direction D
template <int D, typename R>
__device__ __forceinline__ St<R> cell_state( const R c[NV], R p, R ir, R cs )
{
   St<R> s;
   s.d  = c[0];
   s.mn = c[1 + D];
   s.m1 = c[1 + ( D + 1 ) % 3];
   s.m2 = c[1 + ( D + 2 ) % 3];
   s.e  = c[4];
   s.p  = p;  s.ir = ir;  s.cs = cs;
   s.u  = ir*s.mn;
   return s;
}

// cell_state<0> or cell_state<1> chosen per lane (dy = y-normal face) from a cell (conserved, P, 1/rho, c_s)
template <typename R>
__device__ __forceinline__ St<R> cell_state_xy( const R q[8], bool dy )
{
   St<R> s;
   s.d  = q[0];
   s.mn = dy ? q[2] : q[1];
   s.m1 = dy ? q[3] : q[2];
   s.m2 = dy ? q[1] : q[3];
   s.e  = q[4];
   s.p  = q[5];  s.ir = q[6];  s.cs = q[7];
   s.u  = s.ir*s.mn;
   return s;
}

// whether a flux component is NaN (an infinite component is not)
template <typename R>
__device__ __forceinline__ bool any_nan5( const R f[NV] )
{
   return f[0] != f[0] || f[1] != f[1] || f[2] != f[2] || f[3] != f[3] || f[4] != f[4];
}
// float: gfx950's v_maximum3_f32 propagates NaN, so two of them and one unordered compare test all five components
__device__ __forceinline__ bool any_nan5( const float f[NV] )
{
   const float m = __builtin_elementwise_maximum( __builtin_elementwise_maximum( f[0], f[1] ),
                                                  __builtin_elementwise_maximum( f[2], __builtin_elementwise_maximum( f[3], f[4] ) ) );
   return m != m;
}

// Riemann solve across a D-normal face; flux returned in the lab frame.  HLLC falls back to HLLE when a flux
// component is NaN (RSOLVER_RESCUE, CPU_FluidSolver_MHM.cpp:737-786, ComputeFlux.cpp:313-362).
template <int RS, int D, typename R>
__device__ __forceinline__ void riemann( const St<R> &L, const St<R> &Q, const Prm<R> &P, R out[NV] )
{
   R f[NV];
   if ( RS == 4 )
   {
      rs_hllc( L, Q, P.min_pres, f );
      if ( any_nan5( f ) ) rs_hlle( L, Q, f );
   }
   else if ( RS == 3 ) rs_hlle( L, Q, f );
   else                rs_roe( L, Q, P, f );
   out[0] = f[0];
   out[4] = f[4];
   out[1 + D]             = f[1];
   out[1 + ( D + 1 ) % 3] = f[2];
   out[1 + ( D + 2 ) % 3] = f[3];
}

// riemann<RS, 0> or riemann<RS, 1> chosen per lane (dy = y-normal face)
template <int RS, typename R>
__device__ __forceinline__ void riemann_xy( const St<R> &L, const St<R> &Q, const Prm<R> &P, bool dy, R out[NV] )
{
   R f[NV];
   if ( RS == 4 )
   {
      rs_hllc( L, Q, P.min_pres, f );
      if ( any_nan5( f ) ) rs_hlle( L, Q, f );
   }
   else if ( RS == 3 ) rs_hlle( L, Q, f );
   else                rs_roe( L, Q, P, f );
   out[0] = f[0];
   out[4] = f[4];
   out[1] = dy ? f[3] : f[1];
   out[2] = dy ? f[1] : f[2];
   out[3] = dy ? f[2] : f[3];
}

template <typename IO> struct IOLim;
template <> struct IOLim<double> { static constexpr double huge = 1.7976931348623157e308; static constexpr double eps = 2.2204460492503131e-16; };
template <> struct IOLim<float>  { static constexpr float  huge = 3.40282346638528859812e+38f; static constexpr float eps = 1.1920928955078125e-07f; };

// full-step update of one output cell, accumulated in FP64, + the CONS-mode unphysical check on the stored
// values (FullStepUpdate.cpp:106-235, FluUtility.cpp:843-940)
template <typename R, typename IO>
__device__ __forceinline__ bool full_cell( const IO u[NV], const R dfxy[NV], const R fzl[NV], const R fzu[NV],
                                           const Prm<R> &P, IO o[NV] )
{
#  pragma unroll
   for (int v=0; v<NV; v++)
      o[v] = (IO)( (double)u[v] - P.dt_dh*( (double)dfxy[v] + ( (double)fzu[v] - (double)fzl[v] ) ) );
   // each test is one ordered compare, false for NaN: "not finite-and-in-range" covers the NaN test too
   const IO huge = IOLim<IO>::huge, Z = (IO)0;
   bool ok = ( vabs( o[1] ) <= huge ) && ( vabs( o[2] ) <= huge ) && ( vabs( o[3] ) <= huge );
   ok = ok && ( o[0] > Z && o[0] <= huge ) && ( o[4] > Z && o[4] <= huge );
   const IO ei = o[4] - (IO)0.5*( o[1]*o[1] + o[2]*o[2] + o[3]*o[3] )*rcp( o[0] );
   ok = ok && ( ei >= (IO)-3.0*o[4]*IOLim<IO>::eps && ei <= huge );
   return !ok;
}

// Hancock half-step advance of the six face states of one cell (MHM only;
// DataReconstruction.cpp:2063-2160): fluxes of the conserved face states (pressure recovered in FP64),
// U_f -= dt/(2dh) * sum_d (F_{d,R} - F_{d,L}), then either a reset of all six faces to the cell value if
// any face density/energy or input face pressure is invalid, or density and internal-energy floors.
template <typename R, typename IO>
__device__ __forceinline__ void hancock( R fc[6][NV], const R fcp[6][NV], const R ucell[NV], const Prm<R> &P )
{
   R F[6][NV];
#  pragma unroll
   for (int f=0; f<6; f++)
   {
      const int d = f/2;
      const R mn = fc[f][1 + d];
      const double eint = (double)fc[f][4] - 0.5*( (double)fc[f][1]*fc[f][1] + (double)fc[f][2]*fc[f][2]
                                                    + (double)fc[f][3]*fc[f][3] )*rcp_pres<R>( (double)fc[f][0] );
      const R p  = floor_pres( (R)( eint*P.gm1_d ), P.min_pres );
      const R vn = rcp( fc[f][0] )*mn;
      F[f][0] = mn;
      F[f][1] = vn*fc[f][1];
      F[f][2] = vn*fc[f][2];
      F[f][3] = vn*fc[f][3];
      F[f][1 + d] += p;
      F[f][4] = vn*( fc[f][4] + p );
   }
   const R dt_dh2 = (R)P.dt_dh2;
#  pragma unroll
   for (int v=0; v<NV; v++)
   {
      const R dF = dt_dh2*( F[1][v] - F[0][v] + F[3][v] - F[2][v] + F[5][v] - F[4][v] );
#     pragma unroll
      for (int f=0; f<6; f++) fc[f][v] -= dF;
   }
   const R tiny = ( sizeof(R) == 8 ) ? (R)2.2250738585072014e-308 : (R)1.17549435e-38f;
   const R huge = (R)IOLim<IO>::huge;
   bool reset = false;
#  pragma unroll
   for (int f=0; f<6; f++)
   {
      const R d = fc[f][0], e = fc[f][4], pr = fcp[f][4];
      reset |= ( d < tiny || d > huge || d != d ) || ( e < tiny || e > huge || e != e ) ||
               ( pr < (R)0 || pr > huge || pr != pr );
   }
   if ( reset )
   {
#     pragma unroll
      for (int f=0; f<6; f++)
#        pragma unroll
         for (int v=0; v<NV; v++) fc[f][v] = ucell[v];
   }
   else
   {
#     pragma unroll
      for (int f=0; f<6; f++)
      {
         fc[f][0] = vmax( fc[f][0], P.min_dens );
         const double ke  = 0.5*( (double)fc[f][1]*fc[f][1] + (double)fc[f][2]*fc[f][2] + (double)fc[f][3]*fc[f][3] )
                            *rcp_pres<R>( (double)fc[f][0] );
         const double ein = (double)fc[f][4] - ke;
         const double eo  = ( ein == ein ) ? __builtin_fmax( ein, P.min_eint_d ) : ein;
         if ( ein != eo ) fc[f][4] = (R)( (double)fc[f][4] - ein + eo );
      }
   }
}

// primitive -> conserved (Pri2Con, FluUtility.cpp:343-420)
template <typename R>
__device__ __forceinline__ void pri2con( const R w[NV], R c[NV], const Prm<R> &P )
{
   c[0] = w[0];
   c[1] = w[0]*w[1];  c[2] = w[0]*w[2];  c[3] = w[0]*w[3];
   c[4] = (R)0.5*( c[1]*c[1] + c[2]*c[2] + c[3]*c[3] )*rcp( w[0] ) + w[4]*P.inv_gm1;
}

// conserved face state -> interface state rotated to direction D, pressure recovered in FP64
template <int D, typename R>
__device__ __forceinline__ St<R> cons_state( const R c[NV], const Prm<R> &P )
{
   St<R> s;
   const double eint = (double)c[4] - 0.5*( (double)c[1]*c[1] + (double)c[2]*c[2] + (double)c[3]*c[3] )
                                      *rcp_pres<R>( (double)c[0] );
   s.d  = c[0];
   s.mn = c[1 + D];
   s.m1 = c[1 + ( D + 1 ) % 3];
   s.m2 = c[1 + ( D + 2 ) % 3];
   s.e  = c[4];
   s.p  = floor_pres( (R)( eint*P.gm1_d ), P.min_pres );
   s.ir = rcp( c[0] );
   s.u  = s.ir*s.mn;
   s.cs = sound_speed( P.gamma*s.p, s.d, s.ir );
   return s;
}

// max_d(|v_d| + c_s) of one cell, pressure recovered in FP64 (CPU_dtSolver_HydroCFL.cpp:104-190)
template <typename IO>
__device__ __forceinline__ IO cell_cfl( const IO u[NV], double gamma, double gm1, double min_pres )
{
   const double d = (double)u[0], ir = 1.0/d;
   double p = ( (double)u[4] - 0.5*( (double)u[1]*u[1] + (double)u[2]*u[2] + (double)u[3]*u[3] )*ir )*gm1;
   p = ( p == p ) ? __builtin_fmax( p, min_pres ) : p;
   const IO cs = (IO)__builtin_sqrt( gamma*p*ir );
   const IO ri = (IO)ir;
   const IO vx = vabs( u[1] )*ri, vy = vabs( u[2] )*ri, vz = vabs( u[3] )*ri;
   return vmax( vmax( cs + vx, cs + vy ), cs + vz );
}

// Fused CFL of one patch group: each lane holds the maxima of its output column in the two z-halves; the 8
// patches (px + 2 py + 4 pz, x fastest) are reduced with wave64 butterflies and LDS across the 4 waves, in the
// I/O type (the reference computes dt in `real`, CPU_dtSolver_HydroCFL.cpp:62-209).
// Called by every lane of the workgroup (waves 0-3 hold the output columns); writes dt[8] = dh_safety / max.
template <typename IO>
__device__ __forceinline__ void fused_dt( const IO cfl[2], int pxy, IO (*red)[4], IO *dt, double dh_safety )
{
   const int tid = threadIdx.x, w = tid >> 6;
   if ( w < 4 )                      // the output columns live on waves 0-3; further waves hold ghost columns
#  pragma unroll
   for (int pz=0; pz<2; pz++)
#     pragma unroll
      for (int q=0; q<4; q++)
      {
         IO m = ( pxy == q ) ? cfl[pz] : (IO)0;
#        pragma unroll
         for (int o = 32; o > 0; o >>= 1) m = vmax( m, __shfl_xor( m, o, 64 ) );
         if ( ( tid & 63 ) == 0 ) red[q + 4*pz][w] = m;
      }
   __syncthreads();
   if ( tid < 8 )
   {
      const IO m = vmax( vmax( red[tid][0], red[tid][1] ), vmax( red[tid][2], red[tid][3] ) );
      dt[tid] = (IO)( dh_safety / (double)m );
   }
}

} // namespace gamer_amd
