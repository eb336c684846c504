// ORACLE — test infrastructure only, never the product path.
//
// A from-scratch CPU restatement of GAMER's MUSCL-Hancock (MHM / MHM_RP) patch-group fluid update,
// non-MHD, no passive scalars, EOS_GAMMA, PLM reconstruction.  It exists to check the HIP solver:
// only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.
//
// Every floating-point expression keeps the operand order of the reference so the result can be
// compared bit-for-bit with the reference's own CPU build (oracle/_ref, made by oracle/Makefile):
//   driver + half step         src/Model_Hydro/CPU_Hydro/CPU_FluidSolver_MHM.cpp:352-589, 615-949
//   PLM + limiter + Hancock    src/Model_Hydro/CPU_Hydro/CPU_Shared_DataReconstruction.cpp:188-689, 1917-2032, 2063-2160
//   full-step fluxes, store    src/Model_Hydro/CPU_Hydro/CPU_Shared_ComputeFlux.cpp:112-450
//   full-step update + check   src/Model_Hydro/CPU_Hydro/CPU_Shared_FullStepUpdate.cpp:63-271
//   HLLC / HLLE / Roe          src/Model_Hydro/CPU_Hydro/CPU_Shared_RiemannSolver_{HLLC:391-655, HLLE:301-698, Roe:87-711}.cpp
//   conversions and checks     src/Model_Hydro/CPU_Hydro/CPU_Shared_FluUtility.cpp:93-551, 681-1330
//   EOS_GAMMA                  src/EoS/Gamma/CPU_EoS_Gamma.cpp:54-192
//   CFL time step              src/Model_Hydro/CPU_Hydro/CPU_dtSolver_HydroCFL.cpp:57-209
//   closure and AMR sync       src/Fluid/Flu_Close.cpp:83-1297 (CorrectUnphysical, StoreFlux, CorrectFlux, the
//                              patch copy), src/Fluid/Flu_FixUp_Flux.cpp, src/Fluid/Flu_FixUp_Restrict.cpp
//   interpolation              src/Interpolation/Interpolate.cpp, Int_*.cpp
//
// Pinning: tests/test_oracle_ref.py (bit-for-bit against oracle/_ref on seeded inputs: the solver, the CFL
// time step, Interpolate, and the closure / sync functions, whose FP32 flux registers this restatement keeps in
// FP64 and which therefore agree in FP32 to the reference's float rounding),
// tests/test_oracle_golden.py (committed fixtures made from the reference build),
// tests/test_oracle_props.py (uniform state, Riemann consistency, upwinding, limiter closed forms,
// conservation telescoping, axis-permutation symmetry), tests/test_oracle_sync.py and test_oracle_interp.py
// (closed forms and invariants of the closure, sync and interpolation functions).  Every function has a pin.

#include "hydro_oracle.h"
#include <cmath>
#include <cfloat>
#include <cstring>
#include <vector>
#include <omp.h>

namespace {

constexpr int NV  = 5;      // conserved / primitive fields: rho, mx, my, mz, E
constexpr int PS  = 16;     // patch-group interior width
constexpr int PSH = 8;      // patch width

template <typename T> struct Lim;
template <> struct Lim<double> { static double tiny() { return DBL_MIN; } static double huge() { return DBL_MAX; } static double eps() { return DBL_EPSILON; } };
template <> struct Lim<float>  { static float  tiny() { return FLT_MIN; } static float  huge() { return FLT_MAX; } static float  eps() { return FLT_EPSILON; } };

template <typename T> inline T vmax( T a, T b ) { return std::fmax( a, b ); }
template <typename T> inline T vmin( T a, T b ) { return std::fmin( a, b ); }
template <typename T> inline T sgn ( T a )      { return ( a < (T)0 ) ? (T)-1 : (T)+1; }

// EOS_GAMMA auxiliary constants: computed in double, then narrowed (CPU_EoS_Gamma.cpp:54-66)
template <typename T>
struct Gas {
   T gamma, gm1, inv_gm1, inv_gamma;
   explicit Gas( double g ) : gamma( (T)g ), gm1( (T)(g - 1.0) ), inv_gm1( (T)(1.0/(g - 1.0)) ), inv_gamma( (T)(1.0/g) ) {}
};

template <typename T>
struct Opt {
   Gas<T> gas;
   T dt, dh, min_dens, min_pres, min_eint;
   int scheme, rsolver, limiter;
   Opt( double g ) : gas( g ) {}
};

// ---------------- cell conversions ----------------
// cyclic permutation of the momentum so that component 1 is normal to the interface (FluUtility.cpp:93-176)
template <typename T> inline void rotate( T u[], int d, bool fwd )
{
   if ( d == 0 ) return;
   const T a = u[1], b = u[2], c = u[3];
   if ( ( d == 1 ) == fwd ) { u[1] = b; u[2] = c; u[3] = a; }   // (x,y,z) -> (y,z,x)
   else                     { u[1] = c; u[2] = a; u[3] = b; }   // (x,y,z) -> (z,x,y)
}

template <typename T> inline T floor_pres( T p, T pmin ) { return ( p == p ) ? vmax( p, pmin ) : p; }

template <typename T> inline T eint_of( T d, T mx, T my, T mz, T e )
{  return e - (T)0.5*( mx*mx + my*my + mz*mz ) / d; }

template <typename T> inline T pres_of( const Gas<T> &g, T d, T mx, T my, T mz, T e, bool floor, T pmin )
{
   T p = eint_of( d, mx, my, mz, e ) * g.gm1;
   return floor ? floor_pres( p, pmin ) : p;
}

template <typename T> inline T etot_of( T d, T mx, T my, T mz, T eint )
{
   T e = (T)0.5*( mx*mx + my*my + mz*mz ) / d;
   return e + eint;
}

// leave E untouched unless the internal-energy floor actually changes Eint (FluUtility.cpp:787-805)
template <typename T> inline T engy_with_eint_floor( T d, T mx, T my, T mz, T e, T emin )
{
   const T ein  = eint_of( d, mx, my, mz, e );
   const T eout = ( ein == ein ) ? vmax( ein, emin ) : ein;
   return ( ein == eout ) ? e : e - ein + eout;
}

template <typename T> inline void con2pri( const Gas<T> &g, const T c[NV], T w[NV], T pmin )
{
   const T r = (T)1.0 / c[0];
   w[0] = c[0];
   w[1] = c[1]*r;  w[2] = c[2]*r;  w[3] = c[3]*r;
   w[4] = pres_of( g, c[0], c[1], c[2], c[3], c[4], true, pmin );
}

template <typename T> inline void pri2con( const Gas<T> &g, const T w[NV], T c[NV] )
{
   c[0] = w[0];
   c[1] = w[0]*w[1];  c[2] = w[0]*w[2];  c[3] = w[0]*w[3];
   c[4] = etot_of( c[0], c[1], c[2], c[3], w[4]*g.inv_gm1 );
}

// physical flux along x of a state already rotated to x (FluUtility.cpp:458-551)
template <typename T> inline void xflux( const T u[NV], T p, T f[NV] )
{
   const T r  = (T)1.0 / u[0];
   const T vx = r*u[1];
   f[0] = u[1];
   f[1] = vx*u[1] + p;
   f[2] = vx*u[2];
   f[3] = vx*u[3];
   f[4] = vx*( u[4] + p );
}

// flux along direction d of an unrotated state; pressure recovered from the rotated state
template <typename T> inline void dir_flux( const Gas<T> &g, int d, const T in[NV], T pmin, T f[NV] )
{
   T u[NV];
   for (int v=0; v<NV; v++) u[v] = in[v];
   rotate( u, d, true );
   const T p = pres_of( g, u[0], u[1], u[2], u[3], u[4], true, pmin );
   xflux( u, p, f );
   rotate( f, d, false );
}

// ---------------- Riemann solvers (hydro, Davis wave speeds) ----------------
template <typename T> void rs_hlle_x( const Gas<T> &g, const T L[NV], const T R[NV], T pmin, T out[NV] )
{
   const T Z = (T)0.0, ONE = (T)1.0;
   const T rl = ONE / L[0], rr = ONE / R[0];
   const T ul = rl*L[1], ur = rr*R[1];
   const T pl = pres_of( g, L[0], L[1], L[2], L[3], L[4], true, pmin );
   const T pr = pres_of( g, R[0], R[1], R[2], R[3], R[4], true, pmin );
   const T cl = std::sqrt( g.gamma*pl / L[0] );
   const T cr = std::sqrt( g.gamma*pr / R[0] );
   T sl = vmin( ul - cl, ur - cr );
   T sr = vmax( ul + cl, ur + cr );
   sl = vmin( sl, Z );
   sr = vmax( sr, Z );
   T fl[NV], fr[NV];
   xflux( L, pl, fl );
   xflux( R, pr, fr );
   for (int v=0; v<NV; v++) { fl[v] -= sl*L[v];  fr[v] -= sr*R[v]; }
   if ( sl == Z && sr == Z )
      for (int v=0; v<NV; v++) out[v] = fl[v];
   else
   {
      const T inv = ONE / ( sr - sl );
      for (int v=0; v<NV; v++) out[v] = inv*( sr*fl[v] - sl*fr[v] );
   }
}

template <typename T> void rs_hllc_x( const Gas<T> &g, const T L[NV], const T R[NV], T pmin, T out[NV] )
{
   const T Z = (T)0.0, ONE = (T)1.0;
   const T rl = ONE / L[0], rr = ONE / R[0];
   const T ul = rl*L[1], ur = rr*R[1];
   const T pl = pres_of( g, L[0], L[1], L[2], L[3], L[4], true, pmin );
   const T pr = pres_of( g, R[0], R[1], R[2], R[3], R[4], true, pmin );
   const T cl = std::sqrt( g.gamma*pl / L[0] );
   const T cr = std::sqrt( g.gamma*pr / R[0] );
   const T wl1 = ul - cl, wl2 = ur - cr, wr1 = ul + cl, wr2 = ur + cr;
   const T wl = vmin( wl1, wl2 ), wr = vmax( wr1, wr2 );
   const T tl = +L[0]*( ( wl1 < wl2 ) ? cl : ( ul - ur ) + cr );
   const T tr = -R[0]*( ( wr2 > wr1 ) ? cr : ( ul - ur ) + cl );
   const T t2 = ONE / ( tl - tr );
   const T vs = t2*( pl - pr + tl*ul - tr*ur );
   T ps = t2*( tl*( pr + tr*ur ) - tr*( pl + tl*ul ) );
   ps = floor_pres( ps, pmin );

   T f[NV], cf, cs;
   if ( vs >= Z )
   {
      const T s = vmin( wl, Z );
      xflux( L, pl, f );
      for (int v=0; v<NV; v++) f[v] -= s*L[v];
      if ( vs == Z && s == Z ) { cf = ONE; cs = Z; }
      else { const T t4 = ONE / ( vs - s );  cf = t4*vs;  cs = -t4*s*ps; }
   }
   else
   {
      const T s = vmax( wr, Z );
      xflux( R, pr, f );
      for (int v=0; v<NV; v++) f[v] -= s*R[v];
      const T t4 = ONE / ( vs - s );
      cf = t4*vs;  cs = -t4*s*ps;
   }
   for (int v=0; v<NV; v++) out[v] = cf*f[v];
   out[1] += cs;
   out[4] += cs*vs;
}

// check_intermediate: GAMER compiles the intermediate-state check into Hydro_RiemannSolver_Roe only in builds whose
// RSOLVER (or RSOLVER_RESCUE) is ROE (include/CUFLU.h:178-186, CHECK_INTERMEDIATE); in other builds the Roe solve
// of CorrectUnphysical's first-order correction runs without it
template <typename T> void rs_roe_x( const Gas<T> &g, const T L[NV], const T R[NV], T pmin, T out[NV],
                                     bool check_intermediate = true )
{
   const T Z = (T)0.0, ONE = (T)1.0, HALF = (T)0.5;
   const T rl = ONE / L[0], rr = ONE / R[0];
   const T pl = pres_of( g, L[0], L[1], L[2], L[3], L[4], true, pmin );
   const T pr = pres_of( g, R[0], R[1], R[2], R[3], R[4], true, pmin );
   const T hl = rl*( L[4] + pl ), hr = rr*( R[4] + pr );
   const T sql = std::sqrt( L[0] ), sqr = std::sqrt( R[0] );
   const T rho = sql*sqr, irho = ONE/rho;
   const T isql = ONE/sql, isqr = ONE/sqr, isum = ONE/( sql + sqr );
   const T u = isum*( isql*L[1] + isqr*R[1] );
   const T v = isum*( isql*L[2] + isqr*R[2] );
   const T w = isum*( isql*L[3] + isqr*R[3] );
   const T V2 = u*u + v*v + w*w;
   const T H  = isum*( sql*hl + sqr*hr );
   T gpr = g.gm1*( H - HALF*V2 );
   gpr = g.gamma*irho*floor_pres( gpr*rho/g.gamma, pmin );
   const T a2 = gpr, a = std::sqrt( a2 );
   const T ev[NV] = { u - a, u, u, u, u + a };

   T fl[NV], fr[NV];
   xflux( L, pl, fl );
   xflux( R, pr, fr );
   if ( ev[0]    >= Z ) { for (int k=0; k<NV; k++) out[k] = fl[k];  return; }
   if ( ev[NV-1] <= Z ) { for (int k=0; k<NV; k++) out[k] = fr[k];  return; }

   const T rev[NV][NV] = { { ONE,     ONE,       Z,   Z,     ONE   },
                           { u - a,   u,         Z,   Z,     u + a },
                           { v,       v,         ONE, Z,     v     },
                           { w,       w,         Z,   ONE,   w     },
                           { H - u*a, HALF*V2,   v,   w,     H + u*a } };
   T jump[NV], amp[NV];
   for (int k=0; k<NV; k++) jump[k] = R[k] - L[k];
   amp[2] = jump[2] - v*jump[0];
   amp[3] = jump[3] - w*jump[0];
   amp[1] = g.gm1/a2*( jump[0]*( H - u*u ) + u*jump[1] - jump[4] + v*amp[2] + w*amp[3] );
   amp[0] = HALF/a*( jump[0]*( u + a ) - jump[1] - a*amp[1] );
   amp[4] = jump[0] - amp[0] - amp[1];

   // intermediate-state positivity check, falling back to HLLE (Roe.cpp:607-667; CUFLU.h CHECK_INTERMEDIATE)
   T is[NV];
   for (int k=0; k<NV; k++) is[k] = L[k];
   for (int t=0; t<NV-1 && check_intermediate; t++)
   {
      for (int k=0; k<NV; k++) is[k] += amp[t]*rev[k][t];
      if ( ev[t+1] > ev[t] )
      {
         const T ip = pres_of( g, is[0], is[1], is[2], is[3], is[4], false, (T)0 );
         if ( is[0] <= Z || ip <= Z ) { rs_hlle_x( g, L, R, pmin, out );  return; }
      }
   }

   for (int k=0; k<NV; k++) amp[k] *= std::fabs( ev[k] );
   for (int k=0; k<NV; k++)
   {
      out[k] = fl[k] + fr[k];
      for (int t=0; t<NV; t++) out[k] -= amp[t]*rev[k][t];
      out[k] *= HALF;
   }
}

// Riemann solve across a d-normal interface; HLLC falls back to HLLE on a NaN flux component
// (RSOLVER_RESCUE, CPU_FluidSolver_MHM.cpp:737-786, CPU_Shared_ComputeFlux.cpp:313-362)
template <typename T> void riemann( const Gas<T> &g, int rsolver, int d, const T Lin[NV], const T Rin[NV], T pmin, T out[NV] )
{
   T L[NV], R[NV];
   for (int v=0; v<NV; v++) { L[v] = Lin[v];  R[v] = Rin[v]; }
   rotate( L, d, true );
   rotate( R, d, true );
   switch ( rsolver )
   {
      case ORC_HLLC: rs_hllc_x( g, L, R, pmin, out ); break;
      case ORC_HLLE: rs_hlle_x( g, L, R, pmin, out ); break;
      default:       rs_roe_x ( g, L, R, pmin, out ); break;
   }
   rotate( out, d, false );

   if ( rsolver == ORC_HLLC )
      for (int v=0; v<NV; v++)
         if ( out[v] != out[v] )
         {
            for (int k=0; k<NV; k++) { L[k] = Lin[k];  R[k] = Rin[k]; }
            rotate( L, d, true );
            rotate( R, d, true );
            rs_hlle_x( g, L, R, pmin, out );
            rotate( out, d, false );
            break;
         }
}

// ---------------- PLM slope limiter (DataReconstruction.cpp:1917-2032) ----------------
template <typename T> void limit_slope( const T L[NV], const T C[NV], const T R[NV], int lim, T coeff, T s[NV] )
{
   T sl[NV], sr[NV], sc[NV], sa[NV];
   for (int v=0; v<NV; v++)
   {
      sl[v] = C[v] - L[v];
      sr[v] = R[v] - C[v];
      sc[v] = (T)0.5*( sl[v] + sr[v] );
   }
   if ( lim == 4 )
      for (int v=0; v<NV; v++)
         sa[v] = ( sl[v]*sr[v] > (T)0.0 ) ? (T)2.0*sl[v]*sr[v]/( sl[v] + sr[v] ) : (T)0.0;

   for (int v=0; v<NV; v++)
   {
      const T lr = sl[v]*sr[v];
      if ( !( lr > (T)0.0 ) ) { s[v] = (T)0.0;  continue; }
      switch ( lim )
      {
         case 6: s[v] = sc[v];  break;                                                 // central
         case 1: s[v] = (T)2.0*lr/( sl[v] + sr[v] );  break;                           // van Leer
         case 2: {                                                                     // generalised MinMod
            const T a = sl[v]*coeff, b = sr[v]*coeff;
            T m = vmin( std::fabs(a), std::fabs(b) );
            m = vmin( std::fabs(sc[v]), m );
            s[v] = m*sgn( sc[v] );
         } break;
         case 3: s[v] = lr*( sl[v] + sr[v] ) / ( sl[v]*sl[v] + sr[v]*sr[v] );  break;  // van Albada
         default: {                                                                    // van Leer + generalised MinMod
            const T a = sl[v]*coeff, b = sr[v]*coeff;
            T m = vmin( std::fabs(a), std::fabs(b) );
            m = vmin( std::fabs(sc[v]), m );
            m = vmin( std::fabs(sa[v]), m );
            s[v] = m*sgn( sc[v] );
         } break;
      }
   }
}

// ---------------- one patch group ----------------
template <typename T>
struct Scratch {
   std::vector<T> pri;     // [5][N^3]  primitive (MHM) / half-step primitive (MHM_RP, stride 20)
   std::vector<T> hflux;   // [3][5][21^3] half-step fluxes (MHM_RP)
   std::vector<T> fc;      // [6][5][18^3] face-centred conserved states
   std::vector<T> flux;    // [3][5][17^3] full-step fluxes
   Scratch() : pri( NV*22*22*22 ), hflux( 3*NV*21*21*21 ), fc( 6*NV*18*18*18 ), flux( 3*NV*17*17*17 ) {}
};

inline int idx3( int i, int j, int k, int ni, int nj ) { return ( k*nj + j )*ni + i; }

// MHM_RP phase 1a: first-order Riemann fluxes between input cells (CPU_FluidSolver_MHM.cpp:615-798)
template <typename T> void half_fluxes( const Opt<T> &o, const T *u, T *hf )
{
   const int NX = 22, NF = 21, N3F = NF*NF*NF, N3X = NX*NX*NX;
   const int dcv[3] = { 1, NX, NX*NX };
   for (int d=0; d<3; d++)
   {
      const int s[3]  = { d == 0 ? 0 : 1, d == 1 ? 0 : 1, d == 2 ? 0 : 1 };
      const int n[3]  = { d == 0 ? NF : NF-1, d == 1 ? NF : NF-1, d == 2 ? NF : NF-1 };
      for (int k=0; k<n[2]; k++)
      for (int j=0; j<n[1]; j++)
      for (int i=0; i<n[0]; i++)
      {
         const int ic = idx3( i+s[0], j+s[1], k+s[2], NX, NX );
         T L[NV], R[NV], F[NV];
         for (int v=0; v<NV; v++) { L[v] = u[v*N3X + ic];  R[v] = u[v*N3X + ic + dcv[d]]; }
         riemann( o.gas, o.rsolver, d, L, R, o.min_pres, F );
         const int jf = idx3( i, j, k, NF, NF );
         for (int v=0; v<NV; v++) hf[ ( d*NV + v )*N3F + jf ] = F[v];
      }
   }
}

// MHM_RP phase 1b: half-step update, floors, conversion to primitive (CPU_FluidSolver_MHM.cpp:830-949)
template <typename T> void half_update( const Opt<T> &o, const T *u, const T *hf, T *pri )
{
   const int NX = 22, NF = 21, NH = 20, N3F = NF*NF*NF, N3X = NX*NX*NX, N3H = NH*NH*NH;
   const int dfl[3] = { 1, NF, NF*NF };
   const T dt_dh2 = (T)0.5*o.dt/o.dh;
   for (int k=0; k<NH; k++)
   for (int j=0; j<NH; j++)
   for (int i=0; i<NH; i++)
   {
      const int jf = idx3( i, j, k, NF, NF );
      const int ic = idx3( i+1, j+1, k+1, NX, NX );
      T c[NV], w[NV], df[3][NV];
      for (int d=0; d<3; d++)
      for (int v=0; v<NV; v++)
         df[d][v] = hf[ ( d*NV + v )*N3F + jf + dfl[d] ] - hf[ ( d*NV + v )*N3F + jf ];
      for (int v=0; v<NV; v++)
         c[v] = u[v*N3X + ic] - dt_dh2*( df[0][v] + df[1][v] + df[2][v] );
      c[0] = vmax( c[0], o.min_dens );
      c[4] = engy_with_eint_floor( c[0], c[1], c[2], c[3], c[4], o.min_eint );
      con2pri( o.gas, c, w, o.min_pres );
      const int jo = idx3( i, j, k, NH, NH );
      for (int v=0; v<NV; v++) pri[v*N3H + jo] = w[v];
   }
}

// MHM only: Hancock half-step advance of the six face states of one cell (DataReconstruction.cpp:2063-2160)
template <typename T> void hancock( const Opt<T> &o, T fcc[6][NV], const T fcp[6][NV], const T *u, int ic, int N3X )
{
   const T dt_dh2 = (T)0.5*o.dt/o.dh;
   T F[6][NV];
   for (int f=0; f<6; f++) dir_flux( o.gas, f/2, fcc[f], o.min_pres, F[f] );
   for (int v=0; v<NV; v++)
   {
      const T dF = dt_dh2*( F[1][v] - F[0][v] + F[3][v] - F[2][v] + F[5][v] - F[4][v] );
      for (int f=0; f<6; f++) fcc[f][v] -= dF;
   }
   const T tiny = Lim<T>::tiny(), huge = Lim<T>::huge();
   auto bad = [&]( T x, T lo ) { return ( x < lo || x > huge || x != x ); };
   for (int f=0; f<6; f++)
   {
      if ( bad( fcc[f][0], tiny ) || bad( fcc[f][4], tiny ) || bad( fcp[f][4], (T)0.0 ) )
      {
         for (int g=0; g<6; g++)
         for (int v=0; v<NV; v++) fcc[g][v] = u[v*N3X + ic];
         break;
      }
      fcc[f][0] = vmax( fcc[f][0], o.min_dens );
      fcc[f][4] = engy_with_eint_floor( fcc[f][0], fcc[f][1], fcc[f][2], fcc[f][3], fcc[f][4], o.min_eint );
   }
}

// PLM reconstruction of the 18^3 face-centred states (DataReconstruction.cpp:188-689)
template <typename T> void reconstruct( const Opt<T> &o, const T *pri, int NIn, const T *u_mhm, T coeff, T *fc )
{
   const int NFC = 18, N3FC = NFC*NFC*NFC, N3In = NIn*NIn*NIn;
   const int dcc[3] = { 1, NIn, NIn*NIn };
   for (int k=0; k<NFC; k++)
   for (int j=0; j<NFC; j++)
   for (int i=0; i<NFC; i++)
   {
      const int ic = idx3( i+1, j+1, k+1, NIn, NIn );
      T C[NV], fcp[6][NV], fcc[6][NV];
      for (int v=0; v<NV; v++) C[v] = pri[v*N3In + ic];
      for (int d=0; d<3; d++)
      {
         T Lc[NV], Rc[NV], s[NV];
         for (int v=0; v<NV; v++) { Lc[v] = pri[v*N3In + ic - dcc[d]];  Rc[v] = pri[v*N3In + ic + dcc[d]]; }
         limit_slope( Lc, C, Rc, o.limiter, coeff, s );
         T *fl = fcp[2*d], *fr = fcp[2*d+1];
         for (int v=0; v<NV; v++)
         {
            fl[v] = C[v] - (T)0.5*s[v];
            fr[v] = C[v] + (T)0.5*s[v];
         }
         for (int v=0; v<NV; v++)
         {
            T lo = ( C[v] < Lc[v] ) ? C[v] : Lc[v];
            T hi = ( C[v] > Lc[v] ) ? C[v] : Lc[v];
            fl[v] = ( fl[v] > lo ) ? fl[v] : lo;
            fl[v] = ( fl[v] < hi ) ? fl[v] : hi;
            fr[v] = (T)2.0*C[v] - fl[v];
            lo = ( C[v] < Rc[v] ) ? C[v] : Rc[v];
            hi = ( C[v] > Rc[v] ) ? C[v] : Rc[v];
            fr[v] = ( fr[v] > lo ) ? fr[v] : lo;
            fr[v] = ( fr[v] < hi ) ? fr[v] : hi;
            fl[v] = (T)2.0*C[v] - fr[v];
         }
         pri2con( o.gas, fl, fcc[2*d] );
         pri2con( o.gas, fr, fcc[2*d+1] );
      }
      if ( o.scheme == ORC_MHM ) hancock( o, fcc, fcp, u_mhm, ic, N3In );
      const int jf = idx3( i, j, k, NFC, NFC );
      for (int f=0; f<6; f++)
      for (int v=0; v<NV; v++) fc[ ( f*NV + v )*N3FC + jf ] = fcc[f][v];
   }
}

// full-step fluxes between reconstructed faces, NSkip_N=0, NSkip_T=1 (ComputeFlux.cpp:112-376)
template <typename T> void full_fluxes( const Opt<T> &o, const T *fc, T *fl )
{
   const int NFC = 18, NFL = 17, N3FC = NFC*NFC*NFC, N3FL = NFL*NFL*NFL;
   const int dfc[3] = { 1, NFC, NFC*NFC };
   for (int d=0; d<3; d++)
   {
      const int s[3] = { d == 0 ? 0 : 1, d == 1 ? 0 : 1, d == 2 ? 0 : 1 };
      const int n[3] = { d == 0 ? 17 : 16, d == 1 ? 17 : 16, d == 2 ? 17 : 16 };
      for (int k=0; k<n[2]; k++)
      for (int j=0; j<n[1]; j++)
      for (int i=0; i<n[0]; i++)
      {
         const int jc = idx3( i+s[0], j+s[1], k+s[2], NFC, NFC );
         T L[NV], R[NV], F[NV];
         for (int v=0; v<NV; v++)
         {
            L[v] = fc[ ( (2*d+1)*NV + v )*N3FC + jc ];
            R[v] = fc[ ( (2*d  )*NV + v )*N3FC + jc + dfc[d] ];
         }
         riemann( o.gas, o.rsolver, d, L, R, o.min_pres, F );
         const int jf = idx3( i, j, k, NFL, NFL );
         for (int v=0; v<NV; v++) fl[ ( d*NV + v )*N3FL + jf ] = F[v];
      }
   }
}

// coarse-fine interface fluxes: x/y/z planes at 0, 8, 16 (ComputeFlux.cpp:390-450)
template <typename T> void store_int_flux( const T *fl, T *out )
{
   const int NFL = 17, N3FL = NFL*NFL*NFL;
   for (int t=0; t<PS*PS; t++)
   {
      const int a = t % PS, b = t / PS;
      for (int v=0; v<NV; v++)
      for (int p=0; p<3; p++)
      {
         const int pos = p*PSH;
         out[ ( (0+p)*NV + v )*PS*PS + t ] = fl[ ( 0*NV + v )*N3FL + idx3( pos, a, b, NFL, NFL ) ];
         out[ ( (3+p)*NV + v )*PS*PS + t ] = fl[ ( 1*NV + v )*N3FL + idx3( a, pos, b, NFL, NFL ) ];
         out[ ( (6+p)*NV + v )*PS*PS + t ] = fl[ ( 2*NV + v )*N3FL + idx3( a, b, pos, NFL, NFL ) ];
      }
   }
}

// CONS-mode unphysical check, EOS_GAMMA (FluUtility.cpp:843-940)
template <typename T> bool unphysical( const T c[NV] )
{
   const T huge = Lim<T>::huge();
   bool bad = false;
   for (int v=0; v<NV; v++)
   {
      if ( c[v] != c[v] ) bad = true;
      if ( v >= 1 && v <= 3 ) { if ( c[v] < -huge || c[v] > huge ) bad = true; }
      else                    { if ( c[v] <= (T)0.0 || c[v] > huge ) bad = true; }
   }
   const T e = eint_of( c[0], c[1], c[2], c[3], c[4] );
   if ( e < (T)-3.0*c[4]*Lim<T>::eps() || e > huge || e != e ) bad = true;
   return bad;
}

// full-step update of the 16^3 interior; returns true if any cell is unphysical (FullStepUpdate.cpp:63-271)
template <typename T> bool full_update( const Opt<T> &o, const T *u, int NX, int ghost, const T *fl, T *out )
{
   const int NFL = 17, N3FL = NFL*NFL*NFL, N3X = NX*NX*NX;
   const int dfl[3] = { 1, NFL, NFL*NFL };
   const T dt_dh = o.dt/o.dh;
   bool fail = false;
   for (int k=0; k<PS; k++)
   for (int j=0; j<PS; j++)
   for (int i=0; i<PS; i++)
   {
      const int jf = idx3( i, j, k, NFL, NFL );
      const int ic = idx3( i+ghost, j+ghost, k+ghost, NX, NX );
      T df[3][NV], c[NV];
      for (int d=0; d<3; d++)
      for (int v=0; v<NV; v++)
         df[d][v] = fl[ ( d*NV + v )*N3FL + jf + dfl[d] ] - fl[ ( d*NV + v )*N3FL + jf ];
      for (int v=0; v<NV; v++)
         c[v] = u[v*N3X + ic] - dt_dh*( df[0][v] + df[1][v] + df[2][v] );
      const int jo = idx3( i, j, k, PS, PS );
      for (int v=0; v<NV; v++) out[v*PS*PS*PS + jo] = c[v];
      if ( unphysical( c ) ) fail = true;
   }
   return fail;
}

// the per-PG driver, including the MinMod-coefficient retry loop (CPU_FluidSolver_MHM.cpp:365-587)
template <typename T> void solve_pg( const Opt<T> &o, const T *u, T *out, T *iflux, bool store_flux,
                                     T minmod_coeff, int maxiter, Scratch<T> &s )
{
   const bool rp = ( o.scheme == ORC_MHM_RP );
   const int  NX = rp ? 22 : 20;
   const int  ghost = rp ? 3 : 2;

   if ( rp )
   {
      half_fluxes( o, u, s.hflux.data() );
      half_update( o, u, s.hflux.data(), s.pri.data() );
   }
   else
   {
      const int N3X = NX*NX*NX;
      for (int c=0; c<N3X; c++)
      {
         T cc[NV], w[NV];
         for (int v=0; v<NV; v++) cc[v] = u[v*N3X + c];
         con2pri( o.gas, cc, w, o.min_pres );
         for (int v=0; v<NV; v++) s.pri[v*N3X + c] = w[v];
      }
   }

   int it = 0;
   bool fail;
   do {
      T coeff = ( maxiter == 0 ) ? minmod_coeff : minmod_coeff - (T)it*minmod_coeff/(T)maxiter;
      coeff = vmax( coeff, (T)0.0 );
      reconstruct( o, s.pri.data(), rp ? 20 : 20, u, coeff, s.fc.data() );
      full_fluxes( o, s.fc.data(), s.flux.data() );
      if ( store_flux ) store_int_flux( s.flux.data(), iflux );
      fail = full_update( o, u, NX, ghost, s.flux.data(), out );
      it++;
   } while ( fail && it <= maxiter );
}

template <typename T>
int run( const T *in, T *out, T *flux, int npg, double dt, double dh, double gamma, int scheme, int rsolver,
         int store_flux, int limiter, double minmod_coeff, int maxiter, double min_dens, double min_pres,
         double min_eint, int nthreads )
{
   if ( npg < 0 || ( npg > 0 && ( in == nullptr || out == nullptr ) ) ) return 1;
   if ( store_flux && npg > 0 && flux == nullptr ) return 1;
   if ( scheme != ORC_MHM && scheme != ORC_MHM_RP ) return 2;
   if ( rsolver != ORC_HLLC && rsolver != ORC_HLLE && rsolver != ORC_ROE ) return 3;
   if ( limiter != 1 && limiter != 2 && limiter != 3 && limiter != 4 && limiter != 6 ) return 4;
   if ( maxiter < 0 ) return 5;

   Opt<T> o( gamma );
   o.dt = (T)dt;  o.dh = (T)dh;
   o.min_dens = (T)min_dens;  o.min_pres = (T)min_pres;  o.min_eint = (T)min_eint;
   o.scheme = scheme;  o.rsolver = rsolver;  o.limiter = limiter;

   const int NX = ( scheme == ORC_MHM_RP ) ? 22 : 20;
   const long in_sz = (long)NV*NX*NX*NX, out_sz = (long)NV*PS*PS*PS, fl_sz = 9L*NV*PS*PS;
   if ( nthreads > 0 ) omp_set_num_threads( nthreads );

#  pragma omp parallel
   {
      Scratch<T> s;
#     pragma omp for schedule( dynamic, 1 )
      for (int p=0; p<npg; p++)
         solve_pg( o, in + p*in_sz, out + p*out_sz, store_flux ? flux + p*fl_sz : nullptr,
                   store_flux != 0, (T)minmod_coeff, maxiter, s );
   }
   return 0;
}

// ---------------- Flu_Close and the AMR level synchronisation ----------------
// TABLE_02 offsets of the 8 patches in a patch group (src/Tables/Table_02.cpp:22-70)
const int LOCAL_XYZ[8][3] = { {0,0,0}, {1,0,0}, {0,1,0}, {0,0,1}, {1,1,0}, {0,1,1}, {1,0,1}, {1,1,1} };

struct CorrOpt {
   int    mode, rsolver, last_resort_floor, auto_reduce, fixup_flux, check_pres;
   int    roe_check;                              // the build's Roe solver checks intermediate states (RSOLVER == ROE)
   double min_dens, min_pres, min_eint, gm1;      // gm1: EoS aux constant gamma-1, a double (CPU_EoS_Gamma.cpp:54-66)
};

// Unphysical() for HYDRO without dual energy, MHD or unit checks (src/Fluid/Flu_Close.cpp:412-535)
template <typename T> bool close_unphysical( const Gas<T> &g, const T f[NV], bool eint_mode, const CorrOpt &c )
{
   for (int v=0; v<NV; v++) if ( !std::isfinite( f[v] ) ) return true;
   if ( f[0] < (T)c.min_dens ) return true;
   if ( !eint_mode && ( f[4] < (T)c.min_eint || f[4] != f[4] ) ) return true;
   if ( eint_mode )
   {
      const T e = eint_of( f[0], f[1], f[2], f[3], f[4] );
      if ( e < (T)c.min_eint || !std::isfinite( e ) ) return true;
   }
   if ( c.check_pres )
   {
      // EoS_DensEint2Pres of EOS_GAMMA multiplies by the double aux constant and narrows the result
      const T p = (T)( (double)eint_of( f[0], f[1], f[2], f[3], f[4] )*c.gm1 );
      if ( p < (T)c.min_pres || !std::isfinite( p ) ) return true;
   }
   return false;
}

// the first-order interface solve of CorrectUnphysical: the solver itself, without the NaN rescue
// (src/Fluid/Flu_Close.cpp:647-745); roe_check as in rs_roe_x
template <typename T> void riemann_1st( const Gas<T> &g, int rsolver, int d, const T Lin[NV], const T Rin[NV], T pmin, T out[NV],
                                        bool roe_check )
{
   T L[NV], R[NV];
   for (int v=0; v<NV; v++) { L[v] = Lin[v];  R[v] = Rin[v]; }
   rotate( L, d, true );
   rotate( R, d, true );
   switch ( rsolver )
   {
      case ORC_HLLC: rs_hllc_x( g, L, R, pmin, out ); break;
      case ORC_HLLE: rs_hlle_x( g, L, R, pmin, out ); break;
      default:       rs_roe_x ( g, L, R, pmin, out, roe_check ); break;
   }
   rotate( out, d, false );
}

// CorrectUnphysical() for one patch group (src/Fluid/Flu_Close.cpp:542-1297): cells that fail the
// Eint-mode check are recomputed from the input with first-order fluxes (unsplit 3D, then the x->y->z split
// 3D1D sweep over the 3^3 neighbourhood if still unphysical), floored, re-checked, stored, and their first-order
// fluxes replace the coarse-fine flux planes they touch.  cnt[0] counts corrected cells, cnt[1] failures.
template <typename T>
void correct_unphysical_pg( const Gas<T> &g, const CorrOpt &c, const T *in, T *out, T *flux, int NX, int ghost,
                            T dt_dh, long cnt[2] )
{
   const int N3 = NX*NX*NX;
   const int didx[3] = { 1, NX, NX*NX };
   const T pmin = (T)c.min_pres;
   for (int k=0; k<PS; k++)
   for (int j=0; j<PS; j++)
   for (int i=0; i<PS; i++)
   {
      const int ijk[3] = { i, j, k };
      const int io = ( k*PS + j )*PS + i;
      T Out[NV], Update[NV], FL[3][NV], FR[3][NV];
      for (int v=0; v<NV; v++) Out[v] = out[ v*PS*PS*PS + io ];
      if ( !close_unphysical( g, Out, true, c ) ) continue;

      const int ii = ( ( k + ghost )*NX + j + ghost )*NX + i + ghost;
      if ( c.mode != 0 )
      {
         T C[NV], L[3][NV], R[3][NV], dF[3][NV];
         for (int v=0; v<NV; v++)
         {
            C[v] = in[ v*N3 + ii ];
            for (int d=0; d<3; d++) { L[d][v] = in[ v*N3 + ii - didx[d] ];  R[d][v] = in[ v*N3 + ii + didx[d] ]; }
         }
         for (int d=0; d<3; d++)
         {
            riemann_1st( g, c.rsolver, d, L[d], C, pmin, FL[d], c.roe_check != 0 );
            riemann_1st( g, c.rsolver, d, C, R[d], pmin, FR[d], c.roe_check != 0 );
         }
         for (int d=0; d<3; d++) for (int v=0; v<NV; v++) dF[d][v] = FR[d][v] - FL[d][v];
         for (int v=0; v<NV; v++) Update[v] = C[v] - dt_dh*( dF[0][v] + dF[1][v] + dF[2][v] );
      }
      if ( c.mode == 2 && close_unphysical( g, Update, true, c ) )
      {
         T blk[3][3][3][NV];
         for (int z=0; z<3; z++) for (int y=0; y<3; y++) for (int x=0; x<3; x++)
            for (int v=0; v<NV; v++) blk[z][y][x][v] = in[ v*N3 + ii + ( x - 1 ) + ( y - 1 )*NX + ( z - 1 )*NX*NX ];
         const int lo[3][3] = { { 1, 0, 0 }, { 1, 1, 0 }, { 1, 1, 1 } };     // [d][x,y,z]
         const int hi[3][3] = { { 1, 2, 2 }, { 1, 1, 2 }, { 1, 1, 1 } };
         for (int d=0; d<3; d++)
            for (int z=lo[d][2]; z<=hi[d][2]; z++)
            for (int y=lo[d][1]; y<=hi[d][1]; y++)
            for (int x=lo[d][0]; x<=hi[d][0]; x++)
            {
               const int ox = ( d == 0 ), oy = ( d == 1 ), oz = ( d == 2 );
               T fl[NV], fr[NV];
               riemann_1st( g, c.rsolver, d, blk[z-oz][y-oy][x-ox], blk[z][y][x], pmin, fl, c.roe_check != 0 );
               riemann_1st( g, c.rsolver, d, blk[z][y][x], blk[z+oz][y+oy][x+ox], pmin, fr, c.roe_check != 0 );
               for (int v=0; v<NV; v++) blk[z][y][x][v] -= dt_dh*( fr[v] - fl[v] );
               if ( x == 1 && y == 1 && z == 1 )
                  for (int v=0; v<NV; v++) { FL[d][v] = fl[v];  FR[d][v] = fr[v]; }
            }
         for (int v=0; v<NV; v++) Update[v] = blk[1][1][1][v];
      }
      if ( c.mode == 0 ) for (int v=0; v<NV; v++) Update[v] = Out[v];

      if ( !c.auto_reduce && c.last_resort_floor )
      {
         Update[0] = vmax( Update[0], (T)c.min_dens );
         Update[4] = engy_with_eint_floor( Update[0], Update[1], Update[2], Update[3], Update[4], (T)c.min_eint );
      }
      if ( close_unphysical( g, Update, c.auto_reduce != 0, c ) ) { cnt[1]++;  continue; }

      for (int v=0; v<NV; v++) out[ v*PS*PS*PS + io ] = Update[v];
      if ( c.fixup_flux && c.mode != 0 && flux != nullptr )
      {
         const int dim_map[3][2] = { { 1, 2 }, { 0, 2 }, { 0, 1 } };
         for (int d=0; d<3; d++)
         {
            const T *F = nullptr;
            int face = -1;
            if      ( ijk[d] == 0    || ijk[d] == PS/2 )   { F = FL[d];  face = 3*d + ijk[d]/( PS/2 ); }
            else if ( ijk[d] == PS/2 - 1 || ijk[d] == PS - 1 ) { F = FR[d];  face = 3*d + ( ijk[d] + 1 )/( PS/2 ); }
            if ( F == nullptr ) continue;
            const int st = ijk[ dim_map[d][1] ]*PS + ijk[ dim_map[d][0] ];
            for (int v=0; v<NV; v++) flux[ ( face*NV + v )*PS*PS + st ] = F[v];
         }
      }
      cnt[0]++;
   }
}

// StoreFlux() (src/Fluid/Flu_Close.cpp:234-307) into FP64 registers: reg[slot[pg][LocalID][s]][5][8][8] = dt*F
template <typename T>
void store_flux( const T *flux, int npg, const int *slot, double dt, double *reg )
{
   const int H = PS/2;
   for (int p=0; p<npg; p++)
   for (int lid=0; lid<8; lid++)
   for (int s=0; s<6; s++)
   {
      const int r = slot[ ( p*8 + lid )*6 + s ];
      if ( r < 0 ) continue;
      const int *o = LOCAL_XYZ[lid];
      int face, dm, dn;
      if      ( s < 2 ) { face = o[0] + s;      dm = o[2]*H;  dn = o[1]*H; }
      else if ( s < 4 ) { face = o[1] + 1 + s;  dm = o[2]*H;  dn = o[0]*H; }
      else              { face = o[2] + 2 + s;  dm = o[1]*H;  dn = o[0]*H; }
      for (int v=0; v<NV; v++)
      for (int m=0; m<H; m++)
      for (int n=0; n<H; n++)
         reg[ ( ( (long)r*NV + v )*H + m )*H + n ] = dt*(double)flux[ ( ( (long)p*9 + face )*NV + v )*PS*PS + ( m + dm )*PS + n + dn ];
   }
}

// CorrectFlux() (src/Fluid/Flu_Close.cpp:321-395): reg[slot[pg][s]] -= dt/4 * (fine fluxes on PG face s)
template <typename T>
void correct_flux( const T *flux, int npg, const int *slot, double dt, double *reg )
{
   const int H = PS/2;
   const int map[6] = { 0, 2, 3, 5, 6, 8 };
   const double dt4 = 0.25*dt;
   for (int p=0; p<npg; p++)
   for (int s=0; s<6; s++)
   {
      const int r = slot[ p*6 + s ];
      if ( r < 0 ) continue;
      for (int v=0; v<NV; v++)
      for (int m=0; m<PS; m++)
      for (int n=0; n<PS; n++)
         reg[ ( ( (long)r*NV + v )*H + m/2 )*H + n/2 ] -= dt4*(double)flux[ ( ( (long)p*9 + map[s] )*NV + v )*PS*PS + m*PS + n ];
   }
}

// Flu_FixUp_Flux() (src/Fluid/Flu_FixUp_Flux.cpp) on patches [npatch][5][8^3] with FP64 registers
template <typename T>
void fixup_flux( double gm1, T *patch, int npatch, const int *slot, const double *reg, double dh,
                 double min_dens, double min_pres, double min_eint )
{
   const int H = PS/2;
   const double cst[6] = { -1.0/dh, +1.0/dh, -1.0/dh, +1.0/dh, -1.0/dh, +1.0/dh };
   for (int p=0; p<npatch; p++)
   for (int s=0; s<6; s++)
   {
      const int r = slot[ p*6 + s ];
      if ( r < 0 ) continue;
      for (int m=0; m<H; m++)
      for (int n=0; n<H; n++)
      {
         int i, j, k;
         switch ( s )
         {
            case 0:  i = 0;      j = n;      k = m;      break;
            case 1:  i = H-1;    j = n;      k = m;      break;
            case 2:  i = n;      j = 0;      k = m;      break;
            case 3:  i = n;      j = H-1;    k = m;      break;
            case 4:  i = n;      j = m;      k = 0;      break;
            default: i = n;      j = m;      k = H-1;    break;
         }
         T *u = patch + (long)p*NV*H*H*H + ( k*H + j )*H + i;
         T cv[NV];
         for (int v=0; v<NV; v++) cv[v] = (T)( (double)u[ v*H*H*H ] + reg[ ( ( (long)r*NV + v )*H + m )*H + n ]*cst[s] );
         const T e = eint_of( cv[0], cv[1], cv[2], cv[3], cv[4] );
         const T pr = (T)( (double)e*gm1 );
         // MIN_DENS / MIN_EINT / MIN_PRES are doubles in the reference: the comparisons promote
         if ( (double)cv[0] <= min_dens || (double)e <= min_eint || !std::isfinite( e ) ||
              (double)pr <= min_pres || !std::isfinite( pr ) ) continue;
         cv[4] = etot_of( cv[0], cv[1], cv[2], cv[3], e );
         for (int v=0; v<NV; v++) u[ v*H*H*H ] = cv[v];
      }
   }
}

// Flu_FixUp_Restrict() (src/Fluid/Flu_FixUp_Restrict.cpp): father patch fa[fa_idx[q]] = average of the 8 son
// patches son[son0[q] + LocalID], then the internal-energy floor on every father cell
template <typename T>
void fixup_restrict( const T *son, T *fa, int ngroup, const int *son0, const int *fa_idx, double min_eint )
{
   const int H = PS/2, H2 = H/2, P3 = H*H*H;
   for (int q=0; q<ngroup; q++)
   {
      T *F = fa + (long)fa_idx[q]*NV*P3;
      for (int lid=0; lid<8; lid++)
      {
         const T *S = son + (long)( son0[q] + lid )*NV*P3;
         const int *o = LOCAL_XYZ[lid];
         for (int v=0; v<NV; v++)
         for (int k=0; k<H2; k++)
         for (int j=0; j<H2; j++)
         for (int i=0; i<H2; i++)
         {
            const T *s0 = S + v*P3;
            const int I = 2*i, J = 2*j, K = 2*k;
            auto at = [&]( int kk, int jj, int ii ) { return s0[ ( kk*H + jj )*H + ii ]; };
            F[ v*P3 + ( ( k + o[2]*H2 )*H + j + o[1]*H2 )*H + i + o[0]*H2 ] =
               (T)( 0.125*( at( K, J, I ) + at( K, J, I+1 ) + at( K, J+1, I ) + at( K+1, J, I ) +
                            at( K, J+1, I+1 ) + at( K+1, J+1, I ) + at( K+1, J, I+1 ) + at( K+1, J+1, I+1 ) ) );
         }
      }
      for (int c=0; c<P3; c++)
         F[ 4*P3 + c ] = engy_with_eint_floor( F[c], F[P3 + c], F[2*P3 + c], F[3*P3 + c], F[4*P3 + c], (T)min_eint );
   }
}

// copy of patch-group outputs into their 8 patches (src/Fluid/Flu_Close.cpp:133-160)
template <typename T>
void close_patches( const T *pg_out, T *patch, int npg, const int *pid0 )
{
   const int H = PS/2, P3 = H*H*H;
   for (int p=0; p<npg; p++)
   for (int lid=0; lid<8; lid++)
   {
      T *D = patch + (long)( pid0[p] + lid )*NV*P3;
      const int *o = LOCAL_XYZ[lid];
      for (int v=0; v<NV; v++)
      for (int k=0; k<H; k++)
      for (int j=0; j<H; j++)
      for (int i=0; i<H; i++)
         D[ v*P3 + ( k*H + j )*H + i ] = pg_out[ ( (long)p*NV + v )*PS*PS*PS + ( ( o[2]*H + k )*PS + o[1]*H + j )*PS + o[0]*H + i ];
   }
}

// ---------------- coarse -> fine interpolation (src/Interpolation) ----------------
// Int_CQuadratic slope of one split pass (Int_CQuadratic.cpp:53-78): (R-L)/8, optionally limited by the one-sided
// slopes (C-L)/4, (R-C)/4 with the smaller one scaled by MonoCoeff; zero across a sign change of L,R if requested
template <typename T> T cquad_slope( T L, T C, T R, bool mono, T coeff, bool opp0 )
{
   T s = (T)0.125*( R - L );
   if ( mono )
   {
      T l = (T)0.25*( C - L ), r = (T)0.25*( R - C );
      if ( l*r > (T)0.0 )
      {
         const T sg = sgn( l );
         s *= sg;  l *= sg;  r *= sg;
         if ( l < r ) l *= coeff; else r *= coeff;
         s = vmin( l, s );
         s = vmin( r, s );
         s *= sg;
      }
      else s = (T)0.0;
   }
   if ( opp0 && L*R < (T)0.0 ) s = (T)0.0;
   return s;
}

struct Box { int cs[3], c0[3], cr[3], fs[3], f0[3]; };

// interpolation scheme ids (include/Typedef.h:171-180) and their coarse ghost widths (src/Interpolation/Int_Table.cpp)
enum { INT_MINMOD3D = 1, INT_MINMOD1D = 2, INT_VANLEER = 3, INT_CQUAD = 4, INT_QUAD = 5, INT_CQUAR = 6, INT_QUAR = 7 };
inline int int_ghost( int scheme ) { return ( scheme == INT_CQUAR || scheme == INT_QUAR ) ? 2 : 1; }

// One split pass of the tensor-product schemes: the stencil u = {L2, L1, C, R1, R2} along one direction -> the two
// fine values (lo, hi) of the centre cell.  The per-scheme bodies follow the x passes of
//   Int_MinMod3D.cpp:93-103   Int_vanLeer.cpp:93-105   Int_CQuadratic.cpp (cquad_slope above)
//   Int_Quadratic.cpp:104-160  Int_CQuartic.cpp:100-140  Int_Quartic.cpp:100-165
// (the y and z passes of each file repeat the x pass on the previous pass's output).
template <typename T>
void pass1d( int scheme, const T u[5], bool mono, T coeff, bool opp0, T &lo, T &hi )
{
   const T L2 = u[0], L1 = u[1], C = u[2], R1 = u[3], R2 = u[4];
   T s;
   switch ( scheme )
   {
      case INT_MINMOD3D: case INT_VANLEER:
      {
         const T l = C - L1, r = R1 - C;
         if ( r*l <= (T)0.0 ) s = (T)0.0;
         else s = ( scheme == INT_MINMOD3D ) ? (T)0.25*( std::fabs( r ) < std::fabs( l ) ? r : l )
                                             : (T)0.5*l*r/( l + r );
         if ( opp0 && L1*R1 < (T)0.0 ) s = (T)0.0;
         lo = C - s;  hi = C + s;
         return;
      }
      case INT_CQUAD:
         s = cquad_slope( L1, C, R1, mono, coeff, opp0 );
         lo = C - s;  hi = C + s;
         return;
      case INT_CQUAR:
      {
         s = (T)( 3.0/128.0)*L2 + (T)(-22.0/128.0)*L1 + (T)(22.0/128.0)*R1 + (T)(-3.0/128.0)*R2;
         if ( mono )
         {
            T l = (T)0.25*( C - L1 ), r = (T)0.25*( R1 - C );
            if ( l*r > (T)0.0 )
            {
               if ( s*l < (T)0.0 ) s = (T)0.125*( R1 - L1 );
               const T sg = sgn( l );
               s *= sg;  l *= sg;  r *= sg;
               if ( l < r ) l *= coeff; else r *= coeff;
               s = vmin( l, s );
               s = vmin( r, s );
               s *= sg;
            }
            else s = (T)0.0;
         }
         if ( opp0 && L1*R1 < (T)0.0 ) s = (T)0.0;
         lo = C - s;  hi = C + s;
         return;
      }
      default:          // INT_QUAD, INT_QUAR: Lagrange point values at -/+ 1/4, reverted to a limited slope if not monotonic
      {
         if ( scheme == INT_QUAD )
         {
            lo = (T)( 5.0/32.0)*L1 + (T)(30.0/32.0)*C + (T)(-3.0/32.0)*R1;
            hi = (T)(-3.0/32.0)*L1 + (T)(30.0/32.0)*C + (T)( 5.0/32.0)*R1;
         }
         else
         {
            lo = (T)( -45.0/2048.0)*L2 + (T)( 420.0/2048.0)*L1 + (T)(1890.0/2048.0)*C + (T)(-252.0/2048.0)*R1
               + (T)(  35.0/2048.0)*R2;
            hi = (T)(  35.0/2048.0)*L2 + (T)(-252.0/2048.0)*L1 + (T)(1890.0/2048.0)*C + (T)( 420.0/2048.0)*R1
               + (T)( -45.0/2048.0)*R2;
         }
         if ( mono )
         {
            T l = C - L1, r = R1 - C;
            if ( l*r > (T)0.0 )
            {
               const T cmax = ( l > (T)0.0 ) ? R1 : L1, cmin = ( l > (T)0.0 ) ? L1 : R1;
               bool bad = ( vmax( lo, hi ) > cmax || vmin( lo, hi ) < cmin );
               if ( scheme == INT_QUAR ) bad = bad || ( hi - lo )*l < (T)0.0;
               if ( bad )
               {
                  const T c4 = (T)0.25*coeff;
                  s  = (T)0.125*( R1 - L1 );
                  l *= c4;  r *= c4;
                  const T sg = sgn( l );
                  s *= sg;
                  s  = vmin( sg*l, s );
                  s  = vmin( sg*r, s );
                  s *= sg;
                  lo = C - s;  hi = C + s;
               }
            }
            else { lo = C;  hi = C; }
         }
         if ( opp0 && L1*R1 < (T)0.0 ) { lo = C;  hi = C; }
         return;
      }
   }
}

// The tensor-product schemes (all but INT_MINMOD1D): x pass over the coarse rows (with G ghost rows in y and z), then
// y, then z, through the same temporaries as the reference (TDataX, TDataY)
template <typename T>
void int_split( const T *C, T *F, const Box &b, int scheme, int mono_mask, T coeff, bool opp0 )
{
   const int G = int_ghost( scheme );
   const int cdy = b.cs[0], cdz = b.cs[0]*b.cs[1], C3 = cdz*b.cs[2], F3 = b.fs[0]*b.fs[1]*b.fs[2];
   const int nx = 2*b.cr[0], ny = b.cr[1] + 2*G, nz = b.cr[2] + 2*G, my = 2*b.cr[1];
   std::vector<T> X( (size_t)nx*ny*nz ), Y( (size_t)nx*my*nz );
   for (int v=0; v<NV; v++)
   {
      const T *c = C + (size_t)v*C3;
      T *f = F + (size_t)v*F3;
      const bool mono = ( mono_mask >> v ) & 1;
      T u[5], lo, hi;
      for (int z=0; z<nz; z++) for (int y=0; y<ny; y++) for (int x=0; x<b.cr[0]; x++)
      {
         const int i = ( b.c0[2] - G + z )*cdz + ( b.c0[1] - G + y )*cdy + b.c0[0] + x;
         for (int k=0; k<5; k++) u[k] = ( k - 2 >= -G && k - 2 <= G ) ? c[i + k - 2] : (T)0.0;
         pass1d( scheme, u, mono, coeff, opp0, lo, hi );
         X[ ( (size_t)z*ny + y )*nx + 2*x     ] = lo;
         X[ ( (size_t)z*ny + y )*nx + 2*x + 1 ] = hi;
      }
      for (int z=0; z<nz; z++) for (int y=0; y<b.cr[1]; y++) for (int x=0; x<nx; x++)
      {
         for (int k=0; k<5; k++)
            u[k] = ( k - 2 >= -G && k - 2 <= G ) ? X[ ( (size_t)z*ny + y + G + k - 2 )*nx + x ] : (T)0.0;
         pass1d( scheme, u, mono, coeff, opp0, lo, hi );
         Y[ ( (size_t)z*my + 2*y     )*nx + x ] = lo;
         Y[ ( (size_t)z*my + 2*y + 1 )*nx + x ] = hi;
      }
      for (int z=0; z<b.cr[2]; z++) for (int y=0; y<my; y++) for (int x=0; x<nx; x++)
      {
         for (int k=0; k<5; k++)
            u[k] = ( k - 2 >= -G && k - 2 <= G ) ? Y[ ( (size_t)( z + G + k - 2 )*my + y )*nx + x ] : (T)0.0;
         pass1d( scheme, u, mono, coeff, opp0, lo, hi );
         const int o = ( ( b.f0[2] + 2*z )*b.fs[1] + b.f0[1] + y )*b.fs[0] + b.f0[0] + x;
         f[o]                   = lo;
         f[o + b.fs[0]*b.fs[1]] = hi;
      }
   }
}

// Int_MinMod1D (src/Interpolation/Int_MinMod1D.cpp): per-direction minmod slope / 4, eight fine values
template <typename T>
void int_minmod1d( const T *C, T *F, const Box &b, bool opp0 )
{
   const int cd[3] = { 1, b.cs[0], b.cs[0]*b.cs[1] }, C3 = cd[2]*b.cs[2];
   const int fdy = b.fs[0], fdz = b.fs[0]*b.fs[1], F3 = fdz*b.fs[2];
   for (int v=0; v<NV; v++)
   for (int z=0; z<b.cr[2]; z++) for (int y=0; y<b.cr[1]; y++) for (int x=0; x<b.cr[0]; x++)
   {
      const T *c = C + (size_t)v*C3;
      const int i = ( b.c0[2] + z )*cd[2] + ( b.c0[1] + y )*cd[1] + b.c0[0] + x;
      T sl[3];
      for (int d=0; d<3; d++)
      {
         const T l = c[i] - c[i - cd[d]], r = c[i + cd[d]] - c[i];
         sl[d] = ( r*l <= (T)0.0 ) ? (T)0.0 : (T)0.25*( std::fabs( r ) < std::fabs( l ) ? r : l );
         if ( opp0 && c[i - cd[d]]*c[i + cd[d]] < (T)0.0 ) sl[d] = (T)0.0;
      }
      T *f = F + (size_t)v*F3 + ( b.f0[2] + 2*z )*fdz + ( b.f0[1] + 2*y )*fdy + b.f0[0] + 2*x;
      for (int k=0; k<2; k++) for (int j=0; j<2; j++) for (int ii=0; ii<2; ii++)
         f[ k*fdz + j*fdy + ii ] = c[i] + ( k ? sl[2] : -sl[2] ) + ( j ? sl[1] : -sl[1] ) + ( ii ? sl[0] : -sl[0] );
   }
}

// the PRIM-mode check of Interpolate_Iterate: Hydro_IsUnphysical(PRIM) then the internal energy from the EoS
// (src/Model_Hydro/CPU_Hydro/CPU_Shared_FluUtility.cpp:969-1005, src/Interpolation/Interpolate.cpp:234-262)
template <typename T> bool prim_unphysical( const T w[NV], double inv_gm1 )
{
   const T huge = Lim<T>::huge();
   for (int v=0; v<NV; v++)
   {
      if ( w[v] != w[v] ) return true;
      if ( v >= 1 && v <= 3 ) { if ( w[v] < -huge || w[v] > huge ) return true; }
      else                    { if ( w[v] <= (T)0.0 || w[v] > huge ) return true; }
   }
   const T e = w[4]*(T)inv_gm1;                      // Pres*_Gamma_m1 in real (CPU_EoS_Gamma.cpp:151-154)
   return ( e < (T)0.0 || e > huge || e != e );
}

// Interpolate() with Interpolate_Iterate (src/Interpolation/Interpolate.cpp:90-330), INTERP_MASK semantics:
// iteration 0 in conserved variables with INT_MONO_COEFF; iteration 1 (OPT__INT_PRIM) in primitive variables (the
// coarse box is converted in place); later iterations lower the monotonic coefficient by INT_MONO_COEFF/MONO_MAX_ITER.
// Fine cells accepted in an iteration keep their values unless their 2x2x2 block still holds a failed cell.
// Returns 0 on success, 1 if cells are still unphysical after the last iteration.
template <typename T>
int interpolate( T *C, T *F, const Box &b, int scheme, int mono_mask, double mono_coeff, bool opp0, bool int_prim,
                 bool reduce, int mono_max_iter, double gamma, double min_pres )
{
   const Gas<T> g( gamma );
   const double inv_gm1 = 1.0/( gamma - 1.0 );
   auto run = [&]( T *out, T coeff ) {
      if ( scheme == INT_MINMOD1D ) int_minmod1d( C, out, b, opp0 );
      else                          int_split( C, out, b, scheme, mono_mask, coeff, opp0 );
   };
   if ( !reduce && !int_prim ) { run( F, (T)mono_coeff );  return 0; }

   const int C3 = b.cs[0]*b.cs[1]*b.cs[2], F3 = b.fs[0]*b.fs[1]*b.fs[2];
   const int mono_max = reduce ? mono_max_iter : 0, max_iter = int_prim ? mono_max + 1 : mono_max;
   std::vector<T> Ft( (size_t)NV*F3 );
   std::vector<char> mask( F3, 0 );                 // 1 = accepted
   T coeff = (T)mono_coeff;
   bool prim = false;
   for (int it=0; ; it++)
   {
      if ( it == 1 && int_prim )
      {
         for (int i=0; i<C3; i++)
         {
            T c[NV], w[NV];
            for (int v=0; v<NV; v++) c[v] = C[ (size_t)v*C3 + i ];
            con2pri( g, c, w, (T)min_pres );
            for (int v=0; v<NV; v++) C[ (size_t)v*C3 + i ] = w[v];
         }
         prim = true;
      }
      else if ( it > 0 )
      {
         // INT_MINMOD1D / INT_MINMOD3D / INT_VANLEER have no monotonic coefficient to lower (Interpolate.cpp:299-300)
         if ( scheme == INT_MINMOD1D || scheme == INT_MINMOD3D || scheme == INT_VANLEER ) return 1;
         coeff -= (T)mono_coeff/(T)mono_max;
         coeff  = vmax( coeff, (T)0.0 );
      }
      run( Ft.data(), coeff );
      bool any = false;
      for (int i=0; i<F3; i++)
      {
         if ( mask[i] ) continue;
         T t[NV];
         for (int v=0; v<NV; v++) t[v] = Ft[ (size_t)v*F3 + i ];
         const bool bad = prim ? prim_unphysical( t, inv_gm1 ) : unphysical( t );
         if ( bad ) { any = true;  if ( it == max_iter ) return 1;  continue; }
         if ( prim )
         {
            const T w4 = t[4]*(T)inv_gm1;                      // EoS_DensPres2Eint_Gamma: Pres*_Gamma_m1 in real
            T c[NV];
            c[0] = t[0];  c[1] = t[0]*t[1];  c[2] = t[0]*t[2];  c[3] = t[0]*t[3];
            c[4] = etot_of( c[0], c[1], c[2], c[3], w4 );
            for (int v=0; v<NV; v++) t[v] = c[v];
         }
         for (int v=0; v<NV; v++) F[ (size_t)v*F3 + i ] = t[v];
         mask[i] = 1;
      }
      if ( !( any && it < max_iter ) ) return 0;
      for (int k=0; k<b.fs[2]; k+=2) for (int j=0; j<b.fs[1]; j+=2) for (int i=0; i<b.fs[0]; i+=2)
      {
         bool open = false;
         for (int dk=0; dk<2; dk++) for (int dj=0; dj<2; dj++) for (int di=0; di<2; di++)
            open |= !mask[ ( ( k + dk )*b.fs[1] + j + dj )*b.fs[0] + i + di ];
         if ( open )
            for (int dk=0; dk<2; dk++) for (int dj=0; dj<2; dj++) for (int di=0; di<2; di++)
               mask[ ( ( k + dk )*b.fs[1] + j + dj )*b.fs[0] + i + di ] = 0;
      }
   }
}

// same-level part of the patch-group input preparation from per-patch data: window cell (x,y,z) of PG p copies
// patch slot[p][sz][sy][sx] (-1 = not at this level: the cell keeps what the interpolation step wrote)
template <typename T>
void prepare_patches( const T *patch, T *pg_in, int npg, const int *slot, int ghost )
{
   const int H = PS/2, P3 = H*H*H, n = PS + 2*ghost, n3 = n*n*n;
   for (int p=0; p<npg; p++)
   for (int v=0; v<NV; v++)
   for (int z=0; z<n; z++) for (int y=0; y<n; y++) for (int x=0; x<n; x++)
   {
      const int X = x - ghost + H, Y = y - ghost + H, Z = z - ghost + H;
      const int s = slot[ ( ( p*4 + Z/H )*4 + Y/H )*4 + X/H ];
      if ( s < 0 ) continue;
      pg_in[ ( (size_t)p*NV + v )*n3 + ( z*n + y )*n + x ] = patch[ ( (size_t)s*NV + v )*P3 + ( ( Z % H )*H + Y % H )*H + X % H ];
   }
}

} // namespace

// the coarse boxes of InterpolateGhostZone: box cell (x,y,z) <- patch covering cell origin + (x,y,z) of the box's
// tdim table (-1 keeps the box value)
template <typename T>
void gather_boxes( const T *patch, T *box, int nbox, const int size[3], const int origin[3], const int *slot,
                   const int tdim[3] )
{
   const int n3 = size[0]*size[1]*size[2], t3 = tdim[0]*tdim[1]*tdim[2];
   for (int q=0; q<nbox; q++)
   for (int z=0; z<size[2]; z++) for (int y=0; y<size[1]; y++) for (int x=0; x<size[0]; x++)
   {
      const int X = origin[0] + x, Y = origin[1] + y, Z = origin[2] + z;
      const int s = slot[ (size_t)q*t3 + ( Z/8*tdim[1] + Y/8 )*tdim[0] + X/8 ];
      if ( s < 0 ) continue;
      for (int v=0; v<NV; v++)
         box[ ( (size_t)q*NV + v )*n3 + ( z*size[1] + y )*size[0] + x ] =
            patch[ ( (size_t)s*NV + v )*512 + ( ( Z % 8 )*8 + Y % 8 )*8 + X % 8 ];
   }
}

// patch-group input at a refinement boundary: same-level patch where the 4x4x4 slot is >= 0, else the interpolated
// fine box at window cell + fb_origin
template <typename T>
void assemble_patch_groups( const T *patch, const T *fbox, T *pg_in, int npg, const int *slot, int ghost,
                            const int fb_size[3], const int fb_origin[3] )
{
   const int n = 16 + 2*ghost, n3 = n*n*n, f3 = fb_size[0]*fb_size[1]*fb_size[2];
   for (int p=0; p<npg; p++)
   for (int z=0; z<n; z++) for (int y=0; y<n; y++) for (int x=0; x<n; x++)
   {
      const int X = x - ghost + 8, Y = y - ghost + 8, Z = z - ghost + 8;
      const int s = slot[ ( ( p*4 + Z/8 )*4 + Y/8 )*4 + X/8 ];
      for (int v=0; v<NV; v++)
         pg_in[ ( (size_t)p*NV + v )*n3 + ( z*n + y )*n + x ] = ( s >= 0 )
            ? patch[ ( (size_t)s*NV + v )*512 + ( ( Z % 8 )*8 + Y % 8 )*8 + X % 8 ]
            : fbox[ ( (size_t)p*NV + v )*f3 + ( ( z + fb_origin[2] )*fb_size[1] + y + fb_origin[1] )*fb_size[0]
                    + x + fb_origin[0] ];
   }
}

extern "C" {

int oracle_fluid_solver_f64( const double *in, double *out, double *flux, int npg, double dt, double dh,
                             double gamma, int scheme, int rsolver, int store_flux, int limiter,
                             double minmod_coeff, int minmod_maxiter, double min_dens, double min_pres,
                             double min_eint, int nthreads )
{
   return run<double>( in, out, flux, npg, dt, dh, gamma, scheme, rsolver, store_flux, limiter, minmod_coeff,
                       minmod_maxiter, min_dens, min_pres, min_eint, nthreads );
}

int oracle_fluid_solver_f32( const float *in, float *out, float *flux, int npg, double dt, double dh,
                             double gamma, int scheme, int rsolver, int store_flux, int limiter,
                             double minmod_coeff, int minmod_maxiter, double min_dens, double min_pres,
                             double min_eint, int nthreads )
{
   return run<float>( in, out, flux, npg, dt, dh, gamma, scheme, rsolver, store_flux, limiter, minmod_coeff,
                      minmod_maxiter, min_dens, min_pres, min_eint, nthreads );
}

int oracle_riemann_f64( int rsolver, int d, const double L[5], const double R[5], double gamma,
                        double min_pres, double flux[5] )
{
   if ( d < 0 || d > 2 ) return 1;
   if ( rsolver != ORC_HLLC && rsolver != ORC_HLLE && rsolver != ORC_ROE ) return 3;
   Gas<double> g( gamma );
   riemann( g, rsolver, d, L, R, min_pres, flux );
   return 0;
}

int oracle_limit_slope_f64( const double L[5], const double C[5], const double R[5], int limiter,
                            double minmod_coeff, double slope[5] )
{
   if ( limiter != 1 && limiter != 2 && limiter != 3 && limiter != 4 && limiter != 6 ) return 4;
   limit_slope( L, C, R, limiter, minmod_coeff, slope );
   return 0;
}

// per-patch CFL: dt = Safety*dh / max_cells max_d( |v_d| + c_s ) (CPU_dtSolver_HydroCFL.cpp:57-209)
int oracle_dt_cfl_f64( const double *in, double *dt, int npatch, double dh, double safety, double gamma,
                       double min_pres, int nthreads )
{
   if ( npatch < 0 || ( npatch > 0 && ( in == nullptr || dt == nullptr ) ) ) return 1;
   const Gas<double> g( gamma );
   const int N3 = PSH*PSH*PSH;
   const double dhs = safety*dh;
   if ( nthreads > 0 ) omp_set_num_threads( nthreads );
#  pragma omp parallel for schedule( dynamic, 1 )
   for (int p=0; p<npatch; p++)
   {
      const double *u = in + (long)p*NV*N3;
      double mx = 0.0;
      for (int t=0; t<N3; t++)
      {
         const double r  = 1.0 / u[t];
         const double vx = std::fabs( u[1*N3+t] )*r, vy = std::fabs( u[2*N3+t] )*r, vz = std::fabs( u[3*N3+t] )*r;
         const double p_ = pres_of( g, u[t], u[1*N3+t], u[2*N3+t], u[3*N3+t], u[4*N3+t], true, min_pres );
         double cx = std::sqrt( g.gamma*p_ / u[t] );
         const double cy = cx + vy, cz = cx + vz;
         cx += vx;
         mx = vmax( cx, mx );
         mx = vmax( cy, mx );
         mx = vmax( cz, mx );
      }
      dt[p] = dhs/mx;
   }
   return 0;
}

#define ORC_SYNC_API( SUF, T )                                                                                       \
int oracle_correct_unphysical_##SUF( const T *in, T *out, T *flux, int npg, int scheme, double dt, double dh,          \
                                     double gamma, int mode, int rsolver, int last_resort_floor, int auto_reduce,      \
                                     int fixup_flux, int check_pres, double min_dens, double min_pres,               \
                                     double min_eint, int main_rsolver, long counters[2] )                             \
{                                                                                                                      \
   if ( scheme != ORC_MHM && scheme != ORC_MHM_RP ) return 2;                                                          \
   if ( mode < 0 || mode > 2 ) return 3;                                                                               \
   const int NX = ( scheme == ORC_MHM_RP ) ? 22 : 20, ghost = ( NX - PS )/2;                                           \
   Gas<T> g( gamma );                                                                                                  \
   const int roe_check = ( ( main_rsolver == 0 ) ? rsolver : main_rsolver ) == ORC_ROE;                                \
   CorrOpt c = { mode, rsolver, last_resort_floor, auto_reduce, fixup_flux, check_pres, roe_check,                    \
                 min_dens, min_pres, min_eint, gamma - 1.0 };                                                          \
   const T dt_dh = (T)dt/(T)dh;                                                                                        \
   counters[0] = counters[1] = 0;                                                                                      \
   for (int p=0; p<npg; p++)                                                                                           \
      correct_unphysical_pg( g, c, in + (long)p*NV*NX*NX*NX, out + (long)p*NV*PS*PS*PS,                              \
                             flux ? flux + (long)p*9*NV*PS*PS : nullptr, NX, ghost, dt_dh, counters );               \
   return 0;                                                                                                           \
}                                                                                                                      \
int oracle_store_flux_##SUF( const T *flux, int npg, const int *slot, double dt, double *reg )                        \
{ store_flux( flux, npg, slot, dt, reg );  return 0; }                                                                 \
int oracle_correct_flux_##SUF( const T *flux, int npg, const int *slot, double dt, double *reg )                      \
{ correct_flux( flux, npg, slot, dt, reg );  return 0; }                                                               \
int oracle_fixup_flux_##SUF( T *patch, int npatch, const int *slot, const double *reg, double dh, double gamma,      \
                             double min_dens, double min_pres, double min_eint )                                      \
{ fixup_flux( gamma - 1.0, patch, npatch, slot, reg, dh, min_dens, min_pres, min_eint );  return 0; }                \
int oracle_fixup_restrict_##SUF( const T *son, T *fa, int ngroup, const int *son0, const int *fa_idx, double min_eint )\
{ fixup_restrict( son, fa, ngroup, son0, fa_idx, min_eint );  return 0; }                                              \
int oracle_close_patches_##SUF( const T *pg_out, T *patch, int npg, const int *pid0 )                                 \
{ close_patches( pg_out, patch, npg, pid0 );  return 0; }

#define ORC_INTERP_API( SUF, T )                                                                                     \
int oracle_interpolate_##SUF( T *cdata, T *fdata, int nbox, const int csize[3], const int cstart[3],                 \
                              const int crange[3], const int fsize[3], const int fstart[3], int scheme, int mono_mask,\
                              double mono_coeff, int opp0, int int_prim, int reduce, int mono_max_iter, double gamma,  \
                              double min_pres, int *nfail )                                                            \
{                                                                                                                      \
   Box b;                                                                                                              \
   for (int d=0; d<3; d++) { b.cs[d] = csize[d]; b.c0[d] = cstart[d]; b.cr[d] = crange[d]; b.fs[d] = fsize[d]; b.f0[d] = fstart[d]; } \
   const long C3 = (long)csize[0]*csize[1]*csize[2], F3 = (long)fsize[0]*fsize[1]*fsize[2];                          \
   *nfail = 0;                                                                                                         \
   for (int q=0; q<nbox; q++)                                                                                          \
      *nfail += interpolate( cdata + q*NV*C3, fdata + q*NV*F3, b, scheme, mono_mask, mono_coeff, opp0 != 0,            \
                             int_prim != 0, reduce != 0, mono_max_iter, gamma, min_pres );                            \
   return 0;                                                                                                           \
}                                                                                                                      \
int oracle_prepare_patches_##SUF( const T *patch, T *pg_in, int npg, const int *slot, int ghost )                    \
{ prepare_patches( patch, pg_in, npg, slot, ghost );  return 0; }                                                      \
int oracle_gather_boxes_##SUF( const T *patch, T *box, int nbox, const int size[3], const int origin[3],             \
                               const int *slot, const int tdim[3] )                                                   \
{ gather_boxes( patch, box, nbox, size, origin, slot, tdim );  return 0; }                                           \
int oracle_assemble_patch_groups_##SUF( const T *patch, const T *fbox, T *pg_in, int npg, const int *slot,          \
                                        int ghost, const int fb_size[3], const int fb_origin[3] )                     \
{ assemble_patch_groups( patch, fbox, pg_in, npg, slot, ghost, fb_size, fb_origin );  return 0; }

ORC_INTERP_API( f64, double )
ORC_INTERP_API( f32, float )

ORC_SYNC_API( f64, double )
ORC_SYNC_API( f32, float )

} // extern "C"
