// MHM_RP (MUSCL-Hancock with Riemann-solver prediction) patch-group kernel for gfx950.
//
// One 256-lane workgroup updates one patch group: 22^3 input cells (16^3 interior + 3 ghost zones) ->
// 16^3 output cells + 9 coarse-fine flux planes.  The reference runs the same update as eight global-memory
// phases over the whole 22^3 block (src/Model_Hydro/CPU_Hydro/CPU_FluidSolver_MHM.cpp:263-589, ~2.9 MB of
// scratch per patch group in FP64).  Here the workgroup marches along z, one input plane per step, and keeps
// every intermediate on chip:
//
//   * Lane t owns the output column (i,j) = (3 + t%16, 3 + output_row(t)).  Lanes 0-127 (waves 0-1) also own one
//     "arm" ghost column for the half step (arm_column()), lanes 128-167 (wave 2) one single half-step face per
//     plane (face_task()); together they update only the half-step cells the reconstruction reads.  Lanes
//     192-255 (wave 3) each reconstruct one ghost column next to the output block along x or y, and half of them
//     compute the full-step flux on the face between that column and the block (edge_column()).  Everything that
//     couples planes along z — the previous input planes, the half-step z-flux, the last three half-step
//     primitive planes, the lower z-face state and the last full-step z-flux — stays in the owning lane's
//     registers.
//   * Only x/y neighbours go through LDS, one plane at a time: the input plane with its pressure, 1/rho and
//     sound speed, x and y fluxes (half step, then full step), the half-step primitive plane and the upper
//     x/y face states.  A lane-private LDS ring keeps the input cells of the own output column for the planes
//     between load and full-step update, so every input value is read from HBM once (148 KB of LDS per
//     workgroup in FP64, 74 KB in FP32: two workgroups per CU).
//
// Plane schedule at step t (input plane t just loaded; all indices in input coordinates), with the four
// barriers of a step:
//   [A] half-step z-flux at t-1/2, half-step cell update of plane t-1, half-step x/y fluxes of plane t,
//       PLM reconstruction of plane r = t-2
//   [H] x/y half-step flux differences of plane t, half-step plane t-1 into LDS
//   [R] full-step fluxes of plane r (x/y) and at r-1/2 (z)
//   [F] full-step update and output of plane t-3, input plane t+1 into LDS
// [A] is the longest interval and waves 0-1 (arm half-step solves) its critical path; they raise their issue
// priority inside it.  All branches that contain a barrier depend only on t, so every barrier is reached by the
// whole workgroup.

#include <type_traits>

#include "hydro_math.hpp"
#include "hydro_kernels.hpp"

namespace gamer_amd {

namespace {

constexpr int NX  = 22;          // FLU_NXT for MHM_RP
constexpr int NC  = NX*NX;       // columns per plane
constexpr int PS  = 16;          // PS2
constexpr int NK  = 256;         // output columns (lanes of waves 0-3)
#ifndef WPE32
#define WPE32 2                  // target waves per SIMD (register budget) of the 4-wave FP32-bulk kernels
#endif
#ifndef WPE64
#define WPE64 1                  // ... and of the 4-wave FP64 kernels
#endif
#ifndef WPE32_8
#define WPE32_8 4                // ... of the 8-wave FP32-bulk kernels (two workgroups per CU)
#endif
#ifndef WPE64_8
#define WPE64_8 2                // ... of the 8-wave FP64 kernels
#endif


// output row (j - 3) of lane t: wave w takes rows w, w+8, w+4, w+12, 16 lanes each.  The neighbour reads of a
// plane in LDS (row stride NX = 22 words) then put the two rows of each half wave 176 words apart, 16 banks
// modulo 32, so they meet no bank conflict, where rows 4w..4w+3 collided on 6 of 16 lanes per row pair.
__device__ __forceinline__ int output_row( int t )
{
   const int q = ( t & 63 ) >> 4;
   return ( t >> 6 ) + 8*( q & 1 ) + 4*( q >> 1 );
}

// The reconstruction reads half-step cells only in three slabs: the 16^2 output columns over z = 1..20, and the
// "arm" columns (i or j in {1,2,19,20}, the other in 3..18) over z = 3..18 -- 7,168 of the 8,000 cells the
// reference updates (CPU_FluidSolver_MHM.cpp:656-666 takes the whole 20^3 block).  Those cells need 22,656 of
// its 25,200 half-step faces.  The ghost work of waves 0-2 is laid out so that each wave runs one task kind
// over one z range and skips the rest with wave-uniform branches:
//   wave 0: the 64 arms next to the output block (i or j in {2,19}): half-step cell + low x and y faces
//   wave 1: the 64 outer arms (i or j in {1,20}): the same
//   wave 2: 40 single half-step faces (one x or y face per lane, read by an arm or output cell but owned by a
//           column that is neither): the low x face of column (21,j) at z = 3..18 and of column (19,j) at the
//           other planes, where the output cell (18,j) still needs it (and the y analogue); and the low x faces
//           of (19, {1,2,19,20}) and low y faces of ({1,2,19,20}, 19) at z = 3..18
// Arm column of lane t (waves 0-1); false = none.
__device__ __forceinline__ bool arm_column( int t, int &i, int &j )
{
   if ( t >= 128 ) { i = j = 0;  return false; }
   const int k = t & 63, a = 3 + ( k & 15 ), s = ( k >> 4 ) & 1;
   const int e = ( t < 64 ) ? ( s ? 19 : 2 ) : ( s ? 20 : 1 );
   if ( k < 32 ) { i = e;  j = a; }
   else          { i = a;  j = e; }
   return true;
}

// single half-step face of lane t (wave 2) at plane z: the column whose low face it is and whether the face is
// y-normal; false = no face at this plane
__device__ __forceinline__ bool face_task( int t, int z, int &c, bool &dy )
{
   const int k = t - 128;
   const bool mid = ( z >= 3 && z <= 18 );
   int i, j;
   dy = false;
   if ( k < 0 || k >= 40 ) return false;
   if ( k < 32 )
   {
      const int a = 3 + ( k & 15 ), e = mid ? 21 : 19;
      dy = ( k >= 16 );
      i = dy ? a : e;
      j = dy ? e : a;
   }
   else
   {
      const int q = k & 3, e = ( q < 2 ) ? 1 + q : 17 + q;     // 1, 2, 19, 20
      dy = ( k >= 36 );
      i = dy ? e : 19;
      j = dy ? 19 : e;
      if ( !mid ) return false;
   }
   c = i + NX*j;
   return true;
}

// ghost column next to the output block reconstructed by lane t of wave 3 (along y if dy), and whether the
// lane also owns the full-step flux on the face between that column and the block (columns 19 only)
__device__ __forceinline__ bool edge_column( int t, int &i, int &j, bool &dy, bool &face )
{
   const int k = t - 192;
   dy = ( k >= 32 );
   const int e = k & 31;
   face = ( e >= 16 );
   const int a = face ? 19 : 2, b = 3 + ( e & 15 );
   i = dy ? b : a;
   j = dy ? a : b;
   return ( k >= 0 && k < 64 );
}

// roles of a column (i,j) of the 22x22 footprint
struct Col {
   int  c;                   // i + 22 j
   bool w;                   // half-step cell (1..20)^2, owns the half-step z-face below it
   bool hx, hy;              // owns the half-step x (y) face on its low side
   bool rx, ry;              // needs x (y) reconstruction
   bool fx, fy;              // owns the full-step x (y) face on its low side
   bool core;                // output column (3..18)^2
   int  i, j;
};

__device__ __forceinline__ Col make_col( int i, int j )
{
   Col k;
   k.i = i;  k.j = j;  k.c = i + NX*j;
   const bool ii = ( i >= 1 && i <= 20 ), jj = ( j >= 1 && j <= 20 );
   k.w  = ii && jj;
   k.hx = ( i >= 1 && i <= 21 ) && jj;
   k.hy = ( j >= 1 && j <= 21 ) && ii;
   const bool ci = ( i >= 3 && i <= 18 ), cj = ( j >= 3 && j <= 18 );
   k.rx = ( i >= 2 && i <= 19 ) && cj;
   k.ry = ( j >= 2 && j <= 19 ) && ci;
   k.fx = ( i >= 3 && i <= 19 ) && cj;
   k.fy = ( j >= 3 && j <= 19 ) && ci;
   k.core = ci && cj;
   return k;
}

// cell quantities of one input cell: conserved + pressure (recovered in FP64), 1/rho, sound speed
template <typename R, typename IO>
__device__ __forceinline__ void derive_cell( const IO u[NV], R q[8], const Prm<R> &P )
{
   const double d = (double)u[0], mx = (double)u[1], my = (double)u[2], mz = (double)u[3], e = (double)u[4];
   const double eint = e - 0.5*( mx*mx + my*my + mz*mz )*rcp_pres<R>( d );
   const R p = floor_pres( (R)( eint*P.gm1_d ), P.min_pres );
   const R ir = rcp( (R)u[0] );
#  pragma unroll
   for (int v=0; v<NV; v++) q[v] = (R)u[v];
   q[5] = p;
   q[6] = ir;
   q[7] = sound_speed( P.gamma*p, q[0], ir );
}

// half-step cell update, floors and conversion to primitive (CPU_FluidSolver_MHM.cpp:846-933).  The update
// U - dt/(2dh) sum dF runs in the bulk type; the pressure recovery E - |m|^2/(2 rho) and the internal-energy
// floor run in FP64.
template <typename R>
__device__ __forceinline__ void half_cell( const R u[NV], const R dfxy[NV], const R hzl[NV], const R hzu[NV],
                                           const Prm<R> &P, R w[NV] )
{
   R c[NV];
   const R k = (R)P.dt_dh2;
#  pragma unroll
   for (int v=0; v<NV; v++) c[v] = u[v] - k*( dfxy[v] + ( hzu[v] - hzl[v] ) );
   c[0] = vmax( c[0], P.min_dens );
   const double d = (double)c[0], m1 = (double)c[1], m2 = (double)c[2], m3 = (double)c[3];
   const double ke = 0.5*( m1*m1 + m2*m2 + m3*m3 )*rcp_pres<R>( d );
   double e = (double)c[4];
   double ein = e - ke;
   const double eout = ( ein == ein ) ? __builtin_fmax( ein, P.min_eint_d ) : ein;
   if ( ein != eout ) { e = e - ein + eout;  ein = e - ke; }
   const R r = rcp( c[0] );
   w[0] = c[0];
   w[1] = c[1]*r;
   w[2] = c[2]*r;
   w[3] = c[3]*r;
   const double p = ein*P.gm1_d;
   w[4] = (R)( ( p == p ) ? __builtin_fmax( p, P.min_pres_d ) : p );
}

template <typename R>
__device__ __forceinline__ void ld_cell( const R (*a)[NC], int c, R q[8] )
{
#  pragma unroll
   for (int v=0; v<8; v++) q[v] = a[v][c];
}

template <int D, typename R>
__device__ __forceinline__ St<R> st_of( const R q[8] ) { return cell_state<D>( q, q[5], q[6], q[7] ); }

template <int D, typename R>
__device__ __forceinline__ St<R> st_of( const R (*a)[NC], int c ) { R q[8]; ld_cell( a, c, q ); return st_of<D>( q ); }

template <int D, typename R>
__device__ __forceinline__ St<R> st_face( const R (*a)[NC], int v0, int c, const Prm<R> &P )
{
   R w[NV];
#  pragma unroll
   for (int v=0; v<NV; v++) w[v] = a[v0 + v][c];
   return face_state<D>( w, P );
}

template <typename R>
__device__ __forceinline__ St<R> st_face_xy( const R (*a)[NC], int v0, int c, bool dy, const Prm<R> &P )
{
   R w[NV];
#  pragma unroll
   for (int v=0; v<NV; v++) w[v] = a[v0 + v][c];
   return face_state_xy( w, dy, P );
}

// LDS of one workgroup
template <typename R>
struct SmemRP {
   R a[8][NC];                  // input plane (rho,mx,my,mz,E,P,1/rho,c_s)
   R f[2*NV][NC];               // upper x (0-4) and y (5-9) face states of the reconstructed plane
   R b[2*NV][NC];               // half-step or full-step x (0-4) and y (5-9) fluxes
   R w[NV][NC];                 // half-step primitive plane read by the reconstruction
   R u3[3][NV][NK];             // lane-private ring: input cells of the own output column, planes t-3 .. t-1
   int fail, next;
   R cfl[8][4];       // fused CFL: per-patch maxima of the 4 waves (R is the I/O type)
};

// NW = 4: every lane owns an output column and one ghost role (waves 0-1 arms, wave 2 single faces, wave 3 edge
// columns), both marched in one loop.  NW = 8: waves 0-3 own the output columns only and waves 4-7 the ghost
// roles only (in the same wave order), each half in a loop of its own, so a lane keeps only its own half's z
// history in registers and two waves share each SIMD.  Both layouts run the same per-cell arithmetic in the
// same order, so their results are bitwise equal.
template <typename R, typename IO, int RS, int LIM, bool PERSIST, int NW>
__global__ void __launch_bounds__( 64*NW )
__attribute__(( amdgpu_waves_per_eu( NW == 8 ? ( sizeof(R) == 4 ? WPE32_8 : WPE64_8 ) : ( sizeof(R) == 4 ? WPE32 : WPE64 ) ) ))
mhm_rp_kernel( const IO *__restrict__ g_in, IO *__restrict__ g_out, IO *__restrict__ g_flux,
               int npg, int store_flux, int minmod_maxiter, R minmod_coeff, Prm<R> P0,
               IO *__restrict__ g_dt, double dh_safety0, Batch B )
{
   static_assert( sizeof( R ) == sizeof( IO ), "the input ring holds IO values in R storage" );
   static_assert( NW == 4 || NW == 8, "4 or 8 waves per workgroup" );
   __shared__ SmemRP<R> S;

   const int tid  = threadIdx.x;
   const int wave = __builtin_amdgcn_readfirstlane( tid >> 6 );   // wave-uniform role switch
   const int g    = ( NW == 4 ) ? tid : tid - 256;                // ghost-role index (0..255 on the ghost waves)
   const int gw   = ( NW == 4 ) ? wave : wave - 4;                // ghost wave: 0-1 arms, 2 single faces, 3 edges

   const Col K = make_col( 3 + tid % PS, 3 + output_row( tid ) );
   int oi, oj;
   arm_column( g, oi, oj );
   const Col O = make_col( oi, oj );
   int ei, ej;
   bool e_dy, e_face;
   const bool has_e = edge_column( g, ei, ej, e_dy, e_face );
   const int  e_c  = ei + NX*ej;
   const int  e_st = e_dy ? NX : 1;             // neighbour stride of the edge column's direction
   const int  e_vo = e_dy ? NV : 0;             // its face states / fluxes in S.f / S.b
   // lane-constant byte offsets of the global stores: output column in a 16^2 plane, and the coarse-fine flux
   // planes x = 0/8 (ffx), y = 0/8 (ffy) and x or y = 16 (edge lanes) at row K.j, K.i or the edge column
   const int  kxy    = ( K.i - 3 ) + PS*( K.j - 3 );
   const int  fx_off = ( ( ( K.i - 3 )/8 )*NV*PS*PS + ( K.j - 3 ) )*(int)sizeof( IO );
   const int  fy_off = ( ( 3 + ( K.j - 3 )/8 )*NV*PS*PS + ( K.i - 3 ) )*(int)sizeof( IO );
   const int  fe_off = ( ( e_dy ? 5 : 2 )*NV*PS*PS + ( ( e_dy ? ei : ej ) - 3 ) )*(int)sizeof( IO );
   const int  lim = P0.limiter;

   int base = 0;
   for (int item = 0; ; item++)
   {
   // next patch group: blockIdx.x, or the next PG of this workgroup's ticket of the persistent work queue (a ticket
   // covers B.claim consecutive PGs; the barrier before every PG also keeps the previous PG's LDS reads ahead of
   // the next PG's writes)
   int pg;
   if ( PERSIST )
   {
      const int sub = item % B.claim;
      __syncthreads();
      if ( sub == 0 )
      {
         if ( tid == 0 ) S.next = (int)atomicAdd( B.counter, 1u );
         __syncthreads();
         base = __builtin_amdgcn_readfirstlane( S.next )*B.claim;   // uniform: keep pointers and dt in SGPRs
      }
      pg = base + sub;
   }
   else
   {
      if ( item > 0 ) break;
      pg = blockIdx.x;
   }
   if ( pg >= npg ) break;

   constexpr int SZ = sizeof( IO );
   const IO *in  = g_in  + (size_t)pg*NV*NC*NX;
   IO       *out = g_out + (size_t)pg*NV*PS*PS*PS;
   IO       *flx = store_flux ? g_flux + (size_t)pg*9*NV*PS*PS : nullptr;
   const __amdgpu_buffer_rsrc_t b_in  = make_buf( in,  NV*NC*NX*SZ );
   const __amdgpu_buffer_rsrc_t b_out = make_buf( out, NV*PS*PS*PS*SZ );
   const __amdgpu_buffer_rsrc_t b_flx = make_buf( flx, flx ? 9*NV*PS*PS*SZ : 0 );
   Prm<R> P = P0;
   double dh_safety = dh_safety0;
   if ( B.pg_dt_dh )
   {
      const double dt = B.pg_dt_dh[2*pg], dh = B.pg_dt_dh[2*pg + 1];
      P.dt_dh2 = 0.5*dt/dh;  P.dt_dh = dt/dh;  dh_safety = B.safety*dh;
   }

   for (int iter = 0; ; iter++)
   {
      Prm<R> Q = P;
      {
         R cf = ( minmod_maxiter == 0 ) ? minmod_coeff : minmod_coeff - (R)iter*minmod_coeff/(R)minmod_maxiter;
         Q.coeff = vmax( cf, (R)0 );
      }
      if ( tid == 0 ) S.fail = 0;
      IO cfl[2] = { (IO)0, (IO)0 };            // fused CFL: max(|v_d|+c_s) of this column in patches z=0,1

      // the z-march; DK: this lane runs its output column, DG: its ghost role (both compile-time)
      auto march = [&]( auto DK, auto DG )
      {
         constexpr bool doK = decltype( DK )::value, doG = decltype( DG )::value;
         // lane state carried along z (output column K, half-step arm column O)
         R kU[8], kU1[8];                            // input cell of planes t and t-1 (+ pressure, 1/rho, c_s)
         R kHz[NV], kDh[NV];                         // half-step z-flux at t-3/2 ; half-step dFx+dFy of plane t-1
         R kW1[NV], kW2[NV], kW3[NV];                // half-step primitives of planes t-1, t-2, t-3
         R kFzR[NV], kFz[NV], kDf[NV];               // upper z face of plane r-1 ; full z-flux at r-3/2 ; full dFx+dFy of plane r-1
         R oU[8], oU1[8], oHz[NV], oDh[NV], oW[NV];  // arm column

         // prefetch of the next input plane by the ghost-role lanes (two cells per lane, coalesced)
         IO pf[2][NV];
         auto issue = [&]( int t )
         {
#           pragma unroll
            for (int k=0; k<2; k++)
            {
               const int c = g + k*256;
               if ( c < NC )
#                 pragma unroll
                  for (int v=0; v<NV; v++) buf_load( b_in, c*SZ, ( v*NC*NX + t*NC )*SZ, pf[k][v] );
            }
         };
         // input plane (prefetched) -> LDS, with pressure, 1/rho and c_s once per cell
         auto stage = [&]()
         {
#           pragma unroll
            for (int k=0; k<2; k++)
            {
               const int c = g + k*256;
               if ( c < NC )
               {
                  R q[8];
                  derive_cell( pf[k], q, Q );
#                 pragma unroll
                  for (int v=0; v<8; v++) S.a[v][c] = q[v];
               }
            }
         };
         __syncthreads();                            // the previous patch group / iteration is done with S.a
         if constexpr ( doG ) { issue( 0 );  stage();  issue( 1 ); }

         // two steps per iteration let the compiler rename the plane-to-plane register hand-offs (kU1 = kU, the
         // z-flux and face carries) instead of copying them; the Roe kernels stay rolled, their code is twice as large,
         // and so does the 8-wave FP32-bulk kernel, whose 128-register budget the unrolled schedule exceeds more
#        pragma unroll ( RS != 2 && ( NW == 4 || sizeof( R ) == 8 ) ? 2 : 1 )
         for (int t = 0; t < NX; t++)
         {
            const int o = t - 3;                    // output plane of this step
            const int r = t - 2;                    // reconstructed plane of this step
            const int ring = t % 3;                 // u3 slot of input plane t-3 (read at the update); plane t-1
                                                    // goes into slot (t+2) % 3 as soon as it leaves S.a
            __syncthreads();                        // [A] input plane t visible
            // the longest work of this interval issues first: with 4 waves the arm waves 0-1 (output column and
            // arm half step), with 8 the output-column waves 0-3; the others issue behind them
            if ( NW == 4 ? wave < 2 : wave < 4 ) __builtin_amdgcn_s_setprio( 3 ); else __builtin_amdgcn_s_setprio( 1 );
            // arm half step: cells of planes 3..18 (updated at steps 4..19), their z faces (steps 3..19) and their
            // low x/y faces (planes 3..18)
            const bool arm_ld = doG && ( gw < 2 ) && t >= 2 && t <= 19;
            const bool arm_z  = doG && ( gw < 2 ) && t >= 3 && t <= 19;
            const bool arm_xy = doG && ( gw < 2 ) && t >= 3 && t <= 18;

            if constexpr ( doK )
            {
#              pragma unroll
               for (int v=0; v<8; v++) kU1[v] = kU[v];
               ld_cell( S.a, K.c, kU );
#              pragma unroll
               for (int v=0; v<NV; v++) S.u3[( t + 2 ) % 3][v][tid] = kU1[v];     // read back at step t+2
            }
            if constexpr ( doG )
            {
#              pragma unroll
               for (int v=0; v<8; v++) oU1[v] = oU[v];
               if ( arm_ld ) ld_cell( S.a, O.c, oU );
            }

            // ---- half-step z-flux at t-1/2 and half-step cell update of plane t-1
            if ( doK && t >= 1 )
            {
               R fz[NV], w[NV];
               riemann<RS, 2>( st_of<2>( kU1 ), st_of<2>( kU ), Q, fz );
               if ( t >= 2 )
               {
                  half_cell( kU1, kDh, kHz, fz, Q, w );
#                 pragma unroll
                  for (int v=0; v<NV; v++) { kW3[v] = kW2[v];  kW2[v] = kW1[v];  kW1[v] = w[v]; }
               }
#              pragma unroll
               for (int v=0; v<NV; v++) kHz[v] = fz[v];
            }
            if ( arm_z )
            {
               R fz[NV];
               riemann<RS, 2>( st_of<2>( oU1 ), st_of<2>( oU ), Q, fz );
               if ( t >= 4 ) half_cell( oU1, oDh, oHz, fz, Q, oW );
#              pragma unroll
               for (int v=0; v<NV; v++) oHz[v] = fz[v];
            }

            // ---- half-step x and y fluxes of plane t (low face owned by each column)
            R fkx[NV], fox[NV], fky[NV], foy[NV];
            if ( doK && t >= 1 && t <= 20 )
            {
               riemann<RS, 0>( st_of<0>( S.a, K.c - 1 ), st_of<0>( kU ), Q, fkx );
               riemann<RS, 1>( st_of<1>( S.a, K.c - NX ), st_of<1>( kU ), Q, fky );
#              pragma unroll
               for (int v=0; v<NV; v++)
               {
                  S.b[v][K.c] = fkx[v];
                  S.b[NV + v][K.c] = fky[v];
               }
            }
            if ( arm_xy )
            {
               riemann<RS, 0>( st_of<0>( S.a, O.c - 1 ), st_of<0>( oU ), Q, fox );
               riemann<RS, 1>( st_of<1>( S.a, O.c - NX ), st_of<1>( oU ), Q, foy );
#              pragma unroll
               for (int v=0; v<NV; v++)
               {
                  S.b[v][O.c] = fox[v];
                  S.b[NV + v][O.c] = foy[v];
               }
            }
            if ( doG && gw == 2 && t >= 1 && t <= 20 )
            {
               int fc;
               bool fdy;
               if ( face_task( g, t, fc, fdy ) )
               {
                  R ql[8], qr[8], ff[NV];
                  ld_cell( S.a, fc - ( fdy ? NX : 1 ), ql );
                  ld_cell( S.a, fc, qr );
                  riemann_xy<RS>( cell_state_xy( ql, fdy ), cell_state_xy( qr, fdy ), Q, fdy, ff );
#                 pragma unroll
                  for (int v=0; v<NV; v++) S.b[( fdy ? NV : 0 ) + v][fc] = ff[v];
               }
            }

            // ---- PLM reconstruction of plane r (x/y upper faces to LDS, z faces stay in the lane); the x/y
            //      neighbours come from the half-step plane r in S.w
            R kFL[3][NV], kFRz[NV], eFL[NV];
            if ( doK && r >= 2 && r <= 19 )
            {
#              pragma unroll
               for (int v=0; v<NV; v++) plm_face<LIM>( kW3[v], kW2[v], kW1[v], lim, Q.coeff, kFL[2][v], kFRz[v] );
               if ( r >= 3 && r <= 18 )
#                 pragma unroll
                  for (int v=0; v<NV; v++)
                  {
                     R fr;
                     plm_face<LIM>( S.w[v][K.c - 1], kW2[v], S.w[v][K.c + 1], lim, Q.coeff, kFL[0][v], fr );
                     S.f[v][K.c] = fr;
                     plm_face<LIM>( S.w[v][K.c - NX], kW2[v], S.w[v][K.c + NX], lim, Q.coeff, kFL[1][v], fr );
                     S.f[NV + v][K.c] = fr;
                  }
            }
            if ( doG && has_e && r >= 3 && r <= 18 )
#              pragma unroll
               for (int v=0; v<NV; v++)
               {
                  R fr;
                  plm_face<LIM>( S.w[v][e_c - e_st], S.w[v][e_c], S.w[v][e_c + e_st], lim, Q.coeff, eFL[v], fr );
                  S.f[e_vo + v][e_c] = fr;
               }
            __syncthreads();                        // [H] half-step x/y fluxes visible; S.w plane r no longer read
            __builtin_amdgcn_s_setprio( 0 );

            // flux differences; the lane's own low-face fluxes are read back from LDS rather than held in registers
            // across the reconstruction (the stored values, so both launch shapes round alike)
            if ( doK && t >= 1 && t <= 20 )
#              pragma unroll
               for (int v=0; v<NV; v++) kDh[v] = ( S.b[v][K.c + 1] - S.b[v][K.c] ) + ( S.b[NV + v][K.c + NX] - S.b[NV + v][K.c] );
            if ( arm_xy )
#              pragma unroll
               for (int v=0; v<NV; v++) oDh[v] = ( S.b[v][O.c + 1] - S.b[v][O.c] ) + ( S.b[NV + v][O.c + NX] - S.b[NV + v][O.c] );
            // the half-step plane t-1 replaces plane t-2 in LDS for the next reconstruction
            if ( doK && t >= 2 )
#              pragma unroll
               for (int v=0; v<NV; v++) S.w[v][K.c] = kW1[v];
            if ( arm_z && t >= 4 )
#              pragma unroll
               for (int v=0; v<NV; v++) S.w[v][O.c] = oW[v];

            // ---- full-step fluxes: z at r-1/2 (lane-local), x and y of plane r through LDS
            R fzn[NV];
            if ( doK && r >= 3 && r <= 19 )
            {
               riemann<RS, 2>( face_state<2>( kFzR, Q ), face_state<2>( kFL[2], Q ), Q, fzn );
               if ( flx && ( r - 3 ) % 8 == 0 )
#                 pragma unroll
                  for (int v=0; v<NV; v++)
                     buf_store( (IO)fzn[v], b_flx, kxy*SZ, ( ( 6 + ( r - 3 )/8 )*NV + v )*PS*PS*SZ );
            }
            if ( doK && r >= 2 )
#              pragma unroll
               for (int v=0; v<NV; v++) kFzR[v] = kFRz[v];

            R dfn[NV];
            if ( r >= 3 && r <= 18 )
            {
               __syncthreads();                     // [R] the half-step fluxes in S.b are read
               if constexpr ( doK )
               {
                  R ffx[NV], ffy[NV];
                  riemann<RS, 0>( st_face<0>( S.f, 0, K.c - 1, Q ), face_state<0>( kFL[0], Q ), Q, ffx );
                  riemann<RS, 1>( st_face<1>( S.f, NV, K.c - NX, Q ), face_state<1>( kFL[1], Q ), Q, ffy );
#                 pragma unroll
                  for (int v=0; v<NV; v++)
                  {
                     S.b[v][K.c] = ffx[v];
                     S.b[NV + v][K.c] = ffy[v];
                  }
                  if ( flx )
                  {
                     if ( ( K.i - 3 ) % 8 == 0 )
#                       pragma unroll
                        for (int v=0; v<NV; v++)
                           buf_store( (IO)ffx[v], b_flx, fx_off, ( v*PS*PS + PS*( r - 3 ) )*SZ );
                     if ( ( K.j - 3 ) % 8 == 0 )
#                       pragma unroll
                        for (int v=0; v<NV; v++)
                           buf_store( (IO)ffy[v], b_flx, fy_off, ( v*PS*PS + PS*( r - 3 ) )*SZ );
                  }
               }
               if ( doG && has_e && e_face )
               {
                  R fe[NV];
                  riemann_xy<RS>( st_face_xy( S.f, e_vo, e_c - e_st, e_dy, Q ), face_state_xy( eFL, e_dy, Q ), Q, e_dy, fe );
#                 pragma unroll
                  for (int v=0; v<NV; v++) S.b[e_vo + v][e_c] = fe[v];
                  if ( flx )                        // planes x = 16 (2) and y = 16 (5)
#                    pragma unroll
                     for (int v=0; v<NV; v++)
                        buf_store( (IO)fe[v], b_flx, fe_off, ( v*PS*PS + PS*( r - 3 ) )*SZ );
               }
               __syncthreads();                     // [F] full-step fluxes visible
               if constexpr ( doK )                 // own fluxes read back, as for the half step
#                 pragma unroll
                  for (int v=0; v<NV; v++) dfn[v] = ( S.b[v][K.c + 1] - S.b[v][K.c] ) + ( S.b[NV + v][K.c + NX] - S.b[NV + v][K.c] );
            }

            if constexpr ( doK )
            {
               // ---- full-step update and output of plane o = t-3
               if ( o >= 3 && o <= 18 )
               {
                  IO res[NV], uo[NV];
#                 pragma unroll
                  for (int v=0; v<NV; v++) uo[v] = S.u3[ring][v][tid];
                  const bool bad = full_cell( uo, kDf, kFz, fzn, Q, res );
#                 pragma unroll
                  for (int v=0; v<NV; v++) buf_store( res[v], b_out, kxy*SZ, ( v*PS*PS*PS + PS*PS*( o - 3 ) )*SZ );
                  if ( bad ) S.fail = 1;
                  if ( g_dt ) cfl[ ( o - 3 )/8 ] = vmax( cfl[ ( o - 3 )/8 ], cell_cfl( res, (double)Q.gamma, Q.gm1_d, Q.min_pres_d ) );
               }
               if ( r >= 3 && r <= 19 )
#                 pragma unroll
                  for (int v=0; v<NV; v++) kFz[v] = fzn[v];
               if ( r >= 3 && r <= 18 )
#                 pragma unroll
                  for (int v=0; v<NV; v++) kDf[v] = dfn[v];
            }

            // ---- input plane t+1 into LDS: every read of S.a of this step came before barrier [H]
            if ( doG && t + 1 < NX )
            {
               stage();
               if ( t + 2 < NX ) issue( t + 2 );
            }
         } // for t
      };
      using T_ = std::integral_constant<bool, true>;
      using F_ = std::integral_constant<bool, false>;
      if constexpr ( NW == 4 ) march( T_(), T_() );
      else if ( wave < 4 )     march( T_(), F_() );
      else                     march( F_(), T_() );

      if ( g_dt ) fused_dt( cfl, ( K.i - 3 )/8 + 2*( ( K.j - 3 )/8 ), S.cfl, g_dt + (size_t)pg*8, dh_safety );
      __syncthreads();
      const bool redo = ( S.fail != 0 ) && ( iter + 1 <= minmod_maxiter );
      __syncthreads();
      if ( !redo ) break;
   } // MinMod retry
   } // work items
}

} // namespace

template <typename R, typename IO, int RS>
static void launch_rs( const IO *in, IO *out, IO *flux, int npg, int store_flux, int minmod_maxiter,
                       double minmod_coeff, const Prm<R> &P, IO *dt, double dh_safety, const Batch &B, int grid_blocks, hipStream_t stream )
{
   const int nw = ( B.waves == 4 ) ? 4 : 8;
   const dim3 grid( grid_blocks > 0 ? grid_blocks : npg ), block( 64*nw );
   const bool pers = ( B.counter != nullptr );
#  define GA_LAUNCH( LIM_, PERS_, NW_ ) hipLaunchKernelGGL( ( mhm_rp_kernel<R, IO, RS, LIM_, PERS_, NW_> ), grid, block, 0, stream, \
                                                           in, out, flux, npg, store_flux, minmod_maxiter, (R)minmod_coeff, P, dt, dh_safety, B )
#  define GA_SHAPE( LIM_, PERS_ ) { if ( nw == 4 ) GA_LAUNCH( LIM_, PERS_, 4 ); else GA_LAUNCH( LIM_, PERS_, 8 ); }
   if ( kernel_limiter( P.limiter, minmod_coeff ) == 4 ) { if ( pers ) GA_SHAPE( 4, true ) else GA_SHAPE( 4, false ) }
   else                  { if ( pers ) GA_SHAPE( 0, true ) else GA_SHAPE( 0, false ) }
#  undef GA_SHAPE
#  undef GA_LAUNCH
}

template <typename R, typename IO>
int launch_mhm_rp( const IO *in, IO *out, IO *flux, int npg, int rsolver, int store_flux, int minmod_maxiter,
                   double minmod_coeff, const Prm<R> &P, IO *dt, double dh_safety, const Batch &B, int grid_blocks, hipStream_t stream )
{
   if ( npg <= 0 ) return 0;
   switch ( rsolver )
   {
      case 4: launch_rs<R, IO, 4>( in, out, flux, npg, store_flux, minmod_maxiter, minmod_coeff, P, dt, dh_safety, B, grid_blocks, stream ); break;
      case 3: launch_rs<R, IO, 3>( in, out, flux, npg, store_flux, minmod_maxiter, minmod_coeff, P, dt, dh_safety, B, grid_blocks, stream ); break;
      case 2: launch_rs<R, IO, 2>( in, out, flux, npg, store_flux, minmod_maxiter, minmod_coeff, P, dt, dh_safety, B, grid_blocks, stream ); break;
      default: return -1;
   }
   return ( hipGetLastError() == hipSuccess ) ? 0 : -2;
}

template int launch_mhm_rp<double, double>( const double*, double*, double*, int, int, int, int, double, const Prm<double>&, double*, double, const Batch&, int, hipStream_t );
template int launch_mhm_rp<float,  float >( const float*,  float*,  float*,  int, int, int, int, double, const Prm<float>&,  float*, double, const Batch&, int, hipStream_t );

} // namespace gamer_amd
