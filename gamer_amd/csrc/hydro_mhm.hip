// MHM (MUSCL-Hancock with Hancock face prediction, no Riemann prediction step) patch-group kernel for gfx950.
//
// Input per patch group: 20^3 cells (16^3 interior + 2 ghost zones).  Reference: CPU_FluidSolver_MHM.cpp:487-589
// with Con2Pri_Yes, i.e. Hydro_DataReconstruction (PLM + Hydro_HancockPredict, CPU_Shared_DataReconstruction.cpp:
// 188-689, 2063-2160) -> Hydro_ComputeFlux -> Hydro_StoreIntFlux -> Hydro_FullStepUpdate.
//
// Same organisation as the MHM_RP kernel: one 256-lane workgroup per patch group marching along z; lane t owns
// output column (2 + t%16, 2 + row), wave w taking rows w, w+4, w+8, w+12, and, on waves 0-1, one column of the
// reconstruction ring (i or j in {1,18}).
// Every lane that reconstructs a cell keeps that column's primitive planes r-1, r, r+1 and the conserved cell
// value of plane r in registers; x/y neighbours come from a double-buffered primitive plane in LDS, the conserved
// input plane is staged in LDS too, and a lane-private LDS slot keeps the own output column's input until its update
// two planes later (every input value is read from HBM once).
// Plane schedule at step t (input plane t just loaded): primitive conversion of plane t, PLM + Hancock
// prediction of plane r = t-1, full-step fluxes of plane r (x/y) and at r-1/2 (z), full-step update of t-2.

#include "hydro_math.hpp"
#include "hydro_kernels.hpp"

namespace gamer_amd {

namespace {

constexpr int NX = 20;          // FLU_NXT for MHM
constexpr int NC = NX*NX;
constexpr int PS = 16;
constexpr int NT = 256;

template <typename R>
struct SmemM {
   R w[2][NV][NC];    // primitive planes t-1 and t (ping-pong)
   R f[2*NV][NC];     // upper x (0-4) and y (5-9) face states (conserved, after the Hancock prediction)
   R b[2*NV][NC];     // full-step x (0-4) and y (5-9) fluxes
   R u[NV][NC];       // conserved input plane t
   R u2[NV][NT];      // lane-private: input cells of the own output column, plane t-2
   int fail, next;
   R cfl[8][4];       // fused CFL: per-patch maxima of the 4 waves (R is the I/O type)
};

// reconstruction-ring column of lane t (i or j in {1,18}, the corners need nothing); false = none
__device__ __forceinline__ bool ring_column( int t, int &i, int &j )
{
   if ( t < 16 )  { i = 1;          j = 2 + t;       return true; }
   if ( t < 32 )  { i = 18;         j = 2 + t - 16;  return true; }
   if ( t >= 64 && t < 80 ) { i = 2 + t - 64; j = 1;  return true; }
   if ( t >= 80 && t < 96 ) { i = 2 + t - 80; j = 18; return true; }
   i = j = 0;
   return false;
}

// primitive variables of an input cell, pressure recovered in FP64 (Hydro_Con2Pri, FluUtility.cpp:209-320)
template <typename R, typename IO>
__device__ __forceinline__ void prim_of( const IO u[NV], R w[NV], const Prm<R> &P )
{
   const double d = (double)u[0], ir = rcp_pres<R>( d );
   const double eint = (double)u[4] - 0.5*( (double)u[1]*u[1] + (double)u[2]*u[2] + (double)u[3]*u[3] )*ir;
   const R rr = rcp( (R)u[0] );
   w[0] = (R)u[0];
   w[1] = (R)u[1]*rr;  w[2] = (R)u[2]*rr;  w[3] = (R)u[3]*rr;
   w[4] = floor_pres( (R)( eint*P.gm1_d ), P.min_pres );
}

// full PLM reconstruction + Hancock prediction of one cell: fc[6] conserved face states
template <int LIM, typename R, typename IO>
__device__ __forceinline__ void recon_cell( const R (*wl)[NC], int c, const R wm[NV], const R w0[NV], const R wp[NV],
                                            const R u0[NV], const Prm<R> &P, R fc[6][NV] )
{
   R fp[6][NV];
#  pragma unroll
   for (int v=0; v<NV; v++)
   {
      plm_face<LIM>( wl[v][c - 1],  w0[v], wl[v][c + 1],  P.limiter, P.coeff, fp[0][v], fp[1][v] );
      plm_face<LIM>( wl[v][c - NX], w0[v], wl[v][c + NX], P.limiter, P.coeff, fp[2][v], fp[3][v] );
      plm_face<LIM>( wm[v],         w0[v], wp[v],         P.limiter, P.coeff, fp[4][v], fp[5][v] );
   }
#  pragma unroll
   for (int f=0; f<6; f++) pri2con( fp[f], fc[f], P );
   hancock<R, IO>( fc, fp, u0, P );
}

template <typename R, typename IO, int RS, int LIM, bool PERSIST>
__global__ void __launch_bounds__( NT ) __attribute__(( amdgpu_waves_per_eu( sizeof(R) == 4 ? 2 : 1 ) ))
mhm_kernel( const IO *__restrict__ g_in, IO *__restrict__ g_out, IO *__restrict__ g_flux,
            int npg, int store_flux, int minmod_maxiter, R minmod_coeff, Prm<R> P0,
               IO *__restrict__ g_dt, double dh_safety0, Batch B )
{
   static_assert( sizeof( R ) == sizeof( IO ), "the input planes keep IO values in R storage" );
   __shared__ SmemM<R> S;

   const int tid = threadIdx.x;

   // wave w takes output rows w, w+4, w+8, w+12 (16 lanes each), 80 words apart in LDS (row stride NX = 20): the
   // two rows of each half wave lie 16 banks apart modulo 32 and the four rows of the wave cover 64 consecutive
   // banks modulo 64, so a neighbour read of a plane meets no bank conflict; rows 4w..4w+3 partly overlapped
   const int ki = 2 + tid % PS, kj = 2 + ( tid >> 6 ) + 4*( ( tid & 63 ) >> 4 ), kc = ki + NX*kj;
   int oi, oj;
   const bool has_o = ring_column( tid, oi, oj );
   const int  oc = oi + NX*oj;
   const bool ofx = has_o && oi == 18;                      // owns the full-step x face (17,18)
   const bool ofy = has_o && oj == 18;
   // lane-constant byte offsets of the global stores (output column in a 16^2 plane; coarse-fine flux planes
   // x = 0/8, y = 0/8 at row kj / ki; x = 16 and y = 16 for the ring lanes that own those faces)
   const int  kxy    = ( ki - 2 ) + PS*( kj - 2 );
   const int  fx_off = ( ( ( ki - 2 )/8 )*NV*PS*PS + ( kj - 2 ) )*(int)sizeof( IO );
   const int  fy_off = ( ( 3 + ( kj - 2 )/8 )*NV*PS*PS + ( ki - 2 ) )*(int)sizeof( IO );
   const int  ox_off = ( 2*NV*PS*PS + ( oj - 2 ) )*(int)sizeof( IO );
   const int  oy_off = ( 5*NV*PS*PS + ( oi - 2 ) )*(int)sizeof( IO );

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

      R kWm[NV], kW0[NV], kWp[NV], kU0[NV], kUp[NV];        // core column: primitive planes t-2, t-1, t; input t-1, t
      R oWm[NV], oW0[NV], oWp[NV], oU0[NV], oUp[NV];        // ring column
      R kFzR[NV], kFz[NV], kDf[NV];

      IO pf[2][NV];
      auto issue = [&]( int t )
      {
#        pragma unroll
         for (int k=0; k<2; k++)
         {
            const int c = tid + k*NT;
            if ( c < NC )
#              pragma unroll
               for (int v=0; v<NV; v++) buf_load( b_in, c*SZ, ( v*NC*NX + t*NC )*SZ, pf[k][v] );
         }
      };
      issue( 0 );

      // two steps per iteration: the LDS slot (t & 1) becomes a constant and the plane-to-plane register hand-offs
      // are renamed instead of copied; the Roe kernels stay rolled, their code is twice as large
#     pragma unroll ( RS != 2 ? 2 : 1 )
      for (int t = 0; t < NX; t++)
      {
         const int slot = t & 1;
         // ---- plane t -> primitive plane in LDS; own columns' conserved and primitive values to registers
         __syncthreads();
#        pragma unroll
         for (int k=0; k<2; k++)
         {
            const int c = tid + k*NT;
            if ( c < NC )
            {
               R w[NV];
               prim_of( pf[k], w, Q );
#              pragma unroll
               for (int v=0; v<NV; v++) { S.w[slot][v][c] = w[v];  S.u[v][c] = (R)pf[k][v]; }
            }
         }
         if ( t + 1 < NX ) issue( t + 1 );
         const int o = t - 2;
#        pragma unroll
         for (int v=0; v<NV; v++) S.u2[v][tid] = kU0[v];          // input plane t-2 of the own column
         __syncthreads();
#        pragma unroll
         for (int v=0; v<NV; v++)
         {
            kU0[v] = kUp[v];  kUp[v] = S.u[v][kc];
            if ( has_o ) { oU0[v] = oUp[v];  oUp[v] = S.u[v][oc]; }
            kWm[v] = kW0[v];  kW0[v] = kWp[v];  kWp[v] = S.w[slot][v][kc];
            if ( has_o ) { oWm[v] = oW0[v];  oW0[v] = oWp[v];  oWp[v] = S.w[slot][v][oc]; }
         }

         // ---- PLM + Hancock of plane r = t-1
         const int r = t - 1;
         R kF[6][NV], oF[6][NV];
         if ( r >= 1 && r <= 18 )
         {
            recon_cell<LIM, R, IO>( S.w[slot ^ 1], kc, kWm, kW0, kWp, kU0, Q, kF );
#           pragma unroll
            for (int v=0; v<NV; v++) { S.f[v][kc] = kF[1][v];  S.f[NV + v][kc] = kF[3][v]; }
            if ( has_o && r >= 2 && r <= 17 )
            {
               recon_cell<LIM, R, IO>( S.w[slot ^ 1], oc, oWm, oW0, oWp, oU0, Q, oF );
#              pragma unroll
               for (int v=0; v<NV; v++) { S.f[v][oc] = oF[1][v];  S.f[NV + v][oc] = oF[3][v]; }
            }
         }
         __syncthreads();

         // ---- full-step fluxes: z at r-1/2 (lane-local), x and y of plane r through LDS
         R fzn[NV];
         if ( r >= 2 && r <= 18 )
         {
            riemann<RS, 2>( cons_state<2>( kFzR, Q ), cons_state<2>( kF[4], Q ), Q, fzn );
            if ( flx && ( r - 2 ) % 8 == 0 )
#              pragma unroll
               for (int v=0; v<NV; v++)
                  buf_store( (IO)fzn[v], b_flx, kxy*SZ, ( ( 6 + ( r - 2 )/8 )*NV + v )*PS*PS*SZ );
         }
         if ( r >= 1 )
#           pragma unroll
            for (int v=0; v<NV; v++) kFzR[v] = kF[5][v];

         R dfn[NV];
         if ( r >= 2 && r <= 17 )
         {
            R fx[NV], fy[NV], fo[NV], q[NV];
#           pragma unroll
            for (int v=0; v<NV; v++) q[v] = S.f[v][kc - 1];
            riemann<RS, 0>( cons_state<0>( q, Q ), cons_state<0>( kF[0], Q ), Q, fx );
#           pragma unroll
            for (int v=0; v<NV; v++) q[v] = S.f[NV + v][kc - NX];
            riemann<RS, 1>( cons_state<1>( q, Q ), cons_state<1>( kF[2], Q ), Q, fy );
            if ( ofx )
            {
#              pragma unroll
               for (int v=0; v<NV; v++) q[v] = S.f[v][oc - 1];
               riemann<RS, 0>( cons_state<0>( q, Q ), cons_state<0>( oF[0], Q ), Q, fo );
            }
            if ( ofy )
            {
#              pragma unroll
               for (int v=0; v<NV; v++) q[v] = S.f[NV + v][oc - NX];
               riemann<RS, 1>( cons_state<1>( q, Q ), cons_state<1>( oF[2], Q ), Q, fo );
            }
#           pragma unroll
            for (int v=0; v<NV; v++)
            {
               S.b[v][kc] = fx[v];
               S.b[NV + v][kc] = fy[v];
               if ( ofx ) S.b[v][oc] = fo[v];
               if ( ofy ) S.b[NV + v][oc] = fo[v];
            }
            if ( flx )
            {
               if ( ( ki - 2 ) % 8 == 0 )
#                 pragma unroll
                  for (int v=0; v<NV; v++)
                     buf_store( (IO)fx[v], b_flx, fx_off, ( v*PS*PS + PS*( r - 2 ) )*SZ );
               if ( ( kj - 2 ) % 8 == 0 )
#                 pragma unroll
                  for (int v=0; v<NV; v++)
                     buf_store( (IO)fy[v], b_flx, fy_off, ( v*PS*PS + PS*( r - 2 ) )*SZ );
               if ( ofx )
#                 pragma unroll
                  for (int v=0; v<NV; v++) buf_store( (IO)fo[v], b_flx, ox_off, ( v*PS*PS + PS*( r - 2 ) )*SZ );
               if ( ofy )
#                 pragma unroll
                  for (int v=0; v<NV; v++) buf_store( (IO)fo[v], b_flx, oy_off, ( v*PS*PS + PS*( r - 2 ) )*SZ );
            }
            __syncthreads();
#           pragma unroll
            for (int v=0; v<NV; v++) dfn[v] = ( S.b[v][kc + 1] - fx[v] ) + ( S.b[NV + v][kc + NX] - fy[v] );
         }

         // ---- full-step update and output of plane o = t-2
         if ( o >= 2 && o <= 17 )
         {
            IO res[NV], uo[NV];
#           pragma unroll
            for (int v=0; v<NV; v++) uo[v] = S.u2[v][tid];
            const bool bad = full_cell( uo, kDf, kFz, fzn, Q, res );
#           pragma unroll
            for (int v=0; v<NV; v++) buf_store( res[v], b_out, kxy*SZ, ( v*PS*PS*PS + PS*PS*( o - 2 ) )*SZ );
            if ( bad ) S.fail = 1;
            if ( g_dt ) cfl[ ( o - 2 )/8 ] = vmax( cfl[ ( o - 2 )/8 ], cell_cfl( res, (double)Q.gamma, Q.gm1_d, Q.min_pres_d ) );
         }
         if ( r >= 2 && r <= 18 )
#           pragma unroll
            for (int v=0; v<NV; v++) kFz[v] = fzn[v];
         if ( r >= 2 && r <= 17 )
#           pragma unroll
            for (int v=0; v<NV; v++) kDf[v] = dfn[v];
      } // for t

      if ( g_dt ) fused_dt( cfl, ( ki - 2 )/8 + 2*( ( kj - 2 )/8 ), S.cfl, g_dt + (size_t)pg*8, dh_safety );
      __syncthreads();
      const bool redo = ( S.fail != 0 ) && ( iter + 1 <= minmod_maxiter );
      __syncthreads();
      if ( !redo ) break;
   } // MinMod retry
   } // work items
}

template <typename R, typename IO, int RS>
void launch_rs( const IO *in, IO *out, IO *flux, int npg, int store_flux, int minmod_maxiter, double minmod_coeff,
                const Prm<R> &P, IO *dt, double dh_safety, const Batch &B, int grid_blocks, hipStream_t stream )
{
   const dim3 grid( grid_blocks > 0 ? grid_blocks : npg ), block( NT );
   const bool pers = ( B.counter != nullptr );
#  define GA_LAUNCH( LIM_, PERS_ ) hipLaunchKernelGGL( ( mhm_kernel<R, IO, RS, LIM_, PERS_> ), grid, block, 0, stream, \
                                                      in, out, flux, npg, store_flux, minmod_maxiter, (R)minmod_coeff, P, dt, dh_safety, B )
   if ( kernel_limiter( P.limiter, minmod_coeff ) == 4 ) { if ( pers ) GA_LAUNCH( 4, true ); else GA_LAUNCH( 4, false ); }
   else                  { if ( pers ) GA_LAUNCH( 0, true ); else GA_LAUNCH( 0, false ); }
#  undef GA_LAUNCH
}

} // namespace

template <typename R, typename IO>
int launch_mhm( const IO *in, IO *out, IO *flux, int npg, int rsolver, int store_flux, int minmod_maxiter,
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

// The build compiles this file twice, once per precision (MHM_PRECISION 64 or 32), so that each gets the
// instruction scheduler that suits it; without the macro both precisions are instantiated.
#if !defined( MHM_PRECISION ) || MHM_PRECISION == 64
template int launch_mhm<double, double>( const double*, double*, double*, int, int, int, int, double, const Prm<double>&, double*, double, const Batch&, int, hipStream_t );
#endif
#if !defined( MHM_PRECISION ) || MHM_PRECISION == 32
template int launch_mhm<float,  float >( const float*,  float*,  float*,  int, int, int, int, double, const Prm<float>&,  float*, double, const Batch&, int, hipStream_t );
#endif

} // namespace gamer_amd
