// Flu_Close and the AMR level synchronisation on the device (the steps after the patch-group solver):
//
//   correct_unphysical_kernel   CorrectUnphysical (src/Fluid/Flu_Close.cpp:542-1297) with the Unphysical() check of
//                               :412-535: one lane per output cell; cells that pass the check leave at once, the rare
//                               failing cells are recomputed from the solver input with first-order fluxes (unsplit 3D,
//                               then the x->y->z split sweep over the 3^3 neighbourhood for 3D1D), floored, re-checked,
//                               and their first-order fluxes replace the coarse-fine flux planes they touch.
//   store_flux_kernel           StoreFlux (Flu_Close.cpp:234-307): dt * flux of the patch faces -> FP64 flux registers
//   correct_flux_kernel         CorrectFlux (Flu_Close.cpp:321-395): coarse register -= dt/4 * sum of the 4 fine fluxes,
//                               one lane per coarse face cell, the four terms in the reference's order
//   fixup_flux_kernel           Flu_FixUp_Flux (src/Fluid/Flu_FixUp_Flux.cpp): U += register * (-/+ 1/dh) on the cells
//                               of registered patch faces, faces applied in s order per cell, with the positivity test
//   fixup_restrict_kernel       Flu_FixUp_Restrict (src/Fluid/Flu_FixUp_Restrict.cpp): father cell = mean of its 8 son
//                               cells, then the internal-energy floor
//   close_patches_kernel        the copy of patch-group outputs into their 8 patches (Flu_Close.cpp:133-160), one
//                               aligned 4-vector per thread
//
// Patch data use GAMER's per-patch layout [npatch][5][8][8][8]; patch LocalID offsets follow TABLE_02
// (src/Tables/Table_02.cpp:22-70).  Flux registers are FP64 [nreg][5][8][8] in both precisions (the reference keeps
// them in `real`): the flux-correction accumulation is done in double.

#include "hydro_math.hpp"
#include "hydro_kernels.hpp"

namespace gamer_amd {

namespace {

constexpr int PS = 16, H = 8, P3 = H*H*H;

__device__ __forceinline__ void local_xyz( int lid, int o[3] )
{
   o[0] = ( lid == 1 || lid == 4 || lid == 6 || lid == 7 );
   o[1] = ( lid == 2 || lid == 4 || lid == 5 || lid == 7 );
   o[2] = ( lid == 3 || lid == 5 || lid == 6 || lid == 7 );
}

__device__ __forceinline__ int local_id( int ox, int oy, int oz )
{
   const int t[8] = { 0, 1, 2, 4, 3, 6, 5, 7 };       // [z][y][x]
   return t[ ( oz*2 + oy )*2 + ox ];
}

template <typename R> __device__ __forceinline__ bool finite( R x ) { return __builtin_isfinite( x ); }

// Unphysical() for HYDRO without dual energy / MHD / unit checks; `real` arithmetic as the reference
template <typename R>
__device__ __forceinline__ bool close_unphysical( const R f[NV], bool eint_mode, const CorrPrm &c )
{
   bool bad = false;
#  pragma unroll
   for (int v=0; v<NV; v++) bad |= !finite( f[v] );
   if ( bad || f[0] < (R)c.min_dens ) return true;
   if ( !eint_mode && ( f[4] < (R)c.min_eint || f[4] != f[4] ) ) return true;
   const R e = f[4] - (R)0.5*( f[1]*f[1] + f[2]*f[2] + f[3]*f[3] )/f[0];
   if ( eint_mode && ( e < (R)c.min_eint || !finite( e ) ) ) return true;
   if ( c.check_pres )
   {
      const R p = (R)( (double)e*c.gm1 );
      if ( p < (R)c.min_pres || !finite( p ) ) return true;
   }
   return false;
}

// first-order interface solve of the correction: the solver itself, without the NaN rescue
template <int RS, int D, typename R>
__device__ __forceinline__ void rs1( const R L[NV], const R Rr[NV], const Prm<R> &P, R out[NV] )
{
   const St<R> a = cons_state<D>( L, P ), b = cons_state<D>( Rr, P );
   R f[NV];
   if      ( RS == 4 ) rs_hllc( a, b, P.min_pres, f );
   else if ( RS == 3 ) rs_hlle( a, b, f );
   else if ( RS == 1 ) rs_roe<R, false>( a, b, P, f );     // Roe of a build whose RSOLVER is not ROE
   else                rs_roe( a, b, P, f );
   out[0] = f[0];
   out[4] = f[4];
   out[1 + D]             = f[1];
   out[1 + ( D + 1 ) % 3] = f[2];
   out[1 + ( D + 2 ) % 3] = f[3];
}

template <int RS, typename R>
__device__ __forceinline__ void rs1_dir( int d, const R L[NV], const R Rr[NV], const Prm<R> &P, R out[NV] )
{
   if      ( d == 0 ) rs1<RS, 0>( L, Rr, P, out );
   else if ( d == 1 ) rs1<RS, 1>( L, Rr, P, out );
   else               rs1<RS, 2>( L, Rr, P, out );
}

// Four waves per SIMD (128 VGPRs): nearly every lane only loads and tests its cell, and the rare correction path
// (the 3^3 split-sweep block) spills instead of holding the whole kernel at the 2 waves per SIMD its 246 VGPRs
// allowed.  On the 4096-PG bench output with 4 damaged cells per patch group (3D1D, HLLE) the kernel went from
// 0.328 to 0.219 ms (FP32) and 0.271 to 0.242 ms (FP64), bitwise equal; 8 waves per SIMD (64 VGPRs) was slower in FP64.
template <typename R, int RS>
__global__ void __launch_bounds__( 256 ) __attribute__(( amdgpu_waves_per_eu( 4 ) ))
correct_unphysical_kernel( const R *__restrict__ g_in, R *__restrict__ g_out, R *__restrict__ g_flux, int npg,
                           CorrPrm c, Prm<R> P, R dt_dh, int *counters )
{
   const int NX = c.NX, N3 = NX*NX*NX, ghost = c.ghost;
   const int didx[3] = { 1, NX, NX*NX };
   const long total = (long)npg*PS*PS*PS;
   for (long e = blockIdx.x*(long)blockDim.x + threadIdx.x; e < total; e += (long)gridDim.x*blockDim.x)
   {
      const int pg = (int)( e/( PS*PS*PS ) ), io = (int)( e%( PS*PS*PS ) );
      const int ijk[3] = { io % PS, ( io/PS ) % PS, io/( PS*PS ) };
      R *out = g_out + (size_t)pg*NV*PS*PS*PS;
      R Out[NV];
#     pragma unroll
      for (int v=0; v<NV; v++) Out[v] = out[ v*PS*PS*PS + io ];
      if ( !close_unphysical( Out, true, c ) ) continue;

      const R *in = g_in + (size_t)pg*NV*N3;
      const int ii = ( ( ijk[2] + ghost )*NX + ijk[1] + ghost )*NX + ijk[0] + ghost;
      R Update[NV], FL[3][NV], FR[3][NV];
      if ( c.mode != 0 )
      {
         R C[NV], dF[3][NV];
#        pragma unroll
         for (int v=0; v<NV; v++) C[v] = in[ v*N3 + ii ];
         for (int d=0; d<3; d++)
         {
            R L[NV], Rr[NV];
#           pragma unroll
            for (int v=0; v<NV; v++) { L[v] = in[ v*N3 + ii - didx[d] ];  Rr[v] = in[ v*N3 + ii + didx[d] ]; }
            rs1_dir<RS>( d, L, C, P, FL[d] );
            rs1_dir<RS>( d, C, Rr, P, FR[d] );
#           pragma unroll
            for (int v=0; v<NV; v++) dF[d][v] = FR[d][v] - FL[d][v];
         }
#        pragma unroll
         for (int v=0; v<NV; v++) Update[v] = C[v] - dt_dh*( dF[0][v] + dF[1][v] + dF[2][v] );
      }
      if ( c.mode == 2 && close_unphysical( Update, true, c ) )
      {
         R blk[27][NV];                                    // [z][y][x] of the 3^3 neighbourhood
         for (int b=0; b<27; b++)
#           pragma unroll
            for (int v=0; v<NV; v++) blk[b][v] = in[ v*N3 + ii + ( b%3 - 1 ) + ( ( b/3 )%3 - 1 )*NX + ( b/9 - 1 )*NX*NX ];
         const int lo[3][3] = { { 1, 0, 0 }, { 1, 1, 0 }, { 1, 1, 1 } };     // [d][x,y,z]
         const int hi[3][3] = { { 1, 2, 2 }, { 1, 1, 2 }, { 1, 1, 1 } };
         const int bst[3] = { 1, 3, 9 };
         for (int d=0; d<3; d++)
            for (int z=lo[d][2]; z<=hi[d][2]; z++)
            for (int y=lo[d][1]; y<=hi[d][1]; y++)
            for (int x=lo[d][0]; x<=hi[d][0]; x++)
            {
               const int b = ( z*3 + y )*3 + x;
               R fl[NV], fr[NV];
               rs1_dir<RS>( d, blk[b - bst[d]], blk[b], P, fl );
               rs1_dir<RS>( d, blk[b], blk[b + bst[d]], P, fr );
#              pragma unroll
               for (int v=0; v<NV; v++) blk[b][v] -= dt_dh*( fr[v] - fl[v] );
               if ( b == 13 )
#                 pragma unroll
                  for (int v=0; v<NV; v++) { FL[d][v] = fl[v];  FR[d][v] = fr[v]; }
            }
#        pragma unroll
         for (int v=0; v<NV; v++) Update[v] = blk[13][v];
      }
      if ( c.mode == 0 )
#        pragma unroll
         for (int v=0; v<NV; v++) Update[v] = Out[v];

      if ( !c.auto_reduce && c.last_resort_floor )
      {
         Update[0] = vmax( Update[0], (R)c.min_dens );
         const R ein  = Update[4] - (R)0.5*( Update[1]*Update[1] + Update[2]*Update[2] + Update[3]*Update[3] )/Update[0];
         const R eout = ( ein == ein ) ? vmax( ein, (R)c.min_eint ) : ein;
         if ( ein != eout ) Update[4] = Update[4] - ein + eout;
      }
      if ( close_unphysical( Update, c.auto_reduce != 0, c ) ) { atomicAdd( counters + 1, 1 );  continue; }

#     pragma unroll
      for (int v=0; v<NV; v++) out[ v*PS*PS*PS + io ] = Update[v];
      if ( c.fixup_flux && c.mode != 0 && g_flux != nullptr )
      {
         R *flux = g_flux + (size_t)pg*9*NV*PS*PS;
         const int dm[3][2] = { { 1, 2 }, { 0, 2 }, { 0, 1 } };
         for (int d=0; d<3; d++)
         {
            int face = -1;
            bool left = true;
            if      ( ijk[d] == 0     || ijk[d] == H )      { face = 3*d + ijk[d]/H; }
            else if ( ijk[d] == H - 1 || ijk[d] == PS - 1 ) { face = 3*d + ( ijk[d] + 1 )/H;  left = false; }
            if ( face < 0 ) continue;
            const int st = ijk[ dm[d][1] ]*PS + ijk[ dm[d][0] ];
#           pragma unroll
            for (int v=0; v<NV; v++) flux[ ( face*NV + v )*PS*PS + st ] = left ? FL[d][v] : FR[d][v];
         }
      }
      atomicAdd( counters, 1 );
   }
}

template <typename IO>
__global__ void __launch_bounds__( H*H )
store_flux_kernel( const IO *__restrict__ flux, const int *__restrict__ slot, double dt, double *__restrict__ reg )
{
   // block (PG p, LocalID*6 + s): one patch face, its 8 x 8 entries on the threads; most faces have no register,
   // and those blocks leave after one uniform slot read
   const long p = blockIdx.x;
   const int ls = blockIdx.y, s = ls % 6, lid = ls / 6, m = threadIdx.x / H, n = threadIdx.x % H;
   const int r = slot[ p*48 + ls ];
   if ( r < 0 ) return;
   int o[3];
   local_xyz( lid, o );
   int face, dmm, dnn;
   if      ( s < 2 ) { face = o[0] + s;      dmm = o[2]*H;  dnn = o[1]*H; }
   else if ( s < 4 ) { face = o[1] + 1 + s;  dmm = o[2]*H;  dnn = o[0]*H; }
   else              { face = o[2] + 2 + s;  dmm = o[1]*H;  dnn = o[0]*H; }
#  pragma unroll
   for (int v = 0; v < NV; v++)
      reg[ ( ( (long)r*NV + v )*H + m )*H + n ] = dt*(double)flux[ ( ( p*9 + face )*NV + v )*PS*PS + ( m + dmm )*PS + n + dnn ];
}

template <typename IO>
__global__ void correct_flux_kernel( const IO *__restrict__ flux, const int *__restrict__ slot, int npg, double dt,
                                     double *__restrict__ reg )
{
   #pragma clang fp contract(off)      // separate multiply and add, as the reference accumulates
   const int map[6] = { 0, 2, 3, 5, 6, 8 };
   const double dt4 = 0.25*dt;
   const long total = (long)npg*6*NV*H*H;
   for (long e = blockIdx.x*(long)blockDim.x + threadIdx.x; e < total; e += (long)gridDim.x*blockDim.x)
   {
      const int nn = (int)( e % H ), mm = (int)( ( e/H ) % H ), v = (int)( ( e/( H*H ) ) % NV );
      const long q = e/( H*H*NV );                          // pg*6 + s
      const int s = (int)( q % 6 );
      const long p = q/6;
      const int r = slot[q];
      if ( r < 0 ) continue;
      const IO *f = flux + ( ( p*9 + map[s] )*NV + v )*PS*PS;
      double acc = reg[ ( ( (long)r*NV + v )*H + mm )*H + nn ];
      acc -= dt4*(double)f[ ( 2*mm     )*PS + 2*nn     ];
      acc -= dt4*(double)f[ ( 2*mm     )*PS + 2*nn + 1 ];
      acc -= dt4*(double)f[ ( 2*mm + 1 )*PS + 2*nn     ];
      acc -= dt4*(double)f[ ( 2*mm + 1 )*PS + 2*nn + 1 ];
      reg[ ( ( (long)r*NV + v )*H + mm )*H + nn ] = acc;
   }
}

template <typename IO>
__global__ void fixup_flux_kernel( IO *__restrict__ patch, const int *__restrict__ slot, const double *__restrict__ reg,
                                   int npatch, double dh, double gm1, double min_dens, double min_pres, double min_eint )
{
   #pragma clang fp contract(off)      // separate multiply and add, as the reference accumulates
   const long total = (long)npatch*P3;
   for (long e = blockIdx.x*(long)blockDim.x + threadIdx.x; e < total; e += (long)gridDim.x*blockDim.x)
   {
      const long p = e/P3;
      const int c = (int)( e % P3 ), i = c % H, j = ( c/H ) % H, k = c/( H*H );
      if ( i != 0 && i != H-1 && j != 0 && j != H-1 && k != 0 && k != H-1 ) continue;
      IO *u = patch + p*NV*P3 + c;
      for (int s=0; s<6; s++)
      {
         const int r = slot[ p*6 + s ];
         if ( r < 0 ) continue;
         const int d = s >> 1, pos = ( s & 1 ) ? H-1 : 0;
         const int x = ( d == 0 ) ? i : ( d == 1 ) ? j : k;
         if ( x != pos ) continue;
         const int m = ( d == 2 ) ? j : k, n = ( d == 0 ) ? j : i;
         const double cst = ( s & 1 ) ? 1.0/dh : -1.0/dh;
         IO cv[NV];
#        pragma unroll
         for (int v=0; v<NV; v++) cv[v] = (IO)( (double)u[ v*P3 ] + reg[ ( ( (long)r*NV + v )*H + m )*H + n ]*cst );
         const IO ke = (IO)0.5*( cv[1]*cv[1] + cv[2]*cv[2] + cv[3]*cv[3] )/cv[0];
         const IO ei = cv[4] - ke;
         const IO pr = (IO)( (double)ei*gm1 );
         if ( (double)cv[0] <= min_dens || (double)ei <= min_eint || !finite( ei ) ||
              (double)pr <= min_pres || !finite( pr ) ) continue;
         cv[4] = ke + ei;
#        pragma unroll
         for (int v=0; v<NV; v++) u[ v*P3 ] = cv[v];
      }
   }
}

template <typename IO>
__global__ void fixup_restrict_kernel( const IO *__restrict__ son, IO *__restrict__ fa, int ngroup,
                                       const int *__restrict__ son0, const int *__restrict__ fa_idx, double min_eint )
{
   #pragma clang fp contract(off)      // separate multiply and add, as the reference accumulates
   const int H2 = H/2;
   const long total = (long)ngroup*P3;
   for (long e = blockIdx.x*(long)blockDim.x + threadIdx.x; e < total; e += (long)gridDim.x*blockDim.x)
   {
      const long q = e/P3;
      const int c = (int)( e % P3 ), i = c % H, j = ( c/H ) % H, k = c/( H*H );
      const IO *S = son + (long)( son0[q] + local_id( i/H2, j/H2, k/H2 ) )*NV*P3;
      IO *F = fa + (long)fa_idx[q]*NV*P3;
      const int I = 2*( i % H2 ), J = 2*( j % H2 ), K = 2*( k % H2 );
      IO w[NV];
#     pragma unroll
      for (int v=0; v<NV; v++)
      {
         const IO *s0 = S + v*P3;
         auto at = [&]( int kk, int jj, int ii ) { return s0[ ( kk*H + jj )*H + ii ]; };
         w[v] = (IO)( 0.125*( at( K, J, I ) + at( K, J, I+1 ) + at( K, J+1, I ) + at( K+1, J, I ) +
                              at( K, J+1, I+1 ) + at( K+1, J+1, I ) + at( K+1, J, I+1 ) + at( K+1, J+1, I+1 ) ) );
      }
      const IO ein  = w[4] - (IO)0.5*( w[1]*w[1] + w[2]*w[2] + w[3]*w[3] )/w[0];
      const IO eout = ( ein == ein ) ? vmax( ein, (IO)min_eint ) : ein;
      if ( ein != eout ) w[4] = w[4] - ein + eout;
#     pragma unroll
      for (int v=0; v<NV; v++) F[ v*P3 + c ] = w[v];
   }
}

template <typename IO>
__global__ void __launch_bounds__( 256 )
close_patches_kernel( const IO *__restrict__ pg_out, IO *__restrict__ patch, const int *__restrict__ pid0, bool vec )
{
   // block (PG p, field v, planes K = 4 kq .. 4 kq + 3); thread: 4 consecutive cells I = 4 q .. 4 q + 3 of row J of
   // plane K, which stay inside one patch (H = 8), so both sides move as one aligned 4-vector
   typedef IO V4 __attribute__(( ext_vector_type( 4 ) ));
   const int p = blockIdx.x, v = blockIdx.y >> 2, kq = blockIdx.y & 3, t = threadIdx.x;
   const int I = 4*( t & 3 ), J = ( t >> 2 ) & 15, K = 4*kq + ( t >> 6 );
   const IO *src = pg_out + ( (size_t)p*NV + v )*PS*PS*PS + ( K*PS + J )*PS + I;
   IO *dst = patch + ( (size_t)( pid0[p] + local_id( I/H, J/H, K/H ) )*NV + v )*P3 + ( ( K % H )*H + J % H )*H + I % H;
   if ( vec ) *reinterpret_cast<V4*>( dst ) = *reinterpret_cast<const V4*>( src );
   else
#     pragma unroll
      for (int k = 0; k < 4; k++) dst[k] = src[k];
}

int blocks_for( long total )
{
   long b = ( total + 255 )/256;
   return (int)( b < 1 ? 1 : ( b > 65536 ? 65536 : b ) );
}

} // namespace

template <typename R>
int launch_correct_unphysical( const R *in, R *out, R *flux, int npg, const CorrPrm &c, const Prm<R> &P, R dt_dh,
                               int *counters, hipStream_t stream )
{
   if ( npg <= 0 ) return 0;
   const dim3 g( blocks_for( (long)npg*PS*PS*PS ) ), b( 256 );
   switch ( ( c.rsolver == 2 && !c.roe_check ) ? 1 : c.rsolver )
   {
      case 4: hipLaunchKernelGGL( ( correct_unphysical_kernel<R, 4> ), g, b, 0, stream, in, out, flux, npg, c, P, dt_dh, counters ); break;
      case 3: hipLaunchKernelGGL( ( correct_unphysical_kernel<R, 3> ), g, b, 0, stream, in, out, flux, npg, c, P, dt_dh, counters ); break;
      case 2: hipLaunchKernelGGL( ( correct_unphysical_kernel<R, 2> ), g, b, 0, stream, in, out, flux, npg, c, P, dt_dh, counters ); break;
      case 1: hipLaunchKernelGGL( ( correct_unphysical_kernel<R, 1> ), g, b, 0, stream, in, out, flux, npg, c, P, dt_dh, counters ); break;
      default: return -1;
   }
   return ( hipGetLastError() == hipSuccess ) ? 0 : -2;
}

template <typename IO>
int launch_store_flux( const IO *flux, const int *slot, int npg, double dt, double *reg, hipStream_t stream )
{
   if ( npg <= 0 ) return 0;
   hipLaunchKernelGGL( store_flux_kernel<IO>, dim3( npg, 48 ), dim3( H*H ), 0, stream, flux, slot, dt, reg );
   return ( hipGetLastError() == hipSuccess ) ? 0 : -2;
}

template <typename IO>
int launch_correct_flux( const IO *flux, const int *slot, int npg, double dt, double *reg, hipStream_t stream )
{
   if ( npg <= 0 ) return 0;
   hipLaunchKernelGGL( correct_flux_kernel<IO>, dim3( blocks_for( (long)npg*6*NV*H*H ) ), dim3( 256 ), 0, stream, flux, slot, npg, dt, reg );
   return ( hipGetLastError() == hipSuccess ) ? 0 : -2;
}

template <typename IO>
int launch_fixup_flux( IO *patch, const int *slot, const double *reg, int npatch, double dh, double gm1,
                       double min_dens, double min_pres, double min_eint, hipStream_t stream )
{
   if ( npatch <= 0 ) return 0;
   hipLaunchKernelGGL( fixup_flux_kernel<IO>, dim3( blocks_for( (long)npatch*P3 ) ), dim3( 256 ), 0, stream, patch, slot, reg, npatch, dh, gm1,
                       min_dens, min_pres, min_eint );
   return ( hipGetLastError() == hipSuccess ) ? 0 : -2;
}

template <typename IO>
int launch_fixup_restrict( const IO *son, IO *fa, int ngroup, const int *son0, const int *fa_idx, double min_eint,
                           hipStream_t stream )
{
   if ( ngroup <= 0 ) return 0;
   hipLaunchKernelGGL( fixup_restrict_kernel<IO>, dim3( blocks_for( (long)ngroup*P3 ) ), dim3( 256 ), 0, stream, son, fa, ngroup, son0, fa_idx, min_eint );
   return ( hipGetLastError() == hipSuccess ) ? 0 : -2;
}

template <typename IO>
int launch_close_patches( const IO *pg_out, IO *patch, int npg, const int *pid0, hipStream_t stream )
{
   if ( npg <= 0 ) return 0;
   const size_t al = 4*sizeof( IO );
   const bool vec = ( reinterpret_cast<size_t>( pg_out ) % al == 0 ) && ( reinterpret_cast<size_t>( patch ) % al == 0 );
   hipLaunchKernelGGL( close_patches_kernel<IO>, dim3( npg, NV*4 ), dim3( 256 ), 0, stream, pg_out, patch, pid0, vec );
   return ( hipGetLastError() == hipSuccess ) ? 0 : -2;
}

#define GA_SYNC_INST( IO )                                                                                            \
template int launch_correct_unphysical<IO>( const IO*, IO*, IO*, int, const CorrPrm&, const Prm<IO>&, IO, int*, hipStream_t ); \
template int launch_store_flux<IO>( const IO*, const int*, int, double, double*, hipStream_t );                       \
template int launch_correct_flux<IO>( const IO*, const int*, int, double, double*, hipStream_t );                     \
template int launch_fixup_flux<IO>( IO*, const int*, const double*, int, double, double, double, double, double, hipStream_t ); \
template int launch_fixup_restrict<IO>( const IO*, IO*, int, const int*, const int*, double, hipStream_t );           \
template int launch_close_patches<IO>( const IO*, IO*, int, const int*, hipStream_t );
GA_SYNC_INST( double )
GA_SYNC_INST( float )
#undef GA_SYNC_INST

} // namespace gamer_amd
