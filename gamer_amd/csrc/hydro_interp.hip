// Coarse -> fine interpolation of ghost-zone data, and the same-level gather of patch-group inputs from per-patch
// data: the two halves of Prepare_PatchData for the fluid solver at a refinement boundary
// (src/Main/Prepare_PatchData.cpp:166, src/Main/InterpolateGhostZone.cpp, src/Interpolation/Interpolate.cpp:90).
//
//   interpolate_kernel   Interpolate() with Interpolate_Iterate (Interpolate.cpp:90-330) for INT_MINMOD3D, INT_MINMOD1D,
//                        INT_VANLEER, INT_CQUAD, INT_QUAD, INT_CQUAR and INT_QUAR (src/Interpolation/Int_*.cpp): one
//                        workgroup per coarse box, one lane per coarse cell producing its 2x2x2 fine cells.  For the
//                        tensor-product schemes (all but INT_MINMOD1D) the lane evaluates the x pass on the
//                        (2G+1)x(2G+1) coarse (y,z) neighbourhood (G = 1, or 2 for the quartic schemes), then the y
//                        pass, then the z pass: the same operations in the same order as the reference's split
//                        temporaries.  Iterations (conserved
//                        variables, then primitive variables, then lower monotonic coefficients) redo only coarse
//                        cells whose fine cells failed the physicality test: the INTERP_MASK rule (include/Macro.h:947)
//                        at the granularity of 2x2x2 fine blocks.
//   prepare_patches_kernel  the same-level copies: window cell (x,y,z) of a patch group from the patch in its
//                        4x4x4 slot table (-1: no patch at this level, the cell keeps the interpolated value)
//   gather_boxes_kernel  the coarse data of InterpolateGhostZone (src/Main/InterpolateGhostZone.cpp): box cell (x,y,z)
//                        from the patch covering cell origin + (x,y,z) of the box's slot table (e.g. the father patch
//                        and its 26 siblings)
//   assemble_kernel      prepare_patches_kernel with the interpolated fine box as the source of the cells that have no
//                        patch at this level

#include "hydro_math.hpp"
#include "hydro_kernels.hpp"

namespace gamer_amd {

namespace {

constexpr int PS = 16, H = 8, P3 = H*H*H;
constexpr int NT = 256;

template <typename T> __device__ __forceinline__ T sgn( T a ) { return ( a < (T)0 ) ? (T)-1 : (T)+1; }

template <typename T>
__device__ __forceinline__ T cquad_slope( T L, T C, T R, bool mono, T coeff, bool opp0 )
{
   #pragma clang fp contract(off)
   T s = (T)0.125*( R - L );
   if ( mono )
   {
      T l = (T)0.25*( C - L ), r = (T)0.25*( R - C );
      if ( l*r > (T)0 )
      {
         const T sg = sgn( l );
         s *= sg;  l *= sg;  r *= sg;
         if ( l < r ) l *= coeff; else r *= coeff;
         s = vmin( l, s );
         s = vmin( r, s );
         s *= sg;
      }
      else s = (T)0;
   }
   if ( opp0 && L*R < (T)0 ) s = (T)0;
   return s;
}

// One split pass of a tensor-product scheme: stencil u = {L2, L1, C, R1, R2} along one direction -> the fine values
// (lo, hi) of the centre cell (x passes of Int_MinMod3D.cpp:93-103, Int_vanLeer.cpp:93-105, Int_CQuadratic.cpp,
// Int_Quadratic.cpp:104-160, Int_CQuartic.cpp:100-140, Int_Quartic.cpp:100-165; their y and z passes repeat them)
template <typename T, int SCHEME>
__device__ __forceinline__ void pass1d( const T u[5], bool mono, T coeff, bool opp0, T &lo, T &hi )
{
   #pragma clang fp contract(off)
   const T L2 = u[0], L1 = u[1], C = u[2], R1 = u[3], R2 = u[4];
   T s;
   if ( SCHEME == 1 || SCHEME == 3 )               // MinMod-3D, van Leer
   {
      const T l = C - L1, r = R1 - C;
      if ( r*l <= (T)0 ) s = (T)0;
      else s = ( SCHEME == 1 ) ? (T)0.25*( vabs( r ) < vabs( l ) ? r : l ) : (T)0.5*l*r/( l + r );
      if ( opp0 && L1*R1 < (T)0 ) s = (T)0;
      lo = C - s;  hi = C + s;
   }
   else if ( SCHEME == 4 )                         // conservative quadratic
   {
      s = cquad_slope( L1, C, R1, mono, coeff, opp0 );
      lo = C - s;  hi = C + s;
   }
   else if ( SCHEME == 6 )                         // conservative quartic
   {
      s = (T)( 3.0/128.0)*L2 + (T)(-22.0/128.0)*L1 + (T)(22.0/128.0)*R1 + (T)(-3.0/128.0)*R2;
      if ( mono )
      {
         T l = (T)0.25*( C - L1 ), r = (T)0.25*( R1 - C );
         if ( l*r > (T)0 )
         {
            if ( s*l < (T)0 ) s = (T)0.125*( R1 - L1 );
            const T sg = sgn( l );
            s *= sg;  l *= sg;  r *= sg;
            if ( l < r ) l *= coeff; else r *= coeff;
            s = vmin( l, s );
            s = vmin( r, s );
            s *= sg;
         }
         else s = (T)0;
      }
      if ( opp0 && L1*R1 < (T)0 ) s = (T)0;
      lo = C - s;  hi = C + s;
   }
   else                                            // quadratic (5), quartic (7): Lagrange values at -/+ 1/4
   {
      if ( SCHEME == 5 )
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
         if ( l*r > (T)0 )
         {
            const T cmax = ( l > (T)0 ) ? R1 : L1, cmin = ( l > (T)0 ) ? L1 : R1;
            bool bad = ( vmax( lo, hi ) > cmax || vmin( lo, hi ) < cmin );
            if ( SCHEME == 7 ) bad = bad || ( hi - lo )*l < (T)0;
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
      if ( opp0 && L1*R1 < (T)0 ) { lo = C;  hi = C; }
   }
}

// the 2x2x2 fine values of coarse cell i of one field: f[(k*2 + j)*2 + i]
template <typename T, int SCHEME>
__device__ __forceinline__ void interp_cell( const T *__restrict__ c, int i, int dy, int dz, bool mono, T coeff,
                                             bool opp0, T f[8] )
{
   #pragma clang fp contract(off)
   if ( SCHEME != 2 )
   {
      constexpr int G = ( SCHEME == 6 || SCHEME == 7 ) ? 2 : 1, W = 2*G + 1;
      T X[W][W][2], Y[W][2][2], u[5];
#     pragma unroll
      for (int a=0; a<W; a++)
#        pragma unroll
         for (int b=0; b<W; b++)
         {
            const int q = i + ( a - G )*dz + ( b - G )*dy;
#           pragma unroll
            for (int k=0; k<5; k++) u[k] = ( k - 2 >= -G && k - 2 <= G ) ? c[q + k - 2] : (T)0;
            pass1d<T, SCHEME>( u, mono, coeff, opp0, X[a][b][0], X[a][b][1] );
         }
#     pragma unroll
      for (int a=0; a<W; a++)
#        pragma unroll
         for (int x=0; x<2; x++)
         {
#           pragma unroll
            for (int k=0; k<5; k++) u[k] = ( k - 2 >= -G && k - 2 <= G ) ? X[a][k - 2 + G][x] : (T)0;
            pass1d<T, SCHEME>( u, mono, coeff, opp0, Y[a][0][x], Y[a][1][x] );
         }
#     pragma unroll
      for (int y=0; y<2; y++)
#        pragma unroll
         for (int x=0; x<2; x++)
         {
#           pragma unroll
            for (int k=0; k<5; k++) u[k] = ( k - 2 >= -G && k - 2 <= G ) ? Y[k - 2 + G][y][x] : (T)0;
            pass1d<T, SCHEME>( u, mono, coeff, opp0, f[( 0*2 + y )*2 + x], f[( 1*2 + y )*2 + x] );
         }
   }
   else
   {
      const int cd[3] = { 1, dy, dz };
      T sl[3];
#     pragma unroll
      for (int d=0; d<3; d++)
      {
         const T l = c[i] - c[i - cd[d]], r = c[i + cd[d]] - c[i];
         sl[d] = ( r*l <= (T)0 ) ? (T)0 : (T)0.25*( vabs( r ) < vabs( l ) ? r : l );
         if ( opp0 && c[i - cd[d]]*c[i + cd[d]] < (T)0 ) sl[d] = (T)0;
      }
#     pragma unroll
      for (int k=0; k<2; k++)
#        pragma unroll
         for (int j=0; j<2; j++)
#           pragma unroll
            for (int x=0; x<2; x++)
               f[( k*2 + j )*2 + x] = c[i] + ( k ? sl[2] : -sl[2] ) + ( j ? sl[1] : -sl[1] ) + ( x ? sl[0] : -sl[0] );
   }
}

template <typename T> struct Huge;
template <> struct Huge<double> { static constexpr double v = 1.7976931348623157e308; static constexpr double eps = 2.2204460492503131e-16; };
template <> struct Huge<float>  { static constexpr float  v = 3.40282346638528859812e+38f; static constexpr float eps = 1.1920928955078125e-07f; };

// Hydro_IsUnphysical in CONS mode (FluUtility.cpp:843-940)
template <typename T>
__device__ __forceinline__ bool cons_bad( const T u[NV] )
{
   #pragma clang fp contract(off)
   const T huge = Huge<T>::v;
   bool bad = false;
#  pragma unroll
   for (int v=0; v<NV; v++)
   {
      bad |= ( u[v] != u[v] );
      if ( v >= 1 && v <= 3 ) bad |= ( u[v] < -huge || u[v] > huge );
      else                    bad |= ( u[v] <= (T)0 || u[v] > huge );
   }
   const T e = u[4] - (T)0.5*( u[1]*u[1] + u[2]*u[2] + u[3]*u[3] )/u[0];
   bad |= ( e < (T)-3.0*u[4]*Huge<T>::eps || e > huge || e != e );
   return bad;
}

// Hydro_IsUnphysical in PRIM mode plus the interpolated internal energy from the EoS (Interpolate.cpp:234-262)
template <typename T>
__device__ __forceinline__ bool prim_bad( const T w[NV], double inv_gm1 )
{
   const T huge = Huge<T>::v;
   bool bad = false;
#  pragma unroll
   for (int v=0; v<NV; v++)
   {
      bad |= ( w[v] != w[v] );
      if ( v >= 1 && v <= 3 ) bad |= ( w[v] < -huge || w[v] > huge );
      else                    bad |= ( w[v] <= (T)0 || w[v] > huge );
   }
   const T e = w[4]*(T)inv_gm1;                      // P/(gamma-1) in the storage type, as EOS_GAMMA
   return bad || e < (T)0 || e > huge || e != e;
}

// the 2x2x2 fine values of one coarse cell: physicality test (when iterating), conversion back from primitive
// variables, store; false when the test failed (nothing stored)
template <typename T>
__device__ __forceinline__ bool store_fine( T f[NV][8], bool iterate, bool prim, const InterpPrm &p, T *F, int F3, int fo,
                                           int fdy, int fdz )
{
   #pragma clang fp contract(off)
   bool bad = false;
   if ( iterate )
#     pragma unroll
      for (int b=0; b<8; b++)
      {
         T t[NV];
#        pragma unroll
         for (int v=0; v<NV; v++) t[v] = f[v][b];
         bad |= prim ? prim_bad( t, p.inv_gm1 ) : cons_bad( t );
      }
   if ( bad ) return false;
#  pragma unroll
   for (int b=0; b<8; b++)
   {
      T t[NV];
#     pragma unroll
      for (int v=0; v<NV; v++) t[v] = f[v][b];
      if ( prim )
      {
         // Hydro_Pri2Con with EoS_DensPres2Eint
         const T ei = t[4]*(T)p.inv_gm1;
         t[1] = t[0]*t[1];  t[2] = t[0]*t[2];  t[3] = t[0]*t[3];
         t[4] = (T)0.5*( t[1]*t[1] + t[2]*t[2] + t[3]*t[3] )/t[0] + ei;
      }
      const int o = fo + ( b >> 2 )*fdz + ( ( b >> 1 ) & 1 )*fdy + ( b & 1 );
#     pragma unroll
      for (int v=0; v<NV; v++) F[ (size_t)v*F3 + o ] = t[v];
   }
   return true;
}

// x pass of the rows y-1, y, y+1 of coarse plane z' around index q, then the y pass: Y[y-lo/hi][x-lo/hi] of that
// plane, exactly as interp_cell computes its Y[a] (3-point schemes)
template <typename T, int SCHEME>
__device__ __forceinline__ void xy_pass( const T *__restrict__ c, int q, int dy, bool mono, T coeff, bool opp0, T Y[2][2] )
{
   #pragma clang fp contract(off)
   T X[3][2], u[5];
#  pragma unroll
   for (int b=0; b<3; b++)
   {
      const int qq = q + ( b - 1 )*dy;
      u[0] = (T)0;  u[1] = c[qq - 1];  u[2] = c[qq];  u[3] = c[qq + 1];  u[4] = (T)0;
      pass1d<T, SCHEME>( u, mono, coeff, opp0, X[b][0], X[b][1] );
   }
#  pragma unroll
   for (int x=0; x<2; x++)
   {
      u[0] = (T)0;  u[1] = X[0][x];  u[2] = X[1][x];  u[3] = X[2][x];  u[4] = (T)0;
      pass1d<T, SCHEME>( u, mono, coeff, opp0, Y[0][x], Y[1][x] );
   }
}

template <typename T, int SCHEME>
__global__ void __launch_bounds__( NT )
interpolate_kernel( T *__restrict__ g_c, T *__restrict__ g_f, int nbox, InterpPrm p, int *__restrict__ nfail )
{
   #pragma clang fp contract(off)
   __shared__ int s_fail;
   __shared__ unsigned char s_done[INTERP_MAX_CELLS];
   const int q = blockIdx.x, tid = threadIdx.x;
   if ( q >= nbox ) return;
   const int dy = p.cs[0], dz = p.cs[0]*p.cs[1], C3 = dz*p.cs[2];
   const int fdy = p.fs[0], fdz = p.fs[0]*p.fs[1], F3 = fdz*p.fs[2];
   const int ncell = p.cr[0]*p.cr[1]*p.cr[2];
   T *C = g_c + (size_t)q*NV*C3;
   T *F = g_f + (size_t)q*NV*F3;
   const bool iterate = p.int_prim || p.reduce;
   const int mono_max = p.reduce ? p.mono_max_iter : 0, max_iter = p.int_prim ? mono_max + 1 : mono_max;

   for (int c = tid; c < ncell; c += NT) s_done[c] = p.need ? !p.need[ (size_t)q*ncell + c ] : 0;
   T coeff = (T)p.mono_coeff;
   bool prim = false;
   for (int it = 0; ; it++)
   {
      if ( it == 1 && p.int_prim )
      {
         // Hydro_Con2Pri of the whole coarse box, in place (Interpolate.cpp:148-170)
         for (int i = tid; i < C3; i += NT)
         {
            T u[NV];
#           pragma unroll
            for (int v=0; v<NV; v++) u[v] = C[ (size_t)v*C3 + i ];
            const T r = (T)1/u[0];
            const T e = u[4] - (T)0.5*( u[1]*u[1] + u[2]*u[2] + u[3]*u[3] )/u[0];
            T pr = (T)( (double)e*p.gm1 );
            pr = ( pr == pr ) ? vmax( pr, (T)p.min_pres ) : pr;
            C[ (size_t)1*C3 + i ] = u[1]*r;
            C[ (size_t)2*C3 + i ] = u[2]*r;
            C[ (size_t)3*C3 + i ] = u[3]*r;
            C[ (size_t)4*C3 + i ] = pr;
         }
         prim = true;
         __syncthreads();
      }
      else if ( it > 0 )
      {
         coeff -= (T)p.mono_coeff/(T)mono_max;
         coeff  = vmax( coeff, (T)0 );
      }
      if ( tid == 0 ) s_fail = 0;
      __syncthreads();

      constexpr bool MARCH = ( SCHEME == 1 || SCHEME == 3 || SCHEME == 4 || SCHEME == 5 );
      if ( MARCH && it == 0 )
      {
         // first pass of a 3-point tensor-product scheme: one lane per (x, y) column of the coarse range, marching
         // along z with the y-pass results of planes z-1, z, z+1 in registers, so each plane's x and y passes run
         // once per column instead of once per cell that reads them (9 pass1d per cell and field instead of 19).
         // The operations and their order are interp_cell's, so the values are bitwise the same.
         const int ncol = p.cr[0]*p.cr[1], cz = p.cr[0]*p.cr[1];
         for (int col = tid; col < ncol; col += NT)
         {
            const int x = col % p.cr[0], y = col / p.cr[0];
            bool todo = false;
            for (int z = 0; z < p.cr[2]; z++) todo |= !s_done[ col + z*cz ];
            if ( !todo ) continue;
            const int q0 = p.c0[2]*dz + ( p.c0[1] + y )*dy + p.c0[0] + x;      // coarse index of (x, y, z = 0)
            T Yw[NV][3][2][2];
#           pragma unroll
            for (int v=0; v<NV; v++)
            {
               xy_pass<T, SCHEME>( C + (size_t)v*C3, q0 - dz, dy, ( p.mono >> v ) & 1, coeff, p.opp0, Yw[v][0] );
               xy_pass<T, SCHEME>( C + (size_t)v*C3, q0,      dy, ( p.mono >> v ) & 1, coeff, p.opp0, Yw[v][1] );
            }
            for (int z = 0; z < p.cr[2]; z++)
            {
               T f[NV][8], u[5];
#              pragma unroll
               for (int v=0; v<NV; v++)
               {
                  xy_pass<T, SCHEME>( C + (size_t)v*C3, q0 + ( z + 1 )*dz, dy, ( p.mono >> v ) & 1, coeff, p.opp0, Yw[v][2] );
#                 pragma unroll
                  for (int yy=0; yy<2; yy++)
#                    pragma unroll
                     for (int xx=0; xx<2; xx++)
                     {
                        u[0] = (T)0;  u[1] = Yw[v][0][yy][xx];  u[2] = Yw[v][1][yy][xx];  u[3] = Yw[v][2][yy][xx];  u[4] = (T)0;
                        pass1d<T, SCHEME>( u, ( p.mono >> v ) & 1, coeff, p.opp0, f[v][( 0*2 + yy )*2 + xx],
                                           f[v][( 1*2 + yy )*2 + xx] );
                     }
               }
               const int c = col + z*cz;
               if ( !s_done[c] )
               {
                  const int fo = ( p.f0[2] + 2*z )*fdz + ( p.f0[1] + 2*y )*fdy + p.f0[0] + 2*x;
                  if ( store_fine( f, iterate, prim, p, F, F3, fo, fdy, fdz ) ) s_done[c] = 1; else s_fail = 1;
               }
#              pragma unroll
               for (int v=0; v<NV; v++)
#                 pragma unroll
                  for (int yy=0; yy<2; yy++)
#                    pragma unroll
                     for (int xx=0; xx<2; xx++) { Yw[v][0][yy][xx] = Yw[v][1][yy][xx];  Yw[v][1][yy][xx] = Yw[v][2][yy][xx]; }
            }
         }
      }
      else
      for (int c = tid; c < ncell; c += NT)
      {
         if ( s_done[c] ) continue;
         const int x = c % p.cr[0], y = ( c/p.cr[0] ) % p.cr[1], z = c/( p.cr[0]*p.cr[1] );
         const int ic = ( p.c0[2] + z )*dz + ( p.c0[1] + y )*dy + p.c0[0] + x;
         T f[NV][8];
#        pragma unroll
         for (int v=0; v<NV; v++) interp_cell<T, SCHEME>( C + (size_t)v*C3, ic, dy, dz, ( p.mono >> v ) & 1, coeff, p.opp0, f[v] );
         const int fo = ( p.f0[2] + 2*z )*fdz + ( p.f0[1] + 2*y )*fdy + p.f0[0] + 2*x;
         if ( store_fine( f, iterate, prim, p, F, F3, fo, fdy, fdz ) ) s_done[c] = 1; else s_fail = 1;
      }
      __syncthreads();
      const bool any = ( s_fail != 0 );
      if ( !iterate || !any ) break;
      if ( it == max_iter )
      {
         if ( tid == 0 ) atomicAdd( nfail, 1 );
         break;
      }
      __syncthreads();
   }
}

// Coarse boxes of general size: one block per (box, field, z plane), its 256 threads over the plane's cells in linear
// order (one division per cell for (x, y)).  A 32 x 8 thread grid over (x, y) left most lanes idle on the ~14^2
// planes of a refined patch group's coarse box; the linear layout made the two-level AMR step 1.2% faster.
constexpr int CPX = 32, CPY = 8;

// patch-group input window (N = 16 + 2 ghost <= 32 per axis, N even): cell (x,y,z) from the patch
// slot[pg][Z/8][Y/8][X/8], (X,Y,Z) = (x,y,z) - ghost + 8; slot -1: from the fine box fbox at (x,y,z) + fb.origin, or
// kept if fbox == NULL.  Each thread copies EPT consecutive cells of the linear window index e = x + N (y + N z) of
// one (PG, field): every lane is busy, the EPT loads are independent, and the store is one aligned vector (N^3 is a
// multiple of 8, so a group never straddles two windows; an unaligned destination takes scalar stores).  (x,y,z)
// of the first cell come from a float reciprocal: after the +0.5 shift, e/N^2 and r/N sit at least 1/(2 N^2) away
// from an integer, far above the float error for N <= 32.  Measured against a 32 x 8-thread (x, y) block layout
// and a lane-contiguous layout (each instruction over 64 consecutive cells), this one was the fastest.
constexpr int EPT = 4;
template <typename IO> struct Vec4;
template <> struct Vec4<float>  { typedef float  type __attribute__(( ext_vector_type( 4 ) )); };
template <> struct Vec4<double> { typedef double type __attribute__(( ext_vector_type( 4 ) )); };

template <typename IO>
__global__ void __launch_bounds__( NT )
prepare_patches_kernel( const IO *__restrict__ patch, const IO *__restrict__ fbox, IO *__restrict__ pg_in,
                        const int *__restrict__ slot, int ghost, BoxGeom fb )
{
   const int p = blockIdx.x, v = blockIdx.y, n = PS + 2*ghost, n2 = n*n, n3 = n2*n;
   // the patch group's 4^3 slot table in LDS: each cell's patch lookup is then an LDS read, not a second
   // dependent global load in front of the data load
   __shared__ int sl[64];
   if ( threadIdx.x < 64 ) sl[ threadIdx.x ] = slot[ (size_t)p*64 + threadIdx.x ];
   __syncthreads();
   const int e0 = ( blockIdx.z*NT + threadIdx.x )*EPT;
   if ( e0 >= n3 ) return;
   int z = (int)( ( (float)e0 + 0.5f )*( 1.0f/(float)n2 ) );
   const int r = e0 - z*n2;
   int y = (int)( ( (float)r + 0.5f )*( 1.0f/(float)n ) );
   int x = r - y*n;
   const IO  *pv = patch + (size_t)v*P3;
   const IO  *fv = fbox ? fbox + ( (size_t)p*NV + v )*fb.size[2]*fb.size[1]*fb.size[0] : nullptr;
   IO   val[EPT];
   bool put[EPT];
#  pragma unroll
   for (int k = 0; k < EPT; k++)
   {
      const int X = x - ghost + H, Y = y - ghost + H, Z = z - ghost + H;
      const int s = sl[ ( Z >> 3 )*16 + ( Y >> 3 )*4 + ( X >> 3 ) ];
      put[k] = true;
      if ( s >= 0 )  val[k] = pv[ (size_t)s*NV*P3 + ( Z & 7 )*H*H + ( Y & 7 )*H + ( X & 7 ) ];
      else if ( fv ) val[k] = fv[ ( (size_t)( z + fb.origin[2] )*fb.size[1] + y + fb.origin[1] )*fb.size[0] + x + fb.origin[0] ];
      else           put[k] = false;
      // next cell of the window
      x++;
      const bool wx = ( x == n );
      x = wx ? 0 : x;
      y = wx ? y + 1 : y;
      const bool wy = ( y == n );
      y = wy ? 0 : y;
      z = wy ? z + 1 : z;
   }
   IO *dst = pg_in + ( (size_t)p*NV + v )*n3 + e0;
   const bool aligned = ( reinterpret_cast<size_t>( dst ) % sizeof( typename Vec4<IO>::type ) ) == 0;
   if ( aligned && put[0] && put[1] && put[2] && put[3] )
   {
      typename Vec4<IO>::type q;
      q.x = val[0];  q.y = val[1];  q.z = val[2];  q.w = val[3];
      *reinterpret_cast<typename Vec4<IO>::type*>( dst ) = q;
   }
   else
   {
#     pragma unroll
      for (int k = 0; k < EPT; k++) if ( put[k] ) dst[k] = val[k];
   }
}

// box[q][v][z][y][x] (size) = patch[ slot[q][Z/8][Y/8][X/8] ][v][Z%8][Y%8][X%8], (X,Y,Z) = origin + (x,y,z);
// slot -1 keeps the box value
template <typename IO>
__global__ void __launch_bounds__( CPX*CPY )
gather_boxes_kernel( const IO *__restrict__ patch, IO *__restrict__ box, BoxGeom g, const int *__restrict__ slot )
{
   const int q = blockIdx.x, v = blockIdx.y, z = blockIdx.z;
   const int Z = z + g.origin[2];
   const int *sl = slot + (size_t)q*g.tdim[0]*g.tdim[1]*g.tdim[2] + ( Z >> 3 )*g.tdim[1]*g.tdim[0];
   IO *dst = box + ( ( (size_t)q*NV + v )*g.size[2] + z )*g.size[1]*g.size[0];
   const int nxy = g.size[0]*g.size[1];
   for (int e = threadIdx.x + CPX*threadIdx.y; e < nxy; e += CPX*CPY)
   {
      const int y = e / g.size[0], x = e - y*g.size[0];
      const int X = x + g.origin[0], Y = y + g.origin[1];
      const int s = sl[ ( Y >> 3 )*g.tdim[0] + ( X >> 3 ) ];
      if ( s >= 0 ) dst[ e ] = patch[ ( ( (size_t)s*NV + v )*H + ( Z & 7 ) )*H*H + ( Y & 7 )*H + ( X & 7 ) ];
   }
}

} // namespace

template <typename T>
int launch_interpolate( T *cdata, T *fdata, int nbox, const InterpPrm &p, int *nfail, hipStream_t stream )
{
   if ( nbox <= 0 ) return 0;
#  define GA_INT( S ) hipLaunchKernelGGL( ( interpolate_kernel<T, S> ), dim3( nbox ), dim3( NT ), 0, stream, cdata, fdata, nbox, p, nfail )
   switch ( p.scheme )
   {
      case 1: GA_INT( 1 ); break;
      case 2: GA_INT( 2 ); break;
      case 3: GA_INT( 3 ); break;
      case 4: GA_INT( 4 ); break;
      case 5: GA_INT( 5 ); break;
      case 6: GA_INT( 6 ); break;
      case 7: GA_INT( 7 ); break;
      default: return -1;
   }
#  undef GA_INT
   return ( hipGetLastError() == hipSuccess ) ? 0 : -2;
}

template <typename IO>
int launch_prepare_patches( const IO *patch, IO *pg_in, int npg, const int *slot, int ghost, hipStream_t stream )
{
   return launch_assemble<IO>( patch, nullptr, pg_in, npg, slot, ghost, BoxGeom{}, stream );
}

template <typename IO>
int launch_gather_boxes( const IO *patch, IO *box, int nbox, const BoxGeom &g, const int *slot, hipStream_t stream )
{
   if ( nbox <= 0 ) return 0;
   hipLaunchKernelGGL( gather_boxes_kernel<IO>, dim3( nbox, NV, g.size[2] ), dim3( CPX, CPY ), 0, stream, patch, box, g, slot );
   return ( hipGetLastError() == hipSuccess ) ? 0 : -2;
}

template <typename IO>
int launch_assemble( const IO *patch, const IO *fbox, IO *pg_in, int npg, const int *slot, int ghost, const BoxGeom &fb,
                     hipStream_t stream )
{
   if ( npg <= 0 ) return 0;
   const int n = PS + 2*ghost, per = NT*EPT;
   hipLaunchKernelGGL( prepare_patches_kernel<IO>, dim3( npg, NV, ( n*n*n + per - 1 )/per ), dim3( NT ), 0, stream, patch, fbox,
                       pg_in, slot, ghost, fb );
   return ( hipGetLastError() == hipSuccess ) ? 0 : -2;
}

template int launch_gather_boxes<double>( const double*, double*, int, const BoxGeom&, const int*, hipStream_t );
template int launch_gather_boxes<float> ( const float*,  float*,  int, const BoxGeom&, const int*, hipStream_t );
template int launch_assemble<double>( const double*, const double*, double*, int, const int*, int, const BoxGeom&, hipStream_t );
template int launch_assemble<float> ( const float*,  const float*,  float*,  int, const int*, int, const BoxGeom&, hipStream_t );
template int launch_interpolate<double>( double*, double*, int, const InterpPrm&, int*, hipStream_t );
template int launch_interpolate<float> ( float*,  float*,  int, const InterpPrm&, int*, hipStream_t );
template int launch_prepare_patches<double>( const double*, double*, int, const int*, int, hipStream_t );
template int launch_prepare_patches<float> ( const float*,  float*,  int, const int*, int, hipStream_t );

} // namespace gamer_amd
