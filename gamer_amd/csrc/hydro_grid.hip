// Kernels around the patch-group solver: the hydro CFL time step, ghost-zone preparation of patch-group inputs
// from a periodic uniform level, and the scatter of patch-group outputs back to that level.
//
//   dt_cfl_kernel       CPU_dtSolver_HydroCFL / CUFLU_dtSolver_HydroCFL (src/Model_Hydro/CPU_Hydro/
//                       CPU_dtSolver_HydroCFL.cpp:57-209): per 8^3 patch, dt = Safety*dh / max_cells max_d(|v_d|+c_s),
//                       block max-reduction with wave64 shuffles (the reference's CUUTI_BlockReduction_Shuffle
//                       assumes 32-lane warps).
//   gather_pg_kernel    the same-level, periodic case of Prepare_PatchData for the fluid solver
//                       (src/Fluid/Flu_Prepare.cpp:91-94, src/Main/Prepare_PatchData.cpp:166): 16^3 interior +
//                       ghost zones, layout [PG][var][N^3].
//   scatter_pg_kernel   the copy-back of Flu_Close (src/Fluid/Flu_Close.cpp:133-217) into the level array.
// The level array is [var][Z][Y][X], x fastest; patch groups are numbered x fastest, then y, then z.

#include "hydro_math.hpp"
#include "hydro_kernels.hpp"

namespace gamer_amd {

namespace {

constexpr int PS = 16;

__device__ __forceinline__ float  wave_max( float v )
{
#  pragma unroll
   for (int o = 32; o > 0; o >>= 1) v = vmax( v, __shfl_xor( v, o, 64 ) );
   return v;
}
__device__ __forceinline__ double wave_max( double v )
{
#  pragma unroll
   for (int o = 32; o > 0; o >>= 1) v = vmax( v, __shfl_xor( v, o, 64 ) );
   return v;
}

template <typename IO>
__global__ void __launch_bounds__( 256 )
dt_cfl_kernel( const IO *__restrict__ in, IO *__restrict__ dt, int npatch, double dh_safety, double gamma,
               double min_pres )
{
   __shared__ IO red[4];
   const int p = blockIdx.x, tid = threadIdx.x;
   if ( p >= npatch ) return;
   const IO *u0 = in + (size_t)p*NV*512;
   IO m = (IO)0;
#  pragma unroll
   for (int k=0; k<2; k++)
   {
      const int c = tid + 256*k;
      IO u[NV];
#     pragma unroll
      for (int v=0; v<NV; v++) u[v] = u0[ v*512 + c ];
      m = vmax( m, cell_cfl( u, gamma, gamma - 1.0, min_pres ) );
   }
   m = wave_max( m );
   if ( ( tid & 63 ) == 0 ) red[tid >> 6] = m;
   __syncthreads();
   if ( tid == 0 ) dt[p] = (IO)( dh_safety / (double)vmax( vmax( red[0], red[1] ), vmax( red[2], red[3] ) ) );
}

// one lane per destination element of the patch-group inputs
template <typename IO>
__global__ void gather_pg_kernel( const IO *__restrict__ grid, IO *__restrict__ pg_in, int nx, int ny, int nz,
                                  int ghost, long total )
{
   const int n = PS + 2*ghost;
   const long n3 = (long)n*n*n;
   for (long e = blockIdx.x*(long)blockDim.x + threadIdx.x; e < total; e += (long)gridDim.x*blockDim.x)
   {
      const int  c  = (int)( e % n3 );
      const long pv = e / n3;
      const int  v  = (int)( pv % NV );
      const long pg = pv / NV;
      const int gx = nx/PS, gy = ny/PS;
      const int px = (int)( pg % gx ), py = (int)( ( pg / gx ) % gy ), pz = (int)( pg / ( (long)gx*gy ) );
      const int i = c % n, j = ( c / n ) % n, k = c / ( n*n );
      int x = px*PS + i - ghost, y = py*PS + j - ghost, z = pz*PS + k - ghost;
      x = ( x + nx ) % nx;  y = ( y + ny ) % ny;  z = ( z + nz ) % nz;
      pg_in[e] = grid[ ( ( (long)v*nz + z )*ny + y )*nx + x ];
   }
}

template <typename IO>
__global__ void scatter_pg_kernel( const IO *__restrict__ pg_out, IO *__restrict__ grid, int nx, int ny, int nz,
                                   long total )
{
   for (long e = blockIdx.x*(long)blockDim.x + threadIdx.x; e < total; e += (long)gridDim.x*blockDim.x)
   {
      const int  c  = (int)( e % 4096 );
      const long pv = e / 4096;
      const int  v  = (int)( pv % NV );
      const long pg = pv / NV;
      const int gx = nx/PS, gy = ny/PS;
      const int px = (int)( pg % gx ), py = (int)( ( pg / gx ) % gy ), pz = (int)( pg / ( (long)gx*gy ) );
      const int i = c % PS, j = ( c / PS ) % PS, k = c / ( PS*PS );
      grid[ ( ( (long)v*nz + pz*PS + k )*ny + py*PS + j )*nx + px*PS + i ] = pg_out[e];
   }
}

__global__ void reset_counter_kernel( unsigned *c ) { if ( threadIdx.x == 0 ) *c = 0u; }

int grid_blocks( long total )
{
   long b = ( total + 255 )/256;
   return (int)( b < 65536 ? ( b < 1 ? 1 : b ) : 65536 );
}

} // namespace

template <typename IO>
int launch_dt_cfl( const IO *in, IO *dt, int npatch, double dh, double safety, double gamma, double min_pres,
                   hipStream_t stream )
{
   if ( npatch <= 0 ) return 0;
   hipLaunchKernelGGL( dt_cfl_kernel<IO>, dim3( npatch ), dim3( 256 ), 0, stream, in, dt, npatch, safety*dh, gamma, min_pres );
   return ( hipGetLastError() == hipSuccess ) ? 0 : -2;
}

template <typename IO>
int launch_gather( const IO *grid, IO *pg_in, int nx, int ny, int nz, int ghost, hipStream_t stream )
{
   const long n = PS + 2*ghost;
   const long total = (long)( nx/PS )*( ny/PS )*( nz/PS )*NV*n*n*n;
   if ( total <= 0 ) return 0;
   hipLaunchKernelGGL( gather_pg_kernel<IO>, dim3( grid_blocks( total ) ), dim3( 256 ), 0, stream, grid, pg_in, nx, ny, nz, ghost, total );
   return ( hipGetLastError() == hipSuccess ) ? 0 : -2;
}

template <typename IO>
int launch_scatter( const IO *pg_out, IO *grid, int nx, int ny, int nz, hipStream_t stream )
{
   const long total = (long)( nx/PS )*( ny/PS )*( nz/PS )*NV*4096;
   if ( total <= 0 ) return 0;
   hipLaunchKernelGGL( scatter_pg_kernel<IO>, dim3( grid_blocks( total ) ), dim3( 256 ), 0, stream, pg_out, grid, nx, ny, nz, total );
   return ( hipGetLastError() == hipSuccess ) ? 0 : -2;
}

int launch_reset_counter( unsigned *counter, hipStream_t stream )
{
   hipLaunchKernelGGL( reset_counter_kernel, dim3( 1 ), dim3( 64 ), 0, stream, counter );
   return ( hipGetLastError() == hipSuccess ) ? 0 : -2;
}

template int launch_dt_cfl<double>( const double*, double*, int, double, double, double, double, hipStream_t );
template int launch_dt_cfl<float> ( const float*,  float*,  int, double, double, double, double, hipStream_t );
template int launch_gather<double>( const double*, double*, int, int, int, int, hipStream_t );
template int launch_gather<float> ( const float*,  float*,  int, int, int, int, hipStream_t );
template int launch_scatter<double>( const double*, double*, int, int, int, hipStream_t );
template int launch_scatter<float> ( const float*,  float*,  int, int, int, hipStream_t );

} // namespace gamer_amd
