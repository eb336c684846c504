// Test-infrastructure harness for the GAMER-side adapter (gamer_amd/gamer_shim/CUAPI_FluidSolver_gfx950.cpp).
// Compiled against GAMER's headers by tests/Makefile (target `shim`) together with the adapter, and linked to
// gamer_amd/libgamer_amd_hydro.so.  Not product code and no solver arithmetic.  It only
//   (1) defines the run-time globals the adapter reads that GAMER's Main.cpp would define (GAMMA, MPI_Rank and
//       the pinned host-array pointers of include/Global.h:423-425) and an Aux_Error that reports instead of
//       aborting, and
//   (2) drives the adapter in GAMER's order: CUAPI_SetDevice, CUAPI_MemAllocate_Fluid
//       (src/GPU_API/CUAPI_MemAllocate.cu:50), fill the pinned input of buffer 0, CUAPI_Asyn_FluidSolver with the
//       arguments Solver() passes for non-MHD hydro (src/Main/InvokeSolver.cpp:585-599), CUAPI_Synchronize, read the
//       pinned outputs, CUAPI_MemFree_Fluid.
#include "GAMER.h"
#include "gamer_amd_hydro.h"
#include <csetjmp>
#include <cstdarg>
#include <cstdio>
#include <cstring>

double GAMMA    = 5.0/3.0;
int    MPI_Rank = 0;
real (*h_Flu_Array_F_In [2])[FLU_NIN ][ CUBE(FLU_NXT) ];
real (*h_Flu_Array_F_Out[2])[FLU_NOUT][ CUBE(PS2) ];
real (*h_Flux_Array     [2])[9][NFLUX_TOTAL][ SQR(PS2) ];

int CUAPI_MemAllocate_Fluid( const int Flu_NPG, const int Pot_NPG, const int Src_NPG, const int GPU_NStream );

static std::jmp_buf Escape;
static char         Message[1024];

void Aux_Error( const char *File, const int Line, const char *Func, const char *Format, ... )
{
   std::va_list ap;
   va_start( ap, Format );
   const int n = std::snprintf( Message, sizeof( Message ), "%s:%d %s: ", File, Line, Func );
   std::vsnprintf( Message + n, sizeof( Message ) - n, Format, ap );
   va_end( ap );
   std::longjmp( Escape, 1 );
}

// One solver call on npg patch groups through the adapter.  in [npg][5][FLU_NXT^3], out [npg][5][16^3],
// flux [npg][9][5][16^2] in `real`.  Returns 0, or -1 with the Aux_Error message copied to msg.
extern "C" int shim_run( int device, int npg, int nstream, double dt, double dh, int limiter, double minmod_coeff,
                         const real *in, real *out, real *flux, char *msg, int msglen )
{
   if ( setjmp( Escape ) )
   {
      std::snprintf( msg, msglen, "%s", Message );
      hydro_mem_free_fluid();          // what the aborted sequence allocated
      return -1;
   }
   CUAPI_SetDevice( device );
   CUAPI_SetCache();
   if ( CUAPI_MemAllocate_Fluid( npg, 0, 0, nstream ) != GAMER_SUCCESS )
   {
      std::snprintf( msg, msglen, "CUAPI_MemAllocate_Fluid returned GAMER_FAILED" );
      return -1;
   }
   std::memcpy( h_Flu_Array_F_In[0], in, sizeof( real )*(size_t)npg*FLU_NIN*CUBE( FLU_NXT ) );
   MicroPhy_t MicroPhy;
   std::memset( &MicroPhy, 0, sizeof( MicroPhy ) );
   CUAPI_Asyn_FluidSolver( h_Flu_Array_F_In[0], h_Flu_Array_F_Out[0], NULL, NULL, NULL, h_Flux_Array[0], NULL, NULL, NULL,
                           NULL, NULL, npg, (real)dt, (real)dh, true, false, false, (LR_Limiter_t)limiter,
                           (real)minmod_coeff, 0, (real)0, (real)0, false, 0.0, false, EXT_ACC_NONE, MicroPhy,
                           (real)0, (real)0, (real)0, (real)0, 0L, false, 0, false, 0, false, (real)0, nstream, false );
   CUAPI_Synchronize();
   std::memcpy( out,  h_Flu_Array_F_Out[0], sizeof( real )*(size_t)npg*FLU_NOUT*CUBE( PS2 ) );
   std::memcpy( flux, h_Flux_Array[0],      sizeof( real )*(size_t)npg*9*NFLUX_TOTAL*SQR( PS2 ) );
   CUAPI_MemFree_Fluid( nstream );
   return 0;
}
