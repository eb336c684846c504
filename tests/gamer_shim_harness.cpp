// Test-infrastructure harness for the GAMER-side adapter (gamer_amd/gamer_shim/CUAPI_FluidSolver_gfx950.cpp).
// Compiled against GAMER's headers by tests/Makefile (target `shim`) together with the adapter, and linked to
// gamer_amd/libgamer_amd_hydro.so.  Not product code and no solver arithmetic.  It only
//   (1) defines the run-time globals the adapter reads that GAMER's Main.cpp would define (GAMMA, MPI_Rank, the
//       GPU sizing parameters, OUTPUT_DIR, the EoS object and aux arrays and the pinned host-array pointers of
//       include/Global.h), an Aux_Error that reports instead of aborting and a plain Aux_Message, and
//   (2) drives the adapter in GAMER's order: CUAPI_SetDevice, CUAPI_MemAllocate_Fluid
//       (src/GPU_API/CUAPI_MemAllocate.cu:50), fill the pinned input of buffer 0, CUAPI_Asyn_FluidSolver with the
//       arguments Solver() passes for non-MHD hydro (src/Main/InvokeSolver.cpp:585-599), CUAPI_Synchronize, read the
//       pinned outputs, CUAPI_MemFree_Fluid; the dt path as dt_InvokeSolver does (CUAPI_MemAllocate, then
//       CUAPI_Asyn_dtSolver for DT_FLU_SOLVER, src/Miscellaneous/dt_InvokeSolver.cpp); the EoS set-up of EoS_Init
//       (EoS aux array -> CUAPI_SetConstMemory_EoS, src/EoS/EoS_Init.cpp:80-110); CUAPI_DiagnoseDevice.
#include "GAMER.h"
#include "gamer_amd_hydro.h"
#include <csetjmp>
#include <cstdarg>
#include <cstdio>
#include <cstring>

double GAMMA    = 5.0/3.0;
int    MPI_Rank = 0, MPI_NRank = 1;
int    GPU_NSTREAM = 0, FLU_GPU_NPGROUP = 0, SRC_GPU_NPGROUP = 0, OMP_NTHREAD = 1;
char   OUTPUT_DIR[MAX_STRING] = ".";
real (*h_Flu_Array_F_In [2])[FLU_NIN ][ CUBE(FLU_NXT) ];
real (*h_Flu_Array_F_Out[2])[FLU_NOUT][ CUBE(PS2) ];
real (*h_Flux_Array     [2])[9][NFLUX_TOTAL][ SQR(PS2) ];
real (*h_Flu_Array_T    [2])[FLU_NIN_T][ CUBE(PS1) ];
real  *h_dt_Array_T     [2];
EoS_t  EoS;
double EoS_AuxArray_Flt[EOS_NAUX_MAX];
int    EoS_AuxArray_Int[EOS_NAUX_MAX];
real  *h_EoS_Table[EOS_NTABLE_MAX];

// declared locally by GAMER's callers (src/GPU_API/CUAPI_MemAllocate.cu:9-11)
int  CUAPI_MemAllocate_Fluid( const int Flu_NPG, const int Pot_NPG, const int Src_NPG, const int GPU_NStream );
void CUAPI_SetMemSize( int &GPU_NStream, int &Flu_GPU_NPGroup, int &Pot_GPU_NPGroup, int &Che_GPU_NPGroup,
                       int &Src_GPU_NPGroup );

void Aux_Message( FILE *Type, const char *Format, ... )
{
   std::va_list ap;
   va_start( ap, Format );
   std::vfprintf( Type, Format, ap );
   va_end( ap );
}

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


// EoS_Init's GPU step for EOS_GAMMA: the aux array (AuxArray_Flt[0] = gamma, src/EoS/Gamma/CPU_EoS_Gamma.cpp:57) and
// CUAPI_SetConstMemory_EoS, after which the adapter's solver calls use this gamma.  Returns 0 or -1 (msg).
extern "C" int shim_set_eos_gamma( double gamma, char *msg, int msglen )
{
   if ( setjmp( Escape ) )
   {
      std::snprintf( msg, msglen, "%s", Message );
      return -1;
   }
   EoS_AuxArray_Flt[0] = gamma;
   EoS_AuxArray_Flt[1] = gamma - 1.0;
   EoS_AuxArray_Flt[2] = 1.0/( gamma - 1.0 );
   EoS_AuxArray_Flt[3] = 1.0/gamma;
   CUAPI_SetConstMemory_EoS();
   CUAPI_SetConstMemory();
   return 0;
}

// The CFL dt of 8*npg patches in [8 npg][5][8^3] through CUAPI_MemAllocate (FLU_GPU_NPGROUP = npg, GPU_NSTREAM =
// nstream) and CUAPI_Asyn_dtSolver on pinned buffer 1.  Returns 0, or -1 with the Aux_Error message.
extern "C" int shim_dt( int device, int npg, int nstream, double dh, double safety, double min_pres, const real *in,
                        real *dt, char *msg, int msglen )
{
   if ( setjmp( Escape ) )
   {
      std::snprintf( msg, msglen, "%s", Message );
      hydro_mem_free_fluid();
      return -1;
   }
   CUAPI_SetDevice( device );
   GPU_NSTREAM = nstream;  FLU_GPU_NPGROUP = npg;  SRC_GPU_NPGROUP = npg;
   CUAPI_MemAllocate();
   std::memcpy( h_Flu_Array_T[1], in, sizeof( real )*(size_t)8*npg*FLU_NIN_T*CUBE( PS1 ) );
   MicroPhy_t MicroPhy;
   std::memset( &MicroPhy, 0, sizeof( MicroPhy ) );
   CUAPI_Asyn_dtSolver( DT_FLU_SOLVER, h_dt_Array_T[1], h_Flu_Array_T[1], NULL, NULL, NULL, npg, (real)dh, (real)safety,
                        MicroPhy, (real)min_pres, 0L, false, false, EXT_ACC_NONE, 0.0, nstream );
   CUAPI_Synchronize();
   std::memcpy( dt, h_dt_Array_T[1], sizeof( real )*(size_t)8*npg );
   CUAPI_MemFree_Fluid( GPU_NSTREAM );
   return 0;
}

// CUAPI_SetMemSize with automatic sizes (GPU_NSTREAM = FLU_GPU_NPGROUP = -1): the values it sets.
extern "C" int shim_set_mem_size( int *nstream, int *flu_npg, char *msg, int msglen )
{
   if ( setjmp( Escape ) )
   {
      std::snprintf( msg, msglen, "%s", Message );
      return -1;
   }
   int ns = -1, flu = -1, pot = -1, che = -1, src = -1;
   CUAPI_SetMemSize( ns, flu, pot, che, src );
   *nstream = ns;
   *flu_npg = flu;
   return 0;
}

// CUAPI_DiagnoseDevice with OUTPUT_DIR = dir (it appends to dir/Record__Note).
extern "C" int shim_diagnose( const char *dir, char *msg, int msglen )
{
   if ( setjmp( Escape ) )
   {
      std::snprintf( msg, msglen, "%s", Message );
      return -1;
   }
   std::snprintf( OUTPUT_DIR, sizeof( OUTPUT_DIR ), "%s", dir );
   CUAPI_DiagnoseDevice();
   return 0;
}
