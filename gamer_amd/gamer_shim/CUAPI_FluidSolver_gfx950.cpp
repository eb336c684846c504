// GAMER-side adapter: the hydro part of GAMER's GPU API (src/GPU_API/CUAPI_*.cu) implemented on top of the gfx950
// library's C ABI (include/gamer_amd_hydro.h).  A GAMER build that targets MI355X compiles this file with its own
// flags (-DGPU -DMODEL=HYDRO -DFLU_SCHEME=... -DRSOLVER=... [-DFLOAT8]) in place of
//    CUAPI_Asyn_FluidSolver.cu   CUAPI_Asyn_dtSolver.cu       CUAPI_MemAllocate.cu      CUAPI_MemAllocate_Fluid.cu
//    CUAPI_MemFree_Fluid.cu      CUAPI_SetMemSize.cu          CUAPI_SetDevice.cu        CUAPI_SetCache.cu
//    CUAPI_Synchronize.cu        CUAPI_SetConstMemory.cu      CUAPI_SetConstMemory_EoS.cu
//    CUAPI_DiagnoseDevice.cu     EoS/Gamma/GPU_EoS_Gamma.cu (its EoS_SetGPUFunc_Gamma)
// and links -lgamer_amd_hydro.  The signatures are GAMER's own (include/Prototype.h:646-690,
// src/GPU_API/CUAPI_MemAllocate.cu:9-11, src/EoS/Gamma/CPU_EoS_Gamma.cpp:447-449), so the callers
// (src/Main/InvokeSolver.cpp, src/Init/Init_GAMER.cpp, src/EoS/EoS_Init.cpp, src/Miscellaneous/Mis_GetTimeStep.cpp)
// need no change.
//
// Scope: hydro with MHM or MHM_RP + PLM, the EOS_GAMMA equation of state, no MHD, no passive scalars, no dual
// energy, no cosmic rays, no SRHD and no self-gravity (whose Poisson GPU solver this adapter does not provide).
// Any other build configuration stops at compile time below: the library's array layouts are those of 5 fields with
// FLU_NXT = 22 (MHM_RP) or 20 (MHM), and a build whose GAMER arrays had another shape would overflow them.
#include "GAMER.h"
#include "gamer_amd_hydro.h"
#include <unistd.h>

#if ( MODEL != HYDRO  ||  ( FLU_SCHEME != MHM  &&  FLU_SCHEME != MHM_RP ) )
#  error : the gfx950 fluid solver covers HYDRO with FLU_SCHEME MHM or MHM_RP
#endif
#if ( LR_SCHEME != PLM )
#  error : the gfx950 fluid solver covers LR_SCHEME = PLM only (configure.py --slope=PLM)
#endif
#if ( EOS != EOS_GAMMA )
#  error : the gfx950 fluid solver covers EOS = EOS_GAMMA only
#endif
#if ( defined MHD  ||  defined SRHD  ||  defined COSMIC_RAY  ||  defined DUAL_ENERGY )
#  error : the gfx950 fluid solver covers pure hydro (no MHD, SRHD, cosmic rays or dual energy)
#endif
#if ( NCOMP_PASSIVE != 0 )
#  error : the gfx950 fluid solver covers 5 fluid fields without passive scalars
#endif
#ifdef GRAVITY
#  error : self-gravity needs GAMER's Poisson GPU solver, which this adapter does not provide
#endif

static_assert( FLU_NIN == 5 && FLU_NOUT == 5 && NFLUX_TOTAL == 5 && FLU_NIN_T == 5,
               "the library's arrays hold the 5 fluid fields" );
static_assert( FLU_NXT == ( FLU_SCHEME == MHM_RP ? 22 : 20 ) && PS2 == 16 && PS1 == 8,
               "the library's patch groups are 16^3 cells with 3 (MHM_RP) or 2 (MHM) ghost zones" );

#ifdef FLOAT8
static const int Precision = HYDRO_FP64;
#else
static const int Precision = HYDRO_FP32_BULK;
#endif

// adiabatic index of the EoS set by CUAPI_SetConstMemory_EoS (EoS_AuxArray_Flt[0] = GAMMA,
// src/EoS/Gamma/CPU_EoS_Gamma.cpp:54-66); until then the run-time parameter GAMMA
static double EoS_Gamma = -1.0;
static double Gamma() { return ( EoS_Gamma > 1.0 ) ? EoS_Gamma : GAMMA; }

#define GA_CHECK( call )                                                                                   \
   do { if ( (call) != HYDRO_OK )   Aux_Error( ERROR_INFO, "gfx950 fluid solver: %s\n", hydro_last_error() ); } while (0)


// CUAPI_SetDevice.cu: Mode >= 0 selects that device, -1 one device per MPI rank (GPU ID = MPI_Rank), -2 the
// runtime's default device
void CUAPI_SetDevice( const int Mode )
{
   const int Device = ( Mode >= 0 ) ? Mode : ( Mode == -1 ) ? MPI_Rank : 0;
   GA_CHECK( hydro_init( Device ) );
}


// CUAPI_SetCache.cu: no cache preference to set on gfx950
void CUAPI_SetCache()
{
}


// CUAPI_DiagnoseDevice.cu: the device properties of every rank into Record__Note
void CUAPI_DiagnoseDevice()
{
   if ( MPI_Rank == 0 )    Aux_Message( stdout, "%s ...\n", __FUNCTION__ );

   HydroDeviceInfo Info;
   GA_CHECK( hydro_device_info( &Info ) );
   if ( Info.device_count == 0 )    Aux_Error( ERROR_INFO, "no HIP device at MPI_Rank %d !!\n", MPI_Rank );

   char Host[1024];
   gethostname( Host, sizeof( Host ) );
   char FileName[2*MAX_STRING];
   sprintf( FileName, "%s/Record__Note", OUTPUT_DIR );

   if ( MPI_Rank == 0 )
   {
      FILE *Note = fopen( FileName, "a" );
      fprintf( Note, "Device Diagnosis\n" );
      fprintf( Note, "***********************************************************************************\n" );
      fclose( Note );
   }

   for (int YourTurn=0; YourTurn<MPI_NRank; YourTurn++)
   {
      if ( MPI_Rank == YourTurn )
      {
         FILE *Note = fopen( FileName, "a" );
         if ( MPI_Rank != 0 )   fprintf( Note, "\n\n" );
         fprintf( Note, "MPI_Rank = %3d, hostname = %10s, PID = %5d\n\n", MPI_Rank, Host, (int)getpid() );
         fprintf( Note, "GPU Info :\n" );
         fprintf( Note, "Number of GPUs                          : %d\n",      Info.device_count );
         fprintf( Note, "GPU ID                                  : %d\n",      Info.device );
         fprintf( Note, "GPU Name                                : %s\n",      Info.name );
         fprintf( Note, "GPU Architecture                        : %s\n",      Info.arch );
         fprintf( Note, "HIP Driver Version                      : %d\n",      Info.driver_version );
         fprintf( Note, "HIP Runtime Version                     : %d\n",      Info.runtime_version );
         fprintf( Note, "Clock Rate                              : %f GHz\n",  Info.clock_khz/1.0e6 );
         fprintf( Note, "Global Memory Size                      : %ld MB\n",  (long)( Info.global_mem_bytes/1024/1024 ) );
         fprintf( Note, "LDS Size per Block                      : %ld KB\n",  (long)( Info.lds_per_block_bytes/1024 ) );
         fprintf( Note, "Wavefront Size                          : %d\n",      Info.wavefront_size );
         fprintf( Note, "Number of Compute Units                 : %d\n",      Info.compute_units );
         fprintf( Note, "Max Number of Threads per Block         : %d\n",      Info.max_threads_per_block );
         fprintf( Note, "Concurrent Copy and Execution           : %s\n",      Info.async_engines > 0  ? "Yes" : "No" );
         fprintf( Note, "Concurrent Up/Downstream Copies         : %s\n",      Info.async_engines >= 2 ? "Yes" : "No" );
         fprintf( Note, "Concurrent Kernel Execution             : %s\n",      Info.concurrent_kernels ? "Yes" : "No" );
         fprintf( Note, "GPU has ECC Support Enabled             : %s\n",      Info.ecc ? "Yes" : "No" );
         fclose( Note );
      }
#     ifndef SERIAL
      MPI_Barrier( MPI_COMM_WORLD );
#     endif
   }

   if ( MPI_Rank == 0 )
   {
      FILE *Note = fopen( FileName, "a" );
      fprintf( Note, "***********************************************************************************\n" );
      fclose( Note );
      Aux_Message( stdout, "%s ... done\n", __FUNCTION__ );
   }
}


// CUAPI_SetMemSize.cu: GPU_NSTREAM <= 0 -> 4; FLU_GPU_NPGROUP / SRC_GPU_NPGROUP <= 0 -> GPU_NSTREAM x the number of
// compute units, with GAMER's consistency checks
void CUAPI_SetMemSize( int &GPU_NStream, int &Flu_GPU_NPGroup, int &Pot_GPU_NPGroup, int &Che_GPU_NPGroup,
                       int &Src_GPU_NPGroup )
{
   (void)Pot_GPU_NPGroup;
   (void)Che_GPU_NPGroup;
   if ( MPI_Rank == 0 )    Aux_Message( stdout, "%s ...\n", __FUNCTION__ );

   HydroDeviceInfo Info;
   GA_CHECK( hydro_device_info( &Info ) );

   if ( GPU_NStream <= 0 )
   {
      GPU_NStream = ( Info.async_engines > 0 ) ? 4 : 1;
      PRINT_RESET_PARA( GPU_NStream, FORMAT_INT, "--> might be further fine-tuned (GPU_NSTREAM in Input__Parameter)" );
   }
   if ( Flu_GPU_NPGroup <= 0 )
   {
      Flu_GPU_NPGroup = GPU_NStream*Info.compute_units;
      PRINT_RESET_PARA( Flu_GPU_NPGroup, FORMAT_INT, "--> might be further fine-tuned (FLU_GPU_NPGROUP in Input__Parameter)" );
   }
   if ( Src_GPU_NPGroup <= 0 )
   {
      Src_GPU_NPGroup = GPU_NStream*Info.compute_units;
      PRINT_RESET_PARA( Src_GPU_NPGroup, FORMAT_INT, "--> might be further fine-tuned (SRC_GPU_NPGROUP in Input__Parameter)" );
   }

   if ( GPU_NStream < 1 )  Aux_Error( ERROR_INFO, "GPU_NSTREAM (%d) < 1 !!\n", GPU_NStream );
   if ( Flu_GPU_NPGroup % GPU_NStream != 0 )
      Aux_Error( ERROR_INFO, "FLU_GPU_NPGROUP (%d) %% GPU_NSTREAM (%d) != 0 !!\n", Flu_GPU_NPGroup, GPU_NStream );
   if ( Src_GPU_NPGroup % GPU_NStream != 0 )
      Aux_Error( ERROR_INFO, "SRC_GPU_NPGROUP (%d) %% GPU_NSTREAM (%d) != 0 !!\n", Src_GPU_NPGroup, GPU_NStream );
#  ifdef OPENMP
   if ( Flu_GPU_NPGroup < OMP_NTHREAD )
      Aux_Error( ERROR_INFO, "FLU_GPU_NPGROUP (%d) < OMP_NTHREAD (%d) !!\n", Flu_GPU_NPGroup, OMP_NTHREAD );
   if ( Src_GPU_NPGroup < OMP_NTHREAD )
      Aux_Error( ERROR_INFO, "SRC_GPU_NPGROUP (%d) < OMP_NTHREAD (%d) !!\n", Src_GPU_NPGroup, OMP_NTHREAD );
#  endif

   if ( MPI_Rank == 0 )    Aux_Message( stdout, "%s ... done\n", __FUNCTION__ );
}


// CUAPI_MemAllocate_Fluid.cu: device arrays for Flu_NPG patch groups and the pinned [2] host double buffers of the
// fluid solver and of the CFL dt solver.  Out of memory returns GAMER_FAILED so that CUAPI_MemAllocate can retry
// with fewer streams, as with the CUDA build.
int CUAPI_MemAllocate_Fluid( const int Flu_NPG, const int Pot_NPG, const int Src_NPG, const int GPU_NStream )
{
   (void)Pot_NPG;
   (void)Src_NPG;
   const int Status = hydro_mem_allocate_fluid( Flu_NPG, GPU_NStream, FLU_SCHEME, Precision );
   if ( Status == HYDRO_E_NOMEM )   return GAMER_FAILED;
   if ( Status != HYDRO_OK )        Aux_Error( ERROR_INFO, "gfx950 fluid solver: %s\n", hydro_last_error() );

   for (int t=0; t<2; t++)
   {
      void *In = NULL, *Out = NULL, *Flux = NULL, *FluT = NULL, *dt = NULL;
      GA_CHECK( hydro_host_buffers( t, &In, &Out, &Flux ) );
      GA_CHECK( hydro_host_dt_buffers( t, &FluT, &dt ) );
      h_Flu_Array_F_In [t] = ( real (*)[FLU_NIN ][ CUBE(FLU_NXT) ]      )In;
      h_Flu_Array_F_Out[t] = ( real (*)[FLU_NOUT][ CUBE(PS2) ]          )Out;
      h_Flux_Array     [t] = ( real (*)[9][NFLUX_TOTAL][ SQR(PS2) ]     )Flux;
      h_Flu_Array_T    [t] = ( real (*)[FLU_NIN_T][ CUBE(PS1) ]         )FluT;
      h_dt_Array_T     [t] = ( real *                                   )dt;
   }
   return GAMER_SUCCESS;
}


// CUAPI_MemFree_Fluid.cu
void CUAPI_MemFree_Fluid( const int GPU_NStream )
{
   (void)GPU_NStream;
   GA_CHECK( hydro_mem_free_fluid() );
   for (int t=0; t<2; t++)
   {
      h_Flu_Array_F_In [t] = NULL;
      h_Flu_Array_F_Out[t] = NULL;
      h_Flux_Array     [t] = NULL;
      h_Flu_Array_T    [t] = NULL;
      h_dt_Array_T     [t] = NULL;
   }
}


// CUAPI_MemAllocate.cu: set the sizes, allocate, and halve GPU_NSTREAM (resetting the automatic sizes) while the
// allocation runs out of device memory
void CUAPI_MemAllocate()
{
   if ( MPI_Rank == 0 )    Aux_Message( stdout, "%s ...\n", __FUNCTION__ );

   int POT_GPU_NPGROUP = NULL_INT, CHE_GPU_NPGROUP = NULL_INT;
   const bool AutoNPG_Flu = ( FLU_GPU_NPGROUP <= 0 );
   const bool AutoNPG_Src = ( SRC_GPU_NPGROUP <= 0 );

   while ( true )
   {
      CUAPI_SetMemSize( GPU_NSTREAM, FLU_GPU_NPGROUP, POT_GPU_NPGROUP, CHE_GPU_NPGROUP, SRC_GPU_NPGROUP );
      int Status = CUAPI_MemAllocate_Fluid( FLU_GPU_NPGROUP, POT_GPU_NPGROUP, SRC_GPU_NPGROUP, GPU_NSTREAM );
#     ifndef SERIAL
      MPI_Allreduce( MPI_IN_PLACE, &Status, 1, MPI_INT, MPI_BAND, MPI_COMM_WORLD );
#     endif
      if ( Status == GAMER_SUCCESS )   break;

      if ( GPU_NSTREAM > 1 )
      {
         CUAPI_MemFree_Fluid( GPU_NSTREAM );
         const int NStreamOld = GPU_NSTREAM;
         GPU_NSTREAM /= 2;
         if ( AutoNPG_Flu )   FLU_GPU_NPGROUP = -1;
         if ( AutoNPG_Src )   SRC_GPU_NPGROUP = -1;
         if ( MPI_Rank == 0 )
            Aux_Message( stderr, "WARNING : GPU out of memory --> reducing GPU_NSTREAM from %d to %d automatically\n",
                         NStreamOld, GPU_NSTREAM );
      }
      else
         Aux_Error( ERROR_INFO, "GPU out of memory even with GPU_NSTREAM = 1\n"
                    "        --> Try reducing *_GPU_NPGROUP in Input__Parameter manually (or set to -1) !!\n" );
   }

   if ( MPI_Rank == 0 )    Aux_Message( stdout, "%s ... done\n", __FUNCTION__ );
}


// CUAPI_SetConstMemory.cu: without passive scalars and self-gravity (excluded above) there is nothing to copy;
// the library takes every solver constant as a call argument
void CUAPI_SetConstMemory()
{
   if ( MPI_Rank == 0 )    Aux_Message( stdout, "%s ...\n", __FUNCTION__ );
   if ( MPI_Rank == 0 )    Aux_Message( stdout, "%s ... done\n", __FUNCTION__ );
}


// CUAPI_SetConstMemory_EoS.cu: the library evaluates EOS_GAMMA inline from the adiabatic index passed with each
// call, so this records gamma (EoS_AuxArray_Flt[0]) for the solver calls.  The EoS object keeps the host arrays:
// no device code of this build reads it.
void CUAPI_SetConstMemory_EoS()
{
   const double g = EoS_AuxArray_Flt[0];
   if ( !( g > 1.0 ) )    Aux_Error( ERROR_INFO, "gfx950 fluid solver: EoS_AuxArray_Flt[0] = %14.7e is not a gamma > 1 !!\n", g );
   EoS_Gamma = g;
   EoS.AuxArrayDevPtr_Flt = EoS_AuxArray_Flt;
   EoS.AuxArrayDevPtr_Int = EoS_AuxArray_Int;
   EoS.Table              = h_EoS_Table;
}


// GPU_EoS_Gamma.cu (EoS_SetGPUFunc_Gamma): no device EoS functions exist in this build; the pointers stay NULL
void EoS_SetGPUFunc_Gamma( EoS_DE2P_t &EoS_DensEint2Pres_GPUPtr, EoS_DP2E_t &EoS_DensPres2Eint_GPUPtr,
                           EoS_DP2C_t &EoS_DensPres2CSqr_GPUPtr, EoS_DE2T_t &EoS_DensEint2Temp_GPUPtr,
                           EoS_DT2P_t &EoS_DensTemp2Pres_GPUPtr, EoS_DE2S_t &EoS_DensEint2Entr_GPUPtr,
                           EoS_GENE_t &EoS_General_GPUPtr )
{
   EoS_DensEint2Pres_GPUPtr = NULL;
   EoS_DensPres2Eint_GPUPtr = NULL;
   EoS_DensPres2CSqr_GPUPtr = NULL;
   EoS_DensEint2Temp_GPUPtr = NULL;
   EoS_DensTemp2Pres_GPUPtr = NULL;
   EoS_DensEint2Entr_GPUPtr = NULL;
   EoS_General_GPUPtr       = NULL;
}


// CUAPI_Asyn_FluidSolver.cu: asynchronous; the outputs are valid after CUAPI_Synchronize()
void CUAPI_Asyn_FluidSolver( real h_Flu_Array_In[][FLU_NIN ][ CUBE(FLU_NXT) ],
                             real h_Flu_Array_Out[][FLU_NOUT][ CUBE(PS2) ],
                             real h_Mag_Array_In[][NCOMP_MAG][ FLU_NXT_P1*SQR(FLU_NXT) ],
                             real h_Mag_Array_Out[][NCOMP_MAG][ PS2P1*SQR(PS2) ],
                             char h_DE_Array_Out[][ CUBE(PS2) ],
                             real h_Flux_Array[][9][NFLUX_TOTAL][ SQR(PS2) ],
                             real h_Ele_Array[][9][NCOMP_ELE][ PS2P1*PS2 ],
                             const double h_Corner_Array[][3],
                             real h_Pot_Array_USG[][ CUBE(USG_NXT_F) ],
                             const bool h_IsCompletelyRefined[],
                             const bool h_HasWaveCounterpart[][ CUBE(HYB_NXT) ],
                             const int NPatchGroup, const real dt, const real dh,
                             const bool StoreFlux, const bool StoreElectric,
                             const bool XYZ, const LR_Limiter_t LR_Limiter, const real MinMod_Coeff, const int MinMod_MaxIter,
                             const real ELBDM_Eta, real ELBDM_Taylor3_Coeff, const bool ELBDM_Taylor3_Auto,
                             const double Time, const bool UsePot, const OptExtAcc_t ExtAcc, const MicroPhy_t MicroPhy,
                             const real MinDens, const real MinPres, const real MinEint,
                             const real DualEnergySwitch,
                             const long PassiveFloor, const bool NormPassive, const int NNorm,
                             const bool FracPassive, const int NFrac,
                             const bool JeansMinPres, const real JeansMinPres_Coeff,
                             const int GPU_NStream, const bool UseWaveFlag )
{
   if ( UsePot  ||  ExtAcc != EXT_ACC_NONE  ||  JeansMinPres )
      Aux_Error( ERROR_INFO, "gfx950 fluid solver: gravity / external acceleration / Jeans pressure floor are not supported !!\n" );

   HydroSolveArgs a = {};
   a.npg            = NPatchGroup;
   a.dt             = dt;
   a.dh             = dh;
   a.gamma          = Gamma();
   a.scheme         = FLU_SCHEME;
   a.rsolver        = RSOLVER;
   a.limiter        = (int)LR_Limiter;
   a.minmod_coeff   = MinMod_Coeff;
   a.minmod_maxiter = MinMod_MaxIter;
   a.min_dens       = MinDens;
   a.min_pres       = MinPres;
   a.min_eint       = MinEint;
   a.store_flux     = StoreFlux;
   a.precision      = Precision;
   GA_CHECK( hydro_asyn_fluid_solver( h_Flu_Array_In, h_Flu_Array_Out, h_Flux_Array, &a, GPU_NStream ) );
}


// CUAPI_Asyn_dtSolver.cu: asynchronous; h_dt_Array is valid after CUAPI_Synchronize().  The hydro CFL solver
// (DT_FLU_SOLVER) only: without self-gravity there is no DT_GRA_SOLVER.
void CUAPI_Asyn_dtSolver( const Solver_t TSolver, real h_dt_Array[], const real h_Flu_Array[][FLU_NIN_T][ CUBE(PS1) ],
                          const real h_Mag_Array[][NCOMP_MAG][ PS1P1*SQR(PS1) ], const real h_Pot_Array[][ CUBE(GRA_NXT) ],
                          const double h_Corner_Array[][3], const int NPatchGroup, const real dh, const real Safety,
                          const MicroPhy_t MicroPhy, const real MinPres, const long PassiveFloor,
                          const bool P5_Gradient, const bool UsePot,
                          const OptExtAcc_t ExtAcc, const double TargetTime, const int GPU_NStream )
{
   if ( TSolver != DT_FLU_SOLVER )
      Aux_Error( ERROR_INFO, "gfx950 fluid solver: dt solver %d is not supported (DT_FLU_SOLVER only) !!\n", (int)TSolver );
   GA_CHECK( hydro_asyn_dt_solver( h_Flu_Array, h_dt_Array, NPatchGroup, dh, Safety, Gamma(), MinPres, GPU_NStream ) );
}


// CUAPI_Synchronize.cu
void CUAPI_Synchronize()
{
   GA_CHECK( hydro_synchronize() );
}
