// GAMER-side adapter: the fluid-solver part of GAMER's GPU API (src/GPU_API/CUAPI_*.cu) implemented on top of the
// gfx950 library's C ABI (include/gamer_amd_hydro.h).  A GAMER build that targets MI355X compiles this file with
// its own flags (-DGPU -DMODEL=HYDRO -DFLU_SCHEME=... -DRSOLVER=... [-DFLOAT8]) in place of
// CUAPI_Asyn_FluidSolver.cu, CUAPI_MemAllocate_Fluid.cu, CUAPI_MemFree_Fluid.cu, CUAPI_SetDevice.cu,
// CUAPI_Synchronize.cu and CUAPI_SetCache.cu, and links -lgamer_amd_hydro.  The signatures are GAMER's own
// (include/Prototype.h:646-690, src/GPU_API/CUAPI_MemAllocate.cu:9), so the callers (src/Main/InvokeSolver.cpp,
// src/GPU_API/CUAPI_MemAllocate.cu, src/Init/Init_GAMER.cpp) need no change.
//
// Scope: hydro with MHM or MHM_RP, no MHD, no passive scalars, no dual energy (the MHD/ELBDM/dual-energy/USG
// arguments are NULL or false in those builds and are ignored here).  The other GPU solvers (Poisson, source
// terms, dt) keep their GAMER implementations.
#include "GAMER.h"
#include "gamer_amd_hydro.h"

#if ( MODEL != HYDRO  ||  ( FLU_SCHEME != MHM  &&  FLU_SCHEME != MHM_RP ) )
#  error : the gfx950 fluid solver covers HYDRO with FLU_SCHEME MHM or MHM_RP
#endif

#ifdef FLOAT8
static const int Precision = HYDRO_FP64;
#else
static const int Precision = HYDRO_FP32_BULK;
#endif


// CUAPI_SetDevice.cu: Mode >= 0 selects that device, -1 one device per MPI rank (GPU ID = MPI_Rank), -2 the
// runtime's default device
void CUAPI_SetDevice( const int Mode )
{
   const int Device = ( Mode >= 0 ) ? Mode : ( Mode == -1 ) ? MPI_Rank : 0;
   if ( hydro_init( Device ) != HYDRO_OK )   Aux_Error( ERROR_INFO, "gfx950 fluid solver: %s\n", hydro_last_error() );
}


// CUAPI_SetCache.cu: no cache preference to set on gfx950
void CUAPI_SetCache()
{
}


// CUAPI_MemAllocate_Fluid.cu: device arrays for Flu_NPG patch groups per stream and the pinned [2] host double
// buffers of the fluid solver.  Out of memory returns GAMER_FAILED so that CUAPI_MemAllocate can retry with fewer
// streams, as with the CUDA build.
int CUAPI_MemAllocate_Fluid( const int Flu_NPG, const int Pot_NPG, const int Src_NPG, const int GPU_NStream )
{
   (void)Pot_NPG;
   (void)Src_NPG;
   const int Status = hydro_mem_allocate_fluid( Flu_NPG, GPU_NStream, FLU_SCHEME, Precision );
   if ( Status == HYDRO_E_NOMEM )   return GAMER_FAILED;
   if ( Status != HYDRO_OK )        Aux_Error( ERROR_INFO, "gfx950 fluid solver: %s\n", hydro_last_error() );

   for (int t=0; t<2; t++)
   {
      void *In = NULL, *Out = NULL, *Flux = NULL;
      if ( hydro_host_buffers( t, &In, &Out, &Flux ) != HYDRO_OK )
         Aux_Error( ERROR_INFO, "gfx950 fluid solver: %s\n", hydro_last_error() );
      h_Flu_Array_F_In [t] = ( real (*)[FLU_NIN ][ CUBE(FLU_NXT) ]      )In;
      h_Flu_Array_F_Out[t] = ( real (*)[FLU_NOUT][ CUBE(PS2) ]          )Out;
      h_Flux_Array     [t] = ( real (*)[9][NFLUX_TOTAL][ SQR(PS2) ]     )Flux;
   }
   return GAMER_SUCCESS;
}


// CUAPI_MemFree_Fluid.cu
void CUAPI_MemFree_Fluid( const int GPU_NStream )
{
   (void)GPU_NStream;
   if ( hydro_mem_free_fluid() != HYDRO_OK )   Aux_Error( ERROR_INFO, "gfx950 fluid solver: %s\n", hydro_last_error() );
   for (int t=0; t<2; t++)
   {
      h_Flu_Array_F_In [t] = NULL;
      h_Flu_Array_F_Out[t] = NULL;
      h_Flux_Array     [t] = NULL;
   }
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
   HydroSolveArgs a = {};
   a.npg            = NPatchGroup;
   a.dt             = dt;
   a.dh             = dh;
   a.gamma          = GAMMA;
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
   if ( hydro_asyn_fluid_solver( h_Flu_Array_In, h_Flu_Array_Out, h_Flux_Array, &a, GPU_NStream ) != HYDRO_OK )
      Aux_Error( ERROR_INFO, "gfx950 fluid solver: %s\n", hydro_last_error() );
}


// CUAPI_Synchronize.cu
void CUAPI_Synchronize()
{
   if ( hydro_synchronize() != HYDRO_OK )   Aux_Error( ERROR_INFO, "gfx950 fluid solver: %s\n", hydro_last_error() );
}
