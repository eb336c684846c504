// Test-infrastructure harness: links GAMER's own CPU hydro solver sources (compiled in place from the
// reference tree by oracle/Makefile) and exposes them through a flat C entry point, so the oracle
// restatement and the GPU path can be checked against the reference on identical inputs.
//
// This file is NOT product code and contains no solver arithmetic. It only
//   (1) defines the run-time parameters that GAMER's Main.cpp/Global.h would otherwise define
//       (GAMMA, unit constants, EoS aux arrays and CPU function pointers; include/Global.h:70-351),
//   (2) initialises the EOS_GAMMA plug-in exactly as EoS_Init does on a CPU build
//       (src/EoS/EoS_Init.cpp:80-121 -> EoS_Init_Gamma, src/EoS/Gamma/CPU_EoS_Gamma.cpp),
//   (3) allocates the per-PG host scratch arrays as Init_MemAllocate_Fluid does
//       (src/Init/Init_MemAllocate_Fluid.cpp:91-95), and
//   (4) calls CPU_FluidSolver_MHM with the same arguments CPU_FluidSolver passes
//       (src/Fluid/CPU_FluidSolver.cpp:258-266), non-MHD / no gravity / no passive scalars, and
//   (5) calls CPU_dtSolver_HydroCFL as CPU_dtSolver does for DT_FLU_SOLVER (src/Model_Hydro/CPU_Hydro/
//       CPU_dtSolver_HydroCFL.cpp:57-209, src/SelfGravity/CPU_Poisson/CPU_dtSolver.cpp), and
//   (6) calls Interpolate (src/Interpolation/Interpolate.cpp) box by box with the interpolation run-time
//       parameters it reads (INT_MONO_COEFF, MONO_MAX_ITER, MIN_PRES) set from its arguments.

#include "GAMER.h"
#include "CUFLU.h"
#include <omp.h>
#include <new>
#include <cstdlib>
#include <cstring>

// ---- run-time parameters (declared extern in include/Global.h) ----
double GAMMA            = 5.0/3.0;
double MOLECULAR_WEIGHT = 0.6;
double MU_NORM          = 1.0;
double UNIT_L = 1.0, UNIT_M = 1.0, UNIT_T = 1.0, UNIT_V = 1.0, UNIT_D = 1.0, UNIT_E = 1.0, UNIT_P = 1.0;
bool   OPT__UNIT        = false;

// run-time parameters read by Interpolate() (src/Interpolation/Interpolate.cpp; include/Global.h:44-125);
// the interpolation entry point below sets the ones it takes as arguments before each call
int    MPI_Rank                 = 0;
long   PassiveFloorMask         = 0L;
int    PassiveIntFrac_NVar      = 0;
int    PassiveIntFrac_VarIdx[NCOMP_PASSIVE];
bool   OPT__INT_FRAC_PASSIVE_LR = false;
double INT_MONO_COEFF           = 2.0;
int    MONO_MAX_ITER            = 10;
double MIN_PRES                 = 0.0;

double EoS_AuxArray_Flt[EOS_NAUX_MAX];
int    EoS_AuxArray_Int[EOS_NAUX_MAX];
EoS_GUESS_t   EoS_GuessHTilde_CPUPtr   = NULL;
EoS_H2TEM_t   EoS_HTilde2Temp_CPUPtr   = NULL;
EoS_TEM2H_t   EoS_Temp2HTilde_CPUPtr   = NULL;
EoS_DE2P_t    EoS_DensEint2Pres_CPUPtr = NULL;
EoS_DP2E_t    EoS_DensPres2Eint_CPUPtr = NULL;
EoS_DP2C_t    EoS_DensPres2CSqr_CPUPtr = NULL;
EoS_DE2T_t    EoS_DensEint2Temp_CPUPtr = NULL;
EoS_DT2P_t    EoS_DensTemp2Pres_CPUPtr = NULL;
EoS_DE2S_t    EoS_DensEint2Entr_CPUPtr = NULL;
EoS_GENE_t    EoS_General_CPUPtr       = NULL;
real         *h_EoS_Table[EOS_NTABLE_MAX];

void EoS_Init_Gamma();

void CPU_FluidSolver_MHM(
   const real   g_Flu_Array_In [][NCOMP_TOTAL][ CUBE(FLU_NXT) ],
         real   g_Flu_Array_Out[][NCOMP_TOTAL][ CUBE(PS2) ],
   const real   g_Mag_Array_In [][NCOMP_MAG][ FLU_NXT_P1*SQR(FLU_NXT) ],
         real   g_Mag_Array_Out[][NCOMP_MAG][ PS2P1*SQR(PS2) ],
         char   g_DE_Array_Out [][ CUBE(PS2) ],
         real   g_Flux_Array   [][9][NCOMP_TOTAL][ SQR(PS2) ],
         real   g_Ele_Array    [][9][NCOMP_ELE][ PS2P1*PS2 ],
   const double g_Corner_Array [][3],
   const real   g_Pot_Array_USG[][ CUBE(USG_NXT_F) ],
         real   g_PriVar       []   [NCOMP_LR            ][ CUBE(FLU_NXT) ],
         real   g_Slope_PPM    [][3][NCOMP_LR            ][ CUBE(N_SLOPE_PPM) ],
         real   g_FC_Var       [][6][NCOMP_TOTAL_PLUS_MAG][ CUBE(N_FC_VAR) ],
         real   g_FC_Flux      [][3][NCOMP_TOTAL_PLUS_MAG][ CUBE(N_FC_FLUX) ],
         real   g_FC_Mag_Half  [][NCOMP_MAG][ FLU_NXT_P1*SQR(FLU_NXT) ],
         real   g_EC_Ele       [][NCOMP_MAG][ CUBE(N_EC_ELE) ],
   const int NPatchGroup,
   const real dt, const real dh,
   const bool StoreFlux, const bool StoreElectric,
   const LR_Limiter_t LR_Limiter, const real MinMod_Coeff, const int MinMod_MaxIter, const double Time,
   const bool UsePot, const OptExtAcc_t ExtAcc, const ExtAcc_t ExtAcc_Func,
   const double c_ExtAcc_AuxArray[],
   const real MinDens, const real MinPres, const real MinEint,
   const real DualEnergySwitch,
   const long PassiveFloor,
   const bool NormPassive, const int NNorm, const int c_NormIdx[],
   const bool FracPassive, const int NFrac, const int c_FracIdx[],
   const bool JeansMinPres, const real JeansMinPres_Coeff,
   const EoS_t EoS, const MicroPhy_t MicroPhy );

void CPU_dtSolver_HydroCFL( real g_dt_Array[], const real g_Flu_Array[][FLU_NIN_T][ CUBE(PS1) ],
                            const real g_Mag_Array[][NCOMP_MAG][ PS1P1*SQR(PS1) ], const int NPG,
                            const real dh, const real Safety, const real MinPres,
                            const long PassiveFloor, const EoS_t EoS, const MicroPhy_t MicroPhy );

static EoS_t make_eos( double gamma )
{
   GAMMA = gamma;
   EoS_Init_Gamma();
   EoS_t EoS_obj;
   memset( &EoS_obj, 0, sizeof(EoS_obj) );
   EoS_obj.DensEint2Pres_FuncPtr = EoS_DensEint2Pres_CPUPtr;
   EoS_obj.DensPres2Eint_FuncPtr = EoS_DensPres2Eint_CPUPtr;
   EoS_obj.DensPres2CSqr_FuncPtr = EoS_DensPres2CSqr_CPUPtr;
   EoS_obj.DensEint2Temp_FuncPtr = EoS_DensEint2Temp_CPUPtr;
   EoS_obj.DensTemp2Pres_FuncPtr = EoS_DensTemp2Pres_CPUPtr;
   EoS_obj.DensEint2Entr_FuncPtr = EoS_DensEint2Entr_CPUPtr;
   EoS_obj.GuessHTilde_FuncPtr   = EoS_GuessHTilde_CPUPtr;
   EoS_obj.HTilde2Temp_FuncPtr   = EoS_HTilde2Temp_CPUPtr;
   EoS_obj.Temp2HTilde_FuncPtr   = EoS_Temp2HTilde_CPUPtr;
   EoS_obj.General_FuncPtr       = EoS_General_CPUPtr;
   EoS_obj.AuxArrayDevPtr_Flt    = EoS_AuxArray_Flt;
   EoS_obj.AuxArrayDevPtr_Int    = EoS_AuxArray_Int;
   EoS_obj.Table                 = h_EoS_Table;
   return EoS_obj;
}

extern "C" {

// Build-time identity of this reference object: sizeof(real), FLU_NXT, scheme and Riemann solver ids.
int gamer_ref_info( int info[4] )
{
   info[0] = (int)sizeof(real);
   info[1] = FLU_NXT;
   info[2] = FLU_SCHEME;
   info[3] = RSOLVER;
   return 0;
}

// One call of GAMER's CPU_FluidSolver_MHM on NPG patch groups.
//   in   : [npg][5][FLU_NXT^3]   out : [npg][5][16^3]   flux : [npg][9][5][16^2] (may be NULL if !store_flux)
// Returns 0 on success, nonzero on a bad argument or allocation failure.
int gamer_ref_fluid_solver( const void *in, void *out, void *flux, int npg,
                            double dt, double dh, double gamma, int store_flux,
                            int limiter, double minmod_coeff, int minmod_maxiter,
                            double min_dens, double min_pres, double min_eint, int nthreads )
{
   if ( npg <= 0 || in == NULL || out == NULL || ( store_flux && flux == NULL ) )   return 1;

   const EoS_t EoS_obj = make_eos( gamma );

   MicroPhy_t MicroPhy_obj;
   memset( &MicroPhy_obj, 0, sizeof(MicroPhy_obj) );

   real (*h_PriVar)      [NCOMP_LR            ][ CUBE(FLU_NXT)     ] = new (std::nothrow) real [npg]   [NCOMP_LR            ][ CUBE(FLU_NXT)     ];
   real (*h_Slope_PPM)[3][NCOMP_LR            ][ CUBE(N_SLOPE_PPM) ] = NULL;   // unused by PLM
   real (*h_FC_Var)   [6][NCOMP_TOTAL_PLUS_MAG][ CUBE(N_FC_VAR)    ] = new (std::nothrow) real [npg][6][NCOMP_TOTAL_PLUS_MAG][ CUBE(N_FC_VAR)    ];
   real (*h_FC_Flux)  [3][NCOMP_TOTAL_PLUS_MAG][ CUBE(N_FC_FLUX)   ] = new (std::nothrow) real [npg][3][NCOMP_TOTAL_PLUS_MAG][ CUBE(N_FC_FLUX)   ];
   if ( h_PriVar == NULL || h_FC_Var == NULL || h_FC_Flux == NULL )
   {
      delete [] h_PriVar;  delete [] h_FC_Var;  delete [] h_FC_Flux;
      return 2;
   }

   if ( nthreads > 0 )  omp_set_num_threads( nthreads );
   omp_set_schedule( omp_sched_dynamic, 1 );     // GAMER's run-time OpenMP schedule (src/Init/Init_OpenMP.cpp:34)

   CPU_FluidSolver_MHM( (const real (*)[NCOMP_TOTAL][ CUBE(FLU_NXT) ])in,
                        (real (*)[NCOMP_TOTAL][ CUBE(PS2) ])out,
                        NULL, NULL, NULL,
                        (real (*)[9][NCOMP_TOTAL][ SQR(PS2) ])flux,
                        NULL, NULL, NULL,
                        h_PriVar, h_Slope_PPM, h_FC_Var, h_FC_Flux, NULL, NULL,
                        npg, (real)dt, (real)dh, store_flux != 0, false,
                        (LR_Limiter_t)limiter, (real)minmod_coeff, minmod_maxiter, 0.0,
                        false, EXT_ACC_NONE, NULL, NULL,
                        (real)min_dens, (real)min_pres, (real)min_eint,
                        (real)0.0, 0L, false, 0, NULL, false, 0, NULL, false, (real)0.0,
                        EoS_obj, MicroPhy_obj );

   delete [] h_PriVar;  delete [] h_FC_Var;  delete [] h_FC_Flux;
   return 0;
}

// One call of GAMER's CPU_dtSolver_HydroCFL on npatch = 8*NPG patches: in [npatch][5][8^3] -> dt [npatch].
// The reference loops over the 8 patches of each of NPG patch groups, so npatch must be a multiple of 8.
int gamer_ref_dt_cfl( const void *in, void *dt, int npatch, double dh, double safety, double gamma,
                      double min_pres, int nthreads )
{
   if ( npatch <= 0 || npatch % 8 != 0 || in == NULL || dt == NULL )   return 1;
   const EoS_t EoS_obj = make_eos( gamma );
   MicroPhy_t MicroPhy_obj;
   memset( &MicroPhy_obj, 0, sizeof(MicroPhy_obj) );
   if ( nthreads > 0 )  omp_set_num_threads( nthreads );
   omp_set_schedule( omp_sched_dynamic, 1 );
   CPU_dtSolver_HydroCFL( (real *)dt, (const real (*)[FLU_NIN_T][ CUBE(PS1) ])in, NULL, npatch/8,
                          (real)dh, (real)safety, (real)min_pres, 0L, EoS_obj, MicroPhy_obj );
   return 0;
}

// One call of GAMER's Interpolate() per box (src/Interpolation/Interpolate.cpp:90-330, with the schemes of
// src/Interpolation/Int_*.cpp), hydro: NComp = 5 conserved fields, AllCons, no phase unwrapping, no magnetic field.
//   cdata [nbox][5][CSize]  (Interpolate may convert it in place)   fdata [nbox][5][FSize]
//   monotonic: bit v = Monotonic[v]; int_prim / reduce_mono: IntPrim_t / ReduceOrFixMonoCoeff_t values;
//   mono_coeff, mono_max_iter, min_pres: INT_MONO_COEFF, MONO_MAX_ITER and MIN_PRES of the run.
// A box whose interpolation still fails after the last iteration makes the reference abort the process
// (MPI_Exit), so callers pass only inputs it accepts.
int gamer_ref_interpolate( void *cdata, void *fdata, int nbox, const int csize[3], const int cstart[3],
                           const int crange[3], const int fsize[3], const int fstart[3], int scheme, int monotonic,
                           double mono_coeff, int opp_sign0, int int_prim, int reduce_mono, int mono_max_iter,
                           double gamma, double min_pres )
{
   if ( nbox <= 0 || cdata == NULL || fdata == NULL )   return 1;
   make_eos( gamma );
   INT_MONO_COEFF = mono_coeff;
   MONO_MAX_ITER  = mono_max_iter;
   MIN_PRES       = min_pres;
   bool Monotonic[NCOMP_TOTAL];
   for (int v=0; v<NCOMP_TOTAL; v++)   Monotonic[v] = ( monotonic >> v ) & 1;
   const long nc = (long)NCOMP_TOTAL*csize[0]*csize[1]*csize[2];
   const long nf = (long)NCOMP_TOTAL*fsize[0]*fsize[1]*fsize[2];
   for (int b=0; b<nbox; b++)
      Interpolate( (real *)cdata + b*nc, csize, cstart, crange, (real *)fdata + b*nf, fsize, fstart, NCOMP_TOTAL,
                   (IntScheme_t)scheme, false, Monotonic, opp_sign0 != 0, true, (IntPrim_t)int_prim,
                   (ReduceOrFixMonoCoeff_t)reduce_mono, NULL, NULL );
   return 0;
}

} // extern "C"
