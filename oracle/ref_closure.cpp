// Test-infrastructure harness: compiles GAMER's own src/Fluid/Flu_Close.cpp in place (it is #included from the
// reference tree, found through oracle/Makefile's include path) so that its file-local CorrectUnphysical()
// (src/Fluid/Flu_Close.cpp:542-1297) can be called on an array of patch groups, and the oracle's restatement
// (oracle/hydro_oracle.cpp, correct_unphysical_pg) and the GPU closure can be checked against the reference.
//
// This file is NOT product code and contains no solver arithmetic. It only
//   (1) defines the run-time parameters that Flu_Close.cpp and Output_Patch.cpp read and that GAMER's
//       Main.cpp / Flu_AdvanceDt.cpp would otherwise define (include/Global.h:22-125, include/Field.h:104,
//       src/Fluid/Flu_AdvanceDt.cpp:5), beside those ref_harness.cpp already defines,
//   (2) builds the one AMR_t field CorrectUnphysical reads (amr->dh[lv]), and
//   (3) sets OPT__1ST_FLUX_CORR, OPT__1ST_FLUX_CORR_SCHEME, OPT__LAST_RESORT_FLOOR, OPT__FIXUP_FLUX,
//       OPT__CHECK_PRES_AFTER_FLU, AUTO_REDUCE_DT / AutoReduceDt_Continue and MIN_DENS/PRES/EINT from its
//       arguments, initialises EOS_GAMMA as ref_harness.cpp does, and calls CorrectUnphysical as Flu_Close
//       does (src/Fluid/Flu_Close.cpp:98-99), non-MHD, no dual energy, no passive scalars.

#include "Flu_Close.cpp"
#include <omp.h>
#include <new>

// ---- run-time parameters read by Flu_Close.cpp and Output_Patch.cpp ----
bool   AUTO_REDUCE_DT             = false;
bool   AutoReduceDt_Continue      = false;
int    FluStatus_ThisRank         = GAMER_SUCCESS;
long   AdvanceCounter[NLEVEL];
long   NCorrUnphy[NLEVEL];
long   Step                       = 0L;
double Time[NLEVEL];
double MIN_DENS                   = 0.0;
double MIN_EINT                   = 0.0;
bool   OPT__FIXUP_FLUX            = true;
bool   OPT__LAST_RESORT_FLOOR     = true;
int    OPT__CHECK_PRES_AFTER_FLU  = 0;
Opt1stFluxCorr_t OPT__1ST_FLUX_CORR        = FIRST_FLUX_CORR_NONE;
OptRSolver1st_t  OPT__1ST_FLUX_CORR_SCHEME = RSOLVER_1ST_NONE;
char   OUTPUT_DIR[MAX_STRING]     = ".";
char   BlankPlusFormat_Flt[MAX_STRING+1] = " %13.6e";
int    StrLen_Flt                 = 14;
char   FieldLabel[NCOMP_TOTAL][MAX_STRING] = { "Dens", "MomX", "MomY", "MomZ", "Engy" };
AMR_t *amr                        = NULL;

void EoS_Init_Gamma();

extern "C" {

// One call of GAMER's CorrectUnphysical on NPG patch groups (level 0, cell width dh).
//   in   : [npg][5][FLU_NXT^3]   out : [npg][5][16^3] (corrected in place)   flux : [npg][9][5][16^2] or NULL
//   mode : OPT__1ST_FLUX_CORR (0 none, 1 3D, 2 3D1D)   rsolver_1st : OPT__1ST_FLUX_CORR_SCHEME (1 Roe, 2 HLLC, 3 HLLE)
//   counters[0] : NCorrUnphy[0] added by the call (the reference adds it only when no cell failed)
//   counters[1] : 1 if the call left FluStatus_ThisRank at GAMER_FAILED (a cell still unphysical), else 0
// A cell that is still unphysical with auto_reduce == 0 makes the reference write a diagnostic file under
// OUTPUT_DIR and abort the process (Aux_Error -> MPI_Exit), so callers pass auto_reduce = 1 for inputs that
// can fail. Returns 0, or 1 on a bad argument.
int gamer_ref_correct_unphysical( const void *in, void *out, void *flux, int npg, double dt, double dh,
                                  double gamma, int mode, int rsolver_1st, int last_resort_floor, int auto_reduce,
                                  int fixup_flux, int check_pres, double min_dens, double min_pres,
                                  double min_eint, long counters[2], int nthreads )
{
   if ( npg <= 0 || in == NULL || out == NULL || counters == NULL || mode < 0 || mode > 2 )  return 1;
   if ( mode != FIRST_FLUX_CORR_NONE && ( rsolver_1st < RSOLVER_1ST_ROE || rsolver_1st > RSOLVER_1ST_HLLE ) )
      return 1;
   if ( fixup_flux && mode != FIRST_FLUX_CORR_NONE && flux == NULL )  return 1;

   if ( amr == NULL )   amr = new (std::nothrow) AMR_t;
   if ( amr == NULL )   return 1;
   amr->dh[0] = dh;

   GAMMA = gamma;
   EoS_Init_Gamma();

   MIN_DENS                  = min_dens;
   MIN_PRES                  = min_pres;
   MIN_EINT                  = min_eint;
   OPT__1ST_FLUX_CORR        = (Opt1stFluxCorr_t)mode;
   OPT__1ST_FLUX_CORR_SCHEME = ( mode == FIRST_FLUX_CORR_NONE ) ? RSOLVER_1ST_NONE : (OptRSolver1st_t)rsolver_1st;
   OPT__LAST_RESORT_FLOOR    = ( last_resort_floor != 0 );
   OPT__FIXUP_FLUX           = ( fixup_flux != 0 );
   OPT__CHECK_PRES_AFTER_FLU = check_pres;
   AUTO_REDUCE_DT            = ( auto_reduce != 0 );
   AutoReduceDt_Continue     = ( auto_reduce != 0 );
   FluStatus_ThisRank        = GAMER_SUCCESS;
   NCorrUnphy[0]             = 0L;

   int *PID0_List = new (std::nothrow) int [npg];
   if ( PID0_List == NULL )  return 1;
   for (int t=0; t<npg; t++)   PID0_List[t] = 8*t;

   if ( nthreads > 0 )  omp_set_num_threads( nthreads );
   omp_set_schedule( omp_sched_dynamic, 1 );     // GAMER's run-time OpenMP schedule (src/Init/Init_OpenMP.cpp:34)

   CorrectUnphysical( 0, npg, PID0_List,
                      (const real (*)[FLU_NIN][ CUBE(FLU_NXT) ])in,
                      (real (*)[FLU_NOUT][ CUBE(PS2) ])out,
                      NULL,
                      (real (*)[9][NFLUX_TOTAL][ SQR(PS2) ])flux,
                      NULL, NULL, (real)dt );

   delete [] PID0_List;
   counters[0] = NCorrUnphy[0];
   counters[1] = ( FluStatus_ThisRank == GAMER_FAILED ) ? 1 : 0;
   return 0;
}

} // extern "C"
