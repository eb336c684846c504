// Test-infrastructure harness: compiles GAMER's own src/Fluid/Flu_Close.cpp in place (it is #included from the
// reference tree, found through oracle/Makefile's include path) so that its file-local CorrectUnphysical()
// (src/Fluid/Flu_Close.cpp:542-1297), StoreFlux() (:234-307) and CorrectFlux() (:321-395) and its Flu_Close()
// (:83-214) can be called on arrays of patch groups, and links Flu_FixUp_Flux() and Flu_FixUp_Restrict()
// (src/Fluid/Flu_FixUp_Flux.cpp, Flu_FixUp_Restrict.cpp), so that the oracle's restatements
// (oracle/hydro_oracle.cpp) and the GPU closure can be checked against the reference.
//
// This file is NOT product code and contains no solver arithmetic. It only
//   (1) defines the run-time parameters that these files and Output_Patch.cpp read and that GAMER's
//       Main.cpp / Flu_AdvanceDt.cpp would otherwise define (include/Global.h:22-125, include/Field.h:104,
//       src/Fluid/Flu_AdvanceDt.cpp:5), beside those ref_harness.cpp already defines,
//   (2) builds the AMR_t fields they read: amr->dh[lv], and for the flux-register and fix-up calls small
//       patch trees made with GAMER's own AMR_t::pnew / patch_t::fnew (include/AMR.h, include/Patch.h), whose
//       father / sibling links and flux arrays are set from the caller's slot tables,
//   (3) sets OPT__1ST_FLUX_CORR, OPT__1ST_FLUX_CORR_SCHEME, OPT__LAST_RESORT_FLOOR, OPT__FIXUP_FLUX,
//       OPT__CHECK_PRES_AFTER_FLU, AUTO_REDUCE_DT / AutoReduceDt_Continue and MIN_DENS/PRES/EINT from its
//       arguments, initialises EOS_GAMMA as ref_harness.cpp does, and calls the functions as Flu_Close and
//       EvolveLevel do (src/Fluid/Flu_Close.cpp:98-110, src/Main/EvolveLevel.cpp:757, 815), non-MHD, no dual
//       energy, no passive scalars (FixUpVar_Flux = FixUpVar_Restrict = _TOTAL, src/Init/Init_Field.cpp:257-258).

#include "Flu_Close.cpp"
#include <omp.h>
#include <new>
#include <cstring>

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

// the AMR_t of this harness, with the box and level fields that pnew / patch_t::Activate read
static bool amr_ready()
{
   if ( amr == NULL )
   {
      amr = new (std::nothrow) AMR_t;
      if ( amr == NULL )   return false;
      for (int d=0; d<3; d++)   { amr->BoxScale[d] = 2*PS1*( 1<<TOP_LEVEL );  amr->BoxEdgeL[d] = 0.0; }
      // a box of 2 patches a side at level 0: ResPower2 = log2 of the cells per side (src/Init/Init_ResetParameter.cpp:437)
      for (int lv=0; lv<NLEVEL; lv++)   { amr->ResPower2[lv] = 4 + lv;  amr->dh[lv] = 1.0/( 1<<lv );  Time[lv] = 0.0; }
      amr->WithFlux = true;
   }
   return true;
}

// level lv := npatch patches (all at corner 0: only their fluid arrays, links and flux arrays are used)
static void make_level( int lv, int npatch )
{
   amr->Lvdelete( lv, false );
   for (int p=0; p<npatch; p++)   amr->pnew( lv, 0, 0, 0, -1, true, false, false );
   for (int m=1; m<28; m++)   amr->NPatchComma[lv][m] = npatch;
   amr->FluSg[lv] = 0;
}

static void clear_levels()
{
   for (int lv=0; lv<NLEVEL; lv++)   if ( amr->num[lv] > 0 )   amr->Lvdelete( lv, false );
}

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

   if ( !amr_ready() )   return 1;
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

// StoreFlux (src/Fluid/Flu_Close.cpp:234-307) on level-1 patch groups PID0 = 8*pg: a flux array is allocated on
// face s of patch 8*pg+LocalID where slot[pg][LocalID][s] >= 0 and read back into reg[slot] ([nreg][5][8][8]).
int gamer_ref_store_flux( const void *flux, int npg, const int *slot, double dt, double *reg )
{
   if ( npg <= 0 || flux == NULL || slot == NULL || reg == NULL || !amr_ready() )   return 1;
   make_level( 1, 8*npg );
   int *PID0_List = new int [npg];
   for (int t=0; t<npg; t++)   PID0_List[t] = 8*t;
   for (int q=0; q<8*npg; q++)   for (int s=0; s<6; s++)   if ( slot[ 6*q + s ] >= 0 )   amr->patch[0][1][q]->fnew( s, false );
   StoreFlux( 1, (const real (*)[9][NFLUX_TOTAL][ SQR(PS2) ])flux, npg, PID0_List, (real)dt );
   for (int q=0; q<8*npg; q++)   for (int s=0; s<6; s++)
   {
      const int r = slot[ 6*q + s ];
      if ( r < 0 )   continue;
      const real *F = &amr->patch[0][1][q]->flux[s][0][0][0];
      for (int c=0; c<NFLUX_TOTAL*SQR(PS1); c++)   reg[ (long)r*NFLUX_TOTAL*SQR(PS1) + c ] = (double)F[c];
   }
   delete [] PID0_List;
   clear_levels();
   return 0;
}

// CorrectFlux (src/Fluid/Flu_Close.cpp:321-395) of level-1 patch groups PID0 = 8*pg, father PID pg on level 0.
// Where slot[pg][s] >= 0 the father's sibling s is a level-0 patch whose flux array on the mirrored face holds
// reg[slot]; elsewhere the sibling is -2 (a non-periodic boundary, skipped).  reg is updated in place.
int gamer_ref_correct_flux( const void *flux, int npg, const int *slot, double dt, double *reg )
{
   if ( npg <= 0 || flux == NULL || slot == NULL || reg == NULL || !amr_ready() )   return 1;
   const int MirrorSib[6] = { 1, 0, 3, 2, 5, 4 };
   int nnb = 0;
   for (int c=0; c<6*npg; c++)   nnb += ( slot[c] >= 0 );
   make_level( 0, npg + nnb );
   make_level( 1, 8*npg );
   int *PID0_List = new int [npg];
   for (int t=0, nb=npg; t<npg; t++)
   {
      PID0_List[t] = 8*t;
      for (int q=0; q<8; q++)   amr->patch[0][1][8*t+q]->father = t;
      for (int s=0; s<6; s++)
      {
         const int r = slot[ 6*t + s ];
         amr->patch[0][0][t]->sibling[s] = ( r >= 0 ) ? nb : -2;
         if ( r < 0 )   continue;
         amr->patch[0][0][nb]->fnew( MirrorSib[s], false );
         real *F = &amr->patch[0][0][nb]->flux[ MirrorSib[s] ][0][0][0];
         for (int c=0; c<NFLUX_TOTAL*SQR(PS1); c++)   F[c] = (real)reg[ (long)r*NFLUX_TOTAL*SQR(PS1) + c ];
         nb ++;
      }
   }
   AUTO_REDUCE_DT = false;
   CorrectFlux( 1, (const real (*)[9][NFLUX_TOTAL][ SQR(PS2) ])flux, npg, PID0_List, (real)dt );
   for (int t=0; t<npg; t++)   for (int s=0; s<6; s++)
   {
      const int r = slot[ 6*t + s ];
      if ( r < 0 )   continue;
      const real *F = &amr->patch[0][0][ amr->patch[0][0][t]->sibling[s] ]->flux[ MirrorSib[s] ][0][0][0];
      for (int c=0; c<NFLUX_TOTAL*SQR(PS1); c++)   reg[ (long)r*NFLUX_TOTAL*SQR(PS1) + c ] = (double)F[c];
   }
   delete [] PID0_List;
   clear_levels();
   return 0;
}

// Flu_FixUp_Flux (src/Fluid/Flu_FixUp_Flux.cpp) on npatch level-0 patches [npatch][5][8^3], updated in place:
// face s of patch p holds the flux array reg[slot[p][s]] where slot >= 0.
int gamer_ref_fixup_flux( void *patch, int npatch, const int *slot, const double *reg, double dh, double gamma,
                          double min_dens, double min_pres, double min_eint )
{
   if ( npatch <= 0 || patch == NULL || slot == NULL || reg == NULL || !amr_ready() )   return 1;
   GAMMA = gamma;
   EoS_Init_Gamma();
   MIN_DENS = min_dens;  MIN_PRES = min_pres;  MIN_EINT = min_eint;
   make_level( 0, npatch );
   amr->dh[0] = dh;
   real *P = (real *)patch;
   for (int p=0; p<npatch; p++)
   {
      memcpy( amr->patch[0][0][p]->fluid, P + (long)p*NCOMP_TOTAL*CUBE(PS1), sizeof(real)*NCOMP_TOTAL*CUBE(PS1) );
      for (int s=0; s<6; s++)
      {
         const int r = slot[ 6*p + s ];
         if ( r < 0 )   continue;
         amr->patch[0][0][p]->fnew( s, false );
         real *F = &amr->patch[0][0][p]->flux[s][0][0][0];
         for (int c=0; c<NFLUX_TOTAL*SQR(PS1); c++)   F[c] = (real)reg[ (long)r*NFLUX_TOTAL*SQR(PS1) + c ];
      }
   }
   Flu_FixUp_Flux( 0, _TOTAL );
   for (int p=0; p<npatch; p++)
      memcpy( P + (long)p*NCOMP_TOTAL*CUBE(PS1), amr->patch[0][0][p]->fluid, sizeof(real)*NCOMP_TOTAL*CUBE(PS1) );
   clear_levels();
   amr->dh[0] = 1.0;
   return 0;
}

// Flu_FixUp_Restrict (src/Fluid/Flu_FixUp_Restrict.cpp) from level 1 to level 0: group q's sons are the level-1
// patches 8q..8q+7 holding son[son0[q] + LocalID], its father is fa[fa_idx[q]] (nfa level-0 patches, updated).
int gamer_ref_fixup_restrict( const void *son, void *fa, int ngroup, const int *son0, const int *fa_idx, int nfa,
                              double min_eint )
{
   if ( ngroup <= 0 || nfa <= 0 || son == NULL || fa == NULL || son0 == NULL || fa_idx == NULL || !amr_ready() )
      return 1;
   MIN_EINT = min_eint;
   make_level( 0, nfa );
   make_level( 1, 8*ngroup );
   const real *S = (const real *)son;
   real *Fa = (real *)fa;
   for (int p=0; p<nfa; p++)
      memcpy( amr->patch[0][0][p]->fluid, Fa + (long)p*NCOMP_TOTAL*CUBE(PS1), sizeof(real)*NCOMP_TOTAL*CUBE(PS1) );
   for (int q=0; q<ngroup; q++)   for (int l=0; l<8; l++)
   {
      memcpy( amr->patch[0][1][8*q+l]->fluid, S + (long)( son0[q] + l )*NCOMP_TOTAL*CUBE(PS1),
              sizeof(real)*NCOMP_TOTAL*CUBE(PS1) );
      amr->patch[0][1][8*q+l]->father = fa_idx[q];
   }
   Flu_FixUp_Restrict( 0, 0, 0, NULL_INT, NULL_INT, NULL_INT, NULL_INT, _TOTAL, _NONE );
   for (int p=0; p<nfa; p++)
      memcpy( Fa + (long)p*NCOMP_TOTAL*CUBE(PS1), amr->patch[0][0][p]->fluid, sizeof(real)*NCOMP_TOTAL*CUBE(PS1) );
   clear_levels();
   return 0;
}

// Flu_Close (src/Fluid/Flu_Close.cpp:83-214) with no first-order correction, no floors and no flux fix-up, i.e.
// its copy of the patch-group outputs [npg][5][16^3] into the 8 patches PID0 + LocalID of each group
// (:125-212): patch [npatch][5][8^3] is updated in place.  Cells the solver left unphysical only raise the
// AUTO_REDUCE_DT flag here (AutoReduceDt_Continue is on), so the copy is all that changes the patches.
int gamer_ref_close_patches( const void *pg_out, void *patch, int npg, const int *pid0, int npatch )
{
   if ( npg <= 0 || npatch <= 0 || pg_out == NULL || patch == NULL || pid0 == NULL || !amr_ready() )   return 1;
   for (int t=0; t<npg; t++)   if ( pid0[t] < 0 || pid0[t] + 8 > npatch )   return 1;
   OPT__1ST_FLUX_CORR        = FIRST_FLUX_CORR_NONE;
   OPT__1ST_FLUX_CORR_SCHEME = RSOLVER_1ST_NONE;
   OPT__LAST_RESORT_FLOOR    = false;
   OPT__FIXUP_FLUX           = false;
   OPT__CHECK_PRES_AFTER_FLU = 0;
   AUTO_REDUCE_DT = AutoReduceDt_Continue = true;
   MIN_DENS = MIN_PRES = MIN_EINT = 0.0;
   make_level( 1, npatch );
   real *P = (real *)patch;
   for (int p=0; p<npatch; p++)
      memcpy( amr->patch[0][1][p]->fluid, P + (long)p*NCOMP_TOTAL*CUBE(PS1), sizeof(real)*NCOMP_TOTAL*CUBE(PS1) );
   // CorrectUnphysical reads the solver input only for cells it corrects with first-order fluxes (none here)
   Flu_Close( 1, 0, NULL_INT, NULL, NULL, (real (*)[FLU_NOUT][ CUBE(PS2) ])pg_out, NULL, NULL, npg, pid0,
              (const real (*)[FLU_NIN][ CUBE(FLU_NXT) ])pg_out, NULL, 0.0 );
   for (int p=0; p<npatch; p++)
      memcpy( P + (long)p*NCOMP_TOTAL*CUBE(PS1), amr->patch[0][1][p]->fluid, sizeof(real)*NCOMP_TOTAL*CUBE(PS1) );
   clear_levels();
   AUTO_REDUCE_DT = AutoReduceDt_Continue = false;
   return 0;
}

} // extern "C"
