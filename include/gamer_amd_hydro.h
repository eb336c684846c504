/*
 * gamer_amd_hydro.h — C ABI of the MI355X (gfx950) MUSCL-Hancock patch-group fluid solver.
 *
 * Drop-in boundary for GAMER's fluid-solver GPU API (src/GPU_API, include/Prototype.h:646-690).  Plain
 * pointers, sizes and enums only.  Every function returns HYDRO_OK (0) or a negative HYDRO_E* code; the
 * message of the last failure on the calling thread is available from hydro_last_error().
 *
 * Array layouts are those of the reference boundary (include/Macro.h:1197, x fastest):
 *   in   [npg][5][N][N][N]      N = 22 (MHM_RP) or 20 (MHM); fields rho, mx, my, mz, E
 *   out  [npg][5][16][16][16]
 *   flux [npg][9][5][16][16]    x planes at 0/8/16, then y, then z (CPU_Shared_ComputeFlux.cpp:390-450)
 * Element type is double for HYDRO_FP64 and float for HYDRO_FP32_BULK.
 */
#ifndef GAMER_AMD_HYDRO_H
#define GAMER_AMD_HYDRO_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
   HYDRO_OK            =  0,
   HYDRO_E_ARG         = -1,   /* invalid argument (null pointer, bad enum, npg out of range) */
   HYDRO_E_STATE       = -2,   /* call order: not initialised / buffers not allocated */
   HYDRO_E_DEVICE      = -3,   /* HIP runtime error */
   HYDRO_E_NOMEM       = -4    /* device or pinned allocation failed (caller may retry with fewer streams) */
};

/* scheme ids = include/Macro.h:74-75; Riemann solvers = Macro.h:86-88; limiters = include/Typedef.h:188-193 */
enum { HYDRO_MHM = 3, HYDRO_MHM_RP = 4 };
enum { HYDRO_ROE = 2, HYDRO_HLLE = 3, HYDRO_HLLC = 4 };
enum { HYDRO_LIM_VANLEER = 1, HYDRO_LIM_GMINMOD = 2, HYDRO_LIM_ALBADA = 3, HYDRO_LIM_VL_GMINMOD = 4,
       HYDRO_LIM_CENTRAL = 6 };
/* HYDRO_FP64: double storage and arithmetic.  HYDRO_FP32_BULK: float storage and bulk arithmetic, with
   pressure recovery (E - |m|^2/2rho) and the full-step conserved accumulation U - dt/dh sum dF in double. */
enum { HYDRO_FP64 = 0, HYDRO_FP32_BULK = 1 };

/* Arguments of one solver call (the scalar arguments of CUAPI_Asyn_FluidSolver, src/GPU_API/
   CUAPI_Asyn_FluidSolver.cu:266-288, restricted to non-MHD hydro without passive scalars). */
typedef struct HydroSolveArgs {
   const void *in;         /* device pointer, layout above */
   void       *out;        /* device pointer */
   void       *flux;       /* device pointer, may be NULL when store_flux == 0 */
   int    npg;             /* NPatchGroup */
   double dt, dh;
   double gamma;           /* EOS_GAMMA adiabatic index */
   int    scheme;          /* HYDRO_MHM_RP or HYDRO_MHM */
   int    rsolver;         /* HYDRO_HLLC / HYDRO_HLLE / HYDRO_ROE */
   int    limiter;         /* LR_Limiter */
   double minmod_coeff;    /* MinMod_Coeff */
   int    minmod_maxiter;  /* MinMod_MaxIter (>= 0) */
   double min_dens, min_pres, min_eint;
   int    store_flux;      /* StoreFlux */
   int    precision;       /* HYDRO_FP64 or HYDRO_FP32_BULK */
   void  *dt_patch;        /* optional device pointer [npg][8]: fused CFL dt of the updated patches
                              (index px + 2 py + 4 pz within the PG); NULL = not computed */
   double safety;          /* CFL safety factor of the fused dt (DT__FLUID) */
   int    waves_per_group; /* launch shape of the MHM_RP kernel: 0 = default (8), 4 or 8 waves per workgroup
                              (results are bitwise equal); MHM accepts 0 or 4.  No counterpart in the reference,
                              whose block size is a compile-time constant (include/CUFLU.h:302-371). */
} HydroSolveArgs;

/* Select the HIP device (replaces CUAPI_SetDevice, src/GPU_API/CUAPI_SetDevice.cu). */
int hydro_init( int device );

/* Enqueue one solver call on `stream` (a hipStream_t, NULL = default stream) over device-resident arrays.
   Returns after the launch; results are valid once the stream is synchronised.  Replaces the kernel launch
   inside CUAPI_Asyn_FluidSolver (src/GPU_API/CUAPI_Asyn_FluidSolver.cu:557-576). */
int hydro_mhm_solve_async( const HydroSolveArgs *args, void *stream );

/* One launch over patch groups that need different time steps, e.g. the PGs of several AMR levels at once:
   PG p advances with dt = pg_dt_dh[2p], dh = pg_dt_dh[2p+1] (device double array [npg][2]); args->dt/dh are
   ignored when pg_dt_dh != NULL.  With queue_counter (a device unsigned scratch word) the kernel runs
   persistently: groups_per_cu (1..16) workgroups per compute unit pull tickets of pgs_per_ticket (1..64)
   consecutive patch groups from a global work queue whose counter is reset by a kernel on the same stream
   (graph-capturable, no memset).  With queue_counter == NULL there is one workgroup per PG and groups_per_cu,
   pgs_per_ticket are ignored.  No counterpart in the reference, which issues one CUAPI_Asyn_FluidSolver call
   per level and substep (src/Main/EvolveLevel.cpp:203). */
int hydro_mhm_solve_batch_async( const HydroSolveArgs *args, const double *pg_dt_dh, unsigned *queue_counter,
                                 int groups_per_cu, int pgs_per_ticket, void *stream );

/* Number of work-queue counters the library keeps in its own device memory for hydro_mhm_solve_batch_noalloc. */
#define HYDRO_NOALLOC_QUEUES 64

/* hydro_mhm_solve_batch_async in its persistent mode without a caller-allocated counter: the work-queue
   counter is word `queue_slot` (0..HYDRO_NOALLOC_QUEUES-1; anything else returns HYDRO_E_ARG) of a fixed
   256-byte device array that the library holds on every GPU it runs on until the process ends.  The call
   allocates nothing (no hipMalloc, no library allocation); the array's address is looked up by hydro_init (or
   by the first call on a device).  The counter is reset by a kernel on `stream` before every solve, so the
   slot keeps no state between calls; two calls that overlap on one GPU with the same slot race, calls on
   different slots do not.  No counterpart in the reference: it is for callers that cannot allocate device
   memory (code under HIP graph capture, tight serving loops). */
int hydro_mhm_solve_batch_noalloc( const HydroSolveArgs *args, const double *pg_dt_dh, int queue_slot,
                                   int groups_per_cu, int pgs_per_ticket, void *stream );

/* Per-patch hydro CFL time step on device arrays: in[npatch][5][8][8][8] -> dt[npatch],
   dt = safety*dh / max_cells max_d(|v_d| + c_s) (replaces CUAPI_Asyn_dtSolver for DT_FLU_SOLVER,
   src/GPU_API/CUAPI_Asyn_dtSolver.cu:88, kernel src/Model_Hydro/CPU_Hydro/CPU_dtSolver_HydroCFL.cpp:57-209). */
int hydro_dt_cfl_async( const void *in, void *dt, int npatch, double dh, double safety, double gamma,
                        double min_pres, int precision, void *stream );

/* Ghost-zone preparation for one periodic uniform level held on the device as level[5][nz][ny][nx]
   (nx, ny, nz multiples of 16): writes pg_in[npg][5][N^3], N = 16 + 2*ghost, patch groups numbered x fastest
   (the same-level periodic case of Prepare_PatchData for the fluid solver, src/Fluid/Flu_Prepare.cpp:91-94). */
int hydro_prepare_uniform_async( const void *level, void *pg_in, int nx, int ny, int nz, int ghost, int precision,
                                 void *stream );

/* Scatter of patch-group outputs pg_out[npg][5][16^3] back into level[5][nz][ny][nx]
   (the copy-back step of Flu_Close, src/Fluid/Flu_Close.cpp:133-217). */
int hydro_close_uniform_async( const void *pg_out, void *level, int nx, int ny, int nz, int precision, void *stream );

/* ---- Flu_Close and the AMR level synchronisation on device arrays ----
   Patch data use GAMER's per-patch layout patch[npatch][5][8][8][8]; the 8 patches of a patch group are
   PID0 + LocalID with the TABLE_02 offsets (src/Tables/Table_02.cpp:22-70): LocalID 0..7 =
   (0,0,0) (1,0,0) (0,1,0) (0,0,1) (1,1,0) (0,1,1) (1,0,1) (1,1,1).  Flux registers (the patch->flux[s] buffers of
   coarse patches on coarse-fine faces) are double reg[nreg][5][8][8] in both precisions. */

/* Options of CorrectUnphysical (src/Fluid/Flu_Close.cpp:542-1297, run-time parameters of the same names). */
typedef struct HydroCorrectArgs {
   const void *in;            /* the solver input of the call, [npg][5][N^3] */
   void       *out;           /* the solver output [npg][5][16^3], corrected in place */
   void       *flux;          /* [npg][9][5][16^2] or NULL; coarse-fine planes get the first-order fluxes */
   int    npg;
   double dt, dh, gamma;
   int    scheme;             /* HYDRO_MHM_RP or HYDRO_MHM (ghost width of `in`) */
   int    mode;               /* OPT__1ST_FLUX_CORR: 0 none, 1 3D, 2 3D1D */
   int    rsolver;            /* OPT__1ST_FLUX_CORR_SCHEME as HYDRO_ROE / HYDRO_HLLE / HYDRO_HLLC */
   int    last_resort_floor;  /* OPT__LAST_RESORT_FLOOR */
   int    auto_reduce_dt;     /* AutoReduceDt_Continue */
   int    fixup_flux;         /* OPT__FIXUP_FLUX */
   int    check_pres;         /* OPT__CHECK_PRES_AFTER_FLU */
   double min_dens, min_pres, min_eint;
   int    precision;
   int   *counters;           /* device int[2]: [0] += corrected cells, [1] += cells still unphysical (failure) */
   int    main_rsolver;       /* the fluid solver's Riemann solver, GAMER's compile-time RSOLVER (0: same as rsolver).
                                 A first-order Roe solve checks its intermediate states, falling back to HLLE, only
                                 when this is HYDRO_ROE: CHECK_INTERMEDIATE exists only in RSOLVER == ROE builds
                                 (include/CUFLU.h:178-186, CPU_Shared_RiemannSolver_Roe.cpp:606-667). */
} HydroCorrectArgs;

/* Replaces CorrectUnphysical (src/Fluid/Flu_Close.cpp:542-1297) and its Unphysical test (:412-535). */
int hydro_correct_unphysical_async( const HydroCorrectArgs *args, void *stream );

/* StoreFlux (src/Fluid/Flu_Close.cpp:234-307): reg[slot[pg][LocalID][s]] = dt * flux of patch face s
   (slot [npg][8][6], -1 = no register). */
int hydro_store_flux_async( const void *flux, const int *slot, int npg, double dt, double *reg, int precision,
                            void *stream );

/* CorrectFlux (src/Fluid/Flu_Close.cpp:321-395): reg[slot[pg][s]] -= dt/4 * (the 4 fine fluxes of each coarse face
   cell on patch-group face s), slot [npg][6] = the coarse sibling's register, -1 = none. */
int hydro_correct_flux_async( const void *flux, const int *slot, int npg, double dt, double *reg, int precision,
                              void *stream );

/* Flu_FixUp_Flux (src/Fluid/Flu_FixUp_Flux.cpp): cells on registered faces (slot [npatch][6]) get
   U += reg * (-1/dh, +1/dh per face side) unless the result fails the density / Eint / pressure test. */
int hydro_fixup_flux_async( void *patch, const int *slot, const double *reg, int npatch, double dh, double gamma,
                            double min_dens, double min_pres, double min_eint, int precision, void *stream );

/* Flu_FixUp_Restrict (src/Fluid/Flu_FixUp_Restrict.cpp): fa[fa_idx[q]] = mean of the 8 son patches
   son[son0[q] + LocalID], then the internal-energy floor (Hydro_CheckMinEintInEngy). */
int hydro_fixup_restrict_async( const void *son, void *fa, int ngroup, const int *son0, const int *fa_idx,
                                double min_eint, int precision, void *stream );

/* The copy-back of Flu_Close (src/Fluid/Flu_Close.cpp:133-160): patch-group outputs pg_out[npg][5][16^3] into
   the patches pid0[pg] + LocalID. */
int hydro_close_patches_async( const void *pg_out, void *patch, int npg, const int *pid0, int precision,
                               void *stream );

/* ---- Ghost-zone preparation at refinement boundaries ---- */

/* Coarse -> fine interpolation of a batch of equally shaped boxes: Interpolate() with Interpolate_Iterate
   (src/Interpolation/Interpolate.cpp:90-330) for INT_MINMOD3D, INT_MINMOD1D, INT_VANLEER, INT_CQUAD, INT_QUAD,
   INT_CQUAR and INT_QUAR (src/Interpolation/Int_*.cpp).  Boxes are cdata[nbox][5][csize[2]][csize[1]][csize[0]]
   and fdata[nbox][5][fsize...]; coarse cells cstart..cstart+crange-1 (1 coarse ghost cell needed on each side,
   2 for INT_CQUAR / INT_QUAR: Int_Table.cpp) fill fine cells
   fstart..fstart+2*crange-1.  With int_prim (OPT__INT_PRIM) or reduce_mono the result is checked and failing
   2x2x2 blocks are redone in primitive variables, then with lower monotonic coefficients (MONO_MAX_ITER); the
   primitive retry converts cdata in place, as the reference does.  *nfail += number of boxes still unphysical. */
typedef struct HydroInterpArgs {
   void  *cdata, *fdata;
   int    nbox;
   int    csize[3], cstart[3], crange[3], fsize[3], fstart[3];
   int    scheme;              /* 1 INT_MINMOD3D, 2 INT_MINMOD1D, 3 INT_VANLEER, 4 INT_CQUAD, 5 INT_QUAD, 6 INT_CQUAR,
                                  7 INT_QUAR (include/Typedef.h:173-179); reduce_mono is rejected for 1-3 */
   int    monotonic;           /* bit v set: Monotonic[v] */
   double mono_coeff;          /* INT_MONO_COEFF */
   int    opp_sign_0th;        /* OppSign0thOrder */
   int    int_prim;            /* OPT__INT_PRIM */
   int    reduce_mono;         /* ReduceMonoCoeff */
   int    mono_max_iter;       /* MONO_MAX_ITER */
   double gamma, min_pres;
   int    precision;
   int   *nfail;               /* device int */
   const unsigned char *need;  /* optional device [nbox][crange[2]][crange[1]][crange[0]]: 0 = the coarse cell's fine
                                  cells are not used (a same-level patch covers them), skip it; NULL = all cells */
} HydroInterpArgs;
int hydro_interpolate_async( const HydroInterpArgs *args, void *stream );

/* Same-level part of Prepare_PatchData for the fluid solver (src/Main/Prepare_PatchData.cpp:166): cell (x,y,z) of
   patch-group input pg_in[npg][5][N^3] (N = 16 + 2*ghost) is copied from patch slot[pg][sz][sy][sx] (4x4x4 table
   of patch indices covering the window, patch offset -1..2 per axis; -1 = no patch at this level: the cell keeps
   its value, e.g. one written by hydro_interpolate_async). */
int hydro_prepare_patches_async( const void *patch, void *pg_in, int npg, const int *slot, int ghost, int precision,
                                 void *stream );

/* Coarse data of InterpolateGhostZone (src/Main/InterpolateGhostZone.cpp, the CData it hands to Interpolate()):
   box[nbox][5][size[2]][size[1]][size[0]], cell (x,y,z) of box q taken from the patch that covers cell
   origin + (x,y,z) of the box's table slot[q][tdim[2]][tdim[1]][tdim[0]] of patch indices into patch[][5][8^3]
   (-1: cell left unchanged).  For a fine patch group the table is its father patch and the 26 siblings
   (tdim = 3) and the box is coarse cells -2-G .. 9+G around the father (size 12+2G, origin 6-G), G the
   scheme's coarse ghost width. */
int hydro_gather_boxes_async( const void *patch, void *box, int nbox, const int size[3], const int origin[3],
                              const int *slot, const int tdim[3], int precision, void *stream );

/* Patch-group inputs at a refinement boundary (Prepare_PatchData, src/Main/Prepare_PatchData.cpp:166): cells
   with a same-level patch in slot[npg][4][4][4] copy it, as hydro_prepare_patches_async; the others take the
   interpolated fine box fbox[npg][5][fb_size...] at window cell + fb_origin (for the box above, interpolated
   over coarse cells G .. 11+G: fb_size 24, fb_origin 1 with ghost 3). */
int hydro_assemble_patch_groups_async( const void *patch, const void *fbox, void *pg_in, int npg, const int *slot,
                                       int ghost, const int fb_size[3], const int fb_origin[3], int precision,
                                       void *stream );

/* ---- Differentiation of one MHM_RP step (HYDRO_FP64 only) ----
   The step is the map in[npg][5][22^3] -> out[npg][5][16^3] of hydro_mhm_solve_async with the same arguments
   (dt, dh constant; the MinMod retry is not differentiated).  No counterpart in the reference, which has no
   derivative of CPU_FluidSolver_MHM (src/Model_Hydro/CPU_Hydro/CPU_FluidSolver_MHM.cpp:263-589); these are the
   reverse and forward derivatives of that function.  `workspace` is device memory of
   hydro_mhm_adjoint_workspace_bytes(npg) bytes, used by one call at a time. */
size_t hydro_mhm_adjoint_workspace_bytes( int npg );

/* Vector-Jacobian product: lam_in[npg][5][22^3] = (d out / d in)^T lam_out[npg][5][16^3], evaluated at args->in
   (args->out and args->flux are not used). */
int hydro_mhm_adjoint_async( const HydroSolveArgs *args, const void *lam_out, void *lam_in, void *workspace,
                             void *stream );

/* Jacobian-vector product: jv_out[npg][5][16^3] = (d out / d in) v_in[npg][5][22^3], and args->out (if not NULL)
   receives the step's output.  v_in == NULL runs the same phase-wise evaluation without tangents (out only). */
int hydro_mhm_tangent_async( const HydroSolveArgs *args, const void *v_in, void *jv_out, void *workspace,
                             void *stream );

/* Transpose of hydro_prepare_uniform_async: level_lam[5][nz][ny][nx] = sum over all patch-group input elements
   pg_lam[npg][5][N^3] that copy the same level cell (ghost copies included). */
int hydro_prepare_uniform_adjoint_async( const void *pg_lam, void *level_lam, int nx, int ny, int nz, int ghost,
                                         int precision, void *stream );

/* Wait for all work queued by this library (replaces CUAPI_Synchronize, src/GPU_API/CUAPI_Synchronize.cu:11-14). */
int hydro_synchronize( void );

/* Message of the last failure on the calling thread ("" if none). */
const char *hydro_last_error( void );

/* ---- GAMER GPU_API-shaped entry points over host arrays (pinned double buffers + N streams) ---- */

/* Allocate device in/out/flux buffers for flu_npg patch groups and nstream HIP streams, and pinned host
   double buffers [2]; also the CFL dt solver's buffers for the 8*flu_npg patches of those patch groups
   (replaces CUAPI_MemAllocate_Fluid, src/GPU_API/CUAPI_MemAllocate_Fluid.cu:78-347).
   Returns HYDRO_E_NOMEM on allocation failure so the caller can retry with fewer streams
   (src/GPU_API/CUAPI_MemAllocate.cu:41-96). */
int hydro_mem_allocate_fluid( int flu_npg, int nstream, int scheme, int precision );

/* Free everything allocated by hydro_mem_allocate_fluid (replaces CUAPI_MemFree_Fluid). */
int hydro_mem_free_fluid( void );

/* Pinned host buffers [array_id] for the input, output and flux arrays (the h_Flu_Array_In/Out and
   h_Flux_Array double buffers the caller fills and reads between calls). */
int hydro_host_buffers( int array_id, void **h_in, void **h_out, void **h_flux );

/* Pinned host buffers [array_id] of the CFL dt solver: h_flu_t [8*flu_npg][5][8^3] (the fluid data of the
   patches) and h_dt [8*flu_npg] (h_Flu_Array_T / h_dt_Array_T, src/GPU_API/CUAPI_MemAllocate_Fluid.cu:315-316).
   Element type as the allocation's precision. */
int hydro_host_dt_buffers( int array_id, void **h_flu_t, void **h_dt );

/* Asynchronous per-patch hydro CFL dt over host arrays on the allocated streams (replaces CUAPI_Asyn_dtSolver for
   DT_FLU_SOLVER, src/GPU_API/CUAPI_Asyn_dtSolver.cu:88-330): the 8*npg patches are split NPatch/NStream per
   stream with the remainder on the last; per stream H2D copy of h_flu_t, the dt kernel (hydro_dt_cfl_async), D2H
   copy into h_dt.  dt = safety*dh / max_cells max_d(|v_d| + c_s).  Returns before the GPU finishes; the results are
   valid after hydro_synchronize. */
int hydro_asyn_dt_solver( const void *h_flu_t, void *h_dt, int npg, double dh, double safety, double gamma,
                          double min_pres, int gpu_nstream );

/* Properties of the current device (the queries of CUAPI_DiagnoseDevice, src/GPU_API/CUAPI_DiagnoseDevice.cu:14-140,
   and of CUAPI_SetMemSize, src/GPU_API/CUAPI_SetMemSize.cu:19-40). */
typedef struct HydroDeviceInfo {
   char   name[256];
   char   arch[64];                /* e.g. "gfx950:sramecc+:xnack-" */
   int    device, device_count;
   int    compute_units;           /* the multiprocessor count GAMER sizes FLU_GPU_NPGROUP by */
   int    clock_khz;
   size_t global_mem_bytes;
   size_t lds_per_block_bytes;
   int    wavefront_size;
   int    max_threads_per_block;
   int    driver_version, runtime_version;
   int    ecc, async_engines, concurrent_kernels;
} HydroDeviceInfo;
int hydro_device_info( HydroDeviceInfo *info );

/* Asynchronous solve over host arrays: per stream H2D copy, kernel, D2H copy of out (and flux), the
   NPG/NStream split with the remainder on the last stream (src/GPU_API/CUAPI_Asyn_FluidSolver.cu:433-722).
   Host arrays should be the pinned buffers from hydro_host_buffers.  Returns before the GPU finishes.  As in
   GAMER (src/Main/InvokeSolver.cpp:227-267), one call is in flight at a time: both host buffers share the device
   arrays, so call hydro_synchronize before the next call; the other host buffer can be filled or read meanwhile. */
int hydro_asyn_fluid_solver( const void *h_in, void *h_out, void *h_flux, const HydroSolveArgs *args,
                             int gpu_nstream );

#ifdef __cplusplus
}
#endif

#endif /* GAMER_AMD_HYDRO_H */
