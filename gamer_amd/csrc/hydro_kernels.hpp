// Internal launch interface between the C-ABI layer (hydro_api.hip) and the solver kernels.
#pragma once
#include <hip/hip_runtime.h>
#include "hydro_math.hpp"

namespace gamer_amd {

// The dispatcher deals the workgroups of a launch to the 8 XCDs round-robin in dispatch order (x, then y, then z:
// workgroup b runs on XCD b % 8), and each XCD has its own L2.  Kernels whose neighbouring workgroups read the same
// data (planes k-1..k+1 of the adjoint's plane-per-workgroup phases, their neighbouring face stencils) take
// their logical block index from xcd_block(), laid out XCD-major: XCD x runs the logical blocks
// [off_x, off_x + n_x) in dispatch order, n_x = G/8 or G/8 + 1.  It is a bijection of [0, G), so every logical
// block runs exactly once and only the placement changes.  64-bit, as a 3-D grid may hold more than 2^31 blocks.
__device__ __forceinline__ long xcd_block()
{
   constexpr long X = 8;
   const long G = (long)gridDim.x*gridDim.y*gridDim.z;
   const long b = blockIdx.x + (long)gridDim.x*( blockIdx.y + (long)gridDim.y*blockIdx.z );
   const long q = G / X, r = G % X, x = b % X, i = b / X;
   return ( ( x < r ) ? x*( q + 1 ) : r*( q + 1 ) + ( x - r )*q ) + i;
}

// MHM_RP patch-group update: in[npg][5][22^3] -> out[npg][5][16^3] (+ flux[npg][9][5][16^2] if store_flux)
template <typename R, typename IO>
int launch_mhm_rp( const IO *in, IO *out, IO *flux, int npg, int rsolver, int store_flux, int minmod_maxiter,
                   double minmod_coeff, const Prm<R> &P, IO *dt, double dh_safety, const Batch &B, int grid_blocks,
                   hipStream_t stream );

// MHM patch-group update: in[npg][5][20^3] -> out[npg][5][16^3] (+ flux[npg][9][5][16^2] if store_flux)
template <typename R, typename IO>
int launch_mhm( const IO *in, IO *out, IO *flux, int npg, int rsolver, int store_flux, int minmod_maxiter,
                double minmod_coeff, const Prm<R> &P, IO *dt, double dh_safety, const Batch &B, int grid_blocks,
                   hipStream_t stream );

// per-patch CFL dt: in[npatch][5][8^3] -> dt[npatch]
template <typename IO>
int launch_dt_cfl( const IO *in, IO *dt, int npatch, double dh, double safety, double gamma, double min_pres,
                   hipStream_t stream );
// periodic level [5][nz][ny][nx] -> patch-group inputs [npg][5][(16+2g)^3]
template <typename IO>
int launch_gather( const IO *grid, IO *pg_in, int nx, int ny, int nz, int ghost, hipStream_t stream );
// patch-group outputs [npg][5][16^3] -> level [5][nz][ny][nx]
template <typename IO>
int launch_scatter( const IO *pg_out, IO *grid, int nx, int ny, int nz, hipStream_t stream );

// set the work-queue counter to 0 (a kernel, so the reset can live inside a captured graph)
int launch_reset_counter( unsigned *counter, hipStream_t stream );

// options of the device CorrectUnphysical (src/Fluid/Flu_Close.cpp:542-1297)
struct CorrPrm {
   int    mode;                // OPT__1ST_FLUX_CORR: 0 none, 1 3D, 2 3D1D
   int    rsolver;             // first-order solver: HYDRO_ROE / HYDRO_HLLE / HYDRO_HLLC
   int    last_resort_floor, auto_reduce, fixup_flux, check_pres;
   int    roe_check;           // a first-order Roe solve checks its intermediate states (the fluid solver is Roe)
   int    NX, ghost;           // input patch-group width and ghost zones of the solver call
   double min_dens, min_pres, min_eint, gm1;
};
template <typename R>
int launch_correct_unphysical( const R *in, R *out, R *flux, int npg, const CorrPrm &c, const Prm<R> &P, R dt_dh,
                               int *counters, hipStream_t stream );
// flux registers reg[nreg][5][8][8] (FP64); slot tables: store [npg][8][6], correct [npg][6], fix-up [npatch][6]
template <typename IO>
int launch_store_flux( const IO *flux, const int *slot, int npg, double dt, double *reg, hipStream_t stream );
template <typename IO>
int launch_correct_flux( const IO *flux, const int *slot, int npg, double dt, double *reg, hipStream_t stream );
template <typename IO>
int launch_fixup_flux( IO *patch, const int *slot, const double *reg, int npatch, double dh, double gm1,
                       double min_dens, double min_pres, double min_eint, hipStream_t stream );
template <typename IO>
int launch_fixup_restrict( const IO *son, IO *fa, int ngroup, const int *son0, const int *fa_idx, double min_eint,
                           hipStream_t stream );
template <typename IO>
int launch_close_patches( const IO *pg_out, IO *patch, int npg, const int *pid0, hipStream_t stream );

// coarse -> fine interpolation of one batch of equally shaped boxes (src/Interpolation/Interpolate.cpp:90-330)
constexpr int INTERP_MAX_CELLS = 4096;      // coarse cells per box (CRange product)
struct InterpPrm {
   int    cs[3], c0[3], cr[3], fs[3], f0[3];   // CSize, CStart, CRange, FSize, FStart
   int    scheme;                             // IntScheme_t 1-7 (INT_MINMOD3D ... INT_QUAR)
   int    mono;                               // bit v: Monotonic[v]
   int    opp0, int_prim, reduce, mono_max_iter;
   double mono_coeff, gm1, inv_gm1, min_pres;
   const unsigned char *need;                 // optional [nbox][ncell] mask of coarse cells to interpolate
};
template <typename T>
int launch_interpolate( T *cdata, T *fdata, int nbox, const InterpPrm &p, int *nfail, hipStream_t stream );
// same-level part of the patch-group input preparation: slot[npg][4][4][4] of patch indices (-1 = keep)
template <typename IO>
int launch_prepare_patches( const IO *patch, IO *pg_in, int npg, const int *slot, int ghost, hipStream_t stream );
// a box of `size` cells at `origin` inside a tdim[0] x tdim[1] x tdim[2] table of 8^3 patches
struct BoxGeom { int size[3], origin[3], tdim[3]; };
template <typename IO>
int launch_gather_boxes( const IO *patch, IO *box, int nbox, const BoxGeom &g, const int *slot, hipStream_t stream );
template <typename IO>
int launch_assemble( const IO *patch, const IO *fbox, IO *pg_in, int npg, const int *slot, int ghost, const BoxGeom &fb,
                     hipStream_t stream );

// MHM_RP step, FP64: adjoint lam_in = J^T lam_out, and tangent-linear tout = J tin (+ out = U_out; tin == NULL:
// plain phase-wise forward).  ws: adjoint_workspace_bytes(npg) of device memory.
size_t adjoint_workspace_bytes( int npg );
int launch_mhm_rp_adjoint( const double *in, const double *lam_out, double *lam_in, double *ws, int npg, int rsolver,
                           double minmod_coeff, const Prm<double> &P, hipStream_t stream );
int launch_mhm_rp_tangent( const double *in, const double *tin, double *out, double *tout, double *ws, int npg,
                           int rsolver, double minmod_coeff, const Prm<double> &P, hipStream_t stream );
// transpose of launch_gather: level_lam[5][nz][ny][nx] = sum of the patch-group input copies of each cell
template <typename IO>
int launch_prepare_adjoint( const IO *pg_lam, IO *lam, int nx, int ny, int nz, int ghost, hipStream_t stream );

} // namespace gamer_amd
