// C-ABI layer of the gfx950 patch-group fluid solver: argument checking, solver constants, device/stream
// and buffer management.  It replaces GAMER's src/GPU_API fluid path (CUAPI_SetDevice, CUAPI_MemAllocate_Fluid,
// CUAPI_Asyn_FluidSolver, CUAPI_Synchronize, CUAPI_MemFree_Fluid) behind include/gamer_amd_hydro.h.

#include "gamer_amd_hydro.h"
#include "hydro_kernels.hpp"

#include <cstdio>
#include <cstring>
#include <cstdarg>
#include <vector>
#include <mutex>

using namespace gamer_amd;

namespace {

thread_local char t_err[512] = "";

int fail( int code, const char *fmt, ... )
{
   va_list ap;
   va_start( ap, fmt );
   vsnprintf( t_err, sizeof(t_err), fmt, ap );
   va_end( ap );
   return code;
}

#define HIP_TRY( call )                                                                          \
   do { hipError_t e_ = ( call );                                                                \
        if ( e_ != hipSuccess ) return fail( HYDRO_E_DEVICE, "%s: %s", #call, hipGetErrorString( e_ ) ); } while (0)

// GPU_API state: device buffers sized once for flu_npg patch groups, pinned host double buffers, streams
// (the module globals of src/GPU_API/CUAPI_MemAllocate_Fluid.cu:211-334)
struct FluidBuffers {
   int    npg_max = 0, nstream = 0, scheme = HYDRO_MHM_RP, precision = HYDRO_FP64;
   size_t elem = 8;
   void  *d_in = nullptr, *d_out = nullptr, *d_flux = nullptr;
   void  *h_in[2] = { nullptr, nullptr }, *h_out[2] = { nullptr, nullptr }, *h_flux[2] = { nullptr, nullptr };
   // CFL dt solver: 8 patches of 5 x 8^3 cells per patch group (d/h_Flu_Array_T, d/h_dt_Array_T,
   // src/GPU_API/CUAPI_MemAllocate_Fluid.cu:238-239, 315-316)
   void  *d_flu_t = nullptr, *d_dt = nullptr;
   void  *h_flu_t[2] = { nullptr, nullptr }, *h_dt[2] = { nullptr, nullptr };
   std::vector<hipStream_t> streams;
};
FluidBuffers g_buf;
std::mutex   g_mtx;

// Work-queue counters of hydro_mhm_solve_batch_noalloc: one 32-bit word per queue slot, in a file-scope device
// array that every GPU the library runs on holds until the process ends.  Host code never passes `scratch`
// itself to a kernel (on the host it is only a stand-in symbol): each device's copy is looked up once with
// hipGetSymbolAddress, with that device current, and kept in g_scratch[device].
constexpr int MAX_DEVICES = 64;
__device__ __align__( 256 ) unsigned scratch[HYDRO_NOALLOC_QUEUES];
unsigned  *g_scratch[MAX_DEVICES] = {};
std::mutex g_scratch_mtx;

// device address of the current device's scratch array; looked up on the first call per device
int scratch_ptr( unsigned **p )
{
   int dev = 0;
   HIP_TRY( hipGetDevice( &dev ) );
   if ( dev < 0 || dev >= MAX_DEVICES ) return fail( HYDRO_E_DEVICE, "device %d beyond the %d the library tracks", dev, MAX_DEVICES );
   std::lock_guard<std::mutex> lk( g_scratch_mtx );
   if ( g_scratch[dev] == nullptr )
   {
      void *q = nullptr;
      HIP_TRY( hipGetSymbolAddress( &q, HIP_SYMBOL( scratch ) ) );
      if ( q == nullptr ) return fail( HYDRO_E_DEVICE, "hipGetSymbolAddress(scratch) returned NULL" );
      g_scratch[dev] = (unsigned *)q;
   }
   *p = g_scratch[dev];
   return HYDRO_OK;
}

int nx_of( int scheme ) { return scheme == HYDRO_MHM_RP ? 22 : 20; }

int check_args( const HydroSolveArgs *a )
{
   if ( a == nullptr )                                       return fail( HYDRO_E_ARG, "args == NULL" );
   if ( a->npg < 0 )                                         return fail( HYDRO_E_ARG, "npg = %d < 0", a->npg );
   if ( a->npg > 0 && ( a->in == nullptr || a->out == nullptr ) ) return fail( HYDRO_E_ARG, "in/out == NULL" );
   if ( a->store_flux && a->npg > 0 && a->flux == nullptr )  return fail( HYDRO_E_ARG, "store_flux with flux == NULL" );
   if ( a->scheme != HYDRO_MHM_RP && a->scheme != HYDRO_MHM ) return fail( HYDRO_E_ARG, "unsupported scheme %d", a->scheme );
   if ( a->rsolver != HYDRO_HLLC && a->rsolver != HYDRO_HLLE && a->rsolver != HYDRO_ROE )
                                                             return fail( HYDRO_E_ARG, "unsupported Riemann solver %d", a->rsolver );
   const int l = a->limiter;
   if ( l != 1 && l != 2 && l != 3 && l != 4 && l != 6 )      return fail( HYDRO_E_ARG, "unsupported limiter %d", l );
   if ( a->minmod_maxiter < 0 )                              return fail( HYDRO_E_ARG, "minmod_maxiter < 0" );
   if ( !( a->gamma > 1.0 ) )                                return fail( HYDRO_E_ARG, "gamma must be > 1" );
   if ( !( a->dh > 0.0 ) || !( a->dt >= 0.0 ) )              return fail( HYDRO_E_ARG, "dt/dh invalid" );
   if ( a->precision != HYDRO_FP64 && a->precision != HYDRO_FP32_BULK )
                                                             return fail( HYDRO_E_ARG, "unsupported precision %d", a->precision );
   if ( a->npg > ( 1 << 26 ) )                               return fail( HYDRO_E_ARG, "npg too large" );
   if ( a->dt_patch != nullptr && !( a->safety > 0.0 ) )     return fail( HYDRO_E_ARG, "fused dt needs safety > 0" );
   const int w = a->waves_per_group;
   if ( w != 0 && w != 4 && ( w != 8 || a->scheme != HYDRO_MHM_RP ) )
                                                             return fail( HYDRO_E_ARG, "waves_per_group %d: 0, 4 or 8 (MHM: 0 or 4)", w );
   return HYDRO_OK;
}

// solver constants; the EOS_GAMMA aux values are computed in double and narrowed like
// EoS_SetAuxArray_Gamma (src/EoS/Gamma/CPU_EoS_Gamma.cpp:54-66)
template <typename R>
Prm<R> make_prm( const HydroSolveArgs *a )
{
   Prm<R> P;
   P.gamma    = (R)a->gamma;
   P.gm1      = (R)( a->gamma - 1.0 );
   P.inv_gm1  = (R)( 1.0/( a->gamma - 1.0 ) );
   P.min_dens = (R)a->min_dens;
   P.min_pres = (R)a->min_pres;
   P.min_eint = (R)a->min_eint;
   P.coeff    = (R)a->minmod_coeff;
   P.limiter  = a->limiter;
   P.dt_dh2     = 0.5*a->dt/a->dh;        // (real)0.5*dt/dh, CPU_FluidSolver_MHM.cpp:845
   P.dt_dh      = a->dt/a->dh;            // CPU_Shared_FullStepUpdate.cpp:75
   P.gm1_d      = (double)P.gm1;
   P.min_dens_d = (double)P.min_dens;
   P.min_pres_d = (double)P.min_pres;
   P.min_eint_d = (double)P.min_eint;
   return P;
}

int enqueue( const HydroSolveArgs *a, const void *in, void *out, void *flux, void *dt, int npg, hipStream_t s,
             const Batch *batch = nullptr, int grid_blocks = 0 )
{
   Batch B;
   if ( batch ) B = *batch;
   else { B.pg_dt_dh = nullptr;  B.counter = nullptr;  B.safety = 0.0;  B.claim = 1; }
   B.waves = a->waves_per_group;                             // 0: each kernel's default shape
   int rc;
   const bool rp = ( a->scheme == HYDRO_MHM_RP );
   if ( a->precision == HYDRO_FP64 )
      rc = ( rp ? launch_mhm_rp<double, double> : launch_mhm<double, double> )
           ( (const double*)in, (double*)out, (double*)flux, npg, a->rsolver, a->store_flux, a->minmod_maxiter,
             a->minmod_coeff, make_prm<double>( a ), (double*)dt, a->safety*a->dh, B, grid_blocks, s );
   else
      rc = ( rp ? launch_mhm_rp<float, float> : launch_mhm<float, float> )
           ( (const float*)in, (float*)out, (float*)flux, npg, a->rsolver, a->store_flux, a->minmod_maxiter,
             a->minmod_coeff, make_prm<float>( a ), (float*)dt, a->safety*a->dh, B, grid_blocks, s );
   if ( rc != 0 ) return fail( HYDRO_E_DEVICE, "kernel launch failed (%d): %s", rc, hipGetErrorString( hipGetLastError() ) );
   return HYDRO_OK;
}

void free_buffers()
{
   if ( g_buf.d_in )   (void)hipFree( g_buf.d_in );
   if ( g_buf.d_out )  (void)hipFree( g_buf.d_out );
   if ( g_buf.d_flux ) (void)hipFree( g_buf.d_flux );
   if ( g_buf.d_flu_t ) (void)hipFree( g_buf.d_flu_t );
   if ( g_buf.d_dt )    (void)hipFree( g_buf.d_dt );
   for (int k=0; k<2; k++)
   {
      if ( g_buf.h_in[k] )    (void)hipHostFree( g_buf.h_in[k] );
      if ( g_buf.h_out[k] )   (void)hipHostFree( g_buf.h_out[k] );
      if ( g_buf.h_flux[k] )  (void)hipHostFree( g_buf.h_flux[k] );
      if ( g_buf.h_flu_t[k] ) (void)hipHostFree( g_buf.h_flu_t[k] );
      if ( g_buf.h_dt[k] )    (void)hipHostFree( g_buf.h_dt[k] );
   }
   for (hipStream_t s : g_buf.streams) (void)hipStreamDestroy( s );
   g_buf = FluidBuffers();
}

} // namespace

extern "C" {

const char *hydro_last_error( void ) { return t_err; }

int hydro_init( int device )
{
   int n = 0;
   HIP_TRY( hipGetDeviceCount( &n ) );
   if ( device < 0 || device >= n ) return fail( HYDRO_E_ARG, "device %d out of range (%d devices)", device, n );
   HIP_TRY( hipSetDevice( device ) );
   unsigned *p = nullptr;
   return scratch_ptr( &p );                 // set-up of the allocation-free entry point on this device
}

int hydro_mhm_solve_async( const HydroSolveArgs *a, void *stream )
{
   const int rc = check_args( a );
   if ( rc != HYDRO_OK ) return rc;
   return enqueue( a, a->in, a->out, a->flux, a->dt_patch, a->npg, (hipStream_t)stream );
}

int hydro_dt_cfl_async( const void *in, void *dt, int npatch, double dh, double safety, double gamma,
                        double min_pres, int precision, void *stream )
{
   if ( npatch < 0 || ( npatch > 0 && ( in == nullptr || dt == nullptr ) ) ) return fail( HYDRO_E_ARG, "dt_cfl: bad pointers/size" );
   if ( !( dh > 0.0 ) || !( safety > 0.0 ) || !( gamma > 1.0 ) )           return fail( HYDRO_E_ARG, "dt_cfl: dh/safety/gamma invalid" );
   int rc;
   if      ( precision == HYDRO_FP64 )      rc = launch_dt_cfl<double>( (const double*)in, (double*)dt, npatch, dh, safety, gamma, min_pres, (hipStream_t)stream );
   else if ( precision == HYDRO_FP32_BULK ) rc = launch_dt_cfl<float> ( (const float*)in,  (float*)dt,  npatch, dh, safety, gamma, min_pres, (hipStream_t)stream );
   else return fail( HYDRO_E_ARG, "unsupported precision %d", precision );
   return rc == 0 ? HYDRO_OK : fail( HYDRO_E_DEVICE, "dt_cfl launch failed" );
}

static int check_level( int nx, int ny, int nz )
{
   if ( nx <= 0 || ny <= 0 || nz <= 0 || nx % 16 || ny % 16 || nz % 16 ) return fail( HYDRO_E_ARG, "level extents must be positive multiples of 16" );
   return HYDRO_OK;
}

int hydro_prepare_uniform_async( const void *level, void *pg_in, int nx, int ny, int nz, int ghost, int precision, void *stream )
{
   if ( check_level( nx, ny, nz ) ) return HYDRO_E_ARG;
   if ( level == nullptr || pg_in == nullptr ) return fail( HYDRO_E_ARG, "prepare: NULL pointer" );
   if ( ghost < 0 || ghost > 8 ) return fail( HYDRO_E_ARG, "prepare: ghost %d out of range", ghost );
   int rc;
   if      ( precision == HYDRO_FP64 )      rc = launch_gather<double>( (const double*)level, (double*)pg_in, nx, ny, nz, ghost, (hipStream_t)stream );
   else if ( precision == HYDRO_FP32_BULK ) rc = launch_gather<float> ( (const float*)level,  (float*)pg_in,  nx, ny, nz, ghost, (hipStream_t)stream );
   else return fail( HYDRO_E_ARG, "unsupported precision %d", precision );
   return rc == 0 ? HYDRO_OK : fail( HYDRO_E_DEVICE, "prepare launch failed" );
}

int hydro_close_uniform_async( const void *pg_out, void *level, int nx, int ny, int nz, int precision, void *stream )
{
   if ( check_level( nx, ny, nz ) ) return HYDRO_E_ARG;
   if ( level == nullptr || pg_out == nullptr ) return fail( HYDRO_E_ARG, "close: NULL pointer" );
   int rc;
   if      ( precision == HYDRO_FP64 )      rc = launch_scatter<double>( (const double*)pg_out, (double*)level, nx, ny, nz, (hipStream_t)stream );
   else if ( precision == HYDRO_FP32_BULK ) rc = launch_scatter<float> ( (const float*)pg_out,  (float*)level,  nx, ny, nz, (hipStream_t)stream );
   else return fail( HYDRO_E_ARG, "unsupported precision %d", precision );
   return rc == 0 ? HYDRO_OK : fail( HYDRO_E_DEVICE, "close launch failed" );
}

static int check_adjoint_args( const HydroSolveArgs *a, const void *ws )
{
   if ( a == nullptr ) return fail( HYDRO_E_ARG, "args == NULL" );
   HydroSolveArgs b = *a;                 // out and the flux arrays are optional here
   if ( b.out == nullptr ) b.out = (void*)b.in;
   b.store_flux = 0;
   const int rc = check_args( &b );
   if ( rc != HYDRO_OK ) return rc;
   if ( a->scheme != HYDRO_MHM_RP )  return fail( HYDRO_E_ARG, "adjoint/tangent: only the MHM_RP scheme" );
   if ( a->precision != HYDRO_FP64 ) return fail( HYDRO_E_ARG, "adjoint/tangent: only HYDRO_FP64" );
   if ( a->npg > 0 && ws == nullptr ) return fail( HYDRO_E_ARG, "adjoint/tangent: workspace == NULL" );
   return HYDRO_OK;
}

size_t hydro_mhm_adjoint_workspace_bytes( int npg ) { return adjoint_workspace_bytes( npg ); }

int hydro_mhm_adjoint_async( const HydroSolveArgs *a, const void *lam_out, void *lam_in, void *workspace, void *stream )
{
   int rc = check_adjoint_args( a, workspace );
   if ( rc != HYDRO_OK ) return rc;
   if ( a->npg > 0 && ( lam_out == nullptr || lam_in == nullptr ) ) return fail( HYDRO_E_ARG, "adjoint: lam_out/lam_in == NULL" );
   rc = launch_mhm_rp_adjoint( (const double*)a->in, (const double*)lam_out, (double*)lam_in, (double*)workspace, a->npg,
                               a->rsolver, a->minmod_coeff, make_prm<double>( a ), (hipStream_t)stream );
   if ( rc == -3 ) return fail( HYDRO_E_ARG, "adjoint: npg %d does not fit one launch per phase", a->npg );
   return rc == 0 ? HYDRO_OK : fail( HYDRO_E_DEVICE, "adjoint launch failed: %s", hipGetErrorString( hipGetLastError() ) );
}

int hydro_mhm_tangent_async( const HydroSolveArgs *a, const void *v_in, void *jv_out, void *workspace, void *stream )
{
   int rc = check_adjoint_args( a, workspace );
   if ( rc != HYDRO_OK ) return rc;
   if ( a->npg > 0 && v_in != nullptr && jv_out == nullptr ) return fail( HYDRO_E_ARG, "tangent: jv_out == NULL" );
   if ( a->npg > 0 && v_in == nullptr && a->out == nullptr )  return fail( HYDRO_E_ARG, "tangent: v_in and out both NULL" );
   rc = launch_mhm_rp_tangent( (const double*)a->in, (const double*)v_in, (double*)a->out, (double*)jv_out,
                               (double*)workspace, a->npg, a->rsolver, a->minmod_coeff, make_prm<double>( a ),
                               (hipStream_t)stream );
   if ( rc == -3 ) return fail( HYDRO_E_ARG, "tangent: npg %d does not fit one launch per phase", a->npg );
   return rc == 0 ? HYDRO_OK : fail( HYDRO_E_DEVICE, "tangent launch failed: %s", hipGetErrorString( hipGetLastError() ) );
}

int hydro_prepare_uniform_adjoint_async( const void *pg_lam, void *level_lam, int nx, int ny, int nz, int ghost,
                                         int precision, void *stream )
{
   if ( check_level( nx, ny, nz ) ) return HYDRO_E_ARG;
   if ( pg_lam == nullptr || level_lam == nullptr ) return fail( HYDRO_E_ARG, "prepare adjoint: NULL pointer" );
   if ( ghost < 0 || ghost > 8 ) return fail( HYDRO_E_ARG, "prepare adjoint: ghost %d out of range", ghost );
   int rc;
   if      ( precision == HYDRO_FP64 )      rc = launch_prepare_adjoint<double>( (const double*)pg_lam, (double*)level_lam, nx, ny, nz, ghost, (hipStream_t)stream );
   else if ( precision == HYDRO_FP32_BULK ) rc = launch_prepare_adjoint<float> ( (const float*)pg_lam,  (float*)level_lam,  nx, ny, nz, ghost, (hipStream_t)stream );
   else return fail( HYDRO_E_ARG, "unsupported precision %d", precision );
   return rc == 0 ? HYDRO_OK : fail( HYDRO_E_DEVICE, "prepare adjoint launch failed" );
}

int hydro_correct_unphysical_async( const HydroCorrectArgs *a, void *stream )
{
   if ( a == nullptr ) return fail( HYDRO_E_ARG, "args == NULL" );
   if ( a->npg < 0 || a->npg > ( 1 << 26 ) ) return fail( HYDRO_E_ARG, "correct: npg = %d out of range", a->npg );
   if ( a->npg > 0 && ( a->in == nullptr || a->out == nullptr || a->counters == nullptr ) )
      return fail( HYDRO_E_ARG, "correct: in/out/counters == NULL" );
   if ( a->scheme != HYDRO_MHM_RP && a->scheme != HYDRO_MHM ) return fail( HYDRO_E_ARG, "correct: unsupported scheme %d", a->scheme );
   if ( a->mode < 0 || a->mode > 2 ) return fail( HYDRO_E_ARG, "correct: mode %d (0 none, 1 3D, 2 3D1D)", a->mode );
   if ( a->rsolver != HYDRO_HLLC && a->rsolver != HYDRO_HLLE && a->rsolver != HYDRO_ROE )
      return fail( HYDRO_E_ARG, "correct: unsupported Riemann solver %d", a->rsolver );
   if ( a->main_rsolver != 0 && a->main_rsolver != HYDRO_HLLC && a->main_rsolver != HYDRO_HLLE && a->main_rsolver != HYDRO_ROE )
      return fail( HYDRO_E_ARG, "correct: unsupported main Riemann solver %d", a->main_rsolver );
   if ( !( a->gamma > 1.0 ) || !( a->dh > 0.0 ) || !( a->dt >= 0.0 ) ) return fail( HYDRO_E_ARG, "correct: gamma/dt/dh invalid" );
   CorrPrm c;
   c.mode = a->mode;  c.rsolver = a->rsolver;  c.last_resort_floor = a->last_resort_floor;
   c.auto_reduce = a->auto_reduce_dt;  c.fixup_flux = a->fixup_flux;  c.check_pres = a->check_pres;
   c.roe_check = ( ( a->main_rsolver == 0 ) ? a->rsolver : a->main_rsolver ) == HYDRO_ROE;
   c.NX = ( a->scheme == HYDRO_MHM_RP ) ? 22 : 20;  c.ghost = ( c.NX - 16 )/2;
   c.min_dens = a->min_dens;  c.min_pres = a->min_pres;  c.min_eint = a->min_eint;  c.gm1 = a->gamma - 1.0;
   HydroSolveArgs s = {};                     // solver constants for the first-order interface solves
   s.npg = a->npg;  s.dt = a->dt;  s.dh = a->dh;  s.gamma = a->gamma;  s.min_dens = a->min_dens;
   s.min_pres = a->min_pres;  s.min_eint = a->min_eint;  s.limiter = 4;  s.minmod_coeff = 1.5;
   int rc;
   if ( a->precision == HYDRO_FP64 )
      rc = launch_correct_unphysical<double>( (const double*)a->in, (double*)a->out, (double*)a->flux, a->npg, c,
                                              make_prm<double>( &s ), a->dt/a->dh, a->counters, (hipStream_t)stream );
   else if ( a->precision == HYDRO_FP32_BULK )
      rc = launch_correct_unphysical<float>( (const float*)a->in, (float*)a->out, (float*)a->flux, a->npg, c,
                                             make_prm<float>( &s ), (float)a->dt/(float)a->dh, a->counters, (hipStream_t)stream );
   else return fail( HYDRO_E_ARG, "unsupported precision %d", a->precision );
   return rc == 0 ? HYDRO_OK : fail( HYDRO_E_DEVICE, "correct launch failed" );
}

#define GA_PREC_DISPATCH( CALL_D, CALL_F )                                                                           \
   int rc;                                                                                                           \
   if      ( precision == HYDRO_FP64 )      rc = CALL_D;                                                             \
   else if ( precision == HYDRO_FP32_BULK ) rc = CALL_F;                                                             \
   else return fail( HYDRO_E_ARG, "unsupported precision %d", precision );

int hydro_store_flux_async( const void *flux, const int *slot, int npg, double dt, double *reg, int precision, void *stream )
{
   if ( npg < 0 || ( npg > 0 && ( flux == nullptr || slot == nullptr || reg == nullptr ) ) ) return fail( HYDRO_E_ARG, "store_flux: bad pointers/size" );
   GA_PREC_DISPATCH( launch_store_flux<double>( (const double*)flux, slot, npg, dt, reg, (hipStream_t)stream ),
                     launch_store_flux<float> ( (const float*)flux,  slot, npg, dt, reg, (hipStream_t)stream ) )
   return rc == 0 ? HYDRO_OK : fail( HYDRO_E_DEVICE, "store_flux launch failed" );
}

int hydro_correct_flux_async( const void *flux, const int *slot, int npg, double dt, double *reg, int precision, void *stream )
{
   if ( npg < 0 || ( npg > 0 && ( flux == nullptr || slot == nullptr || reg == nullptr ) ) ) return fail( HYDRO_E_ARG, "correct_flux: bad pointers/size" );
   GA_PREC_DISPATCH( launch_correct_flux<double>( (const double*)flux, slot, npg, dt, reg, (hipStream_t)stream ),
                     launch_correct_flux<float> ( (const float*)flux,  slot, npg, dt, reg, (hipStream_t)stream ) )
   return rc == 0 ? HYDRO_OK : fail( HYDRO_E_DEVICE, "correct_flux launch failed" );
}

int hydro_fixup_flux_async( void *patch, const int *slot, const double *reg, int npatch, double dh, double gamma,
                            double min_dens, double min_pres, double min_eint, int precision, void *stream )
{
   if ( npatch < 0 || ( npatch > 0 && ( patch == nullptr || slot == nullptr || reg == nullptr ) ) ) return fail( HYDRO_E_ARG, "fixup_flux: bad pointers/size" );
   if ( !( dh > 0.0 ) || !( gamma > 1.0 ) ) return fail( HYDRO_E_ARG, "fixup_flux: dh/gamma invalid" );
   GA_PREC_DISPATCH( launch_fixup_flux<double>( (double*)patch, slot, reg, npatch, dh, gamma - 1.0, min_dens, min_pres, min_eint, (hipStream_t)stream ),
                     launch_fixup_flux<float> ( (float*)patch,  slot, reg, npatch, dh, gamma - 1.0, min_dens, min_pres, min_eint, (hipStream_t)stream ) )
   return rc == 0 ? HYDRO_OK : fail( HYDRO_E_DEVICE, "fixup_flux launch failed" );
}

int hydro_fixup_restrict_async( const void *son, void *fa, int ngroup, const int *son0, const int *fa_idx, double min_eint,
                                int precision, void *stream )
{
   if ( ngroup < 0 || ( ngroup > 0 && ( son == nullptr || fa == nullptr || son0 == nullptr || fa_idx == nullptr ) ) )
      return fail( HYDRO_E_ARG, "fixup_restrict: bad pointers/size" );
   GA_PREC_DISPATCH( launch_fixup_restrict<double>( (const double*)son, (double*)fa, ngroup, son0, fa_idx, min_eint, (hipStream_t)stream ),
                     launch_fixup_restrict<float> ( (const float*)son,  (float*)fa,  ngroup, son0, fa_idx, min_eint, (hipStream_t)stream ) )
   return rc == 0 ? HYDRO_OK : fail( HYDRO_E_DEVICE, "fixup_restrict launch failed" );
}

int hydro_close_patches_async( const void *pg_out, void *patch, int npg, const int *pid0, int precision, void *stream )
{
   if ( npg < 0 || ( npg > 0 && ( pg_out == nullptr || patch == nullptr || pid0 == nullptr ) ) ) return fail( HYDRO_E_ARG, "close_patches: bad pointers/size" );
   GA_PREC_DISPATCH( launch_close_patches<double>( (const double*)pg_out, (double*)patch, npg, pid0, (hipStream_t)stream ),
                     launch_close_patches<float> ( (const float*)pg_out,  (float*)patch,  npg, pid0, (hipStream_t)stream ) )
   return rc == 0 ? HYDRO_OK : fail( HYDRO_E_DEVICE, "close_patches launch failed" );
}
#undef GA_PREC_DISPATCH

int hydro_interpolate_async( const HydroInterpArgs *a, void *stream )
{
   if ( a == nullptr ) return fail( HYDRO_E_ARG, "args == NULL" );
   if ( a->nbox < 0 || ( a->nbox > 0 && ( a->cdata == nullptr || a->fdata == nullptr || a->nfail == nullptr ) ) )
      return fail( HYDRO_E_ARG, "interpolate: bad pointers/size" );
   if ( a->scheme < 1 || a->scheme > 7 ) return fail( HYDRO_E_ARG, "interpolate: scheme %d (INT_MINMOD3D = 1 ... INT_QUAR = 7)", a->scheme );
   // Interpolate.cpp:299-300: these schemes have no monotonic coefficient to reduce
   if ( ( a->scheme == 1 || a->scheme == 2 || a->scheme == 3 ) && a->reduce_mono )
      return fail( HYDRO_E_ARG, "interpolate: INT_VANLEER/INT_MINMOD3D/INT_MINMOD1D do not support reduce_mono" );
   const int ng = ( a->scheme == 6 || a->scheme == 7 ) ? 2 : 1;      // coarse ghost cells (Int_Table.cpp)
   if ( a->reduce_mono && a->mono_max_iter <= 0 ) return fail( HYDRO_E_ARG, "interpolate: mono_max_iter must be > 0" );
   if ( !( a->gamma > 1.0 ) ) return fail( HYDRO_E_ARG, "interpolate: gamma must be > 1" );
   long ncell = 1;
   for (int d=0; d<3; d++)
   {
      if ( a->crange[d] <= 0 || a->cstart[d] < ng || a->cstart[d] + a->crange[d] > a->csize[d] - ng )
         return fail( HYDRO_E_ARG, "interpolate: coarse range out of the box (%d ghost cells needed) in dim %d", ng, d );
      if ( a->fstart[d] < 0 || a->fstart[d] + 2*a->crange[d] > a->fsize[d] )
         return fail( HYDRO_E_ARG, "interpolate: fine range out of the box in dim %d", d );
      if ( ( a->int_prim || a->reduce_mono ) && a->fstart[d] != 0 )
         return fail( HYDRO_E_ARG, "interpolate: the iterative interpolation needs fstart = 0" );
      ncell *= a->crange[d];
   }
   if ( ncell > INTERP_MAX_CELLS ) return fail( HYDRO_E_ARG, "interpolate: %ld coarse cells > %d per box", ncell, INTERP_MAX_CELLS );
   InterpPrm p;
   for (int d=0; d<3; d++) { p.cs[d] = a->csize[d];  p.c0[d] = a->cstart[d];  p.cr[d] = a->crange[d];  p.fs[d] = a->fsize[d];  p.f0[d] = a->fstart[d]; }
   p.scheme = a->scheme;  p.mono = a->monotonic;  p.opp0 = a->opp_sign_0th;  p.int_prim = a->int_prim;
   p.reduce = a->reduce_mono;  p.mono_max_iter = a->mono_max_iter;  p.mono_coeff = a->mono_coeff;
   p.gm1 = a->gamma - 1.0;  p.inv_gm1 = 1.0/( a->gamma - 1.0 );  p.min_pres = a->min_pres;
   p.need = a->need;
   int rc;
   if      ( a->precision == HYDRO_FP64 )      rc = launch_interpolate<double>( (double*)a->cdata, (double*)a->fdata, a->nbox, p, a->nfail, (hipStream_t)stream );
   else if ( a->precision == HYDRO_FP32_BULK ) rc = launch_interpolate<float> ( (float*)a->cdata,  (float*)a->fdata,  a->nbox, p, a->nfail, (hipStream_t)stream );
   else return fail( HYDRO_E_ARG, "unsupported precision %d", a->precision );
   return rc == 0 ? HYDRO_OK : fail( HYDRO_E_DEVICE, "interpolate launch failed" );
}

int hydro_prepare_patches_async( const void *patch, void *pg_in, int npg, const int *slot, int ghost, int precision,
                                 void *stream )
{
   if ( npg < 0 || ( npg > 0 && ( patch == nullptr || pg_in == nullptr || slot == nullptr ) ) ) return fail( HYDRO_E_ARG, "prepare_patches: bad pointers/size" );
   if ( ghost < 0 || ghost > 8 ) return fail( HYDRO_E_ARG, "prepare_patches: ghost %d out of range 0..8", ghost );
   int rc;
   if      ( precision == HYDRO_FP64 )      rc = launch_prepare_patches<double>( (const double*)patch, (double*)pg_in, npg, slot, ghost, (hipStream_t)stream );
   else if ( precision == HYDRO_FP32_BULK ) rc = launch_prepare_patches<float> ( (const float*)patch,  (float*)pg_in,  npg, slot, ghost, (hipStream_t)stream );
   else return fail( HYDRO_E_ARG, "unsupported precision %d", precision );
   return rc == 0 ? HYDRO_OK : fail( HYDRO_E_DEVICE, "prepare_patches launch failed" );
}

int hydro_gather_boxes_async( const void *patch, void *box, int nbox, const int size[3], const int origin[3],
                              const int *slot, const int tdim[3], int precision, void *stream )
{
   if ( nbox < 0 || size == nullptr || origin == nullptr || tdim == nullptr ||
        ( nbox > 0 && ( patch == nullptr || box == nullptr || slot == nullptr ) ) )
      return fail( HYDRO_E_ARG, "gather_boxes: bad pointers/size" );
   BoxGeom g;
   for (int d=0; d<3; d++)
   {
      if ( size[d] <= 0 || size[d] > 65535 || tdim[d] <= 0 || origin[d] < 0 || origin[d] + size[d] > 8*tdim[d] )
         return fail( HYDRO_E_ARG, "gather_boxes: box [%d, %d) outside the %d-patch table in dim %d", origin[d],
                      origin[d] + size[d], tdim[d], d );
      g.size[d] = size[d];  g.origin[d] = origin[d];  g.tdim[d] = tdim[d];
   }
   int rc;
   if      ( precision == HYDRO_FP64 )      rc = launch_gather_boxes<double>( (const double*)patch, (double*)box, nbox, g, slot, (hipStream_t)stream );
   else if ( precision == HYDRO_FP32_BULK ) rc = launch_gather_boxes<float> ( (const float*)patch,  (float*)box,  nbox, g, slot, (hipStream_t)stream );
   else return fail( HYDRO_E_ARG, "unsupported precision %d", precision );
   return rc == 0 ? HYDRO_OK : fail( HYDRO_E_DEVICE, "gather_boxes launch failed" );
}

int hydro_assemble_patch_groups_async( const void *patch, const void *fbox, void *pg_in, int npg, const int *slot,
                                       int ghost, const int fb_size[3], const int fb_origin[3], int precision,
                                       void *stream )
{
   if ( npg < 0 || fb_size == nullptr || fb_origin == nullptr ||
        ( npg > 0 && ( patch == nullptr || fbox == nullptr || pg_in == nullptr || slot == nullptr ) ) )
      return fail( HYDRO_E_ARG, "assemble_patch_groups: bad pointers/size" );
   if ( ghost < 0 || ghost > 8 ) return fail( HYDRO_E_ARG, "assemble_patch_groups: ghost %d out of range 0..8", ghost );
   BoxGeom fb;
   for (int d=0; d<3; d++)
   {
      if ( fb_origin[d] < 0 || fb_origin[d] + 16 + 2*ghost > fb_size[d] )
         return fail( HYDRO_E_ARG, "assemble_patch_groups: window outside the fine box in dim %d", d );
      fb.size[d] = fb_size[d];  fb.origin[d] = fb_origin[d];  fb.tdim[d] = 4;
   }
   int rc;
   if      ( precision == HYDRO_FP64 )      rc = launch_assemble<double>( (const double*)patch, (const double*)fbox, (double*)pg_in, npg, slot, ghost, fb, (hipStream_t)stream );
   else if ( precision == HYDRO_FP32_BULK ) rc = launch_assemble<float> ( (const float*)patch,  (const float*)fbox,  (float*)pg_in,  npg, slot, ghost, fb, (hipStream_t)stream );
   else return fail( HYDRO_E_ARG, "unsupported precision %d", precision );
   return rc == 0 ? HYDRO_OK : fail( HYDRO_E_DEVICE, "assemble_patch_groups launch failed" );
}

int hydro_mhm_solve_batch_async( const HydroSolveArgs *a, const double *pg_dt_dh, unsigned *queue_counter,
                                 int groups_per_cu, int pgs_per_ticket, void *stream )
{
   int rc = check_args( a );
   if ( rc != HYDRO_OK ) return rc;
   Batch B;
   B.pg_dt_dh = pg_dt_dh;
   B.counter  = queue_counter;
   B.safety   = a->safety;
   B.claim    = 1;
   int grid = 0;
   if ( queue_counter != nullptr )
   {
      if ( groups_per_cu < 1 || groups_per_cu > 16 ) return fail( HYDRO_E_ARG, "batch: groups_per_cu must be 1..16" );
      if ( pgs_per_ticket < 1 || pgs_per_ticket > 64 ) return fail( HYDRO_E_ARG, "batch: pgs_per_ticket must be 1..64" );
      B.claim = pgs_per_ticket;
      int dev = 0, ncu = 0;
      HIP_TRY( hipGetDevice( &dev ) );
      HIP_TRY( hipDeviceGetAttribute( &ncu, hipDeviceAttributeMultiprocessorCount, dev ) );
      grid = groups_per_cu*ncu;
      const int tickets = ( a->npg + B.claim - 1 )/B.claim;      // workgroups beyond this would find the queue empty
      if ( grid > tickets ) grid = tickets;
      if ( a->npg > 0 && launch_reset_counter( queue_counter, (hipStream_t)stream ) != 0 )
         return fail( HYDRO_E_DEVICE, "queue reset launch failed" );
   }
   return enqueue( a, a->in, a->out, a->flux, a->dt_patch, a->npg, (hipStream_t)stream, &B, grid );
}

int hydro_mhm_solve_batch_noalloc( const HydroSolveArgs *a, const double *pg_dt_dh, int queue_slot,
                                   int groups_per_cu, int pgs_per_ticket, void *stream )
{
   if ( queue_slot < 0 || queue_slot >= HYDRO_NOALLOC_QUEUES )
      return fail( HYDRO_E_ARG, "batch_noalloc: queue_slot %d outside 0..%d", queue_slot, HYDRO_NOALLOC_QUEUES - 1 );
   unsigned *q = nullptr;
   const int rc = scratch_ptr( &q );
   if ( rc != HYDRO_OK ) return rc;
   // the slot's counter is reset by a kernel on `stream` before the solve (no memset, no allocation)
   return hydro_mhm_solve_batch_async( a, pg_dt_dh, q + queue_slot, groups_per_cu, pgs_per_ticket, stream );
}

int hydro_synchronize( void )
{
   HIP_TRY( hipDeviceSynchronize() );
   return HYDRO_OK;
}

int hydro_mem_allocate_fluid( int flu_npg, int nstream, int scheme, int precision )
{
   std::lock_guard<std::mutex> lk( g_mtx );
   if ( flu_npg <= 0 || nstream <= 0 || nstream > 64 ) return fail( HYDRO_E_ARG, "flu_npg/nstream invalid" );
   if ( scheme != HYDRO_MHM_RP && scheme != HYDRO_MHM ) return fail( HYDRO_E_ARG, "unsupported scheme %d", scheme );
   if ( precision != HYDRO_FP64 && precision != HYDRO_FP32_BULK ) return fail( HYDRO_E_ARG, "unsupported precision" );
   free_buffers();
   const int nx = nx_of( scheme );
   g_buf.elem = ( precision == HYDRO_FP64 ) ? 8 : 4;
   const size_t in_b   = (size_t)flu_npg*5*nx*nx*nx*g_buf.elem;
   const size_t out_b  = (size_t)flu_npg*5*16*16*16*g_buf.elem;
   const size_t flux_b = (size_t)flu_npg*9*5*16*16*g_buf.elem;
   const size_t flut_b = (size_t)flu_npg*8*5*512*g_buf.elem;
   const size_t dt_b   = (size_t)flu_npg*8*g_buf.elem;
   bool ok = hipMalloc( &g_buf.d_in, in_b ) == hipSuccess &&
             hipMalloc( &g_buf.d_out, out_b ) == hipSuccess &&
             hipMalloc( &g_buf.d_flux, flux_b ) == hipSuccess &&
             hipMalloc( &g_buf.d_flu_t, flut_b ) == hipSuccess &&
             hipMalloc( &g_buf.d_dt, dt_b ) == hipSuccess;
   for (int k=0; k<2 && ok; k++)
      ok = hipHostMalloc( &g_buf.h_in[k], in_b, hipHostMallocDefault ) == hipSuccess &&
           hipHostMalloc( &g_buf.h_out[k], out_b, hipHostMallocDefault ) == hipSuccess &&
           hipHostMalloc( &g_buf.h_flux[k], flux_b, hipHostMallocDefault ) == hipSuccess &&
           hipHostMalloc( &g_buf.h_flu_t[k], flut_b, hipHostMallocDefault ) == hipSuccess &&
           hipHostMalloc( &g_buf.h_dt[k], dt_b, hipHostMallocDefault ) == hipSuccess;
   if ( !ok )
   {
      (void)hipGetLastError();
      free_buffers();
      return fail( HYDRO_E_NOMEM, "allocation of %d patch groups x %d streams failed", flu_npg, nstream );
   }
   g_buf.streams.resize( nstream );
   for (int s=0; s<nstream; s++)
      if ( hipStreamCreateWithFlags( &g_buf.streams[s], hipStreamNonBlocking ) != hipSuccess )
      {
         g_buf.streams.resize( s );
         free_buffers();
         return fail( HYDRO_E_DEVICE, "stream creation failed" );
      }
   g_buf.npg_max = flu_npg;  g_buf.nstream = nstream;  g_buf.scheme = scheme;  g_buf.precision = precision;
   return HYDRO_OK;
}

int hydro_mem_free_fluid( void )
{
   std::lock_guard<std::mutex> lk( g_mtx );
   free_buffers();
   return HYDRO_OK;
}

int hydro_host_buffers( int array_id, void **h_in, void **h_out, void **h_flux )
{
   if ( g_buf.npg_max == 0 ) return fail( HYDRO_E_STATE, "buffers not allocated" );
   if ( array_id < 0 || array_id > 1 ) return fail( HYDRO_E_ARG, "array_id must be 0 or 1" );
   if ( h_in )   *h_in   = g_buf.h_in[array_id];
   if ( h_out )  *h_out  = g_buf.h_out[array_id];
   if ( h_flux ) *h_flux = g_buf.h_flux[array_id];
   return HYDRO_OK;
}

int hydro_host_dt_buffers( int array_id, void **h_flu_t, void **h_dt )
{
   if ( g_buf.npg_max == 0 ) return fail( HYDRO_E_STATE, "buffers not allocated" );
   if ( array_id < 0 || array_id > 1 ) return fail( HYDRO_E_ARG, "array_id must be 0 or 1" );
   if ( h_flu_t ) *h_flu_t = g_buf.h_flu_t[array_id];
   if ( h_dt )    *h_dt    = g_buf.h_dt[array_id];
   return HYDRO_OK;
}
int hydro_asyn_dt_solver( const void *h_flu_t, void *h_dt, int npg, double dh, double safety, double gamma,
                          double min_pres, int gpu_nstream )
{
   if ( g_buf.npg_max == 0 )            return fail( HYDRO_E_STATE, "hydro_mem_allocate_fluid was not called" );
   if ( npg < 0 || npg > g_buf.npg_max ) return fail( HYDRO_E_ARG, "dt solver: npg %d outside 0..%d", npg, g_buf.npg_max );
   if ( npg > 0 && ( h_flu_t == nullptr || h_dt == nullptr ) ) return fail( HYDRO_E_ARG, "dt solver: NULL host array" );
   if ( !( dh > 0.0 ) || !( safety > 0.0 ) || !( gamma > 1.0 ) ) return fail( HYDRO_E_ARG, "dt solver: dh/safety/gamma invalid" );
   if ( gpu_nstream <= 0 || gpu_nstream > g_buf.nstream ) return fail( HYDRO_E_ARG, "gpu_nstream out of range" );
   const int npatch = 8*npg;
   const size_t p_b = (size_t)5*512*g_buf.elem;
   // NPatch/NStream per stream, remainder to the last stream (CUAPI_Asyn_dtSolver.cu:183-197)
   const int base = npatch / gpu_nstream;
   int off = 0;
   for (int s=0; s<gpu_nstream; s++)
   {
      const int n = ( s == gpu_nstream - 1 ) ? npatch - off : base;
      if ( n <= 0 ) continue;
      hipStream_t st = g_buf.streams[s];
      char *dflu = (char*)g_buf.d_flu_t + off*p_b, *ddt = (char*)g_buf.d_dt + off*g_buf.elem;
      HIP_TRY( hipMemcpyAsync( dflu, (const char*)h_flu_t + off*p_b, n*p_b, hipMemcpyHostToDevice, st ) );
      const int rc = ( g_buf.precision == HYDRO_FP64 )
                     ? launch_dt_cfl<double>( (const double*)dflu, (double*)ddt, n, dh, safety, gamma, min_pres, st )
                     : launch_dt_cfl<float> ( (const float*)dflu,  (float*)ddt,  n, dh, safety, gamma, min_pres, st );
      if ( rc != 0 ) return fail( HYDRO_E_DEVICE, "dt solver launch failed" );
      HIP_TRY( hipMemcpyAsync( (char*)h_dt + off*g_buf.elem, ddt, n*g_buf.elem, hipMemcpyDeviceToHost, st ) );
      off += n;
   }
   return HYDRO_OK;
}
int hydro_device_info( HydroDeviceInfo *info )
{
   if ( info == nullptr ) return fail( HYDRO_E_ARG, "info == NULL" );
   std::memset( info, 0, sizeof( *info ) );
   int dev = 0, n = 0, drv = 0, rt = 0;
   HIP_TRY( hipGetDeviceCount( &n ) );
   HIP_TRY( hipGetDevice( &dev ) );
   hipDeviceProp_t p;
   HIP_TRY( hipGetDeviceProperties( &p, dev ) );
   HIP_TRY( hipDriverGetVersion( &drv ) );
   HIP_TRY( hipRuntimeGetVersion( &rt ) );
   std::snprintf( info->name, sizeof( info->name ), "%s", p.name );
   std::snprintf( info->arch, sizeof( info->arch ), "%s", p.gcnArchName );
   info->device                = dev;
   info->device_count          = n;
   info->compute_units         = p.multiProcessorCount;
   info->clock_khz             = p.clockRate;
   info->global_mem_bytes      = p.totalGlobalMem;
   info->lds_per_block_bytes   = p.sharedMemPerBlock;
   info->wavefront_size        = p.warpSize;
   info->max_threads_per_block = p.maxThreadsPerBlock;
   info->driver_version        = drv;
   info->runtime_version       = rt;
   info->ecc                   = p.ECCEnabled;
   info->async_engines         = p.asyncEngineCount;
   info->concurrent_kernels    = p.concurrentKernels;
   return HYDRO_OK;
}
int hydro_asyn_fluid_solver( const void *h_in, void *h_out, void *h_flux, const HydroSolveArgs *args, int gpu_nstream )
{
   // call order first: without buffers no argument can be valid
   if ( g_buf.npg_max == 0 )            return fail( HYDRO_E_STATE, "hydro_mem_allocate_fluid was not called" );
   if ( !args )                         return fail( HYDRO_E_ARG, "args is NULL" );
   HydroSolveArgs a = *args;
   a.in = h_in;  a.out = h_out;  a.flux = h_flux;
   int rc = check_args( &a );
   if ( rc != HYDRO_OK ) return rc;
   if ( a.npg > g_buf.npg_max )         return fail( HYDRO_E_ARG, "npg %d > allocated %d", a.npg, g_buf.npg_max );
   if ( a.scheme != g_buf.scheme || a.precision != g_buf.precision )
                                        return fail( HYDRO_E_ARG, "scheme/precision differ from the allocation" );
   if ( gpu_nstream <= 0 || gpu_nstream > g_buf.nstream ) return fail( HYDRO_E_ARG, "gpu_nstream out of range" );
   if ( a.npg == 0 ) return HYDRO_OK;

   const int nx = nx_of( a.scheme );
   const size_t in_pg = (size_t)5*nx*nx*nx*g_buf.elem, out_pg = (size_t)5*4096*g_buf.elem, fl_pg = (size_t)9*5*256*g_buf.elem;
   // NPG/NStream per stream, remainder to the last stream (CUAPI_Asyn_FluidSolver.cu:433-445)
   const int base = a.npg / gpu_nstream;
   int off = 0;
   for (int s=0; s<gpu_nstream; s++)
   {
      const int n = ( s == gpu_nstream - 1 ) ? a.npg - off : base;
      if ( n <= 0 ) continue;
      hipStream_t st = g_buf.streams[s];
      char *din = (char*)g_buf.d_in + off*in_pg, *dout = (char*)g_buf.d_out + off*out_pg, *dfl = (char*)g_buf.d_flux + off*fl_pg;
      HIP_TRY( hipMemcpyAsync( din, (const char*)h_in + off*in_pg, n*in_pg, hipMemcpyHostToDevice, st ) );
      rc = enqueue( &a, din, dout, dfl, nullptr, n, st );
      if ( rc != HYDRO_OK ) return rc;
      HIP_TRY( hipMemcpyAsync( (char*)h_out + off*out_pg, dout, n*out_pg, hipMemcpyDeviceToHost, st ) );
      if ( a.store_flux )
         HIP_TRY( hipMemcpyAsync( (char*)h_flux + off*fl_pg, dfl, n*fl_pg, hipMemcpyDeviceToHost, st ) );
      off += n;
   }
   return HYDRO_OK;
}

} // extern "C"
