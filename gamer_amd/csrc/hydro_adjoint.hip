// Adjoint (vector-Jacobian product) and tangent-linear (Jacobian-vector product) of one MHM_RP step, FP64.
//
// The step is the map U_in (22^3 x 5 per patch group) -> U_out (16^3 x 5) of the forward kernel
// (src/Model_Hydro/CPU_Hydro/CPU_FluidSolver_MHM.cpp:263-589 with the half-step prediction of :615-949).
// Explicit phases over a per-PG global workspace, each phase one kernel over all patch groups with one face or
// cell per lane:
//
//   forward   F1 + F2, one z plane per workgroup: half-step fluxes (Hydro_RiemannPredict_Flux) in LDS, then the
//                half-step update, floors and primitive W (Hydro_RiemannPredict) -> ws
//             F4 (tangent / plain forward only) PLM faces (Hydro_DataReconstruction) and the full-step flux of
//                each face, then the update of the 16^3 output cells
//   reverse   R4 lambda_F of each full-step face from lambda_out (transpose of Hydro_FullStepUpdate),
//                lambda_FC = J_RS^T lambda_F (transpose of Hydro_ComputeFlux), with the two PLM faces of the
//                face recomputed from W
//             R3 lambda_W = sum over the three stencil cells of J_PLM^T lambda_FC (transpose of the limiter,
//                face clamp and mirror), one line march per direction
//             R2 lambda_U_half = J_Con2Pri^T lambda_W (transpose of the floors and primitive conversion)
//             R1 + R0, one z plane per workgroup: lambda of every half-step face, J_RS^T of the cell-centred
//                half-step solve per side, and lambda_in = lambda_out (core) + lambda_U_half + the half-step face
//                contributions (the z faces' upper sides added by a second, small kernel)
//
// The reverse sweep is written out by hand.  The Riemann solves (HLLC, HLLE and Roe) run backwards as hand-written
// reverse-mode sweeps (hydro_reverse.hpp: solver + interface-state builders, one pass per face for all ten
// inputs).  The PLM reconstruction (face clamps and mirror) is reversed by hand for every
// limiter, with the limited slope's two partials from a two-direction dual evaluation.  The half-step conversion
// with floors takes its local Jacobian from forward-mode dual numbers (hydro_dual.hpp) pushed through the same
// templated device function, one input direction at a time.  All forms take exactly the branches of the forward
// evaluation.
// The reverse phases are gathers: every workspace entry has one writer, so there are no atomics and the result
// is deterministic.
//
// Scope: MHM_RP, FP64, any Riemann solver and limiter of the forward path.  The MinMod retry is not
// differentiated (the first-iteration coefficient is used); dt and dh are constants.
//
// No floating-point contraction in this translation unit (it holds only the derivative kernels, and the
// pragma comes before the solver headers it instantiates).  Every value is then the rounded result of its own
// source expression, with fused multiply-adds only where fmad() is written, so the dual-number evaluation, the
// plain recomputation inside the reverse sweeps and the forward phases all produce bitwise equal values and
// split ties the same way.  Ties are common: a uniform floored region puts the HLLC star pressure on the
// pressure floor up to rounding, and the floor decides whether its derivative passes.

#pragma clang fp contract(off)

// the derivative phases and their dual-number twins take the sound speed as sqrt(gamma P / rho) (see sound_speed)
#define HYDRO_IEEE_SOUND_SPEED
#include "hydro_dual.hpp"
#include "hydro_kernels.hpp"
#include "hydro_reverse.hpp"

namespace gamer_amd {

namespace {

constexpr int NX = 22, N2 = NX*NX, N3 = NX*N2;
constexpr int PS = 16;
constexpr int NT = 256;
constexpr int STRIDE[3] = { 1, NX, N2 };

// workspace layout per patch group, in units of N3 doubles.  Written / last read by:
constexpr int WS_W   = 0;            // [5]       half-step primitives (cells 1..20)            F1+F2 / R3
constexpr int WS_LFC = 5;            // [3][2][5] lambda of the full-step faces                 R4 / R3
constexpr int WS_LW  = 35;           // [5]       lambda_W                                      R3 / R2
constexpr int WS_LU  = 40;           // [5]       lambda of the half-step conserved state       R2 / R0
constexpr int WS_H   = 45;           // [5]       half-step conserved state                     F1+F2 / R2
constexpr int WS_ZGR = 0;            // [5]       upper-side contributions of the half-step z faces  R1 / R0
                                     //           (over W, dead after R3)
constexpr int WS_ADJ = 50;
static_assert( WS_LFC + 30 <= WS_LW && WS_H + NV <= WS_ADJ, "workspace slots overlap" );

__device__ __forceinline__ void coords( int c, int &i, int &j, int &k ) { i = c % NX;  j = ( c / NX ) % NX;  k = c / N2; }
__device__ __forceinline__ bool in_rng( int x, int lo, int hi ) { return x >= lo && x <= hi; }

// face c + 1/2 e_d with normal coordinate n, tangential a, b
__device__ __forceinline__ void split( int d, int i, int j, int k, int &n, int &a, int &b )
{
   n = ( d == 0 ) ? i : ( d == 1 ) ? j : k;
   a = ( d == 0 ) ? j : i;
   b = ( d == 2 ) ? j : k;
}
__device__ __forceinline__ bool half_face( int n, int a, int b ) { return in_rng( n, 0, 20 ) && in_rng( a, 1, 20 ) && in_rng( b, 1, 20 ); }
__device__ __forceinline__ bool core3( int i, int j, int k ) { return in_rng( i, 3, 18 ) && in_rng( j, 3, 18 ) && in_rng( k, 3, 18 ); }
__device__ __forceinline__ int  out_index( int i, int j, int k ) { return ( i - 3 ) + PS*( ( j - 3 ) + PS*( k - 3 ) ); }

// The phases enumerate only the cells or faces they work on (a box [lo, lo + len) per axis, x fastest, so that
// neighbouring lanes read neighbouring cells) instead of striding over the whole 22^3 block and skipping.
struct Box { int lx, ly, lz, nx, ny, nz; };
__device__ __forceinline__ int box_cell( const Box &b, int e, int &i, int &j, int &k )
{
   const int x = e % b.nx, r = e / b.nx, y = r % b.ny, z = r / b.ny;
   i = b.lx + x;  j = b.ly + y;  k = b.lz + z;
   return i + NX*j + N2*k;
}
// d-normal faces c + 1/2 e_d with normal coordinate in [n0, n0 + nn) and both tangential ones in [t0, t0 + nt)
__device__ __forceinline__ Box face_box( int d, int n0, int nn, int t0, int nt )
{
   Box b;
   b.lx = ( d == 0 ) ? n0 : t0;  b.nx = ( d == 0 ) ? nn : nt;
   b.ly = ( d == 1 ) ? n0 : t0;  b.ny = ( d == 1 ) ? nn : nt;
   b.lz = ( d == 2 ) ? n0 : t0;  b.nz = ( d == 2 ) ? nn : nt;
   return b;
}
constexpr int NHF = 21*20*20;        // half-step faces per direction: normal 0..20, tangential 1..20
constexpr int NFF = 17*16*16;        // full-step faces per direction: normal 2..18, tangential 3..18
constexpr Box HALF_CELLS = { 1, 1, 1, 20, 20, 20 }, CORE_CELLS = { 3, 3, 3, 16, 16, 16 };

// ---------------- pieces of the step, generic in the arithmetic type ----------------
// cell-centred interface state of the half step (pressure recovery + floor, 1/rho, sound speed)
template <int D, typename T>
__device__ __forceinline__ St<T> cell_st( const T u[NV], const Prm<T> &P )
{
   const T ir = rcp( u[0] );
   const T eint = u[4] - (T)0.5*( u[1]*u[1] + u[2]*u[2] + u[3]*u[3] )*ir;
   const T p = floor_pres( eint*P.gm1, P.min_pres );
   return cell_state<D>( u, p, ir, vsqrt( P.gamma*p*ir ) );
}

template <int RS, typename T>
__device__ __forceinline__ void half_flux( int d, const T uL[NV], const T uR[NV], const Prm<T> &P, T f[NV] )
{
   if      ( d == 0 ) riemann<RS, 0>( cell_st<0>( uL, P ), cell_st<0>( uR, P ), P, f );
   else if ( d == 1 ) riemann<RS, 1>( cell_st<1>( uL, P ), cell_st<1>( uR, P ), P, f );
   else               riemann<RS, 2>( cell_st<2>( uL, P ), cell_st<2>( uR, P ), P, f );
}

template <int RS, typename T>
__device__ __forceinline__ void full_flux( int d, const T wL[NV], const T wR[NV], const Prm<T> &P, T f[NV] )
{
   if      ( d == 0 ) riemann<RS, 0>( face_state<0>( wL, P ), face_state<0>( wR, P ), P, f );
   else if ( d == 1 ) riemann<RS, 1>( face_state<1>( wL, P ), face_state<1>( wR, P ), P, f );
   else               riemann<RS, 2>( face_state<2>( wL, P ), face_state<2>( wR, P ), P, f );
}

// reverse sweeps of half_flux / full_flux: bL, bR = (d flux / d input side)^T bf, one pass
template <int RS, int D>
__device__ __forceinline__ void half_flux_rev_d( const double uL[NV], const double uR[NV], const Prm<double> &P,
                                                 const double bf[NV], double bL[NV], double bR[NV] )
{
   const StD L = cell_st<D>( uL, P ), Q = cell_st<D>( uR, P );
   StD sL, sQ;
   riemann_rev<RS, D>( L, Q, P, bf, sL, sQ );
   cell_st_rev<D>( uL, L, sL, P, bL );
   cell_st_rev<D>( uR, Q, sQ, P, bR );
}
template <int RS>
__device__ __forceinline__ void half_flux_rev( int d, const double uL[NV], const double uR[NV], const Prm<double> &P,
                                               const double bf[NV], double bL[NV], double bR[NV] )
{
   if      ( d == 0 ) half_flux_rev_d<RS, 0>( uL, uR, P, bf, bL, bR );
   else if ( d == 1 ) half_flux_rev_d<RS, 1>( uL, uR, P, bf, bL, bR );
   else               half_flux_rev_d<RS, 2>( uL, uR, P, bf, bL, bR );
}

template <int RS, int D>
__device__ __forceinline__ void full_flux_rev_d( const double wL[NV], const double wR[NV], const Prm<double> &P,
                                                 const double bf[NV], double bL[NV], double bR[NV] )
{
   const StD L = face_state<D>( wL, P ), Q = face_state<D>( wR, P );
   StD sL, sQ;
   riemann_rev<RS, D>( L, Q, P, bf, sL, sQ );
   face_state_rev<D>( wL, L, sL, P, bL );
   face_state_rev<D>( wR, Q, sQ, P, bR );
}
template <int RS>
__device__ __forceinline__ void full_flux_rev( int d, const double wL[NV], const double wR[NV], const Prm<double> &P,
                                               const double bf[NV], double bL[NV], double bR[NV] )
{
   if      ( d == 0 ) full_flux_rev_d<RS, 0>( wL, wR, P, bf, bL, bR );
   else if ( d == 1 ) full_flux_rev_d<RS, 1>( wL, wR, P, bf, bL, bR );
   else               full_flux_rev_d<RS, 2>( wL, wR, P, bf, bL, bR );
}

// half-step conserved state -> density floor, internal-energy floor, primitive variables with pressure floor
// (CPU_FluidSolver_MHM.cpp:900-933)
template <typename T>
__device__ __forceinline__ void con2pri_floor( const T u[NV], const Prm<T> &P, T w[NV] )
{
   T c[NV];
#  pragma unroll
   for (int v=0; v<NV; v++) c[v] = u[v];
   c[0] = vmax( c[0], P.min_dens );
   const T r  = rcp( c[0] );
   const T ke = (T)0.5*( c[1]*c[1] + c[2]*c[2] + c[3]*c[3] )*r;
   T ein = c[4] - ke;
   const T eout = ( ein == ein ) ? vmax( ein, P.min_eint ) : ein;
   if ( ein != eout ) { c[4] = c[4] - ein + eout;  ein = c[4] - ke; }
   w[0] = c[0];
   w[1] = c[1]*r;
   w[2] = c[2]*r;
   w[3] = c[3]*r;
   const T p = ein*P.gm1;
   w[4] = ( p == p ) ? vmax( p, P.min_pres ) : p;
}

template <int LIM, typename T>
__device__ __forceinline__ void plm( const T &L, const T &C, const T &R, const Prm<T> &P, T &fL, T &fR )
{
   plm_face<LIM>( L, C, R, P.limiter, P.coeff, fL, fR );
}

// ---------------- workspace views ----------------
// The adjoint keeps h (read by R2) in its own slots; the tangent kernels store W as dual
// numbers (two doubles each).
static_assert( 2*NV <= WS_ADJ, "tangent workspace must fit the per-PG slot" );

template <typename T>
struct Fwd {
   T *W;
   __device__ Fwd( T *base ) : W( base + WS_W*N3 ) {}
   __device__ T &w( int v, int c ) { return W[ v*N3 + c ]; }
};

// load cell c of the input (value, and tangent for Dual<1>)
template <typename T> struct Load;
template <> struct Load<double>
{
   __device__ static void cell( const double *in, const double *, int c, double u[NV] )
   {
#     pragma unroll
      for (int v=0; v<NV; v++) u[v] = in[ v*N3 + c ];
   }
};
template <> struct Load< Dual<1> >
{
   __device__ static void cell( const double *in, const double *tin, int c, Dual<1> u[NV] )
   {
#     pragma unroll
      for (int v=0; v<NV; v++) { u[v] = Dual<1>( in[ v*N3 + c ] );  u[v].d[0] = tin[ v*N3 + c ]; }
   }
};

// ---------------- phases: one kernel each, one element per lane ----------------
// Every phase is a gather over the faces or cells of each patch group that reads only what earlier phases
// wrote, so the phases run as separate kernels in stream order, each over a flat grid of bpp(count) workgroups
// per patch group: element e of patch group pg is lane e % NT of workgroup pg*bpp + e / NT.  One workgroup
// marching through all phases of its patch group kept only 4 waves per patch group busy; spread over the grid,
// a phase keeps as many patch groups' workspace traffic in flight as the registers of its own kernel allow.
constexpr int bpp( int count ) { return ( count + NT - 1 ) / NT; }
constexpr int N_F4 = PS*PS*PS, N_R4 = 3*NFF, N_R2 = 8000;

// Consecutive logical blocks of a phase share data (the planes k-1..k+1 of a plane-per-workgroup phase, the W
// stencils of neighbouring face blocks), so every phase takes its block index from xcd_block() (hydro_kernels.hpp).
struct Elem { int pg, e; };
__device__ __forceinline__ Elem elem( int count )
{
   const int b = bpp( count ), l = (int)xcd_block();       // grids_fit keeps the grid below 2^31 blocks
   return { l / b, ( l % b )*NT + (int)threadIdx.x };
}

template <typename T>
__device__ __forceinline__ Prm<T> phase_prm( const Prm<double> &P0, double minmod_coeff )
{
   Prm<T> P = prm_as<T>( P0 );
   P.coeff = vmax( (T)minmod_coeff, (T)0.0 );
   return P;
}

template <typename T> struct Store;
template <> struct Store<double>
{
   __device__ static void put( double *out, double *, size_t o, const double &x ) { out[o] = x; }
};
template <> struct Store< Dual<1> >
{
   __device__ static void put( double *out, double *tout, size_t o, const Dual<1> &x ) { if ( out ) out[o] = x.v;  tout[o] = x.d[0]; }
};

constexpr size_t NIN = (size_t)NV*N3, NOUT = (size_t)NV*PS*PS*PS;

// F1 + F2 for one z plane k per workgroup: the half-step fluxes of the plane's x and y faces and of the z faces
// below and above it (Hydro_RiemannPredict_Flux) go to LDS, then the plane's 20^2 cells take the half-step update
// (x and y flux differences, then z, as the production kernel), the floors and the primitive conversion
// (Hydro_RiemannPredict).  The fluxes never reach memory; the price is that each interior z face is solved by
// both planes it separates (32,800 solves per PG instead of 25,200).  h_off >= 0 (FP64 only): the half-step
// state h is stored too, for R2.
constexpr int F12_NT = 512, F12_NXF = 21*20, F12_NZF = 20*20, F12_NF = 2*F12_NXF + 2*F12_NZF;
template <int RS, typename T>
__global__ void __launch_bounds__( F12_NT )
f12_kernel( const double *__restrict__ g_in, const double *__restrict__ g_tin, double *__restrict__ g_ws, int h_off,
            Prm<double> P0, double minmod_coeff )
{
   __shared__ double s_raw[NV*F12_NF*( sizeof( T )/sizeof( double ) )];
   T *sF = reinterpret_cast<T*>( s_raw );                    // [v][face]
   const int tid = threadIdx.x, lb = (int)xcd_block(), pg = lb / 20, k = 1 + lb % 20;
   const Prm<T> P = phase_prm<T>( P0, minmod_coeff );
   const double *in = g_in + pg*NIN, *tin = g_tin ? g_tin + pg*NIN : nullptr;
   Fwd<T> F( reinterpret_cast<T*>( g_ws + (size_t)pg*WS_ADJ*N3 ) );
   // faces, stored at their low cell: x (normal i = 0..20, j = 1..20), y (i = 1..20, j = 0..20), then the z
   // faces with low cell in plane k - 1 and in plane k (i, j = 1..20); lanes run along x
   for (int f = tid; f < F12_NF; f += F12_NT)
   {
      int d, c;
      if ( f < F12_NXF )        { d = 0;  c = f % 21 + NX*( 1 + f / 21 ) + N2*k; }
      else if ( f < 2*F12_NXF ) { d = 1;  const int r = f - F12_NXF;  c = 1 + r % 20 + NX*( r / 20 ) + N2*k; }
      else
      {
         d = 2;
         const int r = f - 2*F12_NXF, zz = r / F12_NZF, q = r - zz*F12_NZF;
         c = 1 + q % 20 + NX*( 1 + q / 20 ) + N2*( k - 1 + zz );
      }
      T uL[NV], uR[NV], fl[NV];
      Load<T>::cell( in, tin, c, uL );
      Load<T>::cell( in, tin, c + STRIDE[d], uR );
      half_flux<RS>( d, uL, uR, P, fl );
#     pragma unroll
      for (int v=0; v<NV; v++) sF[ v*F12_NF + f ] = fl[v];
   }
   __syncthreads();
   for (int e = tid; e < F12_NZF; e += F12_NT)
   {
      const int i = 1 + e % 20, j = 1 + e / 20, c = i + NX*j + N2*k;
      const int xh = i + 21*( j - 1 ), yh = F12_NXF + ( i - 1 ) + 20*j, zh = 2*F12_NXF + F12_NZF + ( i - 1 ) + 20*( j - 1 );
      T u[NV], h[NV], w[NV];
      Load<T>::cell( in, tin, c, u );
#     pragma unroll
      for (int v=0; v<NV; v++)
      {
         const T *fv = sF + v*F12_NF;
         const T dxy = ( fv[xh] - fv[xh - 1] ) + ( fv[yh] - fv[yh - 20] );
         h[v] = u[v] - P.dt_dh2*( dxy + ( fv[zh] - fv[zh - F12_NZF] ) );
      }
      con2pri_floor( h, P, w );
#     pragma unroll
      for (int v=0; v<NV; v++) F.w( v, c ) = w[v];
      if constexpr ( std::is_same<T, double>::value )
         if ( h_off >= 0 )
         {
            double *H = g_ws + (size_t)pg*WS_ADJ*N3 + (size_t)h_off*N3;
#           pragma unroll
            for (int v=0; v<NV; v++) H[ v*N3 + c ] = h[v];
         }
   }
}

// F4 (tangent / plain forward), in two kernels.  f4_face_kernel: the full-step flux of each of the 13,056 faces
// (3 x 17 x 16 x 16, low cell 2..18 along the normal, tangential 3..18), once per face, with its two PLM faces
// recomputed from the half-step primitives W as the adjoint's R4 does; the fluxes go to the workspace.
// f4_update_kernel: the full-step update of the 16^3 output cells from the fluxes of their faces.
// T = Dual<1>: out = U_out (optional), tout = J v; T = double: out = U_out.
constexpr int WS_TF = 2*NV;          // tangent flux slots: [3][5][NFF] values of T, after W (5 values of T per cell)
static_assert( WS_TF*N3 + 3*NV*NFF*2 <= WS_ADJ*N3, "tangent fluxes must fit the per-PG workspace" );

// index of the d-normal face with low cell (i, j, k) in face_box(d, 2, 17, 3, 16)
__device__ __forceinline__ int full_face_index( int d, int i, int j, int k )
{
   return ( d == 0 ) ? ( i - 2 ) + 17*( ( j - 3 ) + 16*( k - 3 ) )
        : ( d == 1 ) ? ( i - 3 ) + 16*( ( j - 2 ) + 17*( k - 3 ) )
        :              ( i - 3 ) + 16*( ( j - 3 ) + 16*( k - 2 ) );
}

template <int RS, int LIM, typename T>
__global__ void __launch_bounds__( NT )
f4_face_kernel( double *__restrict__ g_ws, Prm<double> P0, double minmod_coeff )
{
   const Elem x = elem( 3*NFF );
   if ( x.e >= 3*NFF ) return;
   const Prm<T> P = phase_prm<T>( P0, minmod_coeff );
   T *ws = reinterpret_cast<T*>( g_ws + (size_t)x.pg*WS_ADJ*N3 );
   Fwd<T> F( ws );
   T *tf = ws + (size_t)WS_TF*N3/( sizeof( T )/sizeof( double ) );
   const int d = x.e / NFF, q = x.e - d*NFF;
   int i, j, k;
   const int c = box_cell( face_box( d, 2, 17, 3, 16 ), q, i, j, k );
   const int s = STRIDE[d];
   T wl[NV], wr[NV], f[NV];
#  pragma unroll
   for (int v=0; v<NV; v++)
   {
      const T wm = F.w( v, c - s ), w0 = F.w( v, c ), w1 = F.w( v, c + s ), w2 = F.w( v, c + 2*s );
      T dummy;
      plm<LIM>( wm, w0, w1, P, dummy, wl[v] );      // R face of the low cell
      plm<LIM>( w0, w1, w2, P, wr[v], dummy );      // L face of the high cell
   }
   full_flux<RS>( d, wl, wr, P, f );
#  pragma unroll
   for (int v=0; v<NV; v++) tf[ ( d*NV + v )*NFF + q ] = f[v];
}

template <typename T>
__global__ void __launch_bounds__( NT )
f4_update_kernel( const double *__restrict__ g_in, const double *__restrict__ g_tin, double *__restrict__ g_out,
                  double *__restrict__ g_tout, double *__restrict__ g_ws, Prm<double> P0, double minmod_coeff )
{
   const Elem x = elem( N_F4 );
   if ( x.e >= N_F4 ) return;
   const Prm<T> P = phase_prm<T>( P0, minmod_coeff );
   const double *in = g_in + x.pg*NIN, *tin = g_tin ? g_tin + x.pg*NIN : nullptr;
   double *out = g_out ? g_out + x.pg*NOUT : nullptr, *tout = g_tout ? g_tout + x.pg*NOUT : nullptr;
   const T *tf = reinterpret_cast<const T*>( g_ws + (size_t)x.pg*WS_ADJ*N3 ) + (size_t)WS_TF*N3/( sizeof( T )/sizeof( double ) );
   int i, j, k;
   const int c = box_cell( CORE_CELLS, x.e, i, j, k );
   T u[NV], dF[3][NV];
   Load<T>::cell( in, tin, c, u );
   for (int d=0; d<3; d++)
   {
      const int qp = full_face_index( d, i, j, k );                                 // face c + 1/2
      const int qm = full_face_index( d, i - ( d == 0 ), j - ( d == 1 ), k - ( d == 2 ) );   // face c - 1/2
#     pragma unroll
      for (int v=0; v<NV; v++) dF[d][v] = tf[ ( d*NV + v )*NFF + qp ] - tf[ ( d*NV + v )*NFF + qm ];
   }
   const size_t o = out_index( i, j, k );
#  pragma unroll
   for (int v=0; v<NV; v++)
      Store<T>::put( out, tout, (size_t)v*PS*PS*PS + o, u[v] - P.dt_dh*( ( dF[0][v] + dF[1][v] ) + dF[2][v] ) );
}

// ---------------- reverse phases ----------------
struct Adj {
   double *LW, *LU, *LFC;
   __device__ Adj( double *base ) : LW( base + WS_LW*N3 ), LU( base + WS_LU*N3 ), LFC( base + WS_LFC*N3 ) {}
   __device__ double &lw( int v, int c )                { return LW[ v*N3 + c ]; }
   __device__ double &lu( int v, int c )                { return LU[ v*N3 + c ]; }
   __device__ double &lfc( int d, int s, int v, int c ) { return LFC[ ( ( d*2 + s )*NV + v )*N3 + c ]; }
};

typedef Dual<1> D1;

// R4 full-step faces: lambda_F = dt/dh (lambda_out(c+e_d) - lambda_out(c)), then J_RS^T per side.  Only the
// faces that feed a full-step flux receive a value; R3 reads the others as zero.
template <int RS, int LIM>
__global__ void __launch_bounds__( NT )
r4_kernel( const double *__restrict__ g_lout, double *__restrict__ g_ws, Prm<double> P0, double minmod_coeff )
{
   const Elem x = elem( N_R4 );
   if ( x.e >= N_R4 ) return;
   const Prm<double> P = phase_prm<double>( P0, minmod_coeff );
   const double *lout = g_lout + x.pg*NOUT;
   double *ws = g_ws + (size_t)x.pg*WS_ADJ*N3;
   Fwd<double> F( ws );
   Adj A( ws );
   const int d = x.e / NFF;
   int i, j, k;
   const int c = box_cell( face_box( d, 2, 17, 3, 16 ), x.e - d*NFF, i, j, k );
   const int s = STRIDE[d];
   int i2 = i, j2 = j, k2 = k;
   if ( d == 0 ) i2++; else if ( d == 1 ) j2++; else k2++;
   double lf[NV];
   const bool lo = core3( i, j, k ), hi = core3( i2, j2, k2 );
#  pragma unroll
   for (int v=0; v<NV; v++)
   {
      const double ll = lo ? lout[ (size_t)v*PS*PS*PS + out_index( i, j, k ) ] : 0.0;
      const double lh = hi ? lout[ (size_t)v*PS*PS*PS + out_index( i2, j2, k2 ) ] : 0.0;
      lf[v] = P.dt_dh*( lh - ll );
   }
   // the R face of c and the L face of c + e_d (F3's values, recomputed from the half-step primitives)
   double wl[NV], wr[NV];
#  pragma unroll
   for (int v=0; v<NV; v++)
   {
      const double wm = F.w( v, c - s ), w0 = F.w( v, c ), w1 = F.w( v, c + s ), w2 = F.w( v, c + 2*s );
      double fl, fr;
      plm<LIM>( wm, w0, w1, P, fl, wl[v] );
      plm<LIM>( w0, w1, w2, P, wr[v], fr );
   }
   // one reverse sweep for both sides
   double gl[NV] = { 0.0, 0.0, 0.0, 0.0, 0.0 }, gr[NV] = { 0.0, 0.0, 0.0, 0.0, 0.0 };
   full_flux_rev<RS>( d, wl, wr, P, lf, gl, gr );
#  pragma unroll
   for (int v=0; v<NV; v++) { A.lfc( d, 1, v, c ) = gl[v];  A.lfc( d, 0, v, c + s ) = gr[v]; }
}

// R3 along y (d = 1) and z (d = 2), after r3x_kernel has written lambda_W along x: one launch per direction.
// Task (v, line): the lane marches the line through the reconstruction stencils (centres 2..19) of field v and
// reverses each one (plm_face_rev), keeping a window of the three cell adjoints the stencil touches, and adds each
// cell's adjoint to lambda_W once the stencil above it is done.  Only the 16^2 lines with both tangential
// coordinates in 3..18 carry face adjoints (the L face of a centre at normal coordinate 3..19, the R face at
// 2..18); lanes run along x, so the loads of a march step are coalesced.
constexpr int R3_LINES = 16*16;
template <int LIM>
__global__ void __launch_bounds__( NT )
r3_kernel( double *__restrict__ g_ws, Prm<double> P0, double minmod_coeff, int d )
{
   const Elem x = elem( NV*R3_LINES );
   if ( x.e >= NV*R3_LINES ) return;
   const Prm<double> P = phase_prm<double>( P0, minmod_coeff );
   double *ws = g_ws + (size_t)x.pg*WS_ADJ*N3;
   Fwd<double> F( ws );
   Adj A( ws );
   const int s = STRIDE[d], v = x.e / R3_LINES, ln = x.e - v*R3_LINES;
   const int a = 3 + ln % 16, b = 3 + ln / 16;                                  // i, and k (d = 1) or j (d = 2)
   const int c0 = ( d == 1 ) ? a + N2*b : a + NX*b;                             // n = 0 of the line
   double wm = F.w( v, c0 + s ), w0 = F.w( v, c0 + 2*s ), am = 0.0, a0 = 0.0;
   for (int n=2; n<=19; n++)
   {
      const int    cc = c0 + n*s;
      const double wp = F.w( v, cc + s );
      const double gl = ( n >= 3 )  ? A.lfc( d, 0, v, cc ) : 0.0;
      const double gr = ( n <= 18 ) ? A.lfc( d, 1, v, cc ) : 0.0;
      double br = 0.0;
      plm_face_rev<LIM>( wm, w0, wp, P.limiter, P.coeff, gl, gr, am, a0, br );
      A.lw( v, cc - s ) += am;
      am = a0;  a0 = br;  wm = w0;  w0 = wp;
   }
   A.lw( v, c0 + 19*s ) += am;
   A.lw( v, c0 + 20*s ) += a0;
}

// R3 along x (d = 0) with the lines staged through LDS.  The x lines of one (z, field) pair are the 20 rows
// j = 1..20 of one z plane, 440 consecutive doubles of each workspace array, so a workgroup loads W and the two
// face-adjoint arrays of R3X_PAIRS pairs coalesced, lane t marches row 1 + t % 20 of pair t / 20 in LDS (the
// adjoint of row cell n - 1 replaces its W, which the march has passed), and the rows go back coalesced.
// Marching lanes over x lines directly puts 64 different rows under every load instruction.
constexpr int R3X_PAIRS = 6, R3X_NT = 128, R3X_RL = 23, R3X_PAIR = 20*R3X_RL, R3X_NPAIR = 20*NV;
constexpr int R3X_BPP = ( R3X_NPAIR + R3X_PAIRS - 1 )/R3X_PAIRS;
static_assert( R3X_PAIRS*20 <= R3X_NT, "one lane per staged row" );

template <int LIM>
__global__ void __launch_bounds__( R3X_NT )
r3x_kernel( double *__restrict__ g_ws, Prm<double> P0, double minmod_coeff )
{
   __shared__ double sW[R3X_PAIRS*R3X_PAIR], sG0[R3X_PAIRS*R3X_PAIR], sG1[R3X_PAIRS*R3X_PAIR];
   const int tid = threadIdx.x, lb = (int)xcd_block(), pg = lb / R3X_BPP, q0 = ( lb % R3X_BPP )*R3X_PAIRS;
   const int np = min( R3X_PAIRS, R3X_NPAIR - q0 );
   const Prm<double> P = phase_prm<double>( P0, minmod_coeff );
   double *ws = g_ws + (size_t)pg*WS_ADJ*N3;
   Fwd<double> F( ws );
   Adj A( ws );
   // pair q: field q / 20, z = 1 + q % 20; its rows j = 1..20 start at cell NX + N2 z, a 16-byte aligned offset
   // (NX and N2 are even), so the block moves as 220 double2 per array, R3X_LD of them per lane in flight
   constexpr int U = 20*NX/2, R3X_LD = 4;
   const int nu = np*U;
   for (int u0 = tid; u0 < nu; u0 += R3X_LD*R3X_NT)
   {
      double2 a[R3X_LD], b[R3X_LD], g[R3X_LD];
#     pragma unroll
      for (int t=0; t<R3X_LD; t++)
      {
         const int u = u0 + t*R3X_NT;
         if ( u >= nu ) continue;
         const int p = u / U, q = q0 + p, v = q / 20, c = NX + N2*( 1 + q % 20 ) + 2*( u - p*U );
         a[t] = *reinterpret_cast<const double2*>( &F.w( v, c ) );
         b[t] = *reinterpret_cast<const double2*>( &A.lfc( 0, 0, v, c ) );
         g[t] = *reinterpret_cast<const double2*>( &A.lfc( 0, 1, v, c ) );
      }
#     pragma unroll
      for (int t=0; t<R3X_LD; t++)
      {
         const int u = u0 + t*R3X_NT;
         if ( u >= nu ) continue;
         const int p = u / U, r = 2*( u - p*U ), l = p*R3X_PAIR + ( r / NX )*R3X_RL + r % NX;   // r, r + 1: one row
         sW[l] = a[t].x;   sW[l + 1] = a[t].y;
         sG0[l] = b[t].x;  sG0[l + 1] = b[t].y;
         sG1[l] = g[t].x;  sG1[l + 1] = g[t].y;
      }
   }
   __syncthreads();
   if ( tid < np*20 )
   {
      const int p = tid / 20, j = 1 + tid % 20, k = 1 + ( q0 + p ) % 20;
      double *row = sW + p*R3X_PAIR + ( j - 1 )*R3X_RL;
      const double *g0 = sG0 + p*R3X_PAIR + ( j - 1 )*R3X_RL, *g1 = sG1 + p*R3X_PAIR + ( j - 1 )*R3X_RL;
      if ( !( in_rng( j, 3, 18 ) && in_rng( k, 3, 18 ) ) )
      {
         for (int n=1; n<=20; n++) row[n] = 0.0;
      }
      else
      {
         double wm = row[1], w0 = row[2], am = 0.0, a0 = 0.0;
         for (int n=2; n<=19; n++)
         {
            const double wp = row[n + 1];
            const double gl = ( n >= 3 )  ? g0[n] : 0.0;
            const double gr = ( n <= 18 ) ? g1[n] : 0.0;
            double br = 0.0;
            plm_face_rev<LIM>( wm, w0, wp, P.limiter, P.coeff, gl, gr, am, a0, br );
            row[n - 1] = am;
            am = a0;  a0 = br;  wm = w0;  w0 = wp;
         }
         row[19] = am;  row[20] = a0;
      }
   }
   __syncthreads();
   for (int e = tid; e < np*400; e += R3X_NT)
   {
      const int p = e / 400, r = e - p*400, q = q0 + p, v = q / 20, j = 1 + r / 20, n = 1 + r % 20;
      A.lw( v, n + NX*j + N2*( 1 + q % 20 ) ) = sW[ p*R3X_PAIR + ( j - 1 )*R3X_RL + n ];
   }
}

// R2 lambda of the half-step conserved state: J_Con2Pri^T lambda_W (floors included), at the state F2 stored
__global__ void __launch_bounds__( NT )
r2_kernel( double *__restrict__ g_ws, Prm<double> P0, double minmod_coeff )
{
   const Elem x = elem( N_R2 );
   if ( x.e >= N_R2 ) return;
   const Prm<double> P = phase_prm<double>( P0, minmod_coeff );
   const Prm<D1> PD = prm_as<D1>( P );
   double *ws = g_ws + (size_t)x.pg*WS_ADJ*N3;
   const double *H = ws + (size_t)WS_H*N3;
   Adj A( ws );
   int i, j, k;
   const int c = box_cell( HALF_CELLS, x.e, i, j, k );
   double h[NV], lw[NV];
#  pragma unroll
   for (int v=0; v<NV; v++)
   {
      h[v]  = H[ v*N3 + c ];
      lw[v] = A.lw( v, c );
   }
   for (int q=0; q<NV; q++)
   {
      D1 hd[NV], w[NV];
#     pragma unroll
      for (int v=0; v<NV; v++) { hd[v] = D1( h[v] );  hd[v].d[0] = ( q == v ) ? 1.0 : 0.0; }
      con2pri_floor( hd, PD, w );
      double g = 0.0;
#     pragma unroll
      for (int v=0; v<NV; v++) g += lw[v]*w[v].d[0];
      A.lu( q, c ) = g;
   }
}

// R1 + R0, one z plane k = 0..21 per workgroup.  R1: lambda_Fh = dt/(2dh) (lambda_U(c+e_d) - lambda_U(c)) of
// every half-step face, then J_RS^T per side; the plane's x and y faces and the z face above it (normal k)
// leave their side contributions in LDS, the z face's upper-side one goes to the workspace (zgr).  R0: lambda_in of the plane's cells = lambda_out (output cells) +
// lambda_U_half (half-step cells) + the face contributions in the order x low, x high, y low, y high, z low;
// r0z_kernel then adds the z high contribution.  Each face is solved once.
// 768 lanes (3 waves per SIMD, <= 170 VGPRs); Roe's reverse solve needs about 210, so it runs 512
constexpr int R10_NXF = 21*20, R10_NZF = 20*20, R10_NF = 2*R10_NXF + R10_NZF;
constexpr int N_R0Z = 21*R10_NZF;
constexpr int r10_nt( int rs ) { return rs == 2 ? 512 : 768; }

template <int RS>
__global__ void __launch_bounds__( r10_nt( RS ) )
r10_kernel( const double *__restrict__ g_in, const double *__restrict__ g_lout, double *__restrict__ g_lin,
            double *__restrict__ g_ws, Prm<double> P0, double minmod_coeff )
{
   constexpr int R10_NT = r10_nt( RS );
   __shared__ double sX[2][NV][R10_NXF], sY[2][NV][R10_NXF], sZ[NV][R10_NZF];     // [low/high side][v][face]
   const int tid = threadIdx.x, lb = (int)xcd_block(), pg = lb / 22, k = lb % 22;
   const Prm<double> P = phase_prm<double>( P0, minmod_coeff );
   const double *in = g_in + pg*NIN, *lout = g_lout + pg*NOUT;
   double *lin = g_lin + pg*NIN, *ws = g_ws + (size_t)pg*WS_ADJ*N3, *zgr = ws + (size_t)WS_ZGR*N3;
   Adj A( ws );
   // faces with low cell in plane k: x (normal i = 0..20, j = 1..20), y (i = 1..20, j = 0..20), z (i, j = 1..20);
   // plane 0 has only its z faces, plane 21 none
   const int nf = ( k >= 1 && k <= 20 ) ? R10_NF : ( k == 0 ) ? R10_NZF : 0, f0 = ( k == 0 ) ? 2*R10_NXF : 0;
   for (int t = tid; t < nf; t += R10_NT)
   {
      const int f = f0 + t;
      int d, c, r;
      if ( f < R10_NXF )        { d = 0;  r = f;                c = f % 21 + NX*( 1 + f / 21 ) + N2*k; }
      else if ( f < 2*R10_NXF ) { d = 1;  r = f - R10_NXF;      c = 1 + r % 20 + NX*( r / 20 ) + N2*k; }
      else                      { d = 2;  r = f - 2*R10_NXF;    c = 1 + r % 20 + NX*( 1 + r / 20 ) + N2*k; }
      int i, j, kk, n, a, b;
      coords( c, i, j, kk );
      split( d, i, j, kk, n, a, b );
      const int s = STRIDE[d];
      const bool lo = in_rng( n, 1, 20 ), hi = in_rng( n + 1, 1, 20 );   // tangential range is 1..20 already
      double lf[NV];
#     pragma unroll
      for (int v=0; v<NV; v++) lf[v] = P.dt_dh2*( ( hi ? A.lu( v, c + s ) : 0.0 ) - ( lo ? A.lu( v, c ) : 0.0 ) );
      double uL[NV], uR[NV];
#     pragma unroll
      for (int v=0; v<NV; v++) { uL[v] = in[ v*N3 + c ];  uR[v] = in[ v*N3 + c + s ]; }
      double gl[NV] = { 0.0, 0.0, 0.0, 0.0, 0.0 }, gr[NV] = { 0.0, 0.0, 0.0, 0.0, 0.0 };
      half_flux_rev<RS>( d, uL, uR, P, lf, gl, gr );
#     pragma unroll
      for (int v=0; v<NV; v++)
      {
         if      ( d == 0 ) { sX[0][v][r] = gl[v];  sX[1][v][r] = gr[v]; }
         else if ( d == 1 ) { sY[0][v][r] = gl[v];  sY[1][v][r] = gr[v]; }
         else               { sZ[v][r] = gl[v];  zgr[ v*N3 + k*R10_NZF + r ] = gr[v]; }
      }
   }
   __syncthreads();
   for (int e = tid; e < NX*NX; e += R10_NT)
   {
      const int i = e % NX, j = e / NX, c = i + NX*j + N2*k;
      const bool core = core3( i, j, k ), half = in_rng( i, 1, 20 ) && in_rng( j, 1, 20 ) && in_rng( k, 1, 20 );
      const bool xl = half_face( i, j, k ), xh = half_face( i - 1, j, k ), yl = half_face( j, i, k ),
                 yh = half_face( j - 1, i, k ), zl = half_face( k, i, j );
      const int rx = i + 21*( j - 1 ), ry = ( i - 1 ) + 20*j, rz = ( i - 1 ) + 20*( j - 1 );
#     pragma unroll
      for (int v=0; v<NV; v++)
      {
         double acc = core ? lout[ (size_t)v*PS*PS*PS + out_index( i, j, k ) ] : 0.0;
         if ( half ) acc += A.lu( v, c );
         if ( xl ) acc += sX[0][v][rx];
         if ( xh ) acc += sX[1][v][rx - 1];
         if ( yl ) acc += sY[0][v][ry];
         if ( yh ) acc += sY[1][v][ry - 20];
         if ( zl ) acc += sZ[v][rz];
         lin[ v*N3 + c ] = acc;
      }
   }
}

// the z high contribution: cell (i, j, k), k = 1..21, i, j = 1..20, is the upper side of z face k - 1/2
__global__ void __launch_bounds__( NT )
r0z_kernel( double *__restrict__ g_lin, const double *__restrict__ g_ws )
{
   const Elem x = elem( N_R0Z );
   if ( x.e >= N_R0Z ) return;
   double *lin = g_lin + x.pg*NIN;
   const double *zgr = g_ws + (size_t)x.pg*WS_ADJ*N3 + (size_t)WS_ZGR*N3;
   const int n = x.e / R10_NZF, q = x.e - n*R10_NZF;                  // face normal n = 0..20
   const int c = 1 + q % 20 + NX*( 1 + q / 20 ) + N2*( n + 1 );
#  pragma unroll
   for (int v=0; v<NV; v++) lin[ v*N3 + c ] += zgr[ v*N3 + n*R10_NZF + q ];
}

// ---------------- transpose of the periodic ghost-zone preparation ----------------
// level_lam[v][z][y][x] = sum over every patch-group input element that copies level cell (x,y,z)
// (gather form: each level cell collects its copies, so there is one writer per element)
__device__ __forceinline__ int copies( int x, int nx, int ghost, int n, int pl[6], int il[6] )
{
   const int G = nx/PS, q = x/PS;
   int cand[3] = { ( q - 1 + G ) % G, q, ( q + 1 ) % G }, m = 0;
   for (int t=0; t<3; t++)
   {
      bool dup = false;
      for (int u=0; u<t; u++) dup |= ( cand[u] == cand[t] );
      if ( dup ) continue;
      const int p = cand[t];
      for (int i = ( ( x - p*PS + ghost ) % nx + nx ) % nx; i < n; i += nx) { pl[m] = p;  il[m] = i;  m++; }
   }
   return m;
}

template <typename IO>
__global__ void prepare_adjoint_kernel( const IO *__restrict__ pg_lam, IO *__restrict__ lam, int nx, int ny, int nz,
                                        int ghost, long total )
{
   const int n = PS + 2*ghost;
   const long n3 = (long)n*n*n;
   const int gx = nx/PS, gy = ny/PS;
   for (long e = blockIdx.x*(long)blockDim.x + threadIdx.x; e < total; e += (long)gridDim.x*blockDim.x)
   {
      const int x = (int)( e % nx ), y = (int)( ( e / nx ) % ny ), z = (int)( ( e / ( (long)nx*ny ) ) % nz );
      const int v = (int)( e / ( (long)nx*ny*nz ) );
      int px[6], ix[6], py[6], iy[6], pz[6], iz[6];
      const int mx = copies( x, nx, ghost, n, px, ix ), my = copies( y, ny, ghost, n, py, iy ), mz = copies( z, nz, ghost, n, pz, iz );
      double acc = 0.0;
      for (int c=0; c<mz; c++)
      for (int b=0; b<my; b++)
      for (int a=0; a<mx; a++)
      {
         const long pg = px[a] + (long)gx*( py[b] + (long)gy*pz[c] );
         acc += (double)pg_lam[ ( pg*NV + v )*n3 + ix[a] + (long)n*( iy[b] + (long)n*iz[c] ) ];
      }
      lam[e] = (IO)acc;
   }
}

} // namespace

size_t adjoint_workspace_bytes( int npg ) { return (size_t)( npg > 0 ? npg : 0 )*WS_ADJ*N3*sizeof( double ); }

namespace {

// one flat grid of bpp(count) workgroups per patch group
void phase_grid( int npg, int count, dim3 &grid ) { grid = dim3( (unsigned)( (long)npg*bpp( count ) ) ); }
// whether every phase's grid of npg patch groups fits one launch (the per-plane R1 + R0 kernel has the most lanes)
bool grids_fit( int npg ) { return (long)npg*22*r10_nt( 4 ) < ( 1L << 32 ) && (long)npg*bpp( N_R4 )*NT < ( 1L << 32 ); }

template <int RS, int LIM>
int adjoint_phases( const double *in, const double *lam_out, double *lam_in, double *ws, int npg, double minmod_coeff,
                    const Prm<double> &P, hipStream_t stream )
{
   const dim3 block( NT );
   dim3 g;
   // forward: F1 + F2 (the PLM faces are recomputed where they are needed instead of stored)
   if ( !grids_fit( npg ) ) return -3;
   hipLaunchKernelGGL( ( f12_kernel<RS, double> ), dim3( (unsigned)( 20*npg ) ), dim3( F12_NT ), 0, stream, in, nullptr, ws, WS_H, P, minmod_coeff );
   // reverse: R4, R3 (x staged through LDS, then y and z), R2, R1 + R0 per z plane, the z faces' upper sides
   phase_grid( npg, N_R4, g );
   hipLaunchKernelGGL( ( r4_kernel<RS, LIM> ), g, block, 0, stream, lam_out, ws, P, minmod_coeff );
   hipLaunchKernelGGL( ( r3x_kernel<LIM> ), dim3( (unsigned)( npg*R3X_BPP ) ), dim3( R3X_NT ), 0, stream, ws, P, minmod_coeff );
   for (int d=1; d<3; d++)
   {
      phase_grid( npg, NV*R3_LINES, g );
      hipLaunchKernelGGL( ( r3_kernel<LIM> ), g, block, 0, stream, ws, P, minmod_coeff, d );
   }
   phase_grid( npg, N_R2, g );
   hipLaunchKernelGGL( r2_kernel, g, block, 0, stream, ws, P, minmod_coeff );
   hipLaunchKernelGGL( ( r10_kernel<RS> ), dim3( (unsigned)( 22*npg ) ), dim3( r10_nt( RS ) ), 0, stream, in, lam_out, lam_in, ws, P, minmod_coeff );
   phase_grid( npg, N_R0Z, g );
   hipLaunchKernelGGL( r0z_kernel, g, block, 0, stream, lam_in, ws );
   return ( hipGetLastError() == hipSuccess ) ? 0 : -2;
}

template <int RS, int LIM, typename T>
int tangent_phases( const double *in, const double *tin, double *out, double *tout, double *ws, int npg,
                    double minmod_coeff, const Prm<double> &P, hipStream_t stream )
{
   const dim3 block( NT );
   dim3 g;
   if ( !grids_fit( npg ) ) return -3;
   hipLaunchKernelGGL( ( f12_kernel<RS, T> ), dim3( (unsigned)( 20*npg ) ), dim3( F12_NT ), 0, stream, in, tin, ws, -1, P, minmod_coeff );
   phase_grid( npg, 3*NFF, g );
   hipLaunchKernelGGL( ( f4_face_kernel<RS, LIM, T> ), g, block, 0, stream, ws, P, minmod_coeff );
   phase_grid( npg, N_F4, g );
   hipLaunchKernelGGL( ( f4_update_kernel<T> ), g, block, 0, stream, in, tin, out, tout, ws, P, minmod_coeff );
   return ( hipGetLastError() == hipSuccess ) ? 0 : -2;
}

} // namespace

int launch_mhm_rp_adjoint( const double *in, const double *lam_out, double *lam_in, double *ws, int npg, int rsolver,
                           double minmod_coeff, const Prm<double> &P, hipStream_t stream )
{
   if ( npg <= 0 ) return 0;
   const bool l4 = ( kernel_limiter( P.limiter, minmod_coeff ) == 4 );
#  define GA_ADJ( RS_ ) ( l4 ? adjoint_phases<RS_, 4>( in, lam_out, lam_in, ws, npg, minmod_coeff, P, stream ) \
                             : adjoint_phases<RS_, 0>( in, lam_out, lam_in, ws, npg, minmod_coeff, P, stream ) )
   switch ( rsolver )
   {
      case 4: return GA_ADJ( 4 );
      case 3: return GA_ADJ( 3 );
      case 2: return GA_ADJ( 2 );
      default: return -1;
   }
#  undef GA_ADJ
}

int launch_mhm_rp_tangent( const double *in, const double *tin, double *out, double *tout, double *ws, int npg,
                           int rsolver, double minmod_coeff, const Prm<double> &P, hipStream_t stream )
{
   if ( npg <= 0 ) return 0;
   const bool l4 = ( kernel_limiter( P.limiter, minmod_coeff ) == 4 ), tl = ( tin != nullptr );
#  define GA_TAN( RS_, T_ ) ( l4 ? tangent_phases<RS_, 4, T_>( in, tin, out, tout, ws, npg, minmod_coeff, P, stream ) \
                                 : tangent_phases<RS_, 0, T_>( in, tin, out, tout, ws, npg, minmod_coeff, P, stream ) )
   switch ( rsolver )
   {
      case 4: return tl ? GA_TAN( 4, D1 ) : GA_TAN( 4, double );
      case 3: return tl ? GA_TAN( 3, D1 ) : GA_TAN( 3, double );
      case 2: return tl ? GA_TAN( 2, D1 ) : GA_TAN( 2, double );
      default: return -1;
   }
#  undef GA_TAN
}

template <typename IO>
int launch_prepare_adjoint( const IO *pg_lam, IO *lam, int nx, int ny, int nz, int ghost, hipStream_t stream )
{
   const long total = (long)NV*nx*ny*nz;
   if ( total <= 0 ) return 0;
   long b = ( total + 255 )/256;
   if ( b > 65536 ) b = 65536;
   hipLaunchKernelGGL( prepare_adjoint_kernel<IO>, dim3( (int)b ), dim3( 256 ), 0, stream, pg_lam, lam, nx, ny, nz, ghost, total );
   return ( hipGetLastError() == hipSuccess ) ? 0 : -2;
}

template int launch_prepare_adjoint<double>( const double*, double*, int, int, int, int, hipStream_t );
template int launch_prepare_adjoint<float> ( const float*,  float*,  int, int, int, int, hipStream_t );

} // namespace gamer_amd
