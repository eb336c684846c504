// Reverse-mode local sweeps (FP64) of the interface-state builders, of the HLLC / HLLE / Roe interface solvers and
// of the PLM face reconstruction, for the adjoint of the MHM_RP step.
//
// Each *_rev function takes the primal inputs, recomputes the forward intermediates and accumulates the input
// adjoints for one output adjoint in a single backward pass: the row vector bf^T J of the piece, which the
// dual-number form (hydro_dual.hpp) needs one evaluation per input direction for (10 for a Riemann solve).
//
// Branches: every min / max / select / floor takes the branch of the forward evaluation, with the tie rules
// of the dual numbers (a tie of min or max goes to the first operand, a NaN second operand loses), so this
// adjoint and the dual-number tangent agree at the kinks too (Davis speeds, HLLC upwind side, pressure
// floors, the HLLE rescue of a NaN HLLC flux).
//
// Forward definitions being reversed:
//   rs_hllc, rs_hlle, rs_roe, riemann, face_state  hydro_math.hpp (CPU_Shared_RiemannSolver_HLLC.cpp:391-655,
//                                                  _HLLE.cpp:301-698, _Roe.cpp:87-711, CPU_FluidSolver_MHM.cpp:737-786)
//   conserved cell -> interface state (cell_st)    hydro_adjoint.hip (Hydro_Con2Pres, FluUtility.cpp:1259-1300)
//   plm_face (limited slope, face clamp, mirror)    hydro_math.hpp (DataReconstruction.cpp:403-430, 1917-2032)
#pragma once
#include "hydro_dual.hpp"

namespace gamer_amd {

typedef St<double> StD;

__device__ __forceinline__ void st_zero( StD &b ) { b.d = b.mn = b.m1 = b.m2 = b.e = b.p = b.ir = b.cs = b.u = 0.0; }

// branch of the dual-number min(a, b) / max(a, b): true = the first operand
__device__ __forceinline__ bool min_first( double a, double b ) { return b != b || a <= b; }
__device__ __forceinline__ bool max_first( double a, double b ) { return b != b || a >= b; }

// F = x-flux(s) - sp*U(s)  (xflux_minus)
__device__ __forceinline__ void xflux_minus_rev( const StD &s, double sp, const double bF[NV], StD &bs, double &bsp )
{
   const double us = s.u - sp;
   bs.d  -= sp*bF[0];
   bs.mn += bF[0] + us*bF[1];
   bs.m1 += us*bF[2];
   bs.m2 += us*bF[3];
   bs.e  += us*bF[4];
   bs.p  += bF[1] + s.u*bF[4];
   bs.u  += s.mn*bF[1] + s.m1*bF[2] + s.m2*bF[3] + ( s.e + s.p )*bF[4];
   bsp   -= s.d*bF[0] + s.mn*bF[1] + s.m1*bF[2] + s.m2*bF[3] + s.e*bF[4];
}

__device__ __forceinline__ void add_upwind( StD &b, const StD &bs )
{
   b.d += bs.d;  b.mn += bs.mn;  b.m1 += bs.m1;  b.m2 += bs.m2;  b.e += bs.e;  b.p += bs.p;  b.u += bs.u;
}

// HLLE: sl = min(min(uL-cL, uR-cR), 0), sr = max(max(uL+cL, uR+cR), 0),
//       f = (sr (F_L - sl U_L) - sl (F_R - sr U_R)) / (sr - sl), or F_L - sl U_L when sl = sr = 0
__device__ __forceinline__ void rs_hlle_rev( const StD &L, const StD &Q, const double bf[NV], StD &bL, StD &bQ )
{
   const double a1 = L.u - L.cs, a2 = Q.u - Q.cs, b1 = L.u + L.cs, b2 = Q.u + Q.cs;
   const bool   k1 = min_first( a1, a2 ), k2 = max_first( b1, b2 );
   const double slr = k1 ? a1 : a2, srr = k2 ? b1 : b2;
   const bool   kl = min_first( slr, 0.0 ), kr = max_first( srr, 0.0 );
   const double sl = kl ? slr : 0.0, sr = kr ? srr : 0.0;
   double fl[NV], fr[NV], bfl[NV], bfr[NV], bsl = 0.0, bsr = 0.0;
   xflux_minus( L, sl, fl );
   xflux_minus( Q, sr, fr );
   if ( sl == 0.0 && sr == 0.0 )
   {
#     pragma unroll
      for (int v=0; v<NV; v++) { bfl[v] = bf[v];  bfr[v] = 0.0; }
   }
   else
   {
      const double inv = rcp( sr - sl );
      double binv = 0.0;
#     pragma unroll
      for (int v=0; v<NV; v++)
      {
         const double bD = bf[v]*inv;
         binv += bf[v]*( sr*fl[v] - sl*fr[v] );
         bsr  += bD*fl[v];
         bsl  -= bD*fr[v];
         bfl[v] =  bD*sr;
         bfr[v] = -bD*sl;
      }
      const double bdf = -inv*inv*binv;
      bsr += bdf;
      bsl -= bdf;
   }
   xflux_minus_rev( L, sl, bfl, bL, bsl );
   xflux_minus_rev( Q, sr, bfr, bQ, bsr );
   const double bslr = kl ? bsl : 0.0, bsrr = kr ? bsr : 0.0;
   if ( k1 ) { bL.u += bslr;  bL.cs -= bslr; } else { bQ.u += bslr;  bQ.cs -= bslr; }
   if ( k2 ) { bL.u += bsrr;  bL.cs += bsrr; } else { bQ.u += bsrr;  bQ.cs += bsrr; }
}

// HLLC (rs_hllc).  Returns false, with nothing accumulated, when the forward flux has a NaN component: the
// caller then differentiates the HLLE rescue instead.
__device__ __forceinline__ bool rs_hllc_rev( const StD &L, const StD &Q, double pmin, const double bf[NV], StD &bL, StD &bQ )
{
   // forward
   const double wl1 = L.u - L.cs, wl2 = Q.u - Q.cs, wr1 = L.u + L.cs, wr2 = Q.u + Q.cs;
   const bool   kl = min_first( wl1, wl2 ), kr = max_first( wr1, wr2 );
   const double wl = kl ? wl1 : wl2, wr = kr ? wr1 : wr2;
   const double du = L.u - Q.u;
   const bool   cl = ( wl1 < wl2 ), cr = ( wr2 > wr1 );
   const double XL = cl ? L.cs : du + Q.cs, XR = cr ? Q.cs : du + L.cs;
   const double tl = L.d*XL, tr = -Q.d*XR;
   const double t2 = rcp( tl - tr );
   const double A  = L.p - Q.p + tl*L.u - tr*Q.u;
   const double B  = tl*( Q.p + tr*Q.u ) - tr*( L.p + tl*L.u );
   const double vs = t2*A, psr = t2*B;
   const bool   pk = !( psr < pmin );
   const double ps = pk ? psr : pmin;
   const bool   left = ( vs >= 0.0 );
   const StD   &s = left ? L : Q;
   const bool   spk = left ? min_first( wl, 0.0 ) : max_first( wr, 0.0 );
   const double sp = spk ? ( left ? wl : wr ) : 0.0;
   double F[NV];
   xflux_minus( s, sp, F );
   const bool   deg = left && ( vs == 0.0 ) && ( sp == 0.0 );
   const double t4 = rcp( vs - sp );
   const double cf = deg ? 1.0 : t4*vs;
   const double cs = deg ? 0.0 : -t4*sp*ps;
   bool nan = false;
#  pragma unroll
   for (int v=0; v<NV; v++)
   {
      const double f = cf*F[v] + ( v == 1 ? cs : v == 4 ? cs*vs : 0.0 );
      nan |= ( f != f );
   }
   if ( nan ) return false;

   // reverse: f_k = cf F_k (+ cs, + cs vs)
   double bF[NV], bcf = 0.0;
#  pragma unroll
   for (int v=0; v<NV; v++) { bF[v] = cf*bf[v];  bcf += bf[v]*F[v]; }
   const double bcs = bf[1] + bf[4]*vs;
   double bvs = bf[4]*cs, bsp = 0.0, bps = 0.0;
   if ( !deg )
   {
      // cf = t4 vs, cs = (-t4) sp ps, t4 = 1/(vs - sp)
      const double bt4 = bcf*vs - bcs*sp*ps;
      bvs += bcf*t4;
      bsp -= bcs*t4*ps;
      bps -= bcs*t4*sp;
      const double bdf = -t4*t4*bt4;
      bvs += bdf;
      bsp -= bdf;
   }
   StD bs;
   st_zero( bs );
   xflux_minus_rev( s, sp, bF, bs, bsp );
   add_upwind( left ? bL : bQ, bs );
   const double bwl = ( spk &&  left ) ? bsp : 0.0;
   const double bwr = ( spk && !left ) ? bsp : 0.0;

   // ps = floor(psr), vs = t2 A, psr = t2 B
   const double bpsr = pk ? bps : 0.0;
   const double bt2 = bvs*A + bpsr*B, bA = bvs*t2, bB = bpsr*t2;
   double btl = 0.0, btr = 0.0;
   // A = L.p - Q.p + tl L.u - tr Q.u
   bL.p += bA;  bQ.p -= bA;
   btl  += bA*L.u;  bL.u += bA*tl;
   btr  -= bA*Q.u;  bQ.u -= bA*tr;
   // B = tl (Q.p + tr Q.u) - tr (L.p + tl L.u)
   btl  += bB*( Q.p + tr*Q.u ) - bB*tr*L.u;
   btr  += bB*tl*Q.u - bB*( L.p + tl*L.u );
   bQ.p += bB*tl;     bQ.u += bB*tl*tr;
   bL.p -= bB*tr;     bL.u -= bB*tr*tl;
   // t2 = 1/(tl - tr)
   const double bd2 = -t2*t2*bt2;
   btl += bd2;
   btr -= bd2;
   // tl = L.d XL, tr = (-Q.d) XR
   bL.d += btl*XL;
   bQ.d -= btr*XR;
   const double bXL = btl*L.d, bXR = -btr*Q.d;
   double bdu = 0.0;
   if ( cl ) bL.cs += bXL; else { bdu += bXL;  bQ.cs += bXL; }
   if ( cr ) bQ.cs += bXR; else { bdu += bXR;  bL.cs += bXR; }
   bL.u += bdu;
   bQ.u -= bdu;
   // wl = min(wl1, wl2), wr = max(wr1, wr2)
   if ( kl ) { bL.u += bwl;  bL.cs -= bwl; } else { bQ.u += bwl;  bQ.cs -= bwl; }
   if ( kr ) { bL.u += bwr;  bL.cs += bwr; } else { bQ.u += bwr;  bQ.cs += bwr; }
   return true;
}

// Roe solver with its positivity fallback to HLLE (rs_roe).  The Roe-averaged state, the eigenvalues, the wave
// amplitudes and the upwinded dissipation are recomputed in the forward's operation order, then reversed.
__device__ __forceinline__ void rs_roe_rev( const StD &L, const StD &Q, const Prm<double> &P, const double bf[NV],
                                            StD &bL, StD &bQ )
{
   const double HALF = 0.5, gm1 = P.gm1;
   // forward
   const double hl = L.ir*( L.e + L.p ), hr = Q.ir*( Q.e + Q.p );
   const double sql = vsqrt( L.d ), sqr = vsqrt( Q.d );
   const double rho = sql*sqr, irho = rcp( rho );
   const double isql = rcp( sql ), isqr = rcp( sqr ), isum = rcp( sql + sqr );
   const double NU = isql*L.mn + isqr*Q.mn, NVv = isql*L.m1 + isqr*Q.m1, NW = isql*L.m2 + isqr*Q.m2;
   const double u = isum*NU, v = isum*NVv, w = isum*NW;
   const double V2 = u*u + v*v + w*w;
   const double NH = sql*hl + sqr*hr, H = isum*NH;
   const double g0 = gm1*( H - HALF*V2 ), rg = rcp( P.gamma ), g1 = g0*rho*rg;
   const bool   kf = !( g1 < P.min_pres );
   const double gf = kf ? g1 : P.min_pres;
   const double a2 = P.gamma*irho*gf, a = vsqrt( a2 );
   const double e0 = u - a, e4 = u + a;

   double bu = 0.0, bv = 0.0, bw = 0.0, bH = 0.0, bV2 = 0.0, ba = 0.0;
   StD bs;
   double bsp = 0.0;
   if ( e0 >= 0.0 || e4 <= 0.0 )                 // supersonic: the upwind physical flux
   {
      st_zero( bs );
      xflux_minus_rev( ( e0 >= 0.0 ) ? L : Q, 0.0, bf, bs, bsp );
      add_upwind( ( e0 >= 0.0 ) ? bL : bQ, bs );
      return;
   }
   const double j0 = Q.d - L.d, j1 = Q.mn - L.mn, j2 = Q.m1 - L.m1, j3 = Q.m2 - L.m2, j4 = Q.e - L.e;
   const double a2_ = j2 - v*j0, a3_ = j3 - w*j0;
   const double c1 = gm1*rcp( a2 ), T1 = j0*( H - u*u ) + u*j1 - j4 + v*a2_ + w*a3_, a1_ = c1*T1;
   const double ra = rcp( a ), c0 = HALF*ra, T0 = j0*( u + a ) - j1 - a*a1_, a0_ = c0*T0;
   const double a4_ = j0 - a0_ - a1_;
   {
      const double d1 = L.d + a0_, m1 = L.mn + a0_*e0, t1 = L.m1 + a0_*v, t2 = L.m2 + a0_*w, en1 = L.e + a0_*( H - u*a );
      const double ie1 = en1 - HALF*( m1*m1 + t1*t1 + t2*t2 )*rcp( d1 );
      bool bad = ( u > e0 ) && ( d1 <= 0.0 || ie1*gm1 <= 0.0 );
      const double d4 = d1 + a1_, m4 = m1 + a1_*u, s4 = t1 + a1_*v + a2_, r4 = t2 + a1_*w + a3_,
                   en4 = en1 + a1_*( HALF*V2 ) + a2_*v + a3_*w;
      const double ie4 = en4 - HALF*( m4*m4 + s4*s4 + r4*r4 )*rcp( d4 );
      bad = bad || ( ( e4 > u ) && ( d4 <= 0.0 || ie4*gm1 <= 0.0 ) );
      if ( bad ) { rs_hlle_rev( L, Q, bf, bL, bQ );  return; }
   }
   const double s0 = __builtin_signbit( e0 ) ? -1.0 : 1.0, su = __builtin_signbit( u ) ? -1.0 : 1.0,
                s4 = __builtin_signbit( e4 ) ? -1.0 : 1.0;
   const double A0 = a0_*( s0*e0 ), A1 = a1_*( su*u ), A2 = a2_*( su*u ), A3 = a3_*( su*u ), A4 = a4_*( s4*e4 );

   // reverse: f = (F_L + F_R - sum_k A_k R_k) / 2
   double bh[NV];
#  pragma unroll
   for (int k=0; k<NV; k++) bh[k] = HALF*bf[k];
   st_zero( bs );
   xflux_minus_rev( L, 0.0, bh, bs, bsp );
   add_upwind( bL, bs );
   st_zero( bs );
   xflux_minus_rev( Q, 0.0, bh, bs, bsp );
   add_upwind( bQ, bs );
   const double bA0 = -( bh[0] + bh[1]*e0 + bh[2]*v + bh[3]*w + bh[4]*( H - u*a ) );
   const double bA1 = -( bh[0] + bh[1]*u  + bh[2]*v + bh[3]*w + bh[4]*( HALF*V2 ) );
   const double bA2 = -( bh[2] + bh[4]*v );
   const double bA3 = -( bh[3] + bh[4]*w );
   const double bA4 = -( bh[0] + bh[1]*e4 + bh[2]*v + bh[3]*w + bh[4]*( H + u*a ) );
   double be0 = -bh[1]*A0, be4 = -bh[1]*A4;
   bu  -= bh[1]*A1 + bh[4]*a*( A4 - A0 );
   bv  -= bh[2]*( A0 + A1 + A4 ) + bh[4]*A2;
   bw  -= bh[3]*( A0 + A1 + A4 ) + bh[4]*A3;
   bH  -= bh[4]*( A0 + A4 );
   ba  -= bh[4]*u*( A4 - A0 );
   bV2 -= bh[4]*HALF*A1;
   // A_k = alpha_k |lambda_k|
   double ba0_ = bA0*( s0*e0 ), ba1_ = bA1*( su*u ), ba2_ = bA2*( su*u ), ba3_ = bA3*( su*u ), ba4_ = bA4*( s4*e4 );
   be0 += bA0*a0_*s0;
   be4 += bA4*a4_*s4;
   bu  += ( bA1*a1_ + bA2*a2_ + bA3*a3_ )*su;
   // alpha_4 = j0 - alpha_0 - alpha_1
   double bj0 = ba4_, bj1 = 0.0, bj2 = 0.0, bj3 = 0.0, bj4 = 0.0;
   ba0_ -= ba4_;
   ba1_ -= ba4_;
   // alpha_0 = (1/(2a)) T0, T0 = j0 (u + a) - j1 - a alpha_1
   const double bT0 = ba0_*c0;
   ba  -= ra*ra*( HALF*ba0_*T0 );
   bj0 += bT0*( u + a );
   bu  += bT0*j0;
   ba  += bT0*( j0 - a1_ );
   bj1 -= bT0;
   ba1_ -= bT0*a;
   // alpha_1 = ((gamma-1)/a^2) T1
   const double bT1 = ba1_*c1, ra2 = rcp( a2 );
   double ba2 = -ra2*ra2*( gm1*ba1_*T1 );
   bj0 += bT1*( H - u*u );
   bH  += bT1*j0;
   bu  += bT1*( j1 - 2.0*u*j0 );
   bj1 += bT1*u;
   bj4 -= bT1;
   bv  += bT1*a2_;
   ba2_ += bT1*v;
   bw  += bT1*a3_;
   ba3_ += bT1*w;
   // alpha_2 = j2 - v j0, alpha_3 = j3 - w j0
   bj3 += ba3_;  bw -= ba3_*j0;  bj0 -= ba3_*w;
   bj2 += ba2_;  bv -= ba2_*j0;  bj0 -= ba2_*v;
   bQ.d  += bj0;  bL.d  -= bj0;
   bQ.mn += bj1;  bL.mn -= bj1;
   bQ.m1 += bj2;  bL.m1 -= bj2;
   bQ.m2 += bj3;  bL.m2 -= bj3;
   bQ.e  += bj4;  bL.e  -= bj4;
   // eigenvalues u -/+ a, a = sqrt(a2), a2 = (gamma/rho) max((gamma-1)(H - V2/2) rho/gamma, MinPres)
   bu += be0 + be4;
   ba += be4 - be0;
   ba2 += ba*( 0.5/a );
   const double birho = ba2*P.gamma*gf, bgf = ba2*P.gamma*irho;
   const double bg1 = kf ? bgf : 0.0;
   const double bg0 = bg1*rho*rg;
   double brho = bg1*g0*rg - irho*irho*birho;
   bH  += bg0*gm1;
   bV2 -= HALF*gm1*bg0;
   // Roe averages
   bu += 2.0*u*bV2;
   bv += 2.0*v*bV2;
   bw += 2.0*w*bV2;
   const double bNH = bH*isum, bNU = bu*isum, bNV = bv*isum, bNW = bw*isum;
   const double bisum = bH*NH + bu*NU + bv*NVv + bw*NW;
   const double bhl = bNH*sql, bhr = bNH*sqr;
   double bsql = bNH*hl, bsqr = bNH*hr;
   const double bisql = bNU*L.mn + bNV*L.m1 + bNW*L.m2, bisqr = bNU*Q.mn + bNV*Q.m1 + bNW*Q.m2;
   bL.mn += bNU*isql;  bL.m1 += bNV*isql;  bL.m2 += bNW*isql;
   bQ.mn += bNU*isqr;  bQ.m1 += bNV*isqr;  bQ.m2 += bNW*isqr;
   const double bss = -isum*isum*bisum;
   bsql += bss - isql*isql*bisql + brho*sqr;
   bsqr += bss - isqr*isqr*bisqr + brho*sql;
   bL.d += bsql*( 0.5/sql );
   bQ.d += bsqr*( 0.5/sqr );
   bL.ir += bhl*( L.e + L.p );  bL.e += bhl*L.ir;  bL.p += bhl*L.ir;
   bQ.ir += bhr*( Q.e + Q.p );  bQ.e += bhr*Q.ir;  bQ.p += bhr*Q.ir;
}

// riemann<RS, D> (HLLC with the HLLE rescue, HLLE, or Roe): bout is the adjoint of the lab-frame flux
template <int RS, int D>
__device__ __forceinline__ void riemann_rev( const StD &L, const StD &Q, const Prm<double> &P, const double bout[NV],
                                             StD &bL, StD &bQ )
{
   static_assert( RS == 2 || RS == 3 || RS == 4, "reverse sweeps exist for Roe, HLLE and HLLC" );
   double bf[NV];
   bf[0] = bout[0];
   bf[1] = bout[1 + D];
   bf[2] = bout[1 + ( D + 1 ) % 3];
   bf[3] = bout[1 + ( D + 2 ) % 3];
   bf[4] = bout[4];
   st_zero( bL );
   st_zero( bQ );
   if ( RS == 2 ) { rs_roe_rev( L, Q, P, bf, bL, bQ );  return; }
   if ( RS == 4 && rs_hllc_rev( L, Q, P.min_pres, bf, bL, bQ ) ) return;
   rs_hlle_rev( L, Q, bf, bL, bQ );
}

// face_state<D>: primitive face (rho, vx, vy, vz, P) -> interface state; bw accumulates
template <int D>
__device__ __forceinline__ void face_state_rev( const double w[NV], const StD &s, const StD &bs, const Prm<double> &P,
                                                double bw[NV] )
{
   constexpr int IN = 1 + D, I1 = 1 + ( D + 1 ) % 3, I2 = 1 + ( D + 2 ) % 3;
   const double K = fmad( w[3], w[3], fmad( w[2], w[2], w[1]*w[1] ) );
   // cs = sqrt(gamma p ir)
   const double bx = bs.cs*( 0.5/s.cs );
   const double bp = bs.p + bx*P.gamma*s.ir;
   const double bir = bs.ir + bx*P.gamma*s.p;
   // e = (0.5 rho) K + P/(gamma-1)
   const double bK = bs.e*0.5*w[0];
   bw[0]  += bs.d + bs.mn*w[IN] + bs.m1*w[I1] + bs.m2*w[I2] + bs.e*0.5*K - s.ir*s.ir*bir;
   bw[IN] += bs.mn*w[0] + bs.u;
   bw[I1] += bs.m1*w[0];
   bw[I2] += bs.m2*w[0];
   bw[1]  += 2.0*bK*w[1];
   bw[2]  += 2.0*bK*w[2];
   bw[3]  += 2.0*bK*w[3];
   bw[4]  += bs.e*P.inv_gm1 + ( ( w[4] < P.min_pres ) ? 0.0 : bp );
}

// conserved cell -> interface state: ir = 1/rho, p = max((gamma-1)(E - |m|^2 ir/2), MinPres),
// cs = sqrt(gamma p ir), u = ir m_n; bu accumulates
template <int D>
__device__ __forceinline__ void cell_st_rev( const double u[NV], const StD &s, const StD &bs, const Prm<double> &P,
                                             double bu[NV] )
{
   constexpr int IN = 1 + D, I1 = 1 + ( D + 1 ) % 3, I2 = 1 + ( D + 2 ) % 3;
   const double S = 0.5*( u[1]*u[1] + u[2]*u[2] + u[3]*u[3] );
   const double eint = u[4] - S*s.ir;
   const double bx = bs.cs*( 0.5/s.cs );
   const double bp = bs.p + bx*P.gamma*s.ir;
   double bir = bs.ir + bx*P.gamma*s.p + bs.u*s.mn;
   const double beint = ( eint*P.gm1 < P.min_pres ) ? 0.0 : bp*P.gm1;
   const double bS = -beint*s.ir;
   bir -= beint*S;
   bu[0]  += bs.d - s.ir*s.ir*bir;
   bu[IN] += bs.mn + bs.u*s.ir;
   bu[I1] += bs.m1;
   bu[I2] += bs.m2;
   bu[1]  += bS*u[1];
   bu[2]  += bS*u[2];
   bu[3]  += bS*u[3];
   bu[4]  += bs.e + beint;
}

// clamp_between(x, a, b) = min(max(x, min(a, b)), max(a, b)) under the dual tie rules: which operand the result
// is (0 = x, 1 = a, 2 = b)
__device__ __forceinline__ int clamp_pick( double x, double a, double b )
{
   const bool   lo_a = min_first( a, b ), hi_a = max_first( a, b );
   const double lo = lo_a ? a : b, hi = hi_a ? a : b;
   const bool   x1_x = max_first( x, lo );
   const double x1 = x1_x ? x : lo;
   if ( min_first( x1, hi ) ) return x1_x ? 0 : ( lo_a ? 1 : 2 );
   return hi_a ? 1 : 2;
}

// plm_face<LIM>: faces fL, fR of the centre C from (L, C, R).  Given their adjoints (bfL, bfR), accumulate the
// adjoints of L, C, R.  The face clamps and the mirror fL = 2C - fR are reversed here; the limited slope's two
// partials come from a two-direction dual evaluation of the same slope function.
template <int LIM>
__device__ __forceinline__ void plm_face_rev( double L, double C, double Rr, int lim, double coeff,
                                              double bfL, double bfR, double &bL, double &bC, double &bR )
{
   Dual<2> sl( C - L ), sr( Rr - C );
   sl.d[0] = 1.0;
   sr.d[1] = 1.0;
   const Dual<2> s = ( LIM == 4 ) ? slope_vl_gminmod( sl, sr, Dual<2>( coeff ) )
                                  : plm_slope<LIM>( sl, sr, Dual<2>( 0.0 ), lim, Dual<2>( coeff ) );
   // forward: fl = clamp(C - s/2, C, L), fr = clamp(2C - fl, C, R), fL = 2C - fr, fR = fr
   const double x1 = C - 0.5*s.v;
   const int    p1 = clamp_pick( x1, C, L );
   const double fl = ( p1 == 0 ) ? x1 : ( p1 == 1 ) ? C : L;
   const int    p2 = clamp_pick( 2.0*C - fl, C, Rr );
   // reverse
   const double bfr = bfR - bfL;
   double bc = 2.0*bfL, bl = 0.0, br = 0.0;
   double bx2 = 0.0, bx1 = 0.0;
   if ( p2 == 0 ) bx2 = bfr; else if ( p2 == 1 ) bc += bfr; else br += bfr;
   bc += 2.0*bx2;
   const double bfl = -bx2;
   if ( p1 == 0 ) bx1 = bfl; else if ( p1 == 1 ) bc += bfl; else bl += bfl;
   bc += bx1;
   const double bs = -0.5*bx1;
   const double bsl = bs*s.d[0], bsr = bs*s.d[1];
   bL += bl - bsl;
   bC += bc + bsl - bsr;
   bR += br + bsr;
}

} // namespace gamer_amd
