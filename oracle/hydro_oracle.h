// Oracle (test infrastructure only): CPU restatement of GAMER's MHM / MHM_RP hydro patch-group solver.
// Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it, and only as the checker.
#pragma once

#ifdef __cplusplus
extern "C" {
#endif

// scheme ids follow include/Macro.h:74-75, Riemann-solver ids include/Macro.h:86-88,
// limiter ids include/Typedef.h:188-193
enum { ORC_MHM = 3, ORC_MHM_RP = 4 };
enum { ORC_ROE = 2, ORC_HLLE = 3, ORC_HLLC = 4 };

// One solver call on npg patch groups (layouts as CPU_FluidSolver_MHM: in[npg][5][N^3], N = 22 (MHM_RP) or
// 20 (MHM); out[npg][5][16^3]; flux[npg][9][5][16^2]).  Returns 0 on success.
int oracle_fluid_solver_f64( const double *in, double *out, double *flux, int npg, double dt, double dh,
                             double gamma, int scheme, int rsolver, int store_flux, int limiter,
                             double minmod_coeff, int minmod_maxiter, double min_dens, double min_pres,
                             double min_eint, int nthreads );
int oracle_fluid_solver_f32( const float *in, float *out, float *flux, int npg, double dt, double dh,
                             double gamma, int scheme, int rsolver, int store_flux, int limiter,
                             double minmod_coeff, int minmod_maxiter, double min_dens, double min_pres,
                             double min_eint, int nthreads );

// Single-interface Riemann solve (conserved L/R states, direction d) and single-stencil slope limiter,
// exposed so tests can pin them against closed forms.
int oracle_riemann_f64( int rsolver, int d, const double L[5], const double R[5], double gamma,
                        double min_pres, double flux[5] );
int oracle_limit_slope_f64( const double L[5], const double C[5], const double R[5], int limiter,
                            double minmod_coeff, double slope[5] );

// Per-patch CFL time step (CPU_dtSolver_HydroCFL): in[npatch][5][8^3] -> dt[npatch]
int oracle_dt_cfl_f64( const double *in, double *dt, int npatch, double dh, double safety, double gamma,
                       double min_pres, int nthreads );

#ifdef __cplusplus
}
#endif
