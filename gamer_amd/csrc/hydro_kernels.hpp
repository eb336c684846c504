// Internal launch interface between the C-ABI layer (hydro_api.hip) and the solver kernels.
#pragma once
#include <hip/hip_runtime.h>
#include "hydro_math.hpp"

namespace gamer_amd {

// MHM_RP patch-group update: in[npg][5][22^3] -> out[npg][5][16^3] (+ flux[npg][9][5][16^2] if store_flux)
template <typename R, typename IO>
int launch_mhm_rp( const IO *in, IO *out, IO *flux, int npg, int rsolver, int store_flux, int minmod_maxiter,
                   double minmod_coeff, const Prm<R> &P, hipStream_t stream );

// MHM patch-group update: in[npg][5][20^3] -> out[npg][5][16^3] (+ flux[npg][9][5][16^2] if store_flux)
template <typename R, typename IO>
int launch_mhm( const IO *in, IO *out, IO *flux, int npg, int rsolver, int store_flux, int minmod_maxiter,
                double minmod_coeff, const Prm<R> &P, hipStream_t stream );

} // namespace gamer_amd
