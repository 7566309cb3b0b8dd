// map_covar.h -- drop-in of lib/map_covar.h: map-covariance diagonals.
#ifndef DS_DROPIN_MAP_COVAR_H
#define DS_DROPIN_MAP_COVAR_H

#include <vector>

#include "pcg.h"
#include "pixel.h"
#include "signal_covar.h"

// lib/map_covar.cc:11-38: sigma_i = sum_j S_ij^2 / N_j (N >> S limit).
void sigma_m_noise_dom(const int n_pix, const NPixel* pixels, const int n_map,
                       const Point* map_coords,
                       const SignalCovarParams* s_params, double* sigma_m);

// lib/map_covar.h:22-40: sigma_i = S_mp (S + N)^{-1} S_pm, one SN-form PCG
// solve per map cell with the column S(map_i, .) as right-hand side.
class MapCovarDiag {
 public:
  const int n_pix;
  const NPixel* const pixels;
  const int n_map;
  const Point* const map_coords;
  const SignalCovarParams* const s_params;

  MapCovarDiag(const int n_pix, const NPixel* const pixels, const int n_map,
               const Point* const map_coords,
               const SignalCovarParams* const s_params)
      : n_pix(n_pix), pixels(pixels), n_map(n_map), map_coords(map_coords),
        s_params(s_params) {}

  PCGResult solve_column_pcg(const int i, const PCGParams& pcg_params,
                             double* const x);
  std::vector<PCGResult> solve_pcg(const PCGParams& pcg_params,
                                   double* const sigma_m);
};

#endif
