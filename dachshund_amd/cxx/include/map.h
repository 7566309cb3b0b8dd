// map.h -- drop-in of lib/map.h: the map projection m = S_mp x on the GPU.
#ifndef DS_DROPIN_MAP_H
#define DS_DROPIN_MAP_H

#include <cmath>
#include <cstdio>
#include <string>
#include <vector>

#include "pixel.h"
#include "signal_covar.h"
#include "wf_vector.h"

// Declared by the reference (lib/map.h:18-39) but never defined there;
// kept so that code naming the type still compiles.
class Map {
 public:
  const int n_pix;
  const Pixel* const pixels;
  const int n_map;
  const Point* const map_coords;
  const SignalCovarParams* const s_params;
  const WFVector* const wf_vector;

  Map(const int n_pix, const Pixel* const pixels, const int n_map,
      const Point* const map_coords, const SignalCovarParams* const s_params,
      const WFVector* const wf_vector)
      : n_pix(n_pix), pixels(pixels), n_map(n_map), map_coords(map_coords),
        s_params(s_params), wf_vector(wf_vector) {}
};

// lib/map.cc:20-43: m_i = sum_j S(map_i, pix_j) x_j.
void smp_product(const int num_pixel_points, const Point* const pixel_coords,
                 const int num_map_points, const Point* const map_coords,
                 const SignalCovarParams* const s_params,
                 const double* const x, double* const m);

#endif
