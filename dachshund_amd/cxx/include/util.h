// util.h -- drop-in of lib/util.h: pixel-file input and the map grid.
#ifndef DS_DROPIN_UTIL_H
#define DS_DROPIN_UTIL_H

#include <string>

#include "map.h"

// lib/util.cc:11-23: read num_pixels raw Pixel records; exit(1) if the file
// cannot be opened.
void read_pixel_data(const std::string path, const int num_pixels,
                     Pixel* const pixels);

// lib/util.cc:25-46: cell centres (i + 1/2) L / N, index (ix*ny + iy)*nz + iz.
void init_uniform_map_coords(const double lx, const double ly, const double lz,
                             const int nx, const int ny, const int nz,
                             Point* const coords);

// lib/util.cc:48-57: report non-finite entries.
void check_finite(const int n, const double* const v);

#endif
