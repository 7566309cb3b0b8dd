// pixel.h -- drop-in of lib/pixel.h: the reference's pixel records
// (standard layout, field order kept: they are the on-disk format and the
// C ABI's record layout) and their in-place-safe conversions.
#ifndef DS_DROPIN_PIXEL_H
#define DS_DROPIN_PIXEL_H

struct Point {
  double x;
  double y;
  double z;
};

// lib/pixel.h:16-22: raw pixel, noise amplitude n and datum d.
struct Pixel {
  double x;
  double y;
  double z;
  double n;
  double d;
};

// lib/pixel.h:27-33: N = n^2.
struct NPixel {
  double x;
  double y;
  double z;
  double N;
  double d;
};

// lib/pixel.h:39-45: w = 1/n, wd = d/n.
struct WPixel {
  double x;
  double y;
  double z;
  double w;
  double wd;
};

static_assert(sizeof(Point) == 24 && sizeof(Pixel) == 40 &&
                  sizeof(NPixel) == 40 && sizeof(WPixel) == 40,
              "pixel records must keep the reference's layout");

// lib/pixel.h:36-50.  Each record is read completely before it is
// written, so every conversion may run in place (p == np).
Point* points_from_pixels_alloc(const int n, Pixel* const pixels);
void pixel_to_npixel(const int n, Pixel* const p, NPixel* const np);
void npixel_to_pixel(const int n, NPixel* const np, Pixel* const p);
void pixel_to_wpixel(const int n, Pixel* const p, WPixel* const wp);
void wpixel_to_pixel(const int n, WPixel* const wp, Pixel* const p);

#endif
