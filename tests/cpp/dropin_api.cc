// dropin_api.cc -- exercises the C++ drop-in API (dachshund_amd/cxx/include,
// the reference's lib/*.h interface) the way the reference's own tests and
// app use it, and writes the results as raw float64 for the Python parity
// test (tests/test_gpu_dropin.py) to compare with the oracle.
//
// Usage: dropin_api.exe <pixels.bin> <n> <out_dir>
// Outputs (float64):
//   w_prod.bin, w_res.bin, w_pc.bin   WFVectorW A_product / A_residual /
//                                     preconditioner_solve of vec = 0.1 i
//   sn_prod.bin                       WFVectorSN::A_product of the same vec
//   w_pcg.bin                         WFVectorW::solve_pcg x (tol 1e-8)
//   generic_pcg.bin                   pcg_solve over WFVectorW's callbacks,
//                                     then x = w .* y (tol 1e-8)
//   cf.bin                            WFVectorW::solve_cf(x, false, true)
//   smp.bin, nd.bin, mcov.bin         smp_product, sigma_m_noise_dom and
//                                     MapCovarDiag::solve_pcg on a 4x4x4 grid
//   scalars.bin                       norm_b, iters, success flags
#include <cmath>
#include <cstdio>
#include <string>
#include <vector>

#include "dachshund.h"

static void put(const std::string& path, const std::vector<double>& v) {
  FILE* f = fopen(path.c_str(), "wb");
  if (!f) {
    fprintf(stderr, "cannot write %s\n", path.c_str());
    exit(2);
  }
  fwrite(v.data(), sizeof(double), v.size(), f);
  fclose(f);
}

int main(int argc, char** argv) {
  if (argc != 4) {
    fprintf(stderr, "usage: %s pixels.bin n out_dir\n", argv[0]);
    return 2;
  }
  const int n = atoi(argv[2]);
  const std::string out = argv[3];
  std::vector<Pixel> pixels(n);
  read_pixel_data(argv[1], n, pixels.data());
  SignalCovarParams s_params(0.05, 2.5, 2.1);
  std::vector<double> vec(n), b(n), scalars;
  for (int i = 0; i < n; ++i) vec[i] = 0.1 * i - 3.0;

  // W form, in place like app/dachshund.cc:230-234
  std::vector<Pixel> wbuf = pixels;
  WPixel* wpix = (WPixel*)wbuf.data();
  pixel_to_wpixel(n, wbuf.data(), wpix);
  {
    WFVectorW wf(n, wpix, &s_params);
    wf.A_product(vec.data(), b.data());
    put(out + "/w_prod.bin", b);
    wf.A_residual(vec.data(), b.data());
    put(out + "/w_res.bin", b);
    wf.preconditioner_solve(vec.data(), b.data());
    put(out + "/w_pc.bin", b);
    scalars.push_back(wf.norm_b());

    PCGParams prm = {n, 1000, 50, 1e-8, false};
    std::vector<double> x(n, 0.0);
    PCGResult r = wf.solve_pcg(prm, x.data());
    put(out + "/w_pcg.bin", x);
    scalars.push_back(r.num_iters);
    scalars.push_back(r.success);

    // the generic solver over the class's bound operators (lib/pcg.h)
    std::vector<double> y(n, 0.0);
    PCGResult g = pcg_solve(prm, wf.norm_b(), wf.this_A_product,
                            wf.this_A_residual, wf.this_pc_solve, y.data());
    for (int i = 0; i < n; ++i) y[i] *= wpix[i].w;
    put(out + "/generic_pcg.bin", y);
    scalars.push_back(g.num_iters);
    scalars.push_back(g.success);

    std::vector<double> xc(n);
    wf.solve_cf(xc.data(), false, true);
    put(out + "/cf.bin", xc);
  }
  // SN form
  std::vector<Pixel> nbuf = pixels;
  NPixel* npix = (NPixel*)nbuf.data();
  pixel_to_npixel(n, nbuf.data(), npix);
  {
    WFVectorSN wf(n, npix, &s_params);
    wf.A_product(vec.data(), b.data());
    put(out + "/sn_prod.bin", b);
    scalars.push_back(wf.norm_b());
  }
  // map side on a 4x4x4 grid over the pixels' box
  double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
  for (int i = 0; i < n; ++i) {
    const double c[3] = {pixels[i].x, pixels[i].y, pixels[i].z};
    for (int a = 0; a < 3; ++a) {
      lo[a] = std::fmin(lo[a], c[a]);
      hi[a] = std::fmax(hi[a], c[a]);
    }
  }
  const int nm = 64;
  std::vector<Point> mc(nm);
  init_uniform_map_coords(hi[0], hi[1], hi[2], 4, 4, 4, mc.data());
  Point* pc = points_from_pixels_alloc(n, pixels.data());
  std::vector<double> m(nm);
  smp_product(n, pc, nm, mc.data(), &s_params, vec.data(), m.data());
  put(out + "/smp.bin", m);
  delete[] pc;
  sigma_m_noise_dom(n, npix, nm, mc.data(), &s_params, m.data());
  put(out + "/nd.bin", m);
  MapCovarDiag mcd(n, npix, 8, mc.data(), &s_params);
  PCGParams mprm = {n, 1000, 50, 1e-6, false};
  std::vector<double> sig(8);
  std::vector<PCGResult> rs = mcd.solve_pcg(mprm, sig.data());
  put(out + "/mcov.bin", sig);
  for (auto& r : rs) scalars.push_back(r.success);
  put(out + "/scalars.bin", scalars);
  printf("dropin_api: ok (total %.2f s)\n", total_timer.elapsed());
  return 0;
}
