// ref_golden.cc -- drives the reference library (compiled from its own
// sources under the reference tree into oracle/_ref/) to write golden
// vectors.  TEST INFRASTRUCTURE ONLY: run in the build container by
// tests/golden/make_golden.py; never shipped, never linked by the product.
//
// All files are raw little-endian float64 unless stated.  Usage:
//   ref_golden table  var_s l_perp l_para out
//   ref_golden covar  var_s l_perp l_para in(npairs x {xp2,xl2}) npairs out
//   ref_golden ops    form n var_s l_perp l_para pix vec out
//   ref_golden pcg    form n var_s l_perp l_para max_iter step_r tol pix x0 out
//   ref_golden cf     form n var_s l_perp l_para exact pix out
//   ref_golden smp    n var_s l_perp l_para pix x lx ly lz nx ny nz out
//   ref_golden nd     n var_s l_perp l_para pix lx ly lz nx ny nz out
//   ref_golden mcov   n var_s l_perp l_para pix max_iter step_r tol lx ly lz nx ny nz out
// form is "sn" or "w"; pix holds reference Pixel records {x,y,z,n,d}.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "dachshund.h"

static std::vector<double> read_f64(const char* path, size_t count) {
  std::vector<double> v(count);
  FILE* f = fopen(path, "rb");
  if (!f) { fprintf(stderr, "cannot open %s\n", path); exit(2); }
  size_t got = fread(v.data(), sizeof(double), count, f);
  fclose(f);
  if (got != count) { fprintf(stderr, "short read %s\n", path); exit(2); }
  return v;
}

static void write_f64(const char* path, const std::vector<double>& v) {
  FILE* f = fopen(path, "wb");
  if (!f) { fprintf(stderr, "cannot write %s\n", path); exit(2); }
  fwrite(v.data(), sizeof(double), v.size(), f);
  fclose(f);
}

int main(int argc, char** argv) {
  if (argc < 2) { fprintf(stderr, "usage: see source\n"); return 2; }
  const std::string mode = argv[1];
  if (mode == "table") {
    SignalCovarParams p(atof(argv[2]), atof(argv[3]), atof(argv[4]));
    std::vector<double> out;
    out.push_back(p.var_s);
    out.push_back(p.gaussian_const_perp);
    out.push_back(p.gaussian_const_para);
    out.push_back((double)p.func_table_n);
    out.push_back(p.func_table_dx_inv);
    for (int i = 0; i < p.func_table_n; ++i) out.push_back(p.func_table[i]);
    write_f64(argv[5], out);
    return 0;
  }
  if (mode == "covar") {
    SignalCovarParams p(atof(argv[2]), atof(argv[3]), atof(argv[4]));
    const size_t np = (size_t)atol(argv[6]);
    std::vector<double> in = read_f64(argv[5], 2 * np);
    std::vector<double> out(2 * np);
    for (size_t k = 0; k < np; ++k) {
      out[2 * k + 0] = signal_covar(in[2 * k], in[2 * k + 1], &p);
      out[2 * k + 1] = signal_covar_exact(in[2 * k], in[2 * k + 1], &p);
    }
    write_f64(argv[7], out);
    return 0;
  }
  const bool has_form = mode == "ops" || mode == "pcg" || mode == "cf";
  int a = 2;
  std::string form;
  if (has_form) form = argv[a++];
  const int n = atoi(argv[a++]);
  const double var_s = atof(argv[a++]);
  const double l_perp = atof(argv[a++]);
  const double l_para = atof(argv[a++]);
  SignalCovarParams p(var_s, l_perp, l_para);

  if (mode == "ops") {
    std::vector<double> pix = read_f64(argv[a++], 5 * (size_t)n);
    std::vector<double> vec = read_f64(argv[a++], n);
    const char* out_path = argv[a++];
    std::vector<double> prod(n), res(n), pc(n), out;
    double nb;
    if (form == "w") {
      pixel_to_wpixel(n, (Pixel*)pix.data(), (WPixel*)pix.data());
      WFVectorW wf(n, (const WPixel*)pix.data(), &p);
      wf.A_product(vec.data(), prod.data());
      wf.A_residual(vec.data(), res.data());
      wf.preconditioner_solve(vec.data(), pc.data());
      nb = wf.norm_b();
    } else {
      pixel_to_npixel(n, (Pixel*)pix.data(), (NPixel*)pix.data());
      WFVectorSN wf(n, (const NPixel*)pix.data(), &p);
      wf.A_product(vec.data(), prod.data());
      wf.A_residual(vec.data(), res.data());
      wf.preconditioner_solve(vec.data(), pc.data());
      nb = wf.norm_b();
    }
    out.insert(out.end(), prod.begin(), prod.end());
    out.insert(out.end(), res.begin(), res.end());
    out.insert(out.end(), pc.begin(), pc.end());
    out.push_back(nb);
    write_f64(out_path, out);
    return 0;
  }
  if (mode == "pcg") {
    PCGParams pp;
    pp.n = n;
    pp.max_iter = atoi(argv[a++]);
    pp.step_r = atoi(argv[a++]);
    pp.tol = atof(argv[a++]);
    pp.verbose = false;
    std::vector<double> pix = read_f64(argv[a++], 5 * (size_t)n);
    std::vector<double> x = read_f64(argv[a++], n);
    const char* out_path = argv[a++];
    PCGResult res;
    if (form == "w") {
      pixel_to_wpixel(n, (Pixel*)pix.data(), (WPixel*)pix.data());
      WFVectorW wf(n, (const WPixel*)pix.data(), &p);
      res = wf.solve_pcg(pp, x.data());
    } else {
      pixel_to_npixel(n, (Pixel*)pix.data(), (NPixel*)pix.data());
      WFVectorSN wf(n, (const NPixel*)pix.data(), &p);
      res = wf.solve_pcg(pp, x.data());
    }
    // layout: success, num_iters, n_norms, norms..., x...
    std::vector<double> out;
    out.push_back(res.success ? 1.0 : 0.0);
    out.push_back((double)res.num_iters);
    out.push_back((double)res.residual_norms.size());
    out.insert(out.end(), res.residual_norms.begin(), res.residual_norms.end());
    out.insert(out.end(), x.begin(), x.end());
    write_f64(out_path, out);
    return 0;
  }
  if (mode == "cf") {
    const bool exact = atoi(argv[a++]) != 0;
    std::vector<double> pix = read_f64(argv[a++], 5 * (size_t)n);
    const char* out_path = argv[a++];
    std::vector<double> x(n);
    if (form == "w") {
      pixel_to_wpixel(n, (Pixel*)pix.data(), (WPixel*)pix.data());
      WFVectorW wf(n, (const WPixel*)pix.data(), &p);
      wf.solve_cf(x.data(), false, exact);
    } else {
      pixel_to_npixel(n, (Pixel*)pix.data(), (NPixel*)pix.data());
      WFVectorSN wf(n, (const NPixel*)pix.data(), &p);
      wf.solve_cf(x.data(), false, exact);
    }
    write_f64(out_path, x);
    return 0;
  }
  // map-side modes share the grid arguments
  std::vector<double> pix = read_f64(argv[a++], 5 * (size_t)n);
  std::vector<double> x;
  if (mode == "smp") x = read_f64(argv[a++], n);
  int max_iter = 0, step_r = 0;
  double tol = 0.0;
  if (mode == "mcov") {
    max_iter = atoi(argv[a++]);
    step_r = atoi(argv[a++]);
    tol = atof(argv[a++]);
  }
  const double lx = atof(argv[a++]), ly = atof(argv[a++]), lz = atof(argv[a++]);
  const int nx = atoi(argv[a++]), ny = atoi(argv[a++]), nz = atoi(argv[a++]);
  const char* out_path = argv[a++];
  const int n_map = nx * ny * nz;
  std::vector<Point> map_coords(n_map);
  init_uniform_map_coords(lx, ly, lz, nx, ny, nz, map_coords.data());
  std::vector<double> m(n_map);
  if (mode == "smp") {
    Point* pts = points_from_pixels_alloc(n, (Pixel*)pix.data());
    smp_product(n, pts, n_map, map_coords.data(), &p, x.data(), m.data());
    delete[] pts;
  } else if (mode == "nd") {
    pixel_to_npixel(n, (Pixel*)pix.data(), (NPixel*)pix.data());
    sigma_m_noise_dom(n, (const NPixel*)pix.data(), n_map, map_coords.data(),
                      &p, m.data());
  } else if (mode == "mcov") {
    pixel_to_npixel(n, (Pixel*)pix.data(), (NPixel*)pix.data());
    MapCovarDiag mc(n, (const NPixel*)pix.data(), n_map, map_coords.data(), &p);
    PCGParams pp;
    pp.n = n; pp.max_iter = max_iter; pp.step_r = step_r; pp.tol = tol;
    pp.verbose = false;
    std::vector<PCGResult> rs = mc.solve_pcg(pp, m.data());
    for (size_t k = 0; k < rs.size(); ++k) {
      m.push_back(rs[k].success ? 1.0 : 0.0);
      m.push_back((double)rs[k].num_iters);
    }
  } else {
    fprintf(stderr, "unknown mode %s\n", mode.c_str());
    return 2;
  }
  write_f64(out_path, m);
  return 0;
}
