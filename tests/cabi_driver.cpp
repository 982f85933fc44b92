// C++ caller of the C ABI with host buffers, in the calling pattern of the
// reference's single-shot operators (create + set medium once, then
// synchronous forward / adjoint calls on caller-owned host arrays):
// nonlinearPropElasticGpu_3D.cpp:66-615 and BornElasticGpu_3D.cpp:70-327.
//
// usage: cabi_driver <dir> <f32|f64>
// Reads a problem written by tests/test_cabi_driver.py into <dir> (grid.txt
// and raw little-endian files), runs the nonlinear forward and adjoint and the
// Born forward and adjoint, and writes their outputs as FP64 raw files.
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "elastic_iso_3d.h"

namespace {

template <typename T>
std::vector<T> read_all(const std::string& path) {
  std::ifstream f(path, std::ios::binary | std::ios::ate);
  if (!f) throw std::runtime_error("cannot open " + path);
  const std::streamsize bytes = f.tellg();
  f.seekg(0);
  std::vector<T> v(static_cast<size_t>(bytes) / sizeof(T));
  f.read(reinterpret_cast<char*>(v.data()), bytes);
  return v;
}

void write_all(const std::string& path, const std::vector<double>& v) {
  std::ofstream f(path, std::ios::binary);
  f.write(reinterpret_cast<const char*>(v.data()), static_cast<std::streamsize>(v.size() * sizeof(double)));
  if (!f) throw std::runtime_error("cannot write " + path);
}

void check(int rc, const char* what) {
  if (rc != ELA_OK) throw std::runtime_error(std::string(what) + ": " + ela_last_error());
}

// Components map to grids as vx->X vy->Y vz->Z sxx,syy,szz->centre sxz->XZ sxy->XY syz->YZ.
const int kCompGrid[9] = {1, 2, 3, 0, 0, 0, 4, 5, 6};

struct Points {
  std::vector<int64_t> flat;
  ela_points p{};
};

// File layout: int64 counts[7], then the positions of the 7 grids.
Points read_points(const std::string& path) {
  Points out;
  out.flat = read_all<int64_t>(path);
  int64_t off = 7;
  for (int g = 0; g < 7; ++g) out.p.n[g] = out.flat[g];
  for (int g = 0; g < 7; ++g) {
    out.p.pos[g] = out.flat.data() + off;
    off += out.p.n[g];
  }
  return out;
}

template <typename T>
struct Traces {
  std::vector<T> data;
  ela_traces t{};
  int64_t off[10]{};
  Traces(const ela_points& pts, int64_t nts) {
    off[0] = 0;
    for (int c = 0; c < 9; ++c) off[c + 1] = off[c] + pts.n[kCompGrid[c]] * nts;
    data.assign(static_cast<size_t>(off[9]), T(0));
    for (int c = 0; c < 9; ++c) t.tr[c] = data.data() + off[c];
  }
};

template <typename T>
std::vector<T> convert(const std::vector<double>& v) {
  return std::vector<T>(v.begin(), v.end());
}

template <typename T>
std::vector<double> widen(const std::vector<T>& v) {
  return std::vector<double>(v.begin(), v.end());
}

template <typename T>
void run(const std::string& dir, int precision) {
  std::ifstream gf(dir + "/grid.txt");
  ela_config cfg{};
  gf >> cfg.nz >> cfg.nx >> cfg.ny >> cfg.nts >> cfg.sub >> cfg.dz >> cfg.dx >> cfg.dy >> cfg.dts >> cfg.min_pad >>
      cfg.alpha_cos;
  if (!gf) throw std::runtime_error("bad grid.txt");
  cfg.precision = precision;
  cfg.device = 0;
  const int64_t n = cfg.nz * cfg.nx * cfg.ny;

  ela_handle* raw = nullptr;
  check(ela_create(&cfg, &raw), "ela_create");
  // the handle is destroyed on every exit path (as the reference's operator destructors deallocate)
  std::unique_ptr<ela_handle, int (*)(ela_handle*)> handle(raw, ela_destroy);
  ela_handle* h = handle.get();
  const std::vector<T> med = convert<T>(read_all<double>(dir + "/med.bin"));
  const std::vector<T> dm = convert<T>(read_all<double>(dir + "/dm.bin"));
  if (static_cast<int64_t>(med.size()) != 8 * n || static_cast<int64_t>(dm.size()) != 8 * n)
    throw std::runtime_error("medium / perturbation size");
  check(ela_set_medium(h, med.data(), ELA_MEM_HOST), "ela_set_medium");

  Points sp = read_points(dir + "/spos.bin"), rp = read_points(dir + "/rpos.bin");
  Traces<T> src(sp.p, cfg.nts), rec(rp.p, cfg.nts);
  src.data = convert<T>(read_all<double>(dir + "/src.bin"));
  rec.data = convert<T>(read_all<double>(dir + "/rec.bin"));
  if (static_cast<int64_t>(src.data.size()) != src.off[9] || static_cast<int64_t>(rec.data.size()) != rec.off[9])
    throw std::runtime_error("trace sizes");
  for (int c = 0; c < 9; ++c) {
    src.t.tr[c] = src.data.data() + src.off[c];
    rec.t.tr[c] = rec.data.data() + rec.off[c];
  }

  Traces<T> d(rp.p, cfg.nts), a(sp.p, cfg.nts), b(rp.p, cfg.nts);
  check(ela_nl_forward(h, &sp.p, &src.t, &rp.p, &d.t, ELA_MEM_HOST), "ela_nl_forward");
  check(ela_nl_adjoint(h, &sp.p, &a.t, &rp.p, &rec.t, ELA_MEM_HOST), "ela_nl_adjoint");
  check(ela_born_forward(h, &sp.p, &src.t, dm.data(), &rp.p, &b.t, ELA_MEM_HOST), "ela_born_forward");
  std::vector<T> img(static_cast<size_t>(8 * n), T(0));
  check(ela_born_adjoint(h, &sp.p, &src.t, img.data(), &rp.p, &rec.t, ELA_MEM_HOST), "ela_born_adjoint");
  check(ela_destroy(handle.release()), "ela_destroy");

  write_all(dir + "/out_nl_fwd.bin", widen(d.data));
  write_all(dir + "/out_nl_adj.bin", widen(a.data));
  write_all(dir + "/out_born_fwd.bin", widen(b.data));
  write_all(dir + "/out_born_adj.bin", widen(img));
}

}  // namespace

int main(int argc, char** argv) {
  if (argc != 3) {
    std::fprintf(stderr, "usage: %s <dir> <f32|f64>\n", argv[0]);
    return 2;
  }
  try {
    const std::string prec = argv[2];
    if (prec == "f32") run<float>(argv[1], ELA_FP32);
    else if (prec == "f64") run<double>(argv[1], ELA_FP64);
    else throw std::runtime_error("precision must be f32 or f64");
  } catch (const std::exception& e) {
    std::fprintf(stderr, "cabi_driver: %s\n", e.what());
    return 1;
  }
  std::printf("cabi_driver: %s done (%s)\n", argv[2], ela_version());
  return 0;
}
