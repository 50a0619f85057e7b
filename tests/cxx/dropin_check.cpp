// Checks of the reference-compatible C++ layer that need no GPU, and a
// record reader front end for the GPU tests.
//   dropin_check selftest <dir>   file formats, line fit, JSON defaults
//   dropin_check rays <out> <bytes per value> <n_upper> <z...> -- <files...>
//       parse muon record files with MutoFile::getMuonData and write the
//       rays (112 bytes each) to <out>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

#include "MutoConfig.h"
#include "MutoFile.h"
#include "MutoMAP.h"
#include "MutoMLSD.h"
#include "MutoPoCA.h"

static int fails = 0;
#define CHECK(c)                                                           \
  do {                                                                     \
    if (!(c)) {                                                            \
      std::cerr << "check failed: " #c " (line " << __LINE__ << ")\n"; \
      ++fails;                                                             \
    }                                                                      \
  } while (0)

// test-only accessors of the protected configuration
struct PeekMLSD : MutoMLSD {
  explicit PeekMLSD(json j) : MutoMLSD(j) {}
  int iters() const { return fNIteration; }
  int pct() const { return fPercentileMean; }
  bool mean() const { return fUpdateWithMean; }
  double exy() const { return fErrorXY; }
  int init() const { return fInitMethod; }
  int prec() const { return fPrecision; }
};

static void selftest(const std::string& dir) {
  // image file round trips, FP32 and FP64, header fields
  VoxelGrid g{5, 4, 3, -10.0, -20.0, -30.0, 2.5, 3.0, 4.0};
  Image img(5, 4, 3);
  for (size_t j = 0; j < img.size(); ++j) img(j) = 1e-7 * (1.0 + 0.37 * j);
  CHECK(img(1, 2, 1) == img((size_t)(1 + 5 * (2 + 4 * 1))));
  for (unsigned prec : {4u, 8u}) {
    ImageHeader h;
    h.method = "mlsd";
    h.grid = g;
    h.nRay = 123456;
    h.nIter = 20;
    h.fPrecision = prec;
    const std::string f = dir + "/img" + std::to_string(prec) + ".bin";
    CHECK(MutoFile::saveImage(img, h, f));
    std::ifstream in(f.c_str(), std::ios::binary | std::ios::ate);
    CHECK((long long)in.tellg() == 64 + (long long)img.size() * prec);
    auto r = MutoFile::loadImage(f);
    CHECK(r.second.method == "mlsd");
    CHECK(r.second.nRay == 123456 && r.second.nIter == 20 && r.second.fPrecision == prec);
    CHECK(r.second.grid.nx == 5 && r.second.grid.ny == 4 && r.second.grid.nz == 3);
    CHECK(r.second.grid.x_min == -10.0 && r.second.grid.dz == 4.0);
    CHECK(r.first.size() == img.size());
    for (size_t j = 0; j < img.size(); ++j)
      CHECK(prec == 8 ? r.first(j) == img(j) : r.first(j) == (double)(float)img(j));
    const std::string fp = dir + "/pure" + std::to_string(prec) + ".bin";
    CHECK(MutoFile::saveImagePure(img, fp, prec));
    Image q = MutoFile::loadImagePure(fp, g, prec);
    CHECK(q.size() == img.size());
    CHECK(MutoFile::loadImagePure(fp, VoxelGrid{5, 4, 4, 0, 0, 0, 1, 1, 1}, prec).size() == 0);
  }
  CHECK(MutoFile::loadImage(dir + "/does_not_exist.bin").first.size() == 0);

  // line fit: points on a known line are fitted exactly; d.z > 0
  {
    const double p0[3] = {12.5, -7.25, 0.0}, d[3] = {0.3, -0.2, -0.9};
    const double z[3] = {1500.0, 1000.0, 500.0};
    double x[3], y[3];
    for (int i = 0; i < 3; ++i) {
      const double t = (z[i] - p0[2]) / d[2];
      x[i] = p0[0] + t * d[0];
      y[i] = p0[1] + t * d[1];
    }
    Vector3 pt, dir3;
    MutoFile::fitLine(x, y, z, 3, 500.0, pt, dir3);
    const double n = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    CHECK(dir3.z() > 0);
    CHECK(std::fabs(dir3.x() + d[0] / n) < 1e-12 && std::fabs(dir3.y() + d[1] / n) < 1e-12);
    CHECK(std::fabs(pt.x() - x[2]) < 1e-9 && std::fabs(pt.y() - y[2]) < 1e-9 && pt.z() == 500.0);
  }

  // record file: 6 layers, 3 upper, one record with a NaN skipped
  {
    const std::vector<MTfloat> zl = {1500, 1000, 500, -500, -1000, -1500};
    std::vector<float> rec;
    for (int r = 0; r < 3; ++r) {
      rec.push_back(3.0f + r);  // e_in
      for (int i = 0; i < 6; ++i) {
        rec.push_back((float)(0.01 * zl[i] + r));  // x
        rec.push_back((float)(-0.02 * zl[i]));     // y
      }
      rec.push_back(2.5f);  // e_out
    }
    rec[14 + 5] = NAN;  // record 1, one coordinate
    const std::string f = dir + "/muons.bin";
    std::ofstream o(f.c_str(), std::ios::binary);
    o.write(reinterpret_cast<const char*>(rec.data()), rec.size() * 4);
    o.close();
    MutoMuonData m = MutoFile::getMuonData({f, dir + "/missing.bin"}, zl, 3);
    CHECK(m.size() == 2);
    if (m.size() == 2) {
      CHECK(m[0].ein == 3.0 && m[1].ein == 5.0 && m[0].eout == 2.5);
      CHECK(std::fabs(m[0].pin.z() - 500.0) < 1e-12 && std::fabs(m[0].pout.z() + 500.0) < 1e-12);
      CHECK(m[0].din.z() > 0 && m[0].dout.z() > 0);
      CHECK(std::fabs(m[0].pin.x() - 5.0) < 1e-4 && std::fabs(m[0].pin.y() + 10.0) < 1e-4);
    }
  }

  // JSON: the reference's keys and defaults
  {
    json j = json::parse(R"({"grid": {"nx": 7, "dx": 2.0}, "num_iteration": 3,
                             "update_with_mean": true, "mean_percentile": 90})");
    PeekMLSD a(j);
    CHECK(a.getCurrentGrid().nx == 7 && a.getCurrentGrid().ny == 100 && a.getCurrentGrid().dx == 2.0);
    CHECK(a.iters() == 3 && a.pct() == 90 && a.mean() && a.exy() == 1.0 && a.init() == 0);
    CHECK(a.prec() == 0);
    PeekMLSD b(json::parse("{}"));
    CHECK(b.getCurrentGrid().nx == 100 && b.getCurrentGrid().x_min == 0.0 && b.iters() == 10);
    bool threw = false;
    try {
      PeekMLSD c(json::parse(R"({"precision": "fp16"})"));
    } catch (const std::invalid_argument&) {
      threw = true;
    }
    CHECK(threw);
    MutoPoCA p(json::parse(R"({"grid": {"nz": 9}})"));
    CHECK(p.getCurrentGrid().nz == 9 && p.getCurrentGrid().nx == 100);
    MutoMAP q(json::parse(R"({"beta": 2.0})"));
    CHECK(q.getCurrentGrid().nx == 100);
  }
}

int main(int argc, char** argv) {
  if (argc >= 3 && std::string(argv[1]) == "selftest") {
    selftest(argv[2]);
    if (fails) return 1;
    std::cout << "dropin selftest ok" << std::endl;
    return 0;
  }
  if (argc >= 7 && std::string(argv[1]) == "rays") {
    const std::string out = argv[2];
    const int prec = std::atoi(argv[3]);
    const int nmid = std::atoi(argv[4]);
    std::vector<MTfloat> z;
    int k = 5;
    for (; k < argc && std::string(argv[k]) != "--"; ++k) z.push_back(std::atof(argv[k]));
    std::vector<std::string> files;
    for (++k; k < argc; ++k) files.push_back(argv[k]);
    MutoMuonData m = MutoFile::getMuonData(files, z, nmid, prec);
    std::ofstream o(out.c_str(), std::ios::binary);
    o.write(reinterpret_cast<const char*>(m.data()), (std::streamsize)(m.size() * sizeof(RayData)));
    std::cout << m.size() << std::endl;
    return o.good() ? 0 : 1;
  }
  std::cerr << "usage: dropin_check selftest <dir> | rays <out> <prec> <n_upper> <z...> -- <files...>\n";
  return 2;
}
