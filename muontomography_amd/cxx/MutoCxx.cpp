// The reference's reconstruction classes (MutoPoCA, MutoPoCut, MutoSLR,
// MutoMLSD, MutoMAP) and
// file I/O (MutoFile) over the C ABI of libmuto: JSON configuration on the
// host, every reconstruction on the GPU.
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iostream>
#include <random>
#include <stdexcept>

#include "MutoFile.h"
#include "MutoMAP.h"
#include "MutoMLSD.h"
#include "MutoPoCA.h"
#include "MutoPoCut.h"
#include "MutoSLR.h"
#include "muto.h"

static_assert(sizeof(VoxelGrid) == sizeof(muto_grid), "VoxelGrid and muto_grid layouts differ");
static_assert(sizeof(RayData) == sizeof(muto_ray), "RayData and muto_ray layouts differ");
static_assert(sizeof(Vector3) == 3 * sizeof(double), "Vector3 must be three doubles");

namespace {

muto_ctx* open_ctx(muto_ctx*& ctx, int device) {
  if (!ctx && muto_ctx_create(device, &ctx) != MUTO_OK) {
    ctx = nullptr;
    throw std::runtime_error("muto: cannot open gfx950 device " + std::to_string(device));
  }
  return ctx;
}

void check(muto_ctx* ctx, int rc, const char* what) {
  if (rc != MUTO_OK) throw std::runtime_error(std::string(what) + ": " + muto_last_error(ctx));
}

muto_grid to_grid(const VoxelGrid& g) {
  muto_grid o;
  std::memcpy(&o, &g, sizeof o);
  o._pad = 0;
  return o;
}

// the "grid" object (src/MutoPoCA.cpp:27-37, src/MutoMLSD.cpp:50-60)
void read_grid(const json& config, VoxelGrid& g) {
  try {
    json grid = config.at("grid");
    g.nx = grid.value("nx", 100);
    g.ny = grid.value("ny", 100);
    g.nz = grid.value("nz", 100);
    g.x_min = grid.value("x_min", 0.0);
    g.y_min = grid.value("y_min", 0.0);
    g.z_min = grid.value("z_min", 0.0);
    g.dx = grid.value("dx", 10.0);
    g.dy = grid.value("dy", 10.0);
    g.dz = grid.value("dz", 10.0);
  } catch (json::out_of_range&) {
  }
}

void default_grid(VoxelGrid& g) {
  g.nx = g.ny = g.nz = 100;
  g.x_min = g.y_min = g.z_min = 0.0;
  g.dx = g.dy = g.dz = 10.0;
}

const muto_ray* as_rays(const MutoMuonData& d) { return reinterpret_cast<const muto_ray*>(d.data()); }

}  // namespace

// ---------------------------------------------------------------- PoCA
MutoPoCA::MutoPoCA()
    : fLengthWeighted(false), fStraightLine(false), fHasEnergy(false), fLogPrint(false),
      fNTotal(0), fNNotParallel(0), fNScatInGrid(0), fAngMag(1.0), fEPS(1.0e-8), fDevice(0),
      fCtx(nullptr) {
  default_grid(fGrid);
}

MutoPoCA::MutoPoCA(json config) : MutoPoCA() {
  read_grid(config, fGrid);
  fLengthWeighted = config.value("length_weighted", false);
  fStraightLine = config.value("straight_line", false);
  fHasEnergy = config.value("has_energy", false);
  fLogPrint = config.value("log_results", false);
  fNTotal = config.value("total_num_rays", 0);
  fAngMag = config.value("angle_unit_mag", 1.0);
  fEPS = config.value("eps", 1.0e-8);
  fDevice = config.value("device", 0);
}

MutoPoCA::~MutoPoCA() {
  if (fCtx) muto_ctx_destroy(fCtx);
}

Image MutoPoCA::reconstruct(const MutoMuonData& rays) {
  open_ctx(fCtx, fDevice);
  muto_poca_cfg c{};
  c.grid = to_grid(fGrid);
  c.length_weighted = fLengthWeighted;
  c.straight_line = fStraightLine;
  c.has_energy = fHasEnergy;
  c.log_results = 0;
  c.total_num_rays = fNTotal;
  c.angle_unit_mag = fAngMag;
  c.eps = fEPS;
  Image img(fGrid.nx, fGrid.ny, fGrid.nz);
  muto_counters cnt{};
  check(fCtx, muto_poca_reconstruct(fCtx, &c, rays.size(), as_rays(rays), img.data(), &cnt),
        "MutoPoCA::reconstruct");
  fNTotal = (int)cnt.n_total;
  fNNotParallel = (int)cnt.n_not_parallel;
  fNScatInGrid = (int)cnt.n_scat_in_grid;
  if (fLogPrint) logResults();
  return img;
}

void MutoPoCA::logResults() {  // src/MutoPoCA.cpp:176-183
  std::cout << "==================================================" << std::endl;
  std::cout << "|| PoCA reconstruction results                  ||" << std::endl;
  std::cout << "==================================================" << std::endl;
  std::cout << "number of rays processed: " << fNTotal << std::endl;
  std::cout << "number of unparalleled rays: " << fNNotParallel << std::endl;
  std::cout << "number of interactions in voxels: " << fNScatInGrid << std::endl;
}

// ---------------------------------------------------------------- PoCut
MutoPoCut::MutoPoCut() : MutoPoCA(), fPercentile(1.0) { fStraightLine = false; }

MutoPoCut::MutoPoCut(json config) : MutoPoCA(config) {
  fStraightLine = false;  // src/MutoPoCut.cpp:13
  fPercentile = config.value("cut_percentile", 1.0);
}

MutoPoCut::~MutoPoCut() {}

Image MutoPoCut::reconstruct(const MutoMuonData& rays) {
  std::cout << "==================================================" << std::endl;
  std::cout << "|| PoCA with threshold for scattering angles    ||" << std::endl;
  std::cout << "==================================================" << std::endl;
  open_ctx(fCtx, fDevice);
  muto_poca_cfg c{};
  c.grid = to_grid(fGrid);
  c.length_weighted = fLengthWeighted;
  c.straight_line = 0;
  c.has_energy = fHasEnergy;
  c.log_results = 0;
  c.total_num_rays = fNTotal;
  c.angle_unit_mag = fAngMag;
  c.eps = fEPS;
  Image img(fGrid.nx, fGrid.ny, fGrid.nz);
  muto_counters cnt{};
  check(fCtx, muto_pocut_reconstruct(fCtx, &c, fPercentile, rays.size(), as_rays(rays), img.data(), &cnt),
        "MutoPoCut::reconstruct");
  fNTotal = (int)cnt.n_total;
  fNNotParallel = (int)cnt.n_not_parallel;
  fNScatInGrid = (int)cnt.n_scat_in_grid;
  if (fLogPrint) logResults();
  return img;
}

// ---------------------------------------------------------------- SLR
MutoSLR::MutoSLR()
    : fLengthWeighted(false), fHasEnergy(false), fLogPrint(false), fNTotal(0), fAngMag(1.0),
      fEPS(1.0e-8), fDevice(0), fCtx(nullptr) {
  default_grid(fGrid);
}

MutoSLR::MutoSLR(json config) : MutoSLR() {  // src/MutoSLR.cpp:21-43
  read_grid(config, fGrid);
  fLengthWeighted = config.value("length_weighted", false);
  fHasEnergy = config.value("has_energy", false);
  fLogPrint = config.value("log_results", false);
  fNTotal = config.value("total_num_rays", 0);
  fAngMag = config.value("angle_unit_mag", 1.0);
  fEPS = config.value("eps", 1.0e-8);
  fDevice = config.value("device", 0);
}

MutoSLR::~MutoSLR() {
  if (fCtx) muto_ctx_destroy(fCtx);
}

Image MutoSLR::reconstruct(const MutoMuonData& rays) {
  open_ctx(fCtx, fDevice);
  muto_poca_cfg c{};
  c.grid = to_grid(fGrid);
  c.length_weighted = fLengthWeighted;
  c.straight_line = 1;
  c.has_energy = fHasEnergy;
  c.log_results = 0;
  c.total_num_rays = fNTotal;
  c.angle_unit_mag = fAngMag;
  c.eps = fEPS;
  Image img(fGrid.nx, fGrid.ny, fGrid.nz);
  muto_counters cnt{};
  check(fCtx, muto_slr_reconstruct(fCtx, &c, rays.size(), as_rays(rays), img.data(), &cnt),
        "MutoSLR::reconstruct");
  fNTotal = (int)cnt.n_total;
  if (fLogPrint) logResults();
  return img;
}

void MutoSLR::logResults() {  // src/MutoSLR.cpp:111-116
  std::cout << "==================================================" << std::endl;
  std::cout << "|| SLR reconstruction results                  ||" << std::endl;
  std::cout << "==================================================" << std::endl;
  std::cout << "number of rays processed: " << fNTotal << std::endl;
}

// ---------------------------------------------------------------- MLSD
MutoMLSD::MutoMLSD()
    : fUpdateWithMean(false), fPercentileMean(100), fStraightLine(false), fEarlyStop(false),
      fHasEnergy(false), fLogPrint(false), fNTotal(0), fNScatInGrid(0), fNIteration(10),
      fInitMethod(0), fInitValue(1.0e-7), fInitImageFile(""), fErrorXY(1.0), fErrorZo(1000.0),
      fErrorZi(1000.0), fAngMag(1.0), fEPS(1.0e-8), fDevice(0), fPrecision(MUTO_FP64),
      fCtx(nullptr) {
  default_grid(fGrid);
}

MutoMLSD::MutoMLSD(json config) : MutoMLSD() {
  read_grid(config, fGrid);
  fUpdateWithMean = config.value("update_with_mean", false);
  fStraightLine = config.value("straight_line", false);
  fHasEnergy = config.value("has_energy", false);
  fEarlyStop = config.value("early_stop", false);
  fLogPrint = config.value("log_results", false);
  fNTotal = config.value("total_num_rays", 0);
  fPercentileMean = config.value("mean_percentile", 100);
  fNIteration = config.value("num_iteration", 10);
  fInitMethod = config.value("initialize_method", 0);
  fInitValue = config.value("initialize_value", 1.0e-7);
  fInitImageFile = config.value("initial_image_file", "");
  fErrorXY = config.value("error_xy", 1.0);
  fErrorZo = config.value("error_dz_outer", 1000.0);
  fErrorZi = config.value("error_dz_inner", 1000.0);
  fAngMag = config.value("angle_unit_mag", 1.0);
  fEPS = config.value("eps", 1.0e-8);
  fDevice = config.value("device", 0);
  const std::string prec = config.value("precision", "fp64");
  if (prec == "fp64")
    fPrecision = MUTO_FP64;
  else if (prec == "fp64_stream")
    fPrecision = MUTO_FP64_STREAM;
  else if (prec == "fp32c")
    fPrecision = MUTO_FP32C;
  else
    throw std::invalid_argument(
        "MutoMLSD: \"precision\" must be \"fp64\", \"fp64_stream\" or \"fp32c\"");
}

MutoMLSD::~MutoMLSD() {
  if (fCtx) muto_ctx_destroy(fCtx);
}

void MutoMLSD::logMessage(const std::string& msg, bool printTime) {  // src/MutoMLSD.cpp:524-533
  if (!fLogPrint) return;
  std::cout << msg;
  if (printTime) {
    const double t = std::chrono::duration_cast<std::chrono::duration<double>>(
                         std::chrono::steady_clock::now() - fTimeStart).count();
    std::cout << " Time: " << t << " seconds. ";
  }
  std::cout << std::endl;
}

void MutoMLSD::logResults() {  // src/MutoMLSD.cpp:535-543
  std::cout << std::endl;
  std::cout << "==================================================" << std::endl;
  std::cout << "|| MLSD reconstruction results                  ||" << std::endl;
  std::cout << "==================================================" << std::endl;
  std::cout << "number of rays processed: " << fNTotal << std::endl;
  std::cout << "number of interactions in voxels: " << fNScatInGrid << std::endl;
  std::cout << "number of iterations: " << fNIteration << std::endl;
}

// src/MutoMLSD.cpp:226-266: file (adopting the file's grid when its
// dimensions differ), N(2, 0.5) 1e-7 floored at 1e-10 from a random_device
// seed (not reproducible, as in the reference), or a constant
Image MutoMLSD::initailizeImage() {
  if (fInitMethod == 0) {
    auto ih = MutoFile::loadImage(fInitImageFile);
    const VoxelGrid in = ih.second.grid;
    if (ih.first.size() == 0) {
      std::cout << "Warning: Initial image file <" << fInitImageFile << "> not exist. " << std::endl;
      std::cout << "\tSetting initialization method to random numbers ..." << std::endl;
      fInitMethod = 1;
    } else if (in.nx != fGrid.nx || in.ny != fGrid.ny || in.nz != fGrid.nz) {
      std::cout << "Warning: Initial image's dimensions do not match the configuration. " << std::endl;
      std::cout << "\tResetting grid configuration as defined in input image  ..." << std::endl;
      fGrid = in;
      return ih.first;
    } else {
      return ih.first;
    }
  }
  Image img(fGrid.nx, fGrid.ny, fGrid.nz);
  if (fInitMethod == 1) {
    std::random_device rdev;
    std::default_random_engine gen{rdev()};
    std::normal_distribution<double> dist(2.0, 0.5);
    for (size_t i = 0; i < img.size(); ++i) {
      const MTfloat v = dist(gen) * 1.0e-7;
      img(i) = 1.0e-10 <= v ? v : 1.0e-10;
    }
  } else {
    img.setConstant(fInitValue);
  }
  return img;
}

void MutoMLSD::fillConfig(muto_mlsd_cfg& c) const {
  c = muto_mlsd_cfg{};
  c.grid = to_grid(fGrid);
  c.update_with_mean = fUpdateWithMean;
  c.straight_line = fStraightLine;
  c.has_energy = fHasEnergy;
  c.early_stop = fEarlyStop;
  c.log_results = fLogPrint;
  c.mean_percentile = fPercentileMean;
  c.num_iteration = fNIteration;
  c.initialize_method = 0;  // the image is prepared here
  c.total_num_rays = fNTotal;
  c.initialize_value = fInitValue;
  c.error_xy = fErrorXY;
  c.error_dz_outer = fErrorZo;
  c.error_dz_inner = fErrorZi;
  c.angle_unit_mag = fAngMag;
  c.eps = fEPS;
  c.beta = 1.0;
  c.p = 1.0;
  c.delta = 0.1;
  c.precision = fPrecision;
  fillPrior(c);
}

Image MutoMLSD::run(const MutoMuonData& rays, bool map) {
  fTimeStart = std::chrono::steady_clock::now();
  open_ctx(fCtx, fDevice);
  Image init = initailizeImage();
  logMessage("Initial image was loaded / generated.", true);
  muto_mlsd_cfg c;
  fillConfig(c);
  Image img(fGrid.nx, fGrid.ny, fGrid.nz);
  fCost.assign(fNIteration > 0 ? fNIteration : 1, NAN);
  muto_counters cnt{};
  const int rc = map ? muto_map_reconstruct(fCtx, &c, rays.size(), as_rays(rays), init.data(),
                                            img.data(), fCost.data(), &cnt)
                     : muto_mlsd_reconstruct(fCtx, &c, rays.size(), as_rays(rays), init.data(),
                                             img.data(), fCost.data(), &cnt);
  check(fCtx, rc, map ? "MutoMAP::reconstruct" : "MutoMLSD::reconstruct");
  fCost.resize((size_t)cnt.iterations_done);
  fNTotal = (int)cnt.n_total;
  fNScatInGrid = (int)cnt.n_scat_in_grid;
  logMessage("EM iterations finished.", true);
  if (fLogPrint) logResults();
  return img;
}

Image MutoMLSD::reconstruct(const MutoMuonData& rays) {
  logMessage("==================================================");
  logMessage("|| MLSD reconstruction start                    ||");
  logMessage("==================================================");
  return run(rays, false);
}

// ----------------------------------------------------------------- MAP
MutoMAP::MutoMAP() : MutoMLSD() { defaultInitialize(); }

MutoMAP::MutoMAP(json config) : MutoMLSD(config) {
  defaultInitialize();
  fBeta = config.value("beta", 1.0);
  fP = config.value("p", 1.0);
  fDelta = config.value("delta", 0.1);
  fAddObjectPrior = config.value("add_object_prior", false);
  fUseAllNeighbours = config.value("all_neighbours", false);
}

MutoMAP::~MutoMAP() {}

void MutoMAP::defaultInitialize() {  // src/MutoMAP.cpp:22-34
  fBeta = 1.0;
  fP = 1.0;
  fDelta = 0.1;
  fUseAllNeighbours = false;
  fAddObjectPrior = false;
}

void MutoMAP::fillPrior(muto_mlsd_cfg& c) const {
  c.beta = fBeta;
  c.p = fP;
  c.delta = fDelta;
  c.add_object_prior = fAddObjectPrior;
  c.all_neighbours = fUseAllNeighbours;
}

Image MutoMAP::reconstruct(const MutoMuonData& rays) {
  logMessage("==================================================");
  logMessage("|| MLSD with regularization start               ||");
  logMessage("==================================================");
  return run(rays, true);
}

// ---------------------------------------------------------------- files
namespace {

bool write_values(std::ostream& f, const Image& img, unsigned prec) {
  if (prec == 8) {
    f.write(reinterpret_cast<const char*>(img.data()), (std::streamsize)(img.size() * 8));
  } else if (prec == 4) {
    std::vector<float> v(img.size());
    for (size_t i = 0; i < v.size(); ++i) v[i] = (float)img(i);
    f.write(reinterpret_cast<const char*>(v.data()), (std::streamsize)(v.size() * 4));
  } else {
    std::cout << "Floating point precision not supported. " << std::endl;
    return false;
  }
  return f.good();
}

bool read_values(std::istream& f, Image& img, unsigned prec) {
  if (prec == 8) {
    f.read(reinterpret_cast<char*>(img.data()), (std::streamsize)(img.size() * 8));
  } else if (prec == 4) {
    std::vector<float> v(img.size());
    f.read(reinterpret_cast<char*>(v.data()), (std::streamsize)(v.size() * 4));
    for (size_t i = 0; i < v.size(); ++i) img(i) = (MTfloat)v[i];
  } else {
    return false;
  }
  return (bool)f;
}

}  // namespace

// src/MutoFile.cpp:85-131
bool MutoFile::saveImage(const Image& img, const ImageHeader& header, std::string filename) {
  std::ofstream f(filename.c_str(), std::ios::out | std::ios::binary);
  if (f.fail()) {
    std::cout << "Writing image to file: " << filename << " failed. " << std::endl;
    return false;
  }
  unsigned char h[64];
  std::memset(h, 0, sizeof h);
  h[0] = 64;  // header length
  h[1] = 1;   // version
  std::memcpy(h + 2, header.method.data(), header.method.size() < 16 ? header.method.size() : 16);
  const uint32_t nray = header.nRay;
  const uint16_t nit = (uint16_t)header.nIter, prec = (uint16_t)header.fPrecision;
  const uint16_t n3[3] = {(uint16_t)header.grid.nx, (uint16_t)header.grid.ny, (uint16_t)header.grid.nz};
  const float g6[6] = {(float)header.grid.x_min, (float)header.grid.y_min, (float)header.grid.z_min,
                       (float)header.grid.dx, (float)header.grid.dy, (float)header.grid.dz};
  std::memcpy(h + 18, &nray, 4);
  std::memcpy(h + 22, &nit, 2);
  std::memcpy(h + 24, &prec, 2);
  std::memcpy(h + 26, n3, 6);
  std::memcpy(h + 32, g6, 24);
  f.write(reinterpret_cast<const char*>(h), 64);
  return write_values(f, img, header.fPrecision);
}

// src/MutoFile.cpp:134-152
bool MutoFile::saveImagePure(const Image& img, std::string filename, MTindex fPrecision) {
  std::ofstream f(filename.c_str(), std::ios::out | std::ios::binary);
  if (f.fail()) {
    std::cout << "Writing image to file: " << filename << " failed. " << std::endl;
    return false;
  }
  return write_values(f, img, fPrecision);
}

// src/MutoFile.cpp:155-201; an unreadable file gives an empty image
std::pair<Image, ImageHeader> MutoFile::loadImage(std::string filename) {
  std::ifstream f(filename.c_str(), std::ios::in | std::ios::binary);
  if (f.fail()) {
    std::cout << "Reading image file: " << filename << " failed. " << std::endl;
    return std::pair<Image, ImageHeader>{};
  }
  unsigned char h[64];
  f.read(reinterpret_cast<char*>(h), 58);
  if (!f) {
    std::cout << "Reading image file: " << filename << " failed. " << std::endl;
    return std::pair<Image, ImageHeader>{};
  }
  ImageHeader header;
  const size_t mlen = strnlen(reinterpret_cast<const char*>(h + 2), 16);
  header.method.assign(reinterpret_cast<const char*>(h + 2), mlen);
  uint32_t nray;
  uint16_t nit, prec, n3[3];
  float g6[6];
  std::memcpy(&nray, h + 18, 4);
  std::memcpy(&nit, h + 22, 2);
  std::memcpy(&prec, h + 24, 2);
  std::memcpy(n3, h + 26, 6);
  std::memcpy(g6, h + 32, 24);
  header.nRay = nray;
  header.nIter = nit;
  header.fPrecision = prec;
  header.grid = VoxelGrid{n3[0], n3[1], n3[2], g6[0], g6[1], g6[2], g6[3], g6[4], g6[5]};
  f.seekg(h[0]);  // end of the header
  Image img(n3[0], n3[1], n3[2]);
  if (!read_values(f, img, prec)) {
    std::cout << "Reading image file: " << filename << " failed. " << std::endl;
    return std::pair<Image, ImageHeader>{};
  }
  return std::pair<Image, ImageHeader>{img, header};
}

// src/MutoFile.cpp:204-245
Image MutoFile::loadImagePure(std::string filename, VoxelGrid grid, MTindex fPrecision) {
  std::ifstream f(filename.c_str(), std::ios::in | std::ios::binary);
  if (f.fail()) {
    std::cout << "Reading image file: " << filename << " failed. " << std::endl;
    return Image{};
  }
  f.seekg(0, std::ios::end);
  const long long size = (long long)f.tellg();
  f.seekg(0, std::ios::beg);
  if ((long long)grid.nx * grid.ny * grid.nz * fPrecision != size) {
    std::cout << "Wrong size of the image file and grid: " << filename << std::endl;
    return Image{};
  }
  Image img(grid.nx, grid.ny, grid.nz);
  if (!read_values(f, img, fPrecision)) {
    std::cout << "Floating point precision not supported. " << std::endl;
    return Image{};
  }
  return img;
}

// Principal axis of n centred points by cyclic Jacobi rotations of their
// 3x3 scatter matrix (the right singular vector of the largest singular
// value of the centred data, which src/MutoFile.cpp:270-271 takes from
// Eigen's JacobiSVD), sign fixed to d.z > 0.
void MutoFile::fitLine(const MTfloat* x, const MTfloat* y, const MTfloat* z, int n, MTfloat zref,
                       Vector3& point, Vector3& dir) {
  double m[3] = {0.0, 0.0, 0.0};
  for (int i = 0; i < n; ++i) {
    m[0] += x[i];
    m[1] += y[i];
    m[2] += z[i];
  }
  for (int k = 0; k < 3; ++k) m[k] /= n;
  double a[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
  for (int i = 0; i < n; ++i) {
    const double d[3] = {x[i] - m[0], y[i] - m[1], z[i] - m[2]};
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) a[r][c] += d[r] * d[c];
  }
  double v[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
  for (int sweep = 0; sweep < 64; ++sweep) {
    const double off = a[0][1] * a[0][1] + a[0][2] * a[0][2] + a[1][2] * a[1][2];
    const double diag = a[0][0] * a[0][0] + a[1][1] * a[1][1] + a[2][2] * a[2][2];
    if (off <= 1e-30 * diag || off == 0.0) break;
    for (int p = 0; p < 2; ++p) {
      for (int q = p + 1; q < 3; ++q) {
        if (a[p][q] == 0.0) continue;
        const double theta = (a[q][q] - a[p][p]) / (2.0 * a[p][q]);
        const double t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
        const double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
        for (int k = 0; k < 3; ++k) {  // A <- J^T A J
          const double akp = a[k][p], akq = a[k][q];
          a[k][p] = c * akp - s * akq;
          a[k][q] = s * akp + c * akq;
        }
        for (int k = 0; k < 3; ++k) {
          const double apk = a[p][k], aqk = a[q][k];
          a[p][k] = c * apk - s * aqk;
          a[q][k] = s * apk + c * aqk;
        }
        for (int k = 0; k < 3; ++k) {
          const double vkp = v[k][p], vkq = v[k][q];
          v[k][p] = c * vkp - s * vkq;
          v[k][q] = s * vkp + c * vkq;
        }
      }
    }
  }
  int best = 0;
  for (int k = 1; k < 3; ++k)
    if (a[k][k] > a[best][best]) best = k;
  double d[3] = {v[0][best], v[1][best], v[2][best]};
  const double nrm = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
  for (int k = 0; k < 3; ++k) d[k] /= nrm;
  if (d[2] < 0.0)
    for (int k = 0; k < 3; ++k) d[k] = -d[k];
  dir = Vector3(d[0], d[1], d[2]);
  const double t = (zref - m[2]) / d[2];  // src/MutoFile.cpp:272-273
  point = Vector3(t * d[0] + m[0], t * d[1] + m[1], zref);
}

// src/MutoFile.cpp:19-82 and 248-304
MutoMuonData MutoFile::getMuonData(std::vector<std::string> flist, std::vector<MTfloat> zlayers,
                                   int nMiddle, int fPrecision) {
  MutoMuonData data;
  if (fPrecision != 4 && fPrecision != 8) {
    std::cerr << "Floating point precision not supported. " << std::endl;
    return data;
  }
  const int L = (int)zlayers.size();
  if (nMiddle < 2 || L - nMiddle < 2) {
    std::cerr << "Error: at least two layers of detectors for both up and down area. " << std::endl;
    return data;
  }
  const size_t each = 2 * (size_t)L + 2;  // e_in, (x, y) per layer, e_out
  std::vector<MTfloat> xs(L), ys(L);
  for (const std::string& name : flist) {
    std::ifstream f(name.c_str(), std::ios::in | std::ios::binary);
    if (f.fail()) {
      std::cout << "Reading muon data error for : " << name << std::endl;
      continue;
    }
    f.seekg(0, std::ios::end);
    const size_t bytes = (size_t)f.tellg();
    f.seekg(0, std::ios::beg);
    const size_t n = bytes / each / (size_t)fPrecision;
    std::vector<MTfloat> raw(each * n);
    if (fPrecision == 8) {
      f.read(reinterpret_cast<char*>(raw.data()), (std::streamsize)(raw.size() * 8));
    } else {
      std::vector<float> tmp(each * n);
      f.read(reinterpret_cast<char*>(tmp.data()), (std::streamsize)(tmp.size() * 4));
      for (size_t i = 0; i < tmp.size(); ++i) raw[i] = tmp[i];
    }
    for (size_t r = 0; r < n; ++r) {
      const MTfloat* it = raw.data() + r * each;
      bool ok = true;
      for (size_t k = 0; k < each; ++k) ok = ok && !std::isnan(it[k]);
      if (!ok) continue;
      RayData ray;
      ray.ein = it[0];
      for (int i = 0; i < L; ++i) {
        xs[i] = it[1 + 2 * i];
        ys[i] = it[2 + 2 * i];
      }
      fitLine(xs.data(), ys.data(), zlayers.data(), nMiddle, zlayers[nMiddle - 1], ray.pin, ray.din);
      fitLine(xs.data() + nMiddle, ys.data() + nMiddle, zlayers.data() + nMiddle, L - nMiddle,
              zlayers[nMiddle], ray.pout, ray.dout);
      ray.eout = it[2 * L + 1];
      data.push_back(ray);
    }
  }
  return data;
}

// getMuonData with the per-record work on the GPU
MutoMuonData MutoFile::getMuonDataGPU(std::vector<std::string> flist, std::vector<MTfloat> zlayers,
                                      int nMiddle, int fPrecision, int device) {
  MutoMuonData data;
  if (fPrecision != 4 && fPrecision != 8) {
    std::cerr << "Floating point precision not supported. " << std::endl;
    return data;
  }
  const int L = (int)zlayers.size();
  if (nMiddle < 2 || L - nMiddle < 2) {
    std::cerr << "Error: at least two layers of detectors for both up and down area. " << std::endl;
    return data;
  }
  muto_ctx* ctx = nullptr;
  open_ctx(ctx, device);
  const size_t each = 2 * (size_t)L + 2;
  try {
    for (const std::string& name : flist) {
      std::ifstream f(name.c_str(), std::ios::in | std::ios::binary);
      if (f.fail()) {
        std::cout << "Reading muon data error for : " << name << std::endl;
        continue;
      }
      f.seekg(0, std::ios::end);
      const size_t bytes = (size_t)f.tellg();
      f.seekg(0, std::ios::beg);
      const size_t n = bytes / each / (size_t)fPrecision;
      std::vector<char> raw(n * each * (size_t)fPrecision);
      f.read(raw.data(), (std::streamsize)raw.size());
      const size_t base = data.size();
      data.resize(base + n);
      size_t kept = 0;
      check(ctx, muto_rays_from_records(ctx, n, L, nMiddle, zlayers.data(), raw.data(), fPrecision,
                                        reinterpret_cast<muto_ray*>(data.data() + base), &kept),
            "MutoFile::getMuonDataGPU");
      data.resize(base + kept);
    }
  } catch (...) {
    muto_ctx_destroy(ctx);
    throw;
  }
  muto_ctx_destroy(ctx);
  return data;
}
