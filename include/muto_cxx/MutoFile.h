/* Drop-in for the reference's MutoFile (include/MutoFile.h,
 * src/MutoFile.cpp): muon record files and the image file format.
 *
 * Image file: a 64-byte header -- u8 header length (64), u8 version (1),
 * 16 chars method, u32 number of rays, u16 iterations, u16 bytes per value
 * (4 or 8), u16 nx, ny, nz, f32 x_min, y_min, z_min, dx, dy, dz, 8 zero
 * bytes -- then the nx*ny*nz values, x fastest (src/MutoFile.cpp:85-131,
 * 155-201), little-endian.
 *
 * Muon records: per muon 2L + 2 values (f32 or f64): e_in, (x, y) on each
 * of the L tracker layers, e_out; records with a NaN are skipped
 * (src/MutoFile.cpp:19-82).  Each side's hits are fitted by the principal
 * axis of the centred points (src/MutoFile.cpp:248-304, there by Eigen's
 * JacobiSVD); the axis sign, which Jacobi SVD leaves arbitrary, is fixed
 * here to d.z > 0. */
#pragma once

#include <string>
#include <utility>
#include <vector>

#include "MutoTypes.h"

class MutoFile {
 public:
  static MutoMuonData getMuonData(std::vector<std::string> flist, std::vector<MTfloat> zlayers,
                                  int nMiddle, int fPrecision = 4);
  /* getMuonData with the NaN filter and the line fits on the GPU (HIP
   * device `device`, through muto_rays_from_records); the files are read on
   * the host.  Same records kept, same rays.  Throws std::runtime_error
   * when the device cannot be reached. */
  static MutoMuonData getMuonDataGPU(std::vector<std::string> flist, std::vector<MTfloat> zlayers,
                                     int nMiddle, int fPrecision = 4, int device = 0);
  static bool saveImage(const Image& img, const ImageHeader& header, std::string filename);
  static bool saveImagePure(const Image& img, std::string filename, MTindex fPrecision = 4);
  static std::pair<Image, ImageHeader> loadImage(std::string filename);
  static Image loadImagePure(std::string filename, VoxelGrid grid, MTindex fPrecision = 4);
  /* One side's line fit from L points (x, y, z): principal axis through the
   * centroid, d.z > 0; the reference point is where the line crosses
   * z = zref. */
  static void fitLine(const MTfloat* x, const MTfloat* y, const MTfloat* z, int n, MTfloat zref,
                      Vector3& point, Vector3& dir);
};
