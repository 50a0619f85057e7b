/* Drop-in types for the reference's include/MutoTypes.h without Eigen.
 * Same names, field order and memory layout: VoxelGrid is 64 bytes
 * (3 x u32 + pad + 6 x f64, = muto_grid), RayData 112 bytes (2 + 4 x 3
 * f64, = muto_ray), and Image stores nx*ny*nz values column-major, x
 * fastest (the Eigen Tensor order of MutoTypes.h:22), so data() can be
 * handed to the C ABI (include/muto.h) as it is. */
#pragma once

#include <cmath>
#include <cstddef>
#include <string>
#include <vector>

typedef double MTfloat;
typedef unsigned int MTindex;

struct VoxelGrid {  // MutoTypes.h:25-37
  MTindex nx;
  MTindex ny;
  MTindex nz;
  MTfloat x_min;
  MTfloat y_min;
  MTfloat z_min;
  MTfloat dx;
  MTfloat dy;
  MTfloat dz;
};

struct VoxelData {  // MutoTypes.h:40-45
  MTindex x;
  MTindex y;
  MTindex z;
  MTfloat length;
};

/* Three doubles with the accessors the reference's callers use on Eigen's
 * Matrix<double, 1, 3>: (i), [i], x()/y()/z(), data(), arithmetic, dot,
 * norm. */
class Vector3 {
 public:
  Vector3() : v_{0.0, 0.0, 0.0} {}
  Vector3(MTfloat x, MTfloat y, MTfloat z) : v_{x, y, z} {}
  MTfloat& operator()(int i) { return v_[i]; }
  MTfloat operator()(int i) const { return v_[i]; }
  MTfloat& operator[](int i) { return v_[i]; }
  MTfloat operator[](int i) const { return v_[i]; }
  MTfloat x() const { return v_[0]; }
  MTfloat y() const { return v_[1]; }
  MTfloat z() const { return v_[2]; }
  MTfloat* data() { return v_; }
  const MTfloat* data() const { return v_; }
  Vector3 operator+(const Vector3& o) const { return Vector3(v_[0] + o.v_[0], v_[1] + o.v_[1], v_[2] + o.v_[2]); }
  Vector3 operator-(const Vector3& o) const { return Vector3(v_[0] - o.v_[0], v_[1] - o.v_[1], v_[2] - o.v_[2]); }
  Vector3 operator*(MTfloat s) const { return Vector3(v_[0] * s, v_[1] * s, v_[2] * s); }
  MTfloat dot(const Vector3& o) const { return (v_[0] * o.v_[0] + v_[1] * o.v_[1]) + v_[2] * o.v_[2]; }
  MTfloat norm() const { return std::sqrt(dot(*this)); }

 private:
  MTfloat v_[3];
};

struct RayData {  // MutoTypes.h:48-55
  MTfloat ein;
  MTfloat eout;
  Vector3 pin;
  Vector3 pout;
  Vector3 din;
  Vector3 dout;
};

typedef std::vector<RayData> MutoMuonData;

/* 3-D image, column-major (x fastest) like Tensor<double, 3>. */
class Image {
 public:
  Image() : n_{0, 0, 0} {}
  Image(MTindex nx, MTindex ny, MTindex nz)
      : n_{nx, ny, nz}, d_((size_t)nx * ny * nz, 0.0) {}
  MTfloat& operator()(MTindex x, MTindex y, MTindex z) { return d_[x + (size_t)n_[0] * (y + (size_t)n_[1] * z)]; }
  MTfloat operator()(MTindex x, MTindex y, MTindex z) const { return d_[x + (size_t)n_[0] * (y + (size_t)n_[1] * z)]; }
  MTfloat& operator()(size_t j) { return d_[j]; }
  MTfloat operator()(size_t j) const { return d_[j]; }
  size_t size() const { return d_.size(); }
  MTindex dimension(int k) const { return n_[k]; }
  MTfloat* data() { return d_.data(); }
  const MTfloat* data() const { return d_.data(); }
  void setConstant(MTfloat v) { d_.assign(d_.size(), v); }

 private:
  MTindex n_[3];
  std::vector<MTfloat> d_;
};

struct ImageHeader {  // MutoTypes.h:59-65
  std::string method;  // the first 16 characters go to the file header
  VoxelGrid grid;
  MTindex nRay;
  MTindex nIter;
  MTindex fPrecision;
};
