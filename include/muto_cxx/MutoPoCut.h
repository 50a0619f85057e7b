/* Drop-in for the reference's MutoPoCut (include/MutoPoCut.h): PoCA whose
 * voxel value keeps the lowest "cut_percentile" fraction (default 1.0) of
 * the scattering values of the voxel (src/MutoPoCut.cpp:4-15), on the GPU
 * through muto_pocut_reconstruct (include/muto.h).  The two-leg path is
 * always used, as in the reference. */
#pragma once

#include "MutoPoCA.h"

class MutoPoCut : public MutoPoCA {
 public:
  MutoPoCut();
  explicit MutoPoCut(json config);
  virtual ~MutoPoCut();
  virtual Image reconstruct(const MutoMuonData&);

 private:
  double fPercentile;
};
