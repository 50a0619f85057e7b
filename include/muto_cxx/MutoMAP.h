/* Drop-in for the reference's MutoMAP (include/MutoMAP.h:16-27): MLSD keys
 * plus "beta", "p", "delta", "add_object_prior", "all_neighbours"
 * (src/MutoMAP.cpp:8-34); the ML step and the per-voxel surrogate step run
 * on the GPU through muto_map_reconstruct. */
#pragma once

#include "MutoMLSD.h"

class MutoMAP : public MutoMLSD {
 public:
  MutoMAP();
  explicit MutoMAP(json config);
  virtual ~MutoMAP();
  virtual Image reconstruct(const MutoMuonData&);

 private:
  MTfloat fBeta;
  MTfloat fP;
  MTfloat fDelta;
  bool fUseAllNeighbours;
  bool fAddObjectPrior;
  void defaultInitialize();

 protected:
  void fillPrior(muto_mlsd_cfg& c) const override;
};
