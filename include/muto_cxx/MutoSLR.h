/* Drop-in for the reference's MutoSLR (include/MutoSLR.h): straight-line
 * reconstruction, same JSON keys and defaults (src/MutoSLR.cpp:4-43), on
 * the GPU through muto_slr_reconstruct (include/muto.h).  Extra key:
 * "device" (HIP device index, default 0). */
#pragma once

#include "MutoConfig.h"
#include "MutoTypes.h"
#include "MutoVReconstruction.h"

struct muto_ctx;

class MutoSLR : public MutoVReconstruction {
 public:
  MutoSLR();
  explicit MutoSLR(json config);
  virtual ~MutoSLR();
  virtual Image reconstruct(const MutoMuonData&);
  VoxelGrid getCurrentGrid() { return fGrid; }

 protected:
  VoxelGrid fGrid;
  bool fLengthWeighted;
  bool fHasEnergy;
  bool fLogPrint;
  int fNTotal;
  MTfloat fAngMag;
  MTfloat fEPS;
  int fDevice;
  muto_ctx* fCtx;
  void logResults();
};
