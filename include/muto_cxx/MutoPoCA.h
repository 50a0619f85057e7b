/* Drop-in for the reference's MutoPoCA (include/MutoPoCA.h:17-31): same
 * name, JSON keys and defaults (src/MutoPoCA.cpp:4-47), reconstruct()
 * running on the GPU through muto_poca_reconstruct (include/muto.h).
 * Extra key: "device" (HIP device index, default 0).  Failures to reach
 * the device or errors from the library throw std::runtime_error. */
#pragma once

#include "MutoConfig.h"
#include "MutoTypes.h"
#include "MutoVReconstruction.h"

struct muto_ctx;

class MutoPoCA : public MutoVReconstruction {
 public:
  MutoPoCA();
  explicit MutoPoCA(json config);
  virtual ~MutoPoCA();
  virtual Image reconstruct(const MutoMuonData&);
  VoxelGrid getCurrentGrid() { return fGrid; }

 protected:
  VoxelGrid fGrid;
  bool fLengthWeighted;
  bool fStraightLine;
  bool fHasEnergy;
  bool fLogPrint;
  int fNTotal;
  int fNNotParallel;
  int fNScatInGrid;
  MTfloat fAngMag;
  MTfloat fEPS;
  int fDevice;
  muto_ctx* fCtx;
  void logResults();
};
