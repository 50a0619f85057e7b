/* Drop-in for the reference's MutoMLSD (include/MutoMLSD.h:27-41): same
 * name, JSON keys and defaults (src/MutoMLSD.cpp:10-84), the initial image
 * handled as in src/MutoMLSD.cpp:226-266 (file with grid adoption, random,
 * constant), the EM iterations on the GPU through muto_mlsd_reconstruct.
 * Extra keys: "device" (default 0) and "precision" ("fp64", the reference's
 * arithmetic in its order; "fp64_stream", FP64 on the brick-major stream,
 * the same W_ij and sums in another association; or "fp32c": FP32 entry
 * terms with FP64/fixed-point sums, mean update only).  Library errors throw
 * std::runtime_error. */
#pragma once

#include <chrono>
#include <string>
#include <vector>

#include "MutoConfig.h"
#include "MutoTypes.h"
#include "MutoVReconstruction.h"

struct muto_ctx;
struct muto_mlsd_cfg;

class MutoMLSD : public MutoVReconstruction {
 public:
  MutoMLSD();
  explicit MutoMLSD(json config);
  virtual ~MutoMLSD();
  virtual Image reconstruct(const MutoMuonData&);
  VoxelGrid getCurrentGrid() { return fGrid; }
  /* cost function value of each iteration run by the last reconstruct() */
  const std::vector<MTfloat>& getCostHistory() const { return fCost; }

 protected:
  VoxelGrid fGrid;
  bool fUpdateWithMean;
  int fPercentileMean;
  bool fStraightLine;
  bool fEarlyStop;
  bool fHasEnergy;
  bool fLogPrint;
  int fNTotal;
  int fNScatInGrid;
  int fNIteration;
  int fInitMethod;
  MTfloat fInitValue;
  std::string fInitImageFile;
  MTfloat fErrorXY;
  MTfloat fErrorZo;
  MTfloat fErrorZi;
  MTfloat fAngMag;
  MTfloat fEPS;
  int fDevice;
  int fPrecision;
  muto_ctx* fCtx;
  std::vector<MTfloat> fCost;
  std::chrono::steady_clock::time_point fTimeStart;

  Image initailizeImage();  // (sic) the reference's name
  void fillConfig(muto_mlsd_cfg& c) const;
  Image run(const MutoMuonData& rays, bool map);
  void logMessage(const std::string& msg, bool printTime = false);
  void logResults();
  virtual void fillPrior(muto_mlsd_cfg&) const {}
};
