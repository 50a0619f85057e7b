/* Drop-in for the reference's include/MutoVReconstruction.h:11-16: the
 * polymorphic reconstruction interface the driver calls. */
#pragma once

#include "MutoTypes.h"

class MutoVReconstruction {
 public:
  MutoVReconstruction() {}
  virtual ~MutoVReconstruction() {}
  virtual Image reconstruct(const MutoMuonData&) = 0;
};
