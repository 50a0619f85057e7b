// Command-line driver in the manner of the reference's MuToTest.cpp:15-83:
// muto_cli <config.json> reads the muon records (line fits on the GPU),
// runs the method named by the configuration's "poca", "mlsd", "map", "slr"
// or "pocut" object on the GPU and writes the image file.  ("slr_iter" is not run: its in-place,
// order-dependent sweep is not part of this library.)
#include <chrono>
#include <iostream>
#include <memory>

#include "MutoConfig.h"
#include "MutoFile.h"
#include "MutoMAP.h"
#include "MutoMLSD.h"
#include "MutoPoCA.h"
#include "MutoPoCut.h"
#include "MutoSLR.h"

int main(int argc, char** argv) {
  MutoConfig config;
  const std::string cfg_file = argc > 1 ? argv[1] : "example.json";
  if (!config.loadFile(cfg_file)) {
    std::cerr << "Configuration file load error for: " << cfg_file << std::endl;
    return 1;
  }
  const auto ts = std::chrono::steady_clock::now();
  try {
    std::vector<std::string> files = config.get("data_files");
    std::vector<MTfloat> zlayers = config.get("detector_height");
    int n_middle = config.get("detector_middle");
    MutoMuonData data = MutoFile::getMuonDataGPU(files, zlayers, n_middle);
    if (data.empty()) {
      std::cerr << "No muon data read." << std::endl;
      return 1;
    }
    std::unique_ptr<MutoVReconstruction> rec;
    std::string method;
    ImageHeader header;
    if (config.exist("poca")) {
      method = "poca";
      auto* r = new MutoPoCA(config.get("poca"));
      rec.reset(r);
      header.grid = r->getCurrentGrid();
      header.nIter = 1;
    } else if (config.exist("mlsd")) {
      method = "mlsd";
      auto* r = new MutoMLSD(config.get("mlsd"));
      rec.reset(r);
      header.nIter = config.get("mlsd")["num_iteration"];
    } else if (config.exist("map")) {
      method = "map";
      auto* r = new MutoMAP(config.get("map"));
      rec.reset(r);
      header.nIter = config.get("map")["num_iteration"];
    } else if (config.exist("slr")) {
      method = "slr";
      auto* r = new MutoSLR(config.get("slr"));
      rec.reset(r);
      header.grid = r->getCurrentGrid();
      header.nIter = 1;
    } else if (config.exist("pocut")) {
      method = "pocut";
      auto* r = new MutoPoCut(config.get("pocut"));
      rec.reset(r);
      header.grid = r->getCurrentGrid();
      header.nIter = 1;
    } else {
      std::cerr << "The configuration names no method this library runs "
                   "(poca, mlsd, map, slr, pocut)." << std::endl;
      return 2;
    }
    Image img = rec->reconstruct(data);
    // MLSD/MAP may adopt the initial image's grid
    if (method == "mlsd" || method == "map")
      header.grid = static_cast<MutoMLSD*>(rec.get())->getCurrentGrid();
    std::string output = config.get("output_file");
    header.fPrecision = 4;
    header.method = method;
    header.nRay = (MTindex)data.size();
    if (!MutoFile::saveImage(img, header, output)) return 1;
  } catch (const std::exception& e) {
    std::cerr << "muto_cli: " << e.what() << std::endl;
    return 1;
  }
  const double t = std::chrono::duration_cast<std::chrono::duration<double>>(
                       std::chrono::steady_clock::now() - ts).count();
  std::cout << "Reconstruction took: " << t << " seconds. " << std::endl;
  return 0;
}
