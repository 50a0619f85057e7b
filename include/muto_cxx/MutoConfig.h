/* Drop-in for the reference's include/MutoConfig.h: a JSON run
 * configuration (nlohmann::json, the single-header json.hpp). */
#pragma once

#include <fstream>
#include <iostream>
#include <string>

#include "json.hpp"

using json = nlohmann::json;

class MutoConfig {
 public:
  MutoConfig() {}
  explicit MutoConfig(std::string filename) {
    if (!loadFile(filename)) std::cerr << "Configuration file load error for: " << filename << std::endl;
  }
  bool loadFile(std::string filename) {
    std::ifstream in(filename.c_str());
    if (in.fail()) return false;
    in >> config_;
    return true;
  }
  std::string getConfigString() { return config_.dump(); }
  json::value_type get(std::string attr) { return config_[attr]; }
  bool exist(std::string key) { return config_.find(key) != config_.end(); }

 private:
  json config_;
};
