// Location of the OBJ fixtures (the reference's test/dataFiles, copied to
// tests/golden/dataFiles): $LBM_TEST_DATA, else relative to the repo root.
#pragma once

#include <cstdlib>
#include <string>

inline std::string data_file(const char* name) {
  const char* dir = std::getenv("LBM_TEST_DATA");
  return std::string(dir ? dir : "tests/golden/dataFiles") + "/" + name;
}
