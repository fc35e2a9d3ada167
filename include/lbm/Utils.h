// simulation/include/Utils.h:1-5
#pragma once

namespace Utils {
// Packs num values of bitLength bits, most significant first
// (simulation/source/Utils.cpp:3-28).
unsigned int compressInts(int* array, const int num, const int bitLength);
}  // namespace Utils
