// simulation/source/Utils.cpp:3-28: packs the first num-1 values shifting
// left by 5 after each, then ORs the last one in; values that do not fit in
// bitLength bits are skipped.
#include "Utils.h"

namespace Utils {
unsigned int compressInts(int* array, const int num, const int bitLength) {
  const unsigned limit = 1u << bitLength;
  unsigned out = 0;
  for (int i = 0; i + 1 < num; ++i) {
    if ((unsigned)array[i] < limit) out = (out | (unsigned)array[i]) << 5;
  }
  if ((unsigned)array[num - 1] < limit) out |= (unsigned)array[num - 1];
  return out;
}
}  // namespace Utils
