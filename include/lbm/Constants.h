// Lattice constants of the reference's CONSTANTS namespace
// (simulation/include/Constants.h:4-9, simulation/source/Constants.cpp:5-7).
#pragma once

namespace CONSTANTS {
extern const float GEOMETRIC_TOLERANCE;
extern const int DIRECTIONS[27][3];
extern const int INVERSE_DIRECTIONS[27];
}  // namespace CONSTANTS
