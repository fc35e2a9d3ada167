// simulation/source/Constants.cpp:5-7: the D3Q27 velocity set, ordered by
// shell (rest, axis, edge, corner) and lexicographically inside a shell,
// and its inverse map.
#include "Constants.h"

namespace CONSTANTS {
const float GEOMETRIC_TOLERANCE = 1e-6f;
const int DIRECTIONS[27][3] = {
    {0, 0, 0},   {-1, 0, 0},  {0, -1, 0},  {0, 0, -1},  {0, 0, 1},
    {0, 1, 0},   {1, 0, 0},   {-1, -1, 0}, {-1, 0, -1}, {-1, 0, 1},
    {-1, 1, 0},  {0, -1, -1}, {0, -1, 1},  {0, 1, -1},  {0, 1, 1},
    {1, -1, 0},  {1, 0, -1},  {1, 0, 1},   {1, 1, 0},   {-1, -1, -1},
    {-1, -1, 1}, {-1, 1, -1}, {-1, 1, 1},  {1, -1, -1}, {1, -1, 1},
    {1, 1, -1},  {1, 1, 1}};
const int INVERSE_DIRECTIONS[27] = {0,  6,  5,  4,  3,  2,  1,  18, 17,
                                    16, 15, 14, 13, 12, 11, 10, 9,  8,
                                    7,  26, 25, 24, 23, 22, 21, 20, 19};
}  // namespace CONSTANTS
