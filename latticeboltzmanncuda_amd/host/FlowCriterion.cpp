// Flow criteria as data (reference: simulation/source/FlowCriterion.cu:5-16).
#include "FlowCriterion.h"

namespace FlowCriterion {

Criterion point(int x, int y, int z) {
  return Criterion{Point, {x, y, z, 0, 0, 0}};
}
Criterion plane(int axis, int coordinate) {
  return Criterion{Plane, {axis, coordinate, 0, 0, 0, 0}};
}
Criterion box(int x0, int y0, int z0, int x1, int y1, int z1) {
  return Criterion{Box, {x0, y0, z0, x1, y1, z1}};
}

bool Criterion::operator()(int x, int y, int z) const {
  switch (kind) {
    case Point:
      return x == params[0] && y == params[1] && z == params[2];
    case Plane: {
      const int c = params[0] == 0 ? x : (params[0] == 1 ? y : z);
      return c == params[1];
    }
    case Box:
      return x >= params[0] && y >= params[1] && z >= params[2] &&
             x <= params[3] && y <= params[4] && z <= params[5];
    default:
      return false;
  }
}

bool testCriteria(int x, int y, int z) { return x == 1 && y == 1 && z == 1; }
bool testPlaneCriteria(int x, int, int) { return x == 0; }

const Criterion ptestCriteria = point(1, 1, 1);
const Criterion ptestPlaneCriteria = plane(0, 0);

}  // namespace FlowCriterion
