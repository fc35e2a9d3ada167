// Flow-injection criteria (reference: simulation/include/FlowCriterion.cuh:6-13,
// simulation/source/FlowCriterion.cu:5-16).  The reference selects nodes
// with __device__ function pointers; here a criterion is plain data that
// the GPU kernel evaluates, so it can be built on the host and copied.
#pragma once

#include <cstdint>

namespace FlowCriterion {

enum Kind : int32_t { None = 0, Point = 1, Plane = 2, Box = 3 };

struct Criterion {
  Kind kind;
  int32_t params[6];  // Point: x,y,z  Plane: axis, coordinate  Box: lo, hi
  bool operator()(int x, int y, int z) const;  // host-side evaluation
};

Criterion point(int x, int y, int z);
Criterion plane(int axis, int coordinate);
Criterion box(int x0, int y0, int z0, int x1, int y1, int z1);

// FlowCriterion.cu:5-8: x == 1 && y == 1 && z == 1
bool testCriteria(int x, int y, int z);
// FlowCriterion.cu:10-13: x == 0
bool testPlaneCriteria(int x, int y, int z);

// The reference's device pointers to the two criteria
// (FlowCriterion.cu:15-16); FlowData takes their address as it does there.
extern const Criterion ptestCriteria;
extern const Criterion ptestPlaneCriteria;

}  // namespace FlowCriterion
