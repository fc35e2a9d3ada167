// FlowData of the reference (simulation/include/FlowSurface.h:9-46): the
// populations injected each step where the criterion holds.
#pragma once

#include "FlowCriterion.h"
#include "LatticePoint.h"

struct FlowData {
  LatticePoint inducedFlow;
  FlowCriterion::Criterion pointCriterion;
  int val = 1;

  FlowData(const FlowCriterion::Criterion* criterion, LatticePoint flow)
      : inducedFlow(flow), pointCriterion(*criterion) {}

  void setCriterion(const FlowCriterion::Criterion* criterion) {
    pointCriterion = *criterion;
  }

  void setInducedFlow(LatticePoint flow) { inducedFlow = flow; }
};
