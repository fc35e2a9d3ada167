// Host exchange records of the reference (simulation/include/LatticePoint.h:3-17).
// LatticePoint keeps the reference's 144-byte layout so that arrays of it
// pass unchanged through lbm_upload_points / lbm_download_points.
#pragma once

#include <cstddef>

struct ReflectionValues {
  int reflection_directions[14];
  float reflection_weight[81];
};

struct LatticePoint {
  int x, y, z;
  float particle_distribution[27];
  bool isReflected;
  bool isInternal;
  ReflectionValues* reflections = nullptr;
  // Device-side record pointer of the reference.  Records live in a table
  // owned by the GPU context here, so downloads leave this NULL.
  ReflectionValues* d_reflections = nullptr;
};

static_assert(sizeof(ReflectionValues) == 380, "ReflectionValues layout");
static_assert(sizeof(LatticePoint) == 144, "LatticePoint layout");
static_assert(offsetof(LatticePoint, particle_distribution) == 12, "layout");
static_assert(offsetof(LatticePoint, isReflected) == 120, "layout");
static_assert(offsetof(LatticePoint, isInternal) == 121, "layout");
static_assert(offsetof(LatticePoint, reflections) == 128, "layout");
static_assert(offsetof(LatticePoint, d_reflections) == 136, "layout");
