// Lattice velocity sets as compile-time tables usable in host and device
// code.  Direction order and inverse map follow the reference's
// simulation/source/Constants.cpp:6-7 (shells rest / axis / edge / corner,
// lexicographic inside a shell); D3Q19 is the first 19 entries of that
// table, D2Q9 its z = 0 subset in the same order.
#pragma once

#include <cstdint>

namespace lbm {

template <int Q>
struct Lat;

template <>
struct Lat<27> {
  static constexpr int Q = 27;
  static constexpr int D = 3;
  static constexpr int E[27][3] = {
      {0, 0, 0},   {-1, 0, 0},  {0, -1, 0},  {0, 0, -1}, {0, 0, 1},
      {0, 1, 0},   {1, 0, 0},   {-1, -1, 0}, {-1, 0, -1}, {-1, 0, 1},
      {-1, 1, 0},  {0, -1, -1}, {0, -1, 1},  {0, 1, -1}, {0, 1, 1},
      {1, -1, 0},  {1, 0, -1},  {1, 0, 1},   {1, 1, 0},  {-1, -1, -1},
      {-1, -1, 1}, {-1, 1, -1}, {-1, 1, 1},  {1, -1, -1}, {1, -1, 1},
      {1, 1, -1},  {1, 1, 1}};
  static constexpr int INV[27] = {0,  6,  5,  4,  3,  2,  1,  18, 17,
                                  16, 15, 14, 13, 12, 11, 10, 9,  8,
                                  7,  26, 25, 24, 23, 22, 21, 20, 19};
  static constexpr double W[27] = {
      8.0 / 27.0,  2.0 / 27.0,  2.0 / 27.0,  2.0 / 27.0,  2.0 / 27.0,
      2.0 / 27.0,  2.0 / 27.0,  1.0 / 54.0,  1.0 / 54.0,  1.0 / 54.0,
      1.0 / 54.0,  1.0 / 54.0,  1.0 / 54.0,  1.0 / 54.0,  1.0 / 54.0,
      1.0 / 54.0,  1.0 / 54.0,  1.0 / 54.0,  1.0 / 54.0,  1.0 / 216.0,
      1.0 / 216.0, 1.0 / 216.0, 1.0 / 216.0, 1.0 / 216.0, 1.0 / 216.0,
      1.0 / 216.0, 1.0 / 216.0};
};

template <>
struct Lat<19> {
  static constexpr int Q = 19;
  static constexpr int D = 3;
  static constexpr int E[19][3] = {
      {0, 0, 0},  {-1, 0, 0},  {0, -1, 0}, {0, 0, -1}, {0, 0, 1},
      {0, 1, 0},  {1, 0, 0},   {-1, -1, 0}, {-1, 0, -1}, {-1, 0, 1},
      {-1, 1, 0}, {0, -1, -1}, {0, -1, 1}, {0, 1, -1}, {0, 1, 1},
      {1, -1, 0}, {1, 0, -1},  {1, 0, 1},  {1, 1, 0}};
  static constexpr int INV[19] = {0,  6,  5,  4,  3,  2,  1,  18, 17, 16,
                                  15, 14, 13, 12, 11, 10, 9,  8,  7};
  static constexpr double W[19] = {
      1.0 / 3.0,  1.0 / 18.0, 1.0 / 18.0, 1.0 / 18.0, 1.0 / 18.0,
      1.0 / 18.0, 1.0 / 18.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0,
      1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0,
      1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0};
};

template <>
struct Lat<9> {
  static constexpr int Q = 9;
  static constexpr int D = 2;
  static constexpr int E[9][3] = {{0, 0, 0},  {-1, 0, 0}, {0, -1, 0},
                                  {0, 1, 0},  {1, 0, 0},  {-1, -1, 0},
                                  {-1, 1, 0}, {1, -1, 0}, {1, 1, 0}};
  static constexpr int INV[9] = {0, 4, 3, 2, 1, 8, 7, 6, 5};
  static constexpr double W[9] = {4.0 / 9.0,  1.0 / 9.0,  1.0 / 9.0,
                                  1.0 / 9.0,  1.0 / 9.0,  1.0 / 36.0,
                                  1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0};
};

// 1/|e|^2 of the D3Q27 set (simulation/source/SimCalcFuncs.cu:14).
constexpr float kSplit27[27] = {
    0.0f,        1.0f,        1.0f,        1.0f,        1.0f,
    1.0f,        1.0f,        0.5f,        0.5f,        0.5f,
    0.5f,        0.5f,        0.5f,        0.5f,        0.5f,
    0.5f,        0.5f,        0.5f,        0.5f,        1.0f / 3.0f,
    1.0f / 3.0f, 1.0f / 3.0f, 1.0f / 3.0f, 1.0f / 3.0f, 1.0f / 3.0f,
    1.0f / 3.0f, 1.0f / 3.0f};

}  // namespace lbm
