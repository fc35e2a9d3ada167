// Minimal test registry with the assertion names of the reference's gtest
// suites (googletest is not available here).  Each TEST body runs inside a
// try block; a thrown exception fails the test.  argv[1], if given, is a
// substring filter on "Suite.Name".
#pragma once

#include <cmath>
#include <cstdio>
#include <cstring>
#include <exception>
#include <functional>
#include <string>
#include <vector>

namespace minitest {

struct Case {
  std::string name;
  std::function<void()> fn;
};

inline std::vector<Case>& registry() {
  static std::vector<Case> r;
  return r;
}
inline int& failures() {
  static int f = 0;
  return f;
}
struct Registrar {
  Registrar(const char* suite, const char* name, std::function<void()> fn) {
    registry().push_back({std::string(suite) + "." + name, std::move(fn)});
  }
};

inline int run_all(int argc, char** argv) {
  const char* filter = argc > 1 ? argv[1] : nullptr;
  int ran = 0, failed = 0;
  for (auto& c : registry()) {
    if (filter && c.name.find(filter) == std::string::npos) continue;
    const int before = failures();
    std::printf("[ RUN      ] %s\n", c.name.c_str());
    std::fflush(stdout);
    try {
      c.fn();
    } catch (const std::exception& e) {
      std::printf("  exception: %s\n", e.what());
      ++failures();
    }
    const bool ok = failures() == before;
    std::printf("[ %s ] %s\n", ok ? "      OK" : " FAILED ", c.name.c_str());
    std::fflush(stdout);
    ++ran;
    if (!ok) ++failed;
  }
  std::printf("[==========] %d tests ran, %d failed\n", ran, failed);
  return failed ? 1 : (ran ? 0 : 2);
}

}  // namespace minitest

#define MT_CAT2(a, b) a##b
#define MT_CAT(a, b) MT_CAT2(a, b)
#define TEST(suite, name)                                                   \
  static void MT_CAT(test_, MT_CAT(suite, name))();                         \
  static minitest::Registrar MT_CAT(reg_, MT_CAT(suite, name))(             \
      #suite, #name, MT_CAT(test_, MT_CAT(suite, name)));                   \
  static void MT_CAT(test_, MT_CAT(suite, name))()

#define MT_FAIL(msg)                                                        \
  do {                                                                      \
    std::printf("  %s:%d: %s\n", __FILE__, __LINE__, msg);                  \
    ++minitest::failures();                                                 \
  } while (0)

#define EXPECT_NEAR(a, b, tol)                                              \
  do {                                                                      \
    const double a_ = (double)(a), b_ = (double)(b);                        \
    if (!(std::fabs(a_ - b_) <= (double)(tol))) {                           \
      char m_[256];                                                         \
      std::snprintf(m_, sizeof m_, "EXPECT_NEAR(%s, %s): %.9g vs %.9g", #a,  \
                    #b, a_, b_);                                            \
      MT_FAIL(m_);                                                          \
    }                                                                       \
  } while (0)
#define EXPECT_EQ(a, b)                                                     \
  do {                                                                      \
    if (!((a) == (b))) MT_FAIL("EXPECT_EQ(" #a ", " #b ")");                \
  } while (0)
#define EXPECT_TRUE(a)                                                      \
  do {                                                                      \
    if (!(a)) MT_FAIL("EXPECT_TRUE(" #a ")");                               \
  } while (0)
#define EXPECT_FALSE(a)                                                     \
  do {                                                                      \
    if (a) MT_FAIL("EXPECT_FALSE(" #a ")");                                 \
  } while (0)
#define RUN_ALL_TESTS() minitest::run_all(argc, argv)
