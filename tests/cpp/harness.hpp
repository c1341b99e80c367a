// Minimal test harness for the C++ API tests (no GoogleTest offline): named
// cases, EXPECT_/ASSERT_ checks with the GoogleTest spellings, typed cases run
// for float and double, and a CPU-only subset for machines without a GPU.
#pragma once

#include <cmath>
#include <cstdio>
#include <exception>
#include <functional>
#include <sstream>
#include <string>
#include <vector>

namespace harness {

struct Case {
    std::string name;
    std::function<void()> body;
    bool needsDevice;
};
inline std::vector<Case>& registry() {
    static std::vector<Case> cases;
    return cases;
}
inline int& failures() {
    static int count = 0;
    return count;
}
struct AssertionAbort {};
struct Registrar {
    Registrar(std::string name, std::function<void()> body, bool needsDevice) {
        registry().push_back({std::move(name), std::move(body), needsDevice});
    }
};
inline void fail(const char* file, int line, const std::string& what) {
    ++failures();
    std::fprintf(stderr, "%s:%d: failure: %s\n", file, line, what.c_str());
}
template <class A, class B>
std::string describe(const char* ea, const char* eb, const A& a, const B& b) {
    std::ostringstream s;
    s << ea << " vs " << eb;
    if constexpr (requires(std::ostream& o) { o << a; o << b; }) s << " (" << a << " vs " << b << ")";
    return s.str();
}

inline int run(int argc, char** argv) {
    bool cpuOnly = false;
    std::string filter;
    for (int i = 1; i < argc; ++i) {
        const std::string arg = argv[i];
        if (arg == "--cpu-only") cpuOnly = true;
        else filter = arg;
    }
    int ran = 0;
    for (const auto& c : registry()) {
        if (cpuOnly && c.needsDevice) continue;
        if (!filter.empty() && c.name.find(filter) == std::string::npos) continue;
        const int before = failures();
        std::printf("[ RUN  ] %s\n", c.name.c_str());
        std::fflush(stdout);
        try {
            c.body();
        }
        catch (const AssertionAbort&) {
        }
        catch (const std::exception& e) {
            fail(__FILE__, __LINE__, std::string("uncaught exception: ") + e.what());
        }
        std::printf("[ %s ] %s\n", failures() == before ? " OK " : "FAIL", c.name.c_str());
        std::fflush(stdout);
        ++ran;
    }
    std::printf("%d cases, %d failed checks\n", ran, failures());
    return failures() ? 1 : 0;
}

}  // namespace harness

#define HX_CAT2(a, b) a##b
#define HX_CAT(a, b) HX_CAT2(a, b)

#define TEST_CASE(name, device)                                                                     \
    static void name();                                                                            \
    static harness::Registrar HX_CAT(reg_, name)(#name, name, device);                             \
    static void name()

// A template case run for float and double.
#define TYPED_CASE(name)                                                                            \
    template <class T>                                                                             \
    static void name();                                                                            \
    static harness::Registrar HX_CAT(regf_, name)(#name "<float>", name<float>, true);             \
    static harness::Registrar HX_CAT(regd_, name)(#name "<double>", name<double>, true);           \
    template <class T>                                                                             \
    static void name()

#define HX_CHECK(cond, text, fatal)                                                                 \
    do {                                                                                           \
        if (!(cond)) {                                                                             \
            harness::fail(__FILE__, __LINE__, text);                                               \
            if (fatal) throw harness::AssertionAbort{};                                            \
        }                                                                                          \
    } while (0)

#define EXPECT_TRUE(c) HX_CHECK((c), #c, false)
#define EXPECT_FALSE(c) HX_CHECK(!(c), "!" #c, false)
#define EXPECT_EQ(a, b) HX_CHECK((a) == (b), harness::describe(#a, #b, (a), (b)), false)
#define EXPECT_NE(a, b) HX_CHECK((a) != (b), harness::describe(#a, #b, (a), (b)), false)
#define EXPECT_LT(a, b) HX_CHECK((a) < (b), harness::describe(#a, #b, (a), (b)), false)
#define EXPECT_LE(a, b) HX_CHECK((a) <= (b), harness::describe(#a, #b, (a), (b)), false)
#define EXPECT_GT(a, b) HX_CHECK((a) > (b), harness::describe(#a, #b, (a), (b)), false)
#define EXPECT_GE(a, b) HX_CHECK((a) >= (b), harness::describe(#a, #b, (a), (b)), false)
#define EXPECT_NEAR(a, b, t)                                                                        \
    HX_CHECK(std::abs(double(a) - double(b)) <= double(t), harness::describe(#a, #b, double(a), double(b)), false)
#define ASSERT_TRUE(c) HX_CHECK((c), #c, true)
#define ASSERT_EQ(a, b) HX_CHECK((a) == (b), harness::describe(#a, #b, (a), (b)), true)
#define ASSERT_GE(a, b) HX_CHECK((a) >= (b), harness::describe(#a, #b, (a), (b)), true)
#define EXPECT_THROW(stmt, type)                                                                    \
    do {                                                                                           \
        bool caught_ = false;                                                                      \
        try {                                                                                      \
            stmt;                                                                                  \
        }                                                                                          \
        catch (const type&) {                                                                      \
            caught_ = true;                                                                        \
        }                                                                                          \
        catch (...) {                                                                              \
        }                                                                                          \
        HX_CHECK(caught_, #stmt " throws " #type, false);                                          \
    } while (0)
#define EXPECT_NO_THROW(stmt)                                                                       \
    do {                                                                                           \
        bool ok_ = true;                                                                           \
        try {                                                                                      \
            stmt;                                                                                  \
        }                                                                                          \
        catch (...) {                                                                              \
            ok_ = false;                                                                           \
        }                                                                                          \
        HX_CHECK(ok_, #stmt " does not throw", false);                                             \
    } while (0)
