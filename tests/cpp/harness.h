// Minimal test harness for the C++ host tests (no external framework).
#ifndef LGD_TEST_HARNESS_H
#define LGD_TEST_HARNESS_H

#include <cstdio>
#include <functional>
#include <string>
#include <vector>

namespace lgdtest {

struct Registry
{
    std::vector<std::pair<std::string, std::function<void()> > > tests;
    int failures = 0;
    static Registry& get()
    {
        static Registry r;
        return r;
    }
};

struct Add
{
    Add(const char *name, std::function<void()> f) { Registry::get().tests.push_back(std::make_pair(name, f)); }
};

inline void fail(const char *file, int line, const std::string& what)
{
    ++Registry::get().failures;
    std::printf("  FAILED %s:%d: %s\n", file, line, what.c_str());
    std::fflush(stdout);
}

inline int runAll(int argc, char **argv)
{
    Registry& r = Registry::get();
    int ran = 0;
    for (std::size_t i = 0; i < r.tests.size(); ++i) {
        if (argc > 1 && r.tests[i].first.find(argv[1]) == std::string::npos) {
            continue;
        }
        const int before = r.failures;
        std::printf("[test] %s\n", r.tests[i].first.c_str());
        std::fflush(stdout);
        try {
            r.tests[i].second();
        } catch (const std::exception& e) {
            fail(__FILE__, __LINE__, std::string("exception: ") + e.what());
        }
        std::printf("[%s] %s\n", r.failures == before ? "ok" : "FAIL", r.tests[i].first.c_str());
        std::fflush(stdout);
        ++ran;
    }
    std::printf("%d tests, %d failures\n", ran, r.failures);
    return r.failures ? 1 : 0;
}

}

#define LGD_TEST(name)                                               \
    static void name();                                              \
    static lgdtest::Add lgd_add_##name(#name, name);                 \
    static void name()

#define CHECK(cond)                                                  \
    do {                                                             \
        if (!(cond)) lgdtest::fail(__FILE__, __LINE__, #cond);       \
    } while (0)

#define CHECK_EQ(a, b)                                                                     \
    do {                                                                                   \
        if (!((a) == (b))) lgdtest::fail(__FILE__, __LINE__, std::string(#a " == " #b));  \
    } while (0)

#define CHECK_THROWS(stmt, type)                                     \
    do {                                                             \
        bool lgd_thrown_ = false;                                    \
        try {                                                        \
            stmt;                                                    \
        } catch (const type&) {                                      \
            lgd_thrown_ = true;                                      \
        }                                                            \
        if (!lgd_thrown_) lgdtest::fail(__FILE__, __LINE__, "expected " #type " from " #stmt); \
    } while (0)

#endif
