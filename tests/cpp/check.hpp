// minimal expectation macros for the C++ drop-in tests (no test framework in the image)
#pragma once
#include <cmath>
#include <cstdio>
#include <cstdlib>

namespace check
{
inline int& failures()
{
    static int f = 0;
    return f;
}
// float 4-ulp comparison of EXPECT_FLOAT_EQ, the reference tests' tolerance
inline bool floatEq(double a, double b)
{
    float fa = float(a), fb = float(b);
    if (fa == fb) return true;
    float diff = std::fabs(fa - fb), mag = std::fmax(std::fabs(fa), std::fabs(fb));
    return diff <= mag * 4.0f * 1.1920929e-07f;
}
}  // namespace check

#define EXPECT_TRUE(c)                                                   \
    do                                                                   \
    {                                                                    \
        if (!(c))                                                        \
        {                                                                \
            std::printf("FAIL %s:%d: %s\n", __FILE__, __LINE__, #c);     \
            ++check::failures();                                         \
        }                                                                \
    } while (0)
#define EXPECT_FALSE(c) EXPECT_TRUE(!(c))
#define EXPECT_LE(a, b) EXPECT_TRUE((a) <= (b))
#define EXPECT_EQ(a, b)                                                                           \
    do                                                                                            \
    {                                                                                             \
        auto va_ = (a);                                                                           \
        auto vb_ = (b);                                                                           \
        if (!(va_ == vb_))                                                                        \
        {                                                                                         \
            std::printf("FAIL %s:%d: %s == %s (%lld vs %lld)\n", __FILE__, __LINE__, #a, #b,      \
                        (long long)va_, (long long)vb_);                                          \
            ++check::failures();                                                                  \
        }                                                                                         \
    } while (0)
#define EXPECT_FLOAT_EQ(a, b)                                                                          \
    do                                                                                                 \
    {                                                                                                  \
        double va_ = (a), vb_ = (b);                                                                   \
        if (!check::floatEq(va_, vb_))                                                                 \
        {                                                                                              \
            std::printf("FAIL %s:%d: %s ~= %s (%.10g vs %.10g)\n", __FILE__, __LINE__, #a, #b, va_, vb_); \
            ++check::failures();                                                                       \
        }                                                                                              \
    } while (0)
#define EXPECT_NEAR(a, b, tol)                                                                         \
    do                                                                                                 \
    {                                                                                                  \
        double va_ = (a), vb_ = (b);                                                                   \
        if (!(std::fabs(va_ - vb_) <= (tol)))                                                          \
        {                                                                                              \
            std::printf("FAIL %s:%d: |%s - %s| <= %g (%.10g vs %.10g)\n", __FILE__, __LINE__, #a, #b,   \
                        double(tol), va_, vb_);                                                        \
            ++check::failures();                                                                       \
        }                                                                                              \
    } while (0)
#define CHECK_DONE()                                                     \
    do                                                                   \
    {                                                                    \
        std::printf("%d failure(s)\n", check::failures());               \
        return check::failures() == 0 ? 0 : 1;                           \
    } while (0)
