// Oracle arithmetic type: TEST INFRASTRUCTURE, never linked into the product.
//
// The oracle is written against `real`.  In the normal build it is `double`,
// so the restatement runs the reference's arithmetic in the reference's order.
// Built with -DOCP_COUNT_FLOPS it becomes a counting wrapper that tallies FP64
// add/sub/mul/div/sqrt (one flop each) and transcendental calls separately;
// that build measures the algorithmic flop count per cell of the flash.
// Built with -DOCP_COUNT_FLOPS -DOCP_ULP_NOISE it also perturbs every
// transcendental by up to one ulp (the conditioning of a cell's phase count
// with respect to libm rounding).
#ifndef ORACLE_REAL_HPP
#define ORACLE_REAL_HPP

#include <cmath>
#include <cstdint>

namespace oracle {

#ifdef OCP_COUNT_FLOPS

struct FlopTally {
    uint64_t flops = 0;  // add, sub, mul, div, sqrt, neg excluded
    uint64_t trans = 0;  // exp, log, pow, acos, cos
    // the parts spent in PhaseStable and in PhaseSplit
    uint64_t staFlops = 0, staTrans = 0, splitFlops = 0, splitTrans = 0;
};
extern thread_local FlopTally g_tally;

#ifdef OCP_ULP_NOISE
// Test-only libm stand-in: every transcendental result moves by one unit in
// the last place, up, down or not at all, drawn from a seeded stream.  A
// correct FP64 libm other than glibc's (the GPU's, for one) differs from it
// by about this much; a cell whose phase count changes under this noise has
// its phase count decided by libm rounding.
extern thread_local uint64_t g_noise;
inline double ulpNoise(double v)
{
    g_noise ^= g_noise << 13;
    g_noise ^= g_noise >> 7;
    g_noise ^= g_noise << 17;
    const int r = static_cast<int>(g_noise % 3);
    return r == 0 ? v : std::nextafter(v, r == 1 ? HUGE_VAL : -HUGE_VAL);
}
#else
inline double ulpNoise(double v) { return v; }
#endif

struct real {
    double v;
    real() : v(0) {}
    real(double x) : v(x) {}
    explicit operator double() const { return v; }
    real& operator+=(const real& o) { ++g_tally.flops; v += o.v; return *this; }
    real& operator-=(const real& o) { ++g_tally.flops; v -= o.v; return *this; }
    real& operator*=(const real& o) { ++g_tally.flops; v *= o.v; return *this; }
    real& operator/=(const real& o) { ++g_tally.flops; v /= o.v; return *this; }
    real operator-() const { return real(-v); }
};
inline real operator+(real a, real b) { ++g_tally.flops; return real(a.v + b.v); }
inline real operator-(real a, real b) { ++g_tally.flops; return real(a.v - b.v); }
inline real operator*(real a, real b) { ++g_tally.flops; return real(a.v * b.v); }
inline real operator/(real a, real b) { ++g_tally.flops; return real(a.v / b.v); }
inline bool operator<(real a, real b) { return a.v < b.v; }
inline bool operator>(real a, real b) { return a.v > b.v; }
inline bool operator<=(real a, real b) { return a.v <= b.v; }
inline bool operator>=(real a, real b) { return a.v >= b.v; }
inline bool operator==(real a, real b) { return a.v == b.v; }
inline bool operator!=(real a, real b) { return a.v != b.v; }
inline real sqrt(real a) { ++g_tally.flops; return real(std::sqrt(a.v)); }
inline real fabs(real a) { return real(std::fabs(a.v)); }
inline real exp(real a) { ++g_tally.trans; return real(ulpNoise(std::exp(a.v))); }
inline real log(real a) { ++g_tally.trans; return real(ulpNoise(std::log(a.v))); }
inline real acos(real a) { ++g_tally.trans; return real(ulpNoise(std::acos(a.v))); }
inline real cos(real a) { ++g_tally.trans; return real(ulpNoise(std::cos(a.v))); }
inline real pow(real a, real b)
{
    // pow(x, 2) is folded to x*x by the compiler in the reference build.
    if (b.v == 2.0) { ++g_tally.flops; return real(a.v * a.v); }
    ++g_tally.trans;
    return real(ulpNoise(std::pow(a.v, b.v)));
}
inline bool isfinite(real a) { return std::isfinite(a.v); }
inline double todbl(real a) { return a.v; }
inline float tofloat(real a) { return static_cast<float>(a.v); }

#else

typedef double real;
using std::isfinite;
inline double todbl(double a) { return a; }
inline float tofloat(double a) { return static_cast<float>(a); }

#endif

}  // namespace oracle

#endif
