// Oracle arithmetic type: TEST INFRASTRUCTURE, never linked into the product.
//
// The oracle is written against `real`.  In the normal build it is `double`,
// so the restatement runs the reference's arithmetic in the reference's order.
// Built with -DOCP_COUNT_FLOPS it becomes a counting wrapper that tallies FP64
// add/sub/mul/div/sqrt (one flop each) and transcendental calls separately;
// that build measures the algorithmic flop count per cell of the flash.
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
inline real exp(real a) { ++g_tally.trans; return real(std::exp(a.v)); }
inline real log(real a) { ++g_tally.trans; return real(std::log(a.v)); }
inline real acos(real a) { ++g_tally.trans; return real(std::acos(a.v)); }
inline real cos(real a) { ++g_tally.trans; return real(std::cos(a.v)); }
inline real pow(real a, real b)
{
    // pow(x, 2) is folded to x*x by the compiler in the reference build.
    if (b.v == 2.0) { ++g_tally.flops; return real(a.v * a.v); }
    ++g_tally.trans;
    return real(std::pow(a.v, b.v));
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
