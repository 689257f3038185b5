// type.h -- scalar types and constants of the HiLeMMS API (the reference's
// Src/type.h:44-52 surface: Real is double, SizeType is std::size_t).
#ifndef TYPE_H
#define TYPE_H
#include <cassert>
#include <cmath>
#include <cstddef>
#include <limits>
#include <vector>

typedef double Real;
typedef std::size_t SizeType;
const Real PI{3.1415926535897932384626433832795};
const Real EPS{std::numeric_limits<Real>::epsilon()};
const Real BOLTZ{1.3806488e-23};
const Real ZERO{(Real)((int)0)};
const int xaxis = 0;
const int yaxis = 1;
#ifdef OPS_3D
const int zaxis = 2;
#endif

template <typename T>
inline void FreeArrayMemory(T* p) {
    delete[] p;
}
template <typename T>
inline bool EssentiallyEqual(const T* a, const T* b, const T eps) {
    return std::fabs(*a - *b) <=
           ((std::fabs(*a) > std::fabs(*b) ? std::fabs(*b) : std::fabs(*a)) * eps);
}
#endif
