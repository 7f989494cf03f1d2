// Shared definitions for the gfx950 state-determination library.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>

namespace xg {

// Error raised by host code; converted to an int status at the C-ABI.
struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define XG_HIP(call)                                                                    \
  do {                                                                                  \
    hipError_t e_ = (call);                                                             \
    if (e_ != hipSuccess)                                                               \
      throw ::xg::Error(-3, std::string(#call) + ": " + hipGetErrorString(e_));         \
  } while (0)

// Number of doubles per material point in each SoA array of the J2 state.
// Tensor components of the symmetric plastic strain are stored once:
// 0:00 1:11 2:22 3:01 4:12 5:20 (J2ThreeDimensional::getStress order).
constexpr int kJ2Comp = 6;
// Upper triangle of the 6x6 (3D) tangent, row-major: (0,0)(0,1)..(0,5)(1,1)..
constexpr int kJ2Tan3d = 21;
// Upper triangle of the 3x3 plane-strain tangent.
constexpr int kJ2TanPs = 6;

// Triangular index of (a,b), a<=b, in an n x n upper triangle stored by rows.
__host__ __device__ constexpr int tri_index(int n, int a, int b) {
  return a * n - (a * (a - 1)) / 2 + (b - a);
}

// J2 material parameters, laid out for a __restrict__ table in global memory.
struct J2Param {
  double K, G, sig0, sigInf, delta, H, eta, tol;  // tol = 1e-10 * sig0
};

inline int ceil_div(long long a, long long b) { return (int)((a + b - 1) / b); }

}  // namespace xg
