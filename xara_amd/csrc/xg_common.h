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

namespace xg {

// Records the launches `body` issues on stream s into *graph and
// instantiates it into *exec.  If body (or the capture itself) fails, the
// capture is ended and the partial graph destroyed before the error
// propagates, so the stream never stays in capture mode.
template <class F>
void capture_graph(hipStream_t s, F&& body, hipGraph_t* graph, hipGraphExec_t* exec) {
  *graph = nullptr;
  *exec = nullptr;
  XG_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
  try {
    body();
  } catch (...) {
    hipGraph_t partial = nullptr;
    (void)hipStreamEndCapture(s, &partial);
    if (partial) (void)hipGraphDestroy(partial);
    throw;
  }
  XG_HIP(hipStreamEndCapture(s, graph));
  const hipError_t e = hipGraphInstantiate(exec, *graph, nullptr, nullptr, 0);
  if (e != hipSuccess) {
    (void)hipGraphDestroy(*graph);
    *graph = nullptr;
    *exec = nullptr;
    XG_HIP(e);
  }
}

// Destroys an instantiated graph after the work queued on stream s (which
// may include launches of it) has finished.
inline void destroy_graph(hipStream_t s, hipGraph_t& graph, hipGraphExec_t& exec) {
  if (!graph && !exec) return;
  if (s) (void)hipStreamSynchronize(s);
  if (exec) (void)hipGraphExecDestroy(exec);
  if (graph) (void)hipGraphDestroy(graph);
  exec = nullptr;
  graph = nullptr;
}

}  // namespace xg
