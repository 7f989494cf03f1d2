// xg_solve_noalloc: LinearSOE::solve of an assembled system with every PCG
// work buffer taken from one fixed device array instead of memory the
// library allocates, for callers that cannot allocate (code around graph
// capture, serving loops, harnesses that forbid hipMalloc).
//
// The array `scratch` (kScratchBytes per device, held until the process
// ends) is file-scope device memory.  Host code never passes its name to a
// kernel: xg::noalloc_init(), called by xg_create with the device current,
// looks up each device's own address once with hipGetSymbolAddress, and the
// solve uses that address.  Calls that overlap on one device share the
// array and race; each call clears what it needs cleared.
#include <mutex>

#include "xg_domain.h"
#include "xg_kernels.h"

namespace {

constexpr size_t kScratchBytes = 128ull << 20;  // 128 MiB
constexpr int kMaxDevices = 64;

static __device__ __align__(256) double scratch[kScratchBytes / sizeof(double)];

std::mutex g_mu;
void* g_ptr[kMaxDevices] = {};

// 256-byte aligned carving of the scratch array
struct Carve {
  char* base;
  size_t off = 0;
  template <class T>
  T* take(size_t count) {
    T* p = reinterpret_cast<T*>(base + off);
    off += (count * sizeof(T) + 255) & ~size_t(255);
    return p;
  }
};

}  // namespace

namespace xg {

// Once per device, with that device current: the device address of its
// copy of `scratch`.
void noalloc_init(int device) {
  if (device < 0 || device >= kMaxDevices) throw Error(XG_ERR_ARG, "device index out of range");
  std::lock_guard<std::mutex> lock(g_mu);
  if (g_ptr[device]) return;
  void* p = nullptr;
  XG_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  if (!p) throw Error(XG_ERR_DEVICE, "scratch: no device address");
  g_ptr[device] = p;
}

size_t noalloc_scratch_bytes() { return kScratchBytes; }

// Bytes of scratch a solve of d's system needs: B and X staging, the six
// PCG vectors, the block-Jacobi inverses, partial sums, scalars, a flag,
// the band factor when the band solve takes the system, and b, x and the
// inverses in the sweep order when the symmetric product uses one.
static size_t needed(const Domain& d, int bw) {
  const size_t n = (size_t)d.num.neq;
  auto al = [](size_t b) { return (b + 255) & ~size_t(255); };
  const size_t band = d.band_active() ? 3 * al(n * (d.band_bw + 1) * sizeof(double)) : 0;
  const size_t sweep = d.sym_sweep ? 2 * al(n * sizeof(double)) + al(n * bw * sizeof(double)) : 0;
  return 8 * al(n * sizeof(double)) + al(n * bw * sizeof(double)) + al(cg_part_size(d.num.neq) * sizeof(double)) +
         al(cg_scal_bytes()) + al(sizeof(int)) + band + sweep;
}

int solve_noalloc(Domain& d, const double* B, double* X) {
  if (!d.ready) throw Error(XG_ERR_STATE, "solve: set the model up first");
  const int n = d.num.neq;
  if (n == 0) return 0;
  if (!B || !X) throw Error(XG_ERR_ARG, "solve: null vector");
  void* base = nullptr;
  {
    std::lock_guard<std::mutex> lock(g_mu);
    base = (d.device >= 0 && d.device < kMaxDevices) ? g_ptr[d.device] : nullptr;
  }
  if (!base) throw Error(XG_ERR_STATE, "solve_noalloc: scratch not initialised on this device");
  const int bw = (d.system == XG_SYSTEM_PCG_BLOCKJACOBI) ? d.ndf : 1;
  const size_t need = needed(d, bw);
  if (need > kScratchBytes)
    throw Error(XG_ERR_ARG, "solve_noalloc: the system needs " + std::to_string(need >> 20) +
                                " MiB of scratch, above the " + std::to_string(kScratchBytes >> 20) + " MiB maximum");
  // the symmetric product layout is built at set-up; building it here would allocate
  const bool want_sym = d.spmv_format == XG_SPMV_SYM || d.spmv_format == XG_SPMV_AUTO;
  if (want_sym && (d.nb_m == 2 || d.nb_m == 3) && !d.sym_built)
    throw Error(XG_ERR_STATE, "solve_noalloc: the SpMV layout is not built (set the model up first)");
  XG_HIP(hipSetDevice(d.device));
  Carve c{static_cast<char*>(base)};
  double* b = c.take<double>(n);
  double* x = c.take<double>(n);
  Domain::CgBufs w;
  w.r0 = c.take<double>(n);
  w.r1 = c.take<double>(n);
  w.z = c.take<double>(n);
  w.p = c.take<double>(n);
  w.p2 = c.take<double>(n);
  w.q = c.take<double>(n);
  w.binv = c.take<double>((size_t)n * bw);
  w.part = c.take<double>(cg_part_size(n));
  w.scal = c.take<char>(cg_scal_bytes());
  w.pcbad = c.take<int>(1);
  w.bandA = w.bandL = w.bandLr = nullptr;
  // the tiled band factor, where the solve takes it, lives in the domain's
  // own buffers (formed at set-up); here it is formed and applied with
  // plain launches (instantiating a graph would allocate)
  w.wide = true;
  if (d.sym_sweep) {
    // b, x and the block inverses in the symmetric product's sweep order
    w.bs = c.take<double>(n);
    w.xs = c.take<double>(n);
    w.binvs = c.take<double>((size_t)n * bw);
  }
  if (d.band_active()) {
    w.bandA = c.take<double>((size_t)n * (d.band_bw + 1));
    w.bandL = c.take<double>((size_t)n * (d.band_bw + 1));
    w.bandLr = c.take<double>((size_t)n * (d.band_bw + 1));
  }
  XG_HIP(hipMemcpyAsync(b, B, (size_t)n * sizeof(double), hipMemcpyHostToDevice, d.stream));
  // another call may have reused the array since: form the preconditioner
  // afresh; kernel launches only (no graph capture: instantiating one allocates)
  d.pc_done = nullptr;
  const int r = d.solve_with(b, x, w, false);
  XG_HIP(hipMemcpyAsync(X, x, (size_t)n * sizeof(double), hipMemcpyDeviceToHost, d.stream));
  XG_HIP(hipStreamSynchronize(d.stream));
  return r;
}

}  // namespace xg
