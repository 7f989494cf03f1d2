// Device buffers owned by a domain.
#pragma once
#include <vector>

#include "xg_common.h"

namespace xg {

template <class T>
struct DVec {
  T* p = nullptr;
  size_t n = 0;
  DVec() = default;
  DVec(const DVec&) = delete;
  DVec& operator=(const DVec&) = delete;
  ~DVec() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  void resize(size_t count) {
    if (count == n && p) return;
    release();
    n = count;
    if (count) XG_HIP(hipMalloc((void**)&p, count * sizeof(T)));
  }
  void zero(hipStream_t s) {
    if (n) XG_HIP(hipMemsetAsync(p, 0, n * sizeof(T), s));
  }
  void upload(const std::vector<T>& h, hipStream_t s) {
    resize(h.size());
    if (n) XG_HIP(hipMemcpyAsync(p, h.data(), n * sizeof(T), hipMemcpyHostToDevice, s));
  }
  void upload(const T* h, size_t count, hipStream_t s) {
    resize(count);
    if (n) XG_HIP(hipMemcpyAsync(p, h, n * sizeof(T), hipMemcpyHostToDevice, s));
  }
  void download(T* h, size_t count, hipStream_t s) const {
    if (count) XG_HIP(hipMemcpyAsync(h, p, count * sizeof(T), hipMemcpyDeviceToHost, s));
    XG_HIP(hipStreamSynchronize(s));
  }
  std::vector<T> download(hipStream_t s) const {
    std::vector<T> h(n);
    download(h.data(), n, s);
    return h;
  }
  void copy_from(const DVec<T>& o, hipStream_t s) {
    resize(o.n);
    if (n) XG_HIP(hipMemcpyAsync(p, o.p, n * sizeof(T), hipMemcpyDeviceToDevice, s));
  }
};

}  // namespace xg
