// noalloc.hip -- pff_accumulate_constant_Q_noalloc: accumulate_constant_Q with
// every scratch buffer taken from one fixed device array, for callers that
// cannot allocate (code under HIP graph capture, tight serving loops).
//
// The array `scratch` below is a file-scope device variable of
// kNoAllocBytes.  Its host-side name is not a device address: host code takes
// the address per device once with hipGetSymbolAddress and uses only that
// pointer (kernel arguments, fills).  Every call on a device shares that
// device's array and finds the last call's contents in it, so the call clears
// what must start at zero (the device words) with a fill kernel; calls that
// overlap on one device (two streams, or two host threads) race on it.
#include <mutex>
#include <string>

#include "pff_internal.hpp"

// the fixed scratch: device words (256 bytes) + the tiled contraction's
// buffers, 48 bytes per cell plus 32 KiB per 64 x 64 tile (56 MiB for a
// 1024 x 1024 grid).  Each GPU the library runs on holds its own copy until
// the process ends.
static constexpr size_t kNoAllocBytes = (size_t)64 << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

namespace pff {
namespace {

constexpr int kMaxDevices = 64;
std::mutex g_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// device address of this device's copy of `scratch` (taken once per device,
// with that device current)
int scratch_ptr(int dev, void** out) {
  if (dev < 0 || dev >= kMaxDevices) return fail(PFF_EINVAL, "device index out of range");
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    // ::scratch: inside namespace pff the bare name is the scratch() function
    PFF_TRY(check_hip(hipGetSymbolAddress(&p, HIP_SYMBOL(::scratch)), "hipGetSymbolAddress(scratch)"));
    if (!p) return fail(PFF_EHIP, "hipGetSymbolAddress(scratch) returned NULL");
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return 0;
}

int ceil_log2_64(int64_t n) {
  int r = 0;
  while ((1ll << r) < n) ++r;
  return r;
}

}  // namespace
}  // namespace pff

using namespace pff;

extern "C" int pff_accumulate_constant_Q_noalloc(void* hip_stream, const pff_grid* g, const int32_t* rcv,
                                                 double value, int32_t area, float* Q, int32_t* converged) {
  GridD d;
  PFF_TRY(check_grid(g, &d));
  if (!rcv) return fail(PFF_EINVAL, "rcv is NULL");
  if (!Q) return fail(PFF_EINVAL, "Q is NULL");
  if (!converged) return fail(PFF_EINVAL, "converged is NULL");
  // layout: device words first, then each slot on a 256-byte boundary
  int slots[32];
  size_t bytes[32];
  int nslots = 0;
  constexpr size_t words = 256;
  static_assert(W_NWORDS * sizeof(int64_t) <= words, "device words do not fit their window");
  const size_t need = words + tiled_upstream_slots(d, slots, bytes, &nslots);
  if (need > kNoAllocBytes)
    return fail(PFF_EINVAL, "accumulate_constant_Q_noalloc: a " + std::to_string(d.nx) + " x " +
                                std::to_string(d.ny) + " grid needs " + std::to_string(need) +
                                " bytes of scratch; the fixed scratch holds " + std::to_string(kNoAllocBytes) +
                                " (48 bytes per cell + 32 KiB per 64 x 64 tile, e.g. 1024 x 1024 cells)");
  int dev = 0;
  PFF_TRY(check_hip(hipGetDevice(&dev), "hipGetDevice"));
  void* base = nullptr;
  PFF_TRY(scratch_ptr(dev, &base));
  // a context on the stack whose scratch slots are windows of the array
  pff_ctx c;
  c.device = dev;
  c.stream = (hipStream_t)hip_stream;
  c.no_alloc = true;
  unsigned char* p = (unsigned char*)base;
  c.dev_words = (int64_t*)p;
  size_t off = words;
  for (int k = 0; k < nslots; ++k) {
    c.buf[slots[k]] = p + off;
    c.cap[slots[k]] = bytes[k];
    off += (bytes[k] + 255) & ~(size_t)255;
  }
  // the words hold the last call's counters: clear them (a kernel, so that
  // the call can be captured in a graph)
  PFF_TRY(launch_fill_i64(&c, W_NWORDS, c.dev_words, 0));
  // Q.fill(value * (DX*DX if area else 1)) in double, stored as f32 (flowfields.py:279)
  const float q0 = (float)(value * (area ? (double)g->dx * (double)g->dx : 1.0));
  // all ceil(log2 N)+1 reduced rounds, no host read; *converged is written
  // on the device (0: the receivers hold a cycle and Q is not meaningful)
  const int cap = ceil_log2_64(d.N) + 1;
  const int unit_weights = 1;
  return tiled_upstream(&c, d, rcv, unit_weights, nullptr, nullptr, q0, Q, nullptr, nullptr, nullptr, nullptr, cap,
                        nullptr, converged);
}
