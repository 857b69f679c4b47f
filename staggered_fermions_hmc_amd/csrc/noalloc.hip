// Allocation-free variant of sfh_polyakov_correlators for callers that cannot
// allocate device memory (code under HIP graph capture, tight serving loops,
// harnesses that forbid hipMalloc): its work buffers live in one static
// device array of this library, `scratch`, whose per-device address is
// looked up once per device when a context is created on it.  Every call on
// one GPU shares that GPU's array, which keeps the last call's contents;
// calls through this library are serialised per device by a lock, and a
// call's results are in the caller's buffer when it returns.  Nothing in the
// array needs to start at zero: every byte a call reads was written by the
// same call.
#include <mutex>
#include <vector>

#include "internal.h"

namespace {

// Layout of the work area (bytes): Polyakov grid of up to 2^18 spatial
// sites, up to 65536 displacements, up to 256 r^2 values (offsets, block
// partials of up to 1024 blocks each, results).
constexpr size_t kGridBytes = (size_t(1) << 18) * sizeof(double2);  // 4 MiB
constexpr size_t kDispMax = 65536;
constexpr size_t kDispBytes = kDispMax * sizeof(int4);  // 1 MiB
constexpr int kCountMax = 256;
constexpr size_t kOffBytes = 4096;  // >= (kCountMax + 1) ints
constexpr int kBlocksMax = 1024;
constexpr size_t kPartBytes = 2 * sizeof(double) * (size_t)kCountMax * kBlocksMax;  // 4 MiB
constexpr size_t kOutBytes = 2 * sizeof(double) * kCountMax;
constexpr size_t kScratchBytes = kGridBytes + kDispBytes + kOffBytes + kPartBytes + kOutBytes;

constexpr int kMaxDevices = 64;
void* g_scratch[kMaxDevices] = {};
std::mutex g_dev_mu[kMaxDevices];
std::mutex g_init_mu;

}  // namespace

// The work area, one copy per GPU (file scope: the host-side name is only a
// handle; its device address comes from hipGetSymbolAddress).
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace sfh {

int noalloc_init(int device) {
  if (device < 0 || device >= kMaxDevices) return SFH_E_INVALID;
  std::lock_guard<std::mutex> lk(g_init_mu);
  if (g_scratch[device]) return SFH_OK;
  void* p = nullptr;
  // the caller (context creation) has made `device` current; ::scratch is
  // the device array (sfh::scratch is the host allocator of this library)
  if (hipGetSymbolAddress(&p, HIP_SYMBOL(::scratch)) != hipSuccess || !p) return SFH_E_HIP;
  g_scratch[device] = p;
  return SFH_OK;
}

}  // namespace sfh

extern "C" int sfh_polyakov_correlators_noalloc(sfh_ctx* c, const sfh_gauge* g, int r2_first, int count,
                                                double* out) {
  if (!c || !g || !out || r2_first < 1 || count < 1) return SFH_E_INVALID;
  std::lock_guard<std::recursive_mutex> guard(c->mu);
  hipSetDevice(c->device);
  if (c->device >= kMaxDevices || !g_scratch[c->device])
    return sfh::set_err(c, SFH_E_HIP, "no-alloc work area: its device address was not found at context creation");
  const int* L = g->g.L;
  const long ns = (long)L[0] * L[1] * L[2];
  if (count > kCountMax || ns > (1L << 18) || sfh::poly_corr_blocks(L) > kBlocksMax)
    return sfh::set_err(c, SFH_E_INVALID,
                        "no-alloc correlators: at most 256 values of r^2 and 2^18 spatial sites (" +
                            std::to_string(count) + " and " + std::to_string(ns) + " asked)");
  std::vector<int4> disp;
  std::vector<int> off;
  sfh::corr_tables(L, r2_first, count, disp, off);
  if (disp.size() > kDispMax)
    return sfh::set_err(c, SFH_E_INVALID,
                        "no-alloc correlators: " + std::to_string(disp.size()) + " displacements exceed 65536");
  unsigned char* base = static_cast<unsigned char*>(g_scratch[c->device]);
  double2* P = reinterpret_cast<double2*>(base);
  int4* d_disp = reinterpret_cast<int4*>(base + kGridBytes);
  int* d_off = reinterpret_cast<int*>(base + kGridBytes + kDispBytes);
  double* d_part = reinterpret_cast<double*>(base + kGridBytes + kDispBytes + kOffBytes);
  double* d_out = reinterpret_cast<double*>(base + kGridBytes + kDispBytes + kOffBytes + kPartBytes);
  std::lock_guard<std::mutex> dev_lock(g_dev_mu[c->device]);
  sfh::polyakov_grid_into(c, g->g, g->U, P);
  return sfh::poly_corr_run(c, L, P, disp, off, d_disp, d_off, d_part, d_out, out);
}
