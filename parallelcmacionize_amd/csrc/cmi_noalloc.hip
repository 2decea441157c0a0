// cmi_noalloc.hip -- cmi_transport_packets_noalloc: the walk engine's
// explicit-list transport for callers that cannot allocate device memory.
//
// cmi_transport_packets stages the caller's host arrays through buffers it
// allocates per call.  This variant stages them through one fixed,
// file-scope device array instead (packet records, start subgrids, the three
// per-packet outputs and the transport counters), so a call allocates
// nothing.  The array holds at most kMaxPackets packets; each GPU the library
// runs on keeps its own copy until the process ends.  Every call on a GPU
// shares that GPU's copy: calls that overlap on one GPU (two host threads,
// two contexts on different streams) race and must be serialised by the
// caller.
#include <mutex>
#include <string>

#include "cmi_ctx.hpp"

namespace {

using cmi::Counters;
using cmi::PhotonRec;

constexpr size_t kMaxPackets = CMI_NOALLOC_MAX_PACKETS;

constexpr size_t align256(size_t v) { return (v + 255) & ~(size_t)255; }

// Layout of the scratch array (byte offsets, each 256-byte aligned).
constexpr size_t kOffRec = 0;
constexpr size_t kOffSub = align256(kOffRec + kMaxPackets * sizeof(PhotonRec));
constexpr size_t kOffStatus = align256(kOffSub + kMaxPackets * sizeof(uint32_t));
constexpr size_t kOffCross = align256(kOffStatus + kMaxPackets * sizeof(int32_t));
constexpr size_t kOffHops = align256(kOffCross + kMaxPackets * sizeof(uint32_t));
constexpr size_t kOffCnt = align256(kOffHops + kMaxPackets * sizeof(uint32_t));
constexpr size_t kScratchBytes = align256(kOffCnt + sizeof(Counters));
static_assert(sizeof(Counters) % sizeof(double) == 0,
              "the counters are cleared as doubles");

// The device memory of this entry point.  Host code never uses this name as
// an address (on the host it is only a stand-in for the symbol): the device
// address comes from hipGetSymbolAddress, per device.
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_mutex;
unsigned char *g_scratch[kMaxDevices] = {};

}  // namespace

namespace cmi {

// Resolve the scratch array's address on the context's device (made current
// here); done once per device, at the first context created on it.
int noalloc_setup(cmi_ctx *c) {
  if (c->device < 0 || c->device >= kMaxDevices)
    return fail(c, CMI_ERR_NODEVICE, "device index beyond the scratch table");
  std::lock_guard<std::mutex> lock(g_mutex);
  if (g_scratch[c->device]) return CMI_OK;
  CMI_TRY_HIP(c, hipSetDevice(c->device));
  void *ptr = nullptr;
  CMI_TRY_HIP(c, hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
  if (!ptr) return fail(c, CMI_ERR_HIP, "hipGetSymbolAddress gave no address");
  g_scratch[c->device] = static_cast<unsigned char *>(ptr);
  return CMI_OK;
}

}  // namespace cmi

extern "C" {

int cmi_transport_packets_noalloc(cmi_ctx *c, cmi_photon *p,
                                  const uint32_t *start_sub, size_t n,
                                  int32_t *status, uint32_t *crossings,
                                  uint32_t *hops, cmi_stats *stats) {
  using namespace cmi;
  if (!c) return CMI_ERR_ARG;
  if (n > kMaxPackets)
    return fail(c, CMI_ERR_ARG,
                "cmi_transport_packets_noalloc: at most " +
                    std::to_string(kMaxPackets) +
                    " packets per call (packets are independent: split the "
                    "list)");
  int rc = prepare_list_transport(c, p, start_sub, n);
  if (rc != CMI_OK) return rc;
  if (n == 0) {
    if (stats) *stats = cmi_stats{};
    return CMI_OK;
  }
  rc = noalloc_setup(c);
  if (rc != CMI_OK) return rc;
  unsigned char *base;
  {
    std::lock_guard<std::mutex> lock(g_mutex);
    base = g_scratch[c->device];
  }
  PhotonRec *dp = reinterpret_cast<PhotonRec *>(base + kOffRec);
  uint32_t *ds = reinterpret_cast<uint32_t *>(base + kOffSub);
  int32_t *dst = reinterpret_cast<int32_t *>(base + kOffStatus);
  uint32_t *dc = reinterpret_cast<uint32_t *>(base + kOffCross);
  uint32_t *dh = reinterpret_cast<uint32_t *>(base + kOffHops);
  Counters *cnt = reinterpret_cast<Counters *>(base + kOffCnt);
  CMI_TRY_HIP(c, hipMemcpyAsync(dp, p, n * sizeof(PhotonRec),
                                hipMemcpyHostToDevice, c->stream));
  CMI_TRY_HIP(c, hipMemcpyAsync(ds, start_sub, n * sizeof(uint32_t),
                                hipMemcpyHostToDevice, c->stream));
  // the previous call's counters are still there: clear them with a fill
  // kernel (no memset node, should a caller capture the stream)
  CMI_TRY_HIP(c, launch_fill(reinterpret_cast<double *>(cnt),
                             sizeof(Counters) / sizeof(double), 0., c->stream));
  CMI_TRY_HIP(c, hipEventRecord(c->ev0, c->stream));
  CMI_TRY_HIP(c, launch_walk_list(c->g, c->f, c->precision, dp, ds, n, dst, dc,
                                  dh, list_reemit_draws(c), cnt, c->stream));
  CMI_TRY_HIP(c, hipEventRecord(c->ev1, c->stream));
  CMI_TRY_HIP(c, hipMemcpyAsync(p, dp, n * sizeof(PhotonRec),
                                hipMemcpyDeviceToHost, c->stream));
  if (status)
    CMI_TRY_HIP(c, hipMemcpyAsync(status, dst, n * sizeof(int32_t),
                                  hipMemcpyDeviceToHost, c->stream));
  if (crossings)
    CMI_TRY_HIP(c, hipMemcpyAsync(crossings, dc, n * sizeof(uint32_t),
                                  hipMemcpyDeviceToHost, c->stream));
  if (hops)
    CMI_TRY_HIP(c, hipMemcpyAsync(hops, dh, n * sizeof(uint32_t),
                                  hipMemcpyDeviceToHost, c->stream));
  CMI_TRY_HIP(c, hipMemcpyAsync(c->h_cnt, cnt, sizeof(Counters),
                                hipMemcpyDeviceToHost, c->stream));
  CMI_TRY_HIP(c, hipStreamSynchronize(c->stream));
  float ms = 0.f;
  CMI_TRY_HIP(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
  fill_ctx_stats(c, stats, n, ms * 1e-3);
  if (c->h_cnt->errors)
    return fail(c, CMI_ERR_DIRECTION,
                "packets stopped on an invalid exit or walk bound");
  return CMI_OK;
}

}  // extern "C"
