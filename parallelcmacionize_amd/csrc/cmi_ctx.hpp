// cmi_ctx.hpp -- the context behind the opaque cmi_ctx handle of cmi.h,
// shared by the C-ABI layer and the engines' host code.
#pragma once

#include <algorithm>
#include <string>
#include <vector>

#include "../../include/cmi.h"
#include "cmi_internal.hpp"

namespace cmi {
struct TilesEngine;
}

struct cmi_ctx {
  int device = 0;
  int num_cus = 0;
  hipStream_t stream = nullptr;
  int precision = CMI_PREC_FP64;
  int engine = CMI_ENGINE_WALK;
  cmi_tuning tuning{};
  double box[6];
  int ncell[3], nsub[3], sc[3], tile[3];
  size_t nsubgrid = 0, cps = 0, ncell_total = 0;
  std::vector<double> anchors;   // [nsubgrid][3]
  std::vector<uint32_t> ngbs;    // [nsubgrid][27]
  cmi::GridDev g{};
  cmi::FieldsDev f{};
  double *d_anchors = nullptr;
  uint32_t *d_ngbs = nullptr;
  cmi::Counters *d_cnt = nullptr;
  cmi::Counters *h_cnt = nullptr;  // pinned
  cmi::SourcesDev *d_src = nullptr;
  cmi::SourcesDev *h_src = nullptr;  // pinned
  double *d_partial = nullptr;
  std::vector<double> h_partial;
  std::vector<double> src_pos, src_L;
  double sigma = 6.3e-22;  // cross section folded into kappa (FP32 mode)
  double reemission = 0.;  // reemission probability
  uint64_t reemit_seed = 0;       // key of the list transports' reemission draws
  uint32_t reemit_iteration = 0;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  double kernel_time_sum = 0.;
  uint64_t kernel_calls = 0;
  int kernel_timing = 1;  // record the tiles engine's per-round process events
  cmi::TilesEngine *tiles = nullptr;
  std::string err;
};

namespace cmi {

inline int fail(cmi_ctx *ctx, int code, const std::string &msg) {
  if (ctx) ctx->err = msg;
  return code;
}

#define CMI_TRY_HIP(ctx, call)                                              \
  do {                                                                      \
    hipError_t e_ = (call);                                                 \
    if (e_ != hipSuccess) {                                                 \
      return ::cmi::fail(ctx, CMI_ERR_HIP,                                  \
                         std::string(#call) + ": " + hipGetErrorString(e_)); \
    }                                                                       \
  } while (0)

// testDensitySubGrid_omp.cpp:2199-2208, 2240-2273 (split over sources with a
// random remainder, distribute() at 1650-1677): floor(L_i / L_tot * N) each,
// the remainder one each to the sources with the smallest Philox keys.
inline void source_counts(const std::vector<double> &L, uint64_t npackets,
                   uint64_t seed, uint32_t iteration,
                   std::vector<uint64_t> &counts) {
  const int nsrc = (int)L.size();
  double Ltot = 0.;
  for (double v : L) Ltot += v;
  counts.assign(nsrc, 0);
  uint64_t added = 0;
  for (int i = 0; i < nsrc; ++i) {
    counts[i] = (uint64_t)(L[i] / Ltot * (double)npackets);
    added += counts[i];
  }
  // the rounded shares can overshoot by a packet here and there: take the
  // excess from the largest shares, so that the split always sums to
  // npackets (the engines index packets by it)
  while (added > npackets) {
    int imax = 0;
    for (int i = 1; i < nsrc; ++i)
      if (counts[i] > counts[imax]) imax = i;
    --counts[imax];
    --added;
  }
  const uint64_t rem = npackets - added;
  if (rem == 0) return;
  std::vector<double> key(nsrc);
  std::vector<int> order(nsrc);
  for (int i = 0; i < nsrc; ++i) {
    const u32x4 o = philox4x32_10(u32x4{(uint32_t)i, 0u, iteration, 0x5EEDu},
                                  (uint32_t)seed, (uint32_t)(seed >> 32));
    key[i] = uniform_open(o.x, o.y);
    order[i] = i;
  }
  std::stable_sort(order.begin(), order.end(),
                   [&](int a, int b) { return key[a] < key[b]; });
  for (uint64_t r = 0; r < rem; ++r) counts[order[r % (uint64_t)nsrc]]++;
}

// Shared by the explicit-list transports (cmi_api.hip).
int prepare_list_transport(cmi_ctx *c, const cmi_photon *p,
                           const uint32_t *start_sub, size_t n);
void fill_ctx_stats(cmi_ctx *c, cmi_stats *st, uint64_t packets,
                    double seconds);

// Reemission draws of the walk engine: list transports (key of
// cmi_set_reemission_draws) and emitted ones (the transport's seed and
// iteration).
ReemitDraws list_reemit_draws(const cmi_ctx *c);

// Device address of the allocation-free entry point's scratch array on the
// context's device (cmi_noalloc.hip); resolved once per device.
int noalloc_setup(cmi_ctx *c);

// Tiles engine (cmi_tiles.hip).
int tiles_transport_packets(cmi_ctx *c, cmi_photon *p,
                            const uint32_t *start_sub, size_t n,
                            int32_t *status, uint32_t *crossings,
                            uint32_t *hops, cmi_stats *stats);
int tiles_transport(cmi_ctx *c, uint64_t npackets, uint64_t seed,
                    uint32_t iteration, double sigma, cmi_stats *stats);
void tiles_destroy(cmi_ctx *c);

}  // namespace cmi
