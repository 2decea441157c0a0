// cmi_api.hip -- the C ABI (include/cmi.h): context, geometry and neighbour
// tables, host<->device layout, and the host side of every entry point.
#include <string.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/cmi.h"
#include "cmi_ctx.hpp"

using namespace cmi;


namespace {

const int kIonizeBlocks = 1024;

// Side offsets in TravelDirection order (TravelDirections.hpp:46-75).
const int kOff[27][3] = {
    {0, 0, 0},   {1, 1, 1},   {1, 1, -1},  {1, -1, 1},  {1, -1, -1},
    {-1, 1, 1},  {-1, 1, -1}, {-1, -1, 1}, {-1, -1, -1}, {0, 1, 1},
    {0, 1, -1},  {0, -1, 1},  {0, -1, -1}, {1, 0, 1},   {1, 0, -1},
    {-1, 0, 1},  {-1, 0, -1}, {1, 1, 0},   {1, -1, 0},  {-1, 1, 0},
    {-1, -1, 0}, {1, 0, 0},   {-1, 0, 0},  {0, 1, 0},   {0, -1, 0},
    {0, 0, 1},   {0, 0, -1}};

int dir_from_offset(int ox, int oy, int oz) {
  for (int d = 0; d < 27; ++d)
    if (kOff[d][0] == ox && kOff[d][1] == oy && kOff[d][2] == oz) return d;
  return -1;
}

int ilog2_exact(int v) {
  int s = 0;
  while ((1 << s) < v) ++s;
  return ((1 << s) == v) ? s : -1;
}

// Device cell index of (subgrid, ix, iy, iz): mirrors cell_addr().
inline size_t dev_index(const cmi_ctx *c, size_t sub, int ix, int iy, int iz) {
  const GridDev &g = c->g;
  const int tx = ix >> g.tshift[0], ty = iy >> g.tshift[1],
            tz = iz >> g.tshift[2];
  const int lx = ix - (tx << g.tshift[0]), ly = iy - (ty << g.tshift[1]),
            lz = iz - (tz << g.tshift[2]);
  const size_t t = ((size_t)tx * g.ntile[1] + ty) * g.ntile[2] + tz;
  const size_t l = ((((size_t)lx << g.tshift[1]) + ly) << g.tshift[2]) + lz;
  return sub * (size_t)g.cps + t * (size_t)g.tile_cells + l;
}

// Reference order (subgrid-major, ix*ny*nz+iy*nz+iz) <-> device order.
template <class T>
void to_device_order(const cmi_ctx *c, const T *ref, std::vector<T> &dev) {
  dev.resize(c->ncell_total);
  size_t r = 0;
  for (size_t s = 0; s < c->nsubgrid; ++s)
    for (int ix = 0; ix < c->sc[0]; ++ix)
      for (int iy = 0; iy < c->sc[1]; ++iy)
        for (int iz = 0; iz < c->sc[2]; ++iz) dev[dev_index(c, s, ix, iy, iz)] = ref[r++];
}

template <class T>
void from_device_order(const cmi_ctx *c, const std::vector<T> &dev, T *ref) {
  size_t r = 0;
  for (size_t s = 0; s < c->nsubgrid; ++s)
    for (int ix = 0; ix < c->sc[0]; ++ix)
      for (int iy = 0; iy < c->sc[1]; ++iy)
        for (int iz = 0; iz < c->sc[2]; ++iz) ref[r++] = dev[dev_index(c, s, ix, iy, iz)];
}

// Largest power of two <= cap that divides v.
int default_tile(int v, int cap) {
  int t = 1;
  while (t * 2 <= cap && v % (t * 2) == 0) t *= 2;
  return t;
}

int set_tile(cmi_ctx *c, const int tile[3]) {
  for (int a = 0; a < 3; ++a) {
    const int s = ilog2_exact(tile[a]);
    // at most 64 cells per axis and 8192 per tile: the tiles engine keeps a
    // flight's cell (as a byte offset) and per-axis step counts in one
    // 32-bit word
    if (s < 0 || tile[a] > c->sc[a] || c->sc[a] % tile[a] != 0 || tile[a] > 64)
      return fail(c, CMI_ERR_ARG,
                  "tile sizes must be powers of two up to 64 dividing the subgrid");
  }
  if (tile[0] * tile[1] * tile[2] > 8192)
    return fail(c, CMI_ERR_ARG, "at most 8192 cells per tile");
  for (int a = 0; a < 3; ++a) {
    c->tile[a] = tile[a];
    c->g.tshift[a] = ilog2_exact(tile[a]);
    c->g.ntile[a] = c->sc[a] / tile[a];
  }
  c->g.tile_cells = tile[0] * tile[1] * tile[2];
  return CMI_OK;
}

double delta0(const cmi_ctx *c) { return c->g.cell_size[0]; }

int upload_sources(cmi_ctx *c, uint64_t npackets, uint64_t seed,
                   uint32_t iteration) {
  if (c->src_L.empty()) return fail(c, CMI_ERR_STATE, "no sources set");
  std::vector<uint64_t> counts;
  source_counts(c->src_L, npackets, seed, iteration, counts);
  SourcesDev &S = *c->h_src;
  S.nsrc = (int)c->src_L.size();
  uint64_t acc = 0;
  for (int s = 0; s < S.nsrc; ++s) {
    acc += counts[s];
    S.cum[s] = acc;
    unsigned int si[3];
    for (int a = 0; a < 3; ++a) {
      S.pos[s][a] = c->src_pos[3 * s + a];
      si[a] = (unsigned int)((S.pos[s][a] - c->box[a]) / c->box[3 + a] *
                             c->nsub[a]);
      // a source a rounding step below the upper face of the box would
      // otherwise index one subgrid past the end
      if (si[a] >= (unsigned int)c->nsub[a]) si[a] = c->nsub[a] - 1;
    }
    S.start_sub[s] = si[0] * c->nsub[1] * c->nsub[2] + si[1] * c->nsub[2] + si[2];
  }
  CMI_TRY_HIP(c, hipMemcpyAsync(c->d_src, c->h_src, sizeof(SourcesDev),
                                hipMemcpyHostToDevice, c->stream));
  return CMI_OK;
}

int check_packets_fp32(cmi_ctx *c, const cmi_photon *p, size_t n) {
  for (size_t i = 0; i < n; ++i) {
    if (p[i].sigma != c->sigma || p[i].weight != 1.)
      return fail(c, CMI_ERR_ARG,
                  "FP32 mode folds sigma into kappa and assumes weight 1: every "
                  "packet must carry the context cross section and weight 1");
  }
  return CMI_OK;
}

int ensure_sigma(cmi_ctx *c, double sigma) {
  if (c->precision == CMI_PREC_FP32 && sigma != c->sigma) {
    c->sigma = sigma;
    CMI_TRY_HIP(c, launch_kappa(c->f, c->ncell_total, c->sigma * delta0(c),
                                c->stream));
  }
  return CMI_OK;
}

void fill_stats(cmi_ctx *c, cmi_stats *st, uint64_t packets, double seconds,
                uint32_t rounds) {
  if (!st) return;
  st->reemitted = c->h_cnt->reemitted;
  st->walkout_crossings = 0;  // the walk engine has no walk-out launch
  st->walkout_hops = 0;
  st->packets = packets;
  st->crossings = c->h_cnt->crossings;
  st->hops = c->h_cnt->hops;
  st->absorbed = c->h_cnt->absorbed;
  st->escaped = c->h_cnt->escaped;
  st->errors = c->h_cnt->errors;
  st->rounds = rounds;
  st->reserved = 0;
  st->seconds = seconds;
}

}  // namespace

namespace cmi {
// Checks and set-up shared by the explicit-list transports: arguments, start
// subgrids in range, and in FP32 mode the packets' cross section folded into
// kappa and weight 1.
int prepare_list_transport(cmi_ctx *c, const cmi_photon *p,
                           const uint32_t *start_sub, size_t n) {
  if (!c || (n && (!p || !start_sub))) return CMI_ERR_ARG;
  for (size_t i = 0; i < n; ++i)
    if (start_sub[i] >= c->nsubgrid)
      return fail(c, CMI_ERR_ARG, "start subgrid out of range");
  if (n == 0) return CMI_OK;
  (void)hipSetDevice(c->device);
  if (c->precision == CMI_PREC_FP32) {
    int rc = ensure_sigma(c, p[0].sigma);
    if (rc == CMI_OK) rc = check_packets_fp32(c, p, n);
    if (rc != CMI_OK) return rc;
  }
  return CMI_OK;
}

void fill_ctx_stats(cmi_ctx *c, cmi_stats *st, uint64_t packets,
                    double seconds) {
  fill_stats(c, st, packets, seconds, 0);
}

ReemitDraws list_reemit_draws(const cmi_ctx *c) {
  ReemitDraws re{};
  re.p = c->reemission;
  re.key0 = (uint32_t)c->reemit_seed;
  re.key1 = (uint32_t)(c->reemit_seed >> 32);
  re.iteration = c->reemit_iteration;
  return re;
}

// n_H of a context in the reference layout (for the result writers)
int get_density_ref(cmi_ctx *c, double *nH) {
  (void)hipSetDevice(c->device);
  CMI_TRY_HIP(c, hipStreamSynchronize(c->stream));
  std::vector<double> d(c->ncell_total);
  CMI_TRY_HIP(c, hipMemcpy(d.data(), c->f.n, d.size() * sizeof(double),
                           hipMemcpyDeviceToHost));
  from_device_order(c, d, nH);
  return CMI_OK;
}
}  // namespace cmi

extern "C" {

int cmi_output_to_input_direction(int d) {
  if (d < 0 || d >= 27) return -1;
  return dir_from_offset(-kOff[d][0], -kOff[d][1], -kOff[d][2]);
}

int cmi_direction_from_mask(int mask) {
  if (mask < 0 || mask > 63) return -1;
  int o[3];
  for (int a = 0; a < 3; ++a) {
    const int hi = (mask >> (5 - 2 * a)) & 1, lo = (mask >> (4 - 2 * a)) & 1;
    if (hi && lo) return -1;
    o[a] = hi - lo;
  }
  return dir_from_offset(o[0], o[1], o[2]);
}

const char *cmi_last_error(const cmi_ctx *ctx) {
  return ctx ? ctx->err.c_str() : "null context";
}

int cmi_create(const double box[6], const int ncell[3], const int nsub[3],
               int precision, cmi_ctx **out) {
  if (!out || !box || !ncell || !nsub) return CMI_ERR_ARG;
  *out = nullptr;
  if (precision != CMI_PREC_FP64 && precision != CMI_PREC_FP32)
    return CMI_ERR_ARG;
  for (int a = 0; a < 3; ++a) {
    if (ncell[a] <= 0 || nsub[a] <= 0 || ncell[a] % nsub[a] != 0 ||
        !(box[3 + a] > 0.))
      return CMI_ERR_ARG;
    if (ncell[a] / nsub[a] > 1024) return CMI_ERR_ARG;
  }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0)
    return CMI_ERR_NODEVICE;
  cmi_ctx *c = new cmi_ctx();
  c->precision = precision;
  hipGetDevice(&c->device);
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, c->device) != hipSuccess) {
    delete c;
    return CMI_ERR_HIP;
  }
  c->num_cus = prop.multiProcessorCount;
  memcpy(c->box, box, sizeof(c->box));
  for (int a = 0; a < 3; ++a) {
    c->ncell[a] = ncell[a];
    c->nsub[a] = nsub[a];
    c->sc[a] = ncell[a] / nsub[a];
  }
  c->nsubgrid = (size_t)nsub[0] * nsub[1] * nsub[2];
  c->cps = (size_t)c->sc[0] * c->sc[1] * c->sc[2];
  c->ncell_total = c->cps * c->nsubgrid;
  // geometry exactly as testDensitySubGrid_omp.cpp:1805-1829 and
  // DensitySubGrid.hpp:516-527 compute it
  const double side[3] = {box[3] / nsub[0], box[4] / nsub[1], box[5] / nsub[2]};
  GridDev &g = c->g;
  for (int a = 0; a < 3; ++a) {
    g.nsub[a] = nsub[a];
    g.sc[a] = c->sc[a];
    g.cell_size[a] = side[a] / c->sc[a];
    g.inv_cell_size[a] = c->sc[a] / side[a];
  }
  for (int a = 0; a < 3; ++a) g.unit_scale[a] = g.cell_size[0] / g.cell_size[a];
  g.cps = (int)c->cps;
  g.nsubgrid = (int)c->nsubgrid;
  int tile[3];
  for (int a = 0; a < 3; ++a) tile[a] = c->sc[a] >= 1 ? default_tile(c->sc[a], 16) : 1;
  if (set_tile(c, tile) != CMI_OK) {
    delete c;
    return CMI_ERR_ARG;
  }
  c->anchors.resize(3 * c->nsubgrid);
  c->ngbs.assign(27 * c->nsubgrid, CMI_NEIGHBOUR_OUTSIDE);
  for (size_t index = 0; index < c->nsubgrid; ++index) {
    const int ix = (int)(index / (nsub[1] * nsub[2]));
    const int iy = (int)((index - (size_t)ix * nsub[1] * nsub[2]) / nsub[2]);
    const int iz = (int)(index - (size_t)ix * nsub[1] * nsub[2] - (size_t)iy * nsub[2]);
    c->anchors[3 * index + 0] = box[0] + ix * side[0];
    c->anchors[3 * index + 1] = box[1] + iy * side[1];
    c->anchors[3 * index + 2] = box[2] + iz * side[2];
    for (int nix = -1; nix < 2; ++nix)
      for (int niy = -1; niy < 2; ++niy)
        for (int niz = -1; niz < 2; ++niz) {
          const int cix = ix + nix, ciy = iy + niy, ciz = iz + niz;
          if (cix >= 0 && cix < nsub[0] && ciy >= 0 && ciy < nsub[1] &&
              ciz >= 0 && ciz < nsub[2]) {
            c->ngbs[27 * index + dir_from_offset(nix, niy, niz)] =
                (uint32_t)(cix * nsub[1] * nsub[2] + ciy * nsub[2] + ciz);
          }
        }
  }
  hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipMalloc(&c->d_anchors, c->anchors.size() * sizeof(double));
  if (e == hipSuccess) e = hipMalloc(&c->d_ngbs, c->ngbs.size() * sizeof(uint32_t));
  if (e == hipSuccess) e = hipMalloc(&c->f.n, c->ncell_total * sizeof(double));
  if (e == hipSuccess) e = hipMalloc(&c->f.xn, c->ncell_total * sizeof(double));
  if (e == hipSuccess) e = hipMalloc(&c->f.J, c->ncell_total * sizeof(double));
  if (e == hipSuccess && precision == CMI_PREC_FP32)
    e = hipMalloc(&c->f.kappa, c->ncell_total * sizeof(float));
  if (e == hipSuccess) e = hipMalloc(&c->d_cnt, sizeof(Counters));
  if (e == hipSuccess) e = hipHostMalloc(&c->h_cnt, sizeof(Counters));
  if (e == hipSuccess) e = hipMalloc(&c->d_src, sizeof(SourcesDev));
  if (e == hipSuccess) e = hipHostMalloc(&c->h_src, sizeof(SourcesDev));
  if (e == hipSuccess) e = hipMalloc(&c->d_partial, kIonizeBlocks * sizeof(double));
  if (e == hipSuccess) e = hipEventCreate(&c->ev0);
  if (e == hipSuccess) e = hipEventCreate(&c->ev1);
  if (e != hipSuccess) {
    cmi_destroy(c);
    return CMI_ERR_ALLOC;
  }
  c->h_partial.resize(kIonizeBlocks);
  g.anchors = c->d_anchors;
  g.ngbs = c->d_ngbs;
  e = hipMemcpy(c->d_anchors, c->anchors.data(), c->anchors.size() * sizeof(double),
                hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = hipMemcpy(c->d_ngbs, c->ngbs.data(), c->ngbs.size() * sizeof(uint32_t),
                  hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    cmi_destroy(c);
    return CMI_ERR_HIP;
  }
  // DensitySubGrid.hpp:544-551: n = 1e8, x_n = 1e-6, J = 0
  std::vector<double> n0(c->ncell_total, 1.e8), x0(c->ncell_total, 1.e-6);
  int rc = cmi_set_fields(c, n0.data(), x0.data());
  if (rc == CMI_OK) rc = cmi_reset_intensity(c);
  // the allocation-free entry point's scratch address on this device
  if (rc == CMI_OK) rc = noalloc_setup(c);
  if (rc != CMI_OK) {
    cmi_destroy(c);
    return rc;
  }
  *out = c;
  return CMI_OK;
}

void cmi_destroy(cmi_ctx *c) {
  if (!c) return;
  (void)hipSetDevice(c->device);
  if (c->stream) hipStreamSynchronize(c->stream);
  tiles_destroy(c);
  hipFree(c->d_anchors);
  hipFree(c->d_ngbs);
  hipFree(c->f.n);
  hipFree(c->f.xn);
  hipFree(c->f.J);
  hipFree(c->f.kappa);
  hipFree(c->d_cnt);
  if (c->h_cnt) hipHostFree(c->h_cnt);
  hipFree(c->d_src);
  if (c->h_src) hipHostFree(c->h_src);
  hipFree(c->d_partial);
  if (c->ev0) hipEventDestroy(c->ev0);
  if (c->ev1) hipEventDestroy(c->ev1);
  if (c->stream) hipStreamDestroy(c->stream);
  delete c;
}

int cmi_set_engine(cmi_ctx *c, int engine, const cmi_tuning *t) {
  if (!c) return CMI_ERR_ARG;
  if (engine != CMI_ENGINE_WALK && engine != CMI_ENGINE_TILES)
    return fail(c, CMI_ERR_ARG, "unknown engine");
  (void)hipSetDevice(c->device);
  if (t) {
    if (t->tile[0] || t->tile[1] || t->tile[2]) {
      int tile[3];
      for (int a = 0; a < 3; ++a) tile[a] = t->tile[a] ? t->tile[a] : c->tile[a];
      if (tile[0] != c->tile[0] || tile[1] != c->tile[1] || tile[2] != c->tile[2]) {
        // re-layout the fields for the new tile shape
        std::vector<double> n(c->ncell_total), x(c->ncell_total), J(c->ncell_total);
        int rc = cmi_get_fields(c, J.data(), x.data());
        if (rc != CMI_OK) return rc;
        std::vector<double> nd(c->ncell_total);
        CMI_TRY_HIP(c, hipMemcpy(nd.data(), c->f.n, c->ncell_total * sizeof(double),
                                 hipMemcpyDeviceToHost));
        from_device_order(c, nd, n.data());
        const int old_tile[3] = {c->tile[0], c->tile[1], c->tile[2]};
        rc = set_tile(c, tile);
        if (rc != CMI_OK) {
          set_tile(c, old_tile);
          return rc;
        }
        rc = cmi_set_fields(c, n.data(), x.data());
        if (rc == CMI_OK) rc = cmi_set_intensity(c, J.data());
        if (rc != CMI_OK) return rc;
      }
    }
    c->tuning = *t;
  }
  c->engine = engine;
  return CMI_OK;
}

int cmi_set_fields(cmi_ctx *c, const double *nH, const double *xn) {
  if (!c) return CMI_ERR_ARG;
  (void)hipSetDevice(c->device);
  // the copies below go through the null stream, which does not wait for
  // the context's non-blocking stream: work still queued there (a
  // restriction or prolongation, a warm start) must not land after them
  CMI_TRY_HIP(c, hipStreamSynchronize(c->stream));
  std::vector<double> d;
  if (nH) {
    to_device_order(c, nH, d);
    CMI_TRY_HIP(c, hipMemcpy(c->f.n, d.data(), d.size() * sizeof(double),
                             hipMemcpyHostToDevice));
  }
  if (xn) {
    to_device_order(c, xn, d);
    CMI_TRY_HIP(c, hipMemcpy(c->f.xn, d.data(), d.size() * sizeof(double),
                             hipMemcpyHostToDevice));
  }
  if (c->precision == CMI_PREC_FP32) {
    CMI_TRY_HIP(c, launch_kappa(c->f, c->ncell_total, c->sigma * delta0(c),
                                c->stream));
    CMI_TRY_HIP(c, hipStreamSynchronize(c->stream));
  }
  return CMI_OK;
}

int cmi_set_intensity(cmi_ctx *c, const double *J) {
  if (!c || !J) return CMI_ERR_ARG;
  (void)hipSetDevice(c->device);
  CMI_TRY_HIP(c, hipStreamSynchronize(c->stream));  // as in cmi_set_fields
  std::vector<double> d;
  to_device_order(c, J, d);
  if (c->precision == CMI_PREC_FP32) {
    const double inv = 1. / (c->sigma * delta0(c));
    for (double &v : d) v *= inv;
  }
  CMI_TRY_HIP(c, hipMemcpy(c->f.J, d.data(), d.size() * sizeof(double),
                           hipMemcpyHostToDevice));
  return CMI_OK;
}

int cmi_get_fields(cmi_ctx *c, double *J, double *xn) {
  if (!c) return CMI_ERR_ARG;
  (void)hipSetDevice(c->device);
  CMI_TRY_HIP(c, hipStreamSynchronize(c->stream));
  std::vector<double> d(c->ncell_total);
  if (J) {
    CMI_TRY_HIP(c, hipMemcpy(d.data(), c->f.J, d.size() * sizeof(double),
                             hipMemcpyDeviceToHost));
    if (c->precision == CMI_PREC_FP32) {
      const double s = c->sigma * delta0(c);
      for (double &v : d) v *= s;
    }
    from_device_order(c, d, J);
  }
  if (xn) {
    CMI_TRY_HIP(c, hipMemcpy(d.data(), c->f.xn, d.size() * sizeof(double),
                             hipMemcpyDeviceToHost));
    from_device_order(c, d, xn);
  }
  return CMI_OK;
}

int cmi_reset_intensity(cmi_ctx *c) {
  if (!c) return CMI_ERR_ARG;
  (void)hipSetDevice(c->device);
  CMI_TRY_HIP(c, launch_fill(c->f.J, c->ncell_total, 0., c->stream));
  CMI_TRY_HIP(c, hipStreamSynchronize(c->stream));
  return CMI_OK;
}

}  // extern "C"

namespace {
// one subgrid, explicit packets, one of the reference's three walk variants
int walk_batch(cmi_ctx *c, uint32_t subgrid, int entry_dir, cmi_photon *p, size_t n,
               int32_t *exit_dir, uint32_t *crossings, int kind) {
  if (!c) return CMI_ERR_ARG;
  if (c->precision != CMI_PREC_FP64)
    return fail(c, CMI_ERR_ARG, "per-subgrid walks need an FP64 context");
  if (subgrid >= c->nsubgrid || entry_dir < 0 || entry_dir >= 27 ||
      (n && (!p || !exit_dir || !crossings)))
    return fail(c, CMI_ERR_ARG, "bad subgrid, entry direction or buffers");
  if (n == 0) return CMI_OK;
  (void)hipSetDevice(c->device);
  PhotonRec *dp = nullptr;
  int32_t *de = nullptr;
  uint32_t *dc = nullptr;
  hipError_t e = hipMalloc(&dp, n * sizeof(PhotonRec));
  if (e == hipSuccess) e = hipMalloc(&de, n * sizeof(int32_t));
  if (e == hipSuccess) e = hipMalloc(&dc, n * sizeof(uint32_t));
  if (e == hipSuccess) e = hipMemcpyAsync(dp, p, n * sizeof(PhotonRec),
                                          hipMemcpyHostToDevice, c->stream);
  if (e == hipSuccess)
    e = launch_interact_batch(c->g, c->f.n, c->f.xn, c->f.J, (int)subgrid,
                              entry_dir, dp, n, de, dc, c->stream, kind);
  if (e == hipSuccess) e = hipMemcpyAsync(p, dp, n * sizeof(PhotonRec),
                                          hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(exit_dir, de, n * sizeof(int32_t),
                                          hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(crossings, dc, n * sizeof(uint32_t),
                                          hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  hipFree(dp);
  hipFree(de);
  hipFree(dc);
  if (e != hipSuccess)
    return fail(c, CMI_ERR_HIP, std::string("interact_batch: ") + hipGetErrorString(e));
  for (size_t i = 0; i < n; ++i)
    if (exit_dir[i] < 0)
      return fail(c, CMI_ERR_DIRECTION, "a packet hit the walk bound");
  return CMI_OK;
}
}  // namespace

extern "C" {

int cmi_interact_batch(cmi_ctx *c, uint32_t subgrid, int entry_dir, cmi_photon *p,
                       size_t n, int32_t *exit_dir, uint32_t *crossings) {
  return walk_batch(c, subgrid, entry_dir, p, n, exit_dir, crossings, 0);
}

int cmi_propagate_batch(cmi_ctx *c, uint32_t subgrid, int entry_dir, cmi_photon *p,
                        size_t n, int32_t *exit_dir, uint32_t *crossings) {
  return walk_batch(c, subgrid, entry_dir, p, n, exit_dir, crossings, 1);
}

int cmi_optical_depth_batch(cmi_ctx *c, uint32_t subgrid, int entry_dir, cmi_photon *p,
                            size_t n, int32_t *exit_dir, uint32_t *crossings) {
  return walk_batch(c, subgrid, entry_dir, p, n, exit_dir, crossings, 2);
}

int cmi_transport_packets(cmi_ctx *c, cmi_photon *p, const uint32_t *start_sub,
                          size_t n, int32_t *status, uint32_t *crossings,
                          uint32_t *hops, cmi_stats *stats) {
  const int prc = prepare_list_transport(c, p, start_sub, n);
  if (prc != CMI_OK) return prc;
  if (n == 0) {
    if (stats) memset(stats, 0, sizeof(*stats));
    return CMI_OK;
  }
  if (c->engine == CMI_ENGINE_TILES) {
    if (c->reemission > 0.)
      return fail(c, CMI_ERR_ARG,
                  "reemission on explicit packet lists runs on the walk engine");
    return tiles_transport_packets(c, p, start_sub, n, status, crossings, hops,
                                   stats);
  }
  PhotonRec *dp = nullptr;
  uint32_t *ds = nullptr, *dc = nullptr, *dh = nullptr;
  int32_t *dst = nullptr;
  hipError_t e = hipMalloc(&dp, n * sizeof(PhotonRec));
  if (e == hipSuccess) e = hipMalloc(&ds, n * sizeof(uint32_t));
  if (e == hipSuccess) e = hipMalloc(&dc, n * sizeof(uint32_t));
  if (e == hipSuccess) e = hipMalloc(&dh, n * sizeof(uint32_t));
  if (e == hipSuccess) e = hipMalloc(&dst, n * sizeof(int32_t));
  if (e == hipSuccess) e = hipMemcpyAsync(dp, p, n * sizeof(PhotonRec),
                                          hipMemcpyHostToDevice, c->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(ds, start_sub, n * sizeof(uint32_t),
                                          hipMemcpyHostToDevice, c->stream);
  if (e == hipSuccess) e = hipMemsetAsync(c->d_cnt, 0, sizeof(Counters), c->stream);
  if (e == hipSuccess) e = hipEventRecord(c->ev0, c->stream);
  if (e == hipSuccess)
    e = launch_walk_list(c->g, c->f, c->precision, dp, ds, n, dst, dc, dh,
                         list_reemit_draws(c), c->d_cnt, c->stream);
  if (e == hipSuccess) e = hipEventRecord(c->ev1, c->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(p, dp, n * sizeof(PhotonRec),
                                          hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess && status)
    e = hipMemcpyAsync(status, dst, n * sizeof(int32_t), hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess && crossings)
    e = hipMemcpyAsync(crossings, dc, n * sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess && hops)
    e = hipMemcpyAsync(hops, dh, n * sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(c->h_cnt, c->d_cnt, sizeof(Counters),
                                          hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  float ms = 0.f;
  if (e == hipSuccess) e = hipEventElapsedTime(&ms, c->ev0, c->ev1);
  hipFree(dp);
  hipFree(ds);
  hipFree(dc);
  hipFree(dh);
  hipFree(dst);
  if (e != hipSuccess)
    return fail(c, CMI_ERR_HIP, std::string("transport_packets: ") + hipGetErrorString(e));
  fill_stats(c, stats, n, ms * 1e-3, 0);
  if (c->h_cnt->errors)
    return fail(c, CMI_ERR_DIRECTION, "packets stopped on an invalid exit or walk bound");
  return CMI_OK;
}

int cmi_set_sources(cmi_ctx *c, const double *pos, const double *L, int nsrc) {
  if (!c || !pos || !L || nsrc <= 0 || nsrc > kMaxSources)
    return fail(c, CMI_ERR_ARG, "need 1..64 sources");
  for (int s = 0; s < nsrc; ++s) {
    if (!(L[s] > 0.)) return fail(c, CMI_ERR_ARG, "luminosities must be > 0");
    for (int a = 0; a < 3; ++a) {
      if (!(pos[3 * s + a] >= c->box[a] && pos[3 * s + a] < c->box[a] + c->box[3 + a]))
        return fail(c, CMI_ERR_ARG, "source outside the box");
    }
  }
  c->src_pos.assign(pos, pos + 3 * nsrc);
  c->src_L.assign(L, L + nsrc);
  return CMI_OK;
}

double cmi_total_luminosity(const cmi_ctx *c) {
  if (!c) return 0.;
  double t = 0.;
  for (double v : c->src_L) t += v;
  return t;
}

int cmi_emit(cmi_ctx *c, uint64_t npackets, uint64_t first, uint64_t count,
             uint64_t seed, uint32_t iteration, double sigma, cmi_photon *out,
             uint32_t *start_sub) {
  if (!c || (count && !out) || first + count > npackets) return CMI_ERR_ARG;
  if (count == 0) return CMI_OK;
  (void)hipSetDevice(c->device);
  int rc = upload_sources(c, npackets, seed, iteration);
  if (rc != CMI_OK) return rc;
  PhotonRec *dp = nullptr;
  uint32_t *ds = nullptr;
  hipError_t e = hipMalloc(&dp, count * sizeof(PhotonRec));
  if (e == hipSuccess) e = hipMalloc(&ds, count * sizeof(uint32_t));
  if (e == hipSuccess)
    e = launch_emit(c->d_src, first, count, iteration, seed, sigma, dp, ds, c->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(out, dp, count * sizeof(PhotonRec),
                                          hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess && start_sub)
    e = hipMemcpyAsync(start_sub, ds, count * sizeof(uint32_t),
                       hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  hipFree(dp);
  hipFree(ds);
  if (e != hipSuccess)
    return fail(c, CMI_ERR_HIP, std::string("emit: ") + hipGetErrorString(e));
  return CMI_OK;
}

int cmi_set_reemission(cmi_ctx *c, double probability) {
  if (!c) return CMI_ERR_ARG;
  if (!(probability >= 0. && probability < 1.))
    return fail(c, CMI_ERR_ARG, "reemission probability must be in [0, 1)");
  c->reemission = probability;
  return CMI_OK;
}

int cmi_set_reemission_draws(cmi_ctx *c, uint64_t seed, uint32_t iteration) {
  if (!c) return CMI_ERR_ARG;
  c->reemit_seed = seed;
  c->reemit_iteration = iteration;
  return CMI_OK;
}

int cmi_transport(cmi_ctx *c, uint64_t npackets, uint64_t seed,
                  uint32_t iteration, double sigma, cmi_stats *stats) {
  if (!c) return CMI_ERR_ARG;
  (void)hipSetDevice(c->device);
  int rc = ensure_sigma(c, sigma);
  if (rc != CMI_OK) return rc;
  rc = upload_sources(c, npackets, seed, iteration);
  if (rc != CMI_OK) return rc;
  if (c->engine == CMI_ENGINE_TILES)
    return tiles_transport(c, npackets, seed, iteration, sigma, stats);
  ReemitDraws re{};
  re.p = c->reemission;
  re.key0 = (uint32_t)seed;
  re.key1 = (uint32_t)(seed >> 32);
  re.iteration = iteration;
  const int blocks = c->tuning.walk_blocks > 0 ? c->tuning.walk_blocks
                                               : c->num_cus * 8;
  CMI_TRY_HIP(c, hipMemsetAsync(c->d_cnt, 0, sizeof(Counters), c->stream));
  CMI_TRY_HIP(c, hipEventRecord(c->ev0, c->stream));
  CMI_TRY_HIP(c, launch_walk_emit(c->g, c->f, c->precision, c->d_src, npackets,
                                  iteration, seed, sigma, blocks, re, c->d_cnt,
                                  c->stream));
  CMI_TRY_HIP(c, hipEventRecord(c->ev1, c->stream));
  CMI_TRY_HIP(c, hipMemcpyAsync(c->h_cnt, c->d_cnt, sizeof(Counters),
                                hipMemcpyDeviceToHost, c->stream));
  CMI_TRY_HIP(c, hipStreamSynchronize(c->stream));
  float ms = 0.f;
  CMI_TRY_HIP(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
  c->kernel_time_sum += ms * 1e-3;
  c->kernel_calls += 1;
  fill_stats(c, stats, npackets, ms * 1e-3, 0);
  if (c->h_cnt->errors)
    return fail(c, CMI_ERR_DIRECTION, "packets stopped on an invalid exit or walk bound");
  return CMI_OK;
}

int cmi_update_ionization(cmi_ctx *c, double luminosity, uint64_t npackets,
                          double alphaH, int form, double *l1_change) {
  if (!c || npackets == 0 || !(alphaH > 0.) || (form != 0 && form != 1))
    return fail(c, CMI_ERR_ARG, "bad ionization arguments");
  (void)hipSetDevice(c->device);
  // DensitySubGrid.hpp:1635-1636: L / N * (1/dx)(1/dy)(1/dz), left to right
  const double jfac = luminosity / (double)npackets * c->g.inv_cell_size[0] *
                      c->g.inv_cell_size[1] * c->g.inv_cell_size[2];
  const double jscale = (c->precision == CMI_PREC_FP64) ? 1. : c->sigma * delta0(c);
  const int write_kappa = c->precision == CMI_PREC_FP32;
  CMI_TRY_HIP(c, launch_ionize(c->f, c->ncell_total, jfac, jscale, alphaH, form,
                               c->sigma * delta0(c), write_kappa, c->d_partial,
                               kIonizeBlocks, c->stream));
  CMI_TRY_HIP(c, hipMemcpyAsync(c->h_partial.data(), c->d_partial,
                                kIonizeBlocks * sizeof(double),
                                hipMemcpyDeviceToHost, c->stream));
  CMI_TRY_HIP(c, hipStreamSynchronize(c->stream));
  if (l1_change) {
    double s = 0.;
    for (double v : c->h_partial) s += v;
    *l1_change = s / (double)c->ncell_total;
  }
  return CMI_OK;
}

int cmi_get_neighbours(const cmi_ctx *c, uint32_t subgrid, uint32_t ngbs[27]) {
  if (!c || !ngbs || subgrid >= c->nsubgrid) return CMI_ERR_ARG;
  memcpy(ngbs, &c->ngbs[27 * (size_t)subgrid], 27 * sizeof(uint32_t));
  return CMI_OK;
}

int cmi_get_sizes(const cmi_ctx *c, uint64_t *nsubgrid, uint64_t *ncell) {
  if (!c) return CMI_ERR_ARG;
  if (nsubgrid) *nsubgrid = c->nsubgrid;
  if (ncell) *ncell = c->ncell_total;
  return CMI_OK;
}

int cmi_kernel_timing(cmi_ctx *c, double *avg, uint64_t *calls, int reset) {
  if (!c) return CMI_ERR_ARG;
  if (avg) *avg = c->kernel_calls ? c->kernel_time_sum / c->kernel_calls : 0.;
  if (calls) *calls = c->kernel_calls;
  if (reset) {
    c->kernel_time_sum = 0.;
    c->kernel_calls = 0;
  }
  return CMI_OK;
}

int cmi_set_kernel_timing(cmi_ctx *c, int on) {
  if (!c) return CMI_ERR_ARG;
  c->kernel_timing = on ? 1 : 0;
  return CMI_OK;
}

int cmi_synchronize(cmi_ctx *c) {
  if (!c) return CMI_ERR_ARG;
  (void)hipSetDevice(c->device);
  CMI_TRY_HIP(c, hipStreamSynchronize(c->stream));
  return CMI_OK;
}

}  // extern "C"
