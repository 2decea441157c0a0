// cmi_walk.hip -- per-packet transport kernels ("walk" engine).
//
// One thread owns one packet and walks it through subgrid after subgrid until
// it is absorbed or leaves the box; lanes that finish refill themselves from
// a wave-aggregated ticket counter (persistent threads, counter-based
// emission), so a wave keeps 64 packets in flight regardless of how uneven
// their path lengths are.  Deposits go to the FP64 J array with hardware
// float64 atomics.
//
// Two arithmetic modes:
//   Mode64 -- the reference arithmetic of DensitySubGrid::interact
//             (DensitySubGrid.hpp:1180-1323) in FP64, positions relative to the
//             subgrid anchor in metres, n and x_n read separately; with
//             contraction off this reproduces the reference walk bit for bit.
//   Mode32 -- FP32 geometry in cell units (position relative to the subgrid in
//             cells, direction scaled by Delta_0/Delta_a) with the optical
//             depth per unit length kappa = sigma n x_n Delta_0 stored as FP32;
//             J accumulates the path length in units of Delta_0 (FP64).
#include <float.h>

#include "cmi_device.hpp"
#include "cmi_internal.hpp"

namespace cmi {

// std::min of <algorithm>: (b < a) ? b : a.
template <class T>
__device__ __forceinline__ T ref_min(T a, T b) {
  return (b < a) ? b : a;
}

template <class T>
struct RealLimits;
template <>
struct RealLimits<double> {
  __device__ static constexpr double big() { return DBL_MAX; }
};
template <>
struct RealLimits<float> {
  __device__ static constexpr float big() { return FLT_MAX; }
};

// One packet inside one subgrid visit.
template <class T>
struct Walker {
  T pos[3];  // relative to the subgrid anchor (Mode64: m, Mode32: cells)
  T dir[3];
  T inv[3];
  T tau_target;  // remaining optical depth at the start of this visit
  T tau_done;    // optical depth accumulated during this visit
  int idx[3];
  int sub;
};

// Result of a crossing step.
enum StepResult : int { kContinue = 0, kAbsorbed = 1, kLeft = 2 };

__device__ __forceinline__ bool inside3(const int idx[3], const int n[3]) {
  return idx[0] >= 0 && idx[0] < n[0] && idx[1] >= 0 && idx[1] < n[1] &&
         idx[2] >= 0 && idx[2] < n[2];
}

// Entry cell (DensitySubGrid.hpp:247-469): clamp along the axes the entry side
// names, truncate position * inv_cell_size along the free ones.
template <class T>
__device__ __forceinline__ void start_index(Walker<T> &w, int in_dir,
                                            const int sc[3], const T inv_cs[3]) {
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const int o = c_dir.offset[in_dir][a];
    w.idx[a] = (o == 0) ? (int)(w.pos[a] * inv_cs[a]) : (o < 0 ? 0 : sc[a] - 1);
  }
}

// ---- Mode64: one loop iteration of DensitySubGrid::interact --------------
// kKind selects the reference's walk variant: kInteract
// (DensitySubGrid::interact, 1180-1323: deposit, stop at the target optical
// depth), kPropagate (propagate, 1333-1472: the same walk without deposits)
// or kOpticalDepth (compute_optical_depth, 1482-1604: no deposits, no stop,
// the optical depth of the whole path accumulated).
enum WalkKind { kInteract = 0, kPropagate = 1, kOpticalDepth = 2 };

template <int kKind>
struct Mode64T {
  using real = double;
  const double *n;
  const double *xn;
  double *J;

  __device__ __forceinline__ int step(Walker<double> &w, const GridDev &g,
                                      double sigma, double weight) const {
#pragma clang fp contract(off)
    const size_t c = cell_addr(g, w.sub, w.idx[0], w.idx[1], w.idx[2]);
    double lo[3], hi[3], l[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      lo[a] = w.idx[a] * g.cell_size[a];
      hi[a] = (w.idx[a] + 1.) * g.cell_size[a];
      l[a] = (w.dir[a] > 0.)   ? (hi[a] - w.pos[a]) * w.inv[a]
             : (w.dir[a] < 0.) ? (lo[a] - w.pos[a]) * w.inv[a]
                                : DBL_MAX;
    }
    double lmin = ref_min(l[0], ref_min(l[1], l[2]));
    double lminsigma = lmin * sigma;
    const double tau = lminsigma * n[c] * xn[c];
    w.tau_done += tau;
    const bool absorbed = kKind != kOpticalDepth && w.tau_done >= w.tau_target;
    if (absorbed) {
      const double correction = (w.tau_done - w.tau_target) / tau;
      lmin *= (1. - correction);
      lminsigma = lmin * sigma;
    } else {
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        if (l[a] == lmin) w.idx[a] += (w.dir[a] > 0.) ? 1 : -1;
      }
    }
    if (kKind == kInteract) atomicAdd(&J[c], lminsigma * weight);
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      w.pos[a] = (l[a] == lmin) ? ((w.dir[a] > 0.) ? hi[a] : lo[a])
                                : w.pos[a] + lmin * w.dir[a];
    }
    if (kKind == kOpticalDepth) return inside3(w.idx, g.sc) ? kContinue : kLeft;
    if (absorbed) return kAbsorbed;
    // a NaN optical depth ends the reference loop too (its condition
    // tau_done < tau_target fails); the packet then counts as absorbed when
    // its index is still inside
    if (!(w.tau_done < w.tau_target))
      return inside3(w.idx, g.sc) ? kAbsorbed : kLeft;
    return inside3(w.idx, g.sc) ? kContinue : kLeft;
  }
};

using Mode64 = Mode64T<kInteract>;

// ---- Mode32: the same step in FP32 cell units ----------------------------
struct Mode32 {
  using real = float;
  const float *kappa;
  double *J;

  __device__ __forceinline__ int step(Walker<float> &w, const GridDev &g,
                                      float, float) const {
    const size_t c = cell_addr(g, w.sub, w.idx[0], w.idx[1], w.idx[2]);
    float lo[3], hi[3], l[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      lo[a] = (float)w.idx[a];
      hi[a] = (float)(w.idx[a] + 1);
      l[a] = (w.dir[a] > 0.f)   ? (hi[a] - w.pos[a]) * w.inv[a]
             : (w.dir[a] < 0.f) ? (lo[a] - w.pos[a]) * w.inv[a]
                                 : FLT_MAX;
    }
    float lmin = ref_min(l[0], ref_min(l[1], l[2]));
    const float tau = lmin * kappa[c];
    w.tau_done += tau;
    const bool absorbed = w.tau_done >= w.tau_target;
    if (absorbed) {
      const float correction = (w.tau_done - w.tau_target) / tau;
      lmin *= (1.f - correction);
    } else {
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        if (l[a] == lmin) w.idx[a] += (w.dir[a] > 0.f) ? 1 : -1;
      }
    }
    atomicAdd(&J[c], (double)lmin);
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      w.pos[a] = (l[a] == lmin) ? ((w.dir[a] > 0.f) ? hi[a] : lo[a])
                                : w.pos[a] + lmin * w.dir[a];
    }
    if (absorbed) return kAbsorbed;
    if (!(w.tau_done < w.tau_target))
      return inside3(w.idx, g.sc) ? kAbsorbed : kLeft;
    return inside3(w.idx, g.sc) ? kContinue : kLeft;
  }
};

// Set a walker up for its first visit from an absolute FP64 packet.
__device__ __forceinline__ void load_packet(Walker<double> &w, const GridDev &g,
                                            int sub, const double pos[3],
                                            const double dir[3], double tau,
                                            int in_dir) {
#pragma clang fp contract(off)
  w.sub = sub;
  const double *an = g.anchors + 3 * (size_t)sub;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    w.dir[a] = dir[a];
    w.inv[a] = 1. / dir[a];
    w.pos[a] = pos[a] - an[a];
  }
  w.tau_target = tau;
  w.tau_done = 0.;
  start_index(w, in_dir, g.sc, g.inv_cell_size);
}

__device__ __forceinline__ void load_packet(Walker<float> &w, const GridDev &g,
                                            int sub, const double pos[3],
                                            const double dir[3], double tau,
                                            int in_dir) {
  w.sub = sub;
  const double *an = g.anchors + 3 * (size_t)sub;
  const float one[3] = {1.f, 1.f, 1.f};
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    w.dir[a] = (float)(dir[a] * g.unit_scale[a]);
    w.inv[a] = 1.f / w.dir[a];
    w.pos[a] = (float)((pos[a] - an[a]) * g.inv_cell_size[a]);
  }
  w.tau_target = (float)tau;
  w.tau_done = 0.f;
  start_index(w, in_dir, g.sc, one);
}

// Absolute position (m) of a walker.
__device__ __forceinline__ void abs_position(const Walker<double> &w,
                                             const GridDev &g, double out[3]) {
#pragma clang fp contract(off)
  const double *an = g.anchors + 3 * (size_t)w.sub;
#pragma unroll
  for (int a = 0; a < 3; ++a) out[a] = w.pos[a] + an[a];
}

__device__ __forceinline__ void abs_position(const Walker<float> &w,
                                             const GridDev &g, double out[3]) {
  const double *an = g.anchors + 3 * (size_t)w.sub;
#pragma unroll
  for (int a = 0; a < 3; ++a) out[a] = an[a] + (double)w.pos[a] * g.cell_size[a];
}

// End of a subgrid visit that left the subgrid: route to the neighbour
// (testDensitySubGrid_omp.cpp:1149-1329 with TravelDirections.hpp:89-152).
// Returns 0 if the packet moved to a neighbour, 1 if it left the box, -1 on
// an invalid exit mask.
__device__ __forceinline__ int hop(Walker<double> &w, const GridDev &g) {
#pragma clang fp contract(off)
  const int d = exit_direction(w.idx[0], w.idx[1], w.idx[2], g.sc[0], g.sc[1],
                               g.sc[2]);
  if (d <= 0) return -1;
  const uint32_t ngb = g.ngbs[(size_t)w.sub * TD_NUMBER + d];
  if (ngb == 0xffffffffu) return 1;
  double p[3];
  abs_position(w, g, p);  // photon.set_position(position + _anchor)
  const double *an = g.anchors + 3 * (size_t)ngb;
#pragma unroll
  for (int a = 0; a < 3; ++a) w.pos[a] = p[a] - an[a];
  w.tau_target = w.tau_target - w.tau_done;
  w.tau_done = 0.;
  w.sub = (int)ngb;
  start_index(w, opposite_direction(d), g.sc, g.inv_cell_size);
  return 0;
}

__device__ __forceinline__ int hop(Walker<float> &w, const GridDev &g) {
  const int d = exit_direction(w.idx[0], w.idx[1], w.idx[2], g.sc[0], g.sc[1],
                               g.sc[2]);
  if (d <= 0) return -1;
  const uint32_t ngb = g.ngbs[(size_t)w.sub * TD_NUMBER + d];
  if (ngb == 0xffffffffu) return 1;
  // anchors differ by exactly one subgrid along every axis the side names:
  // the shift in cell units is an exact integer
#pragma unroll
  for (int a = 0; a < 3; ++a)
    w.pos[a] = w.pos[a] - (float)(c_dir.offset[d][a] * g.sc[a]);
  w.tau_target = w.tau_target - w.tau_done;
  w.tau_done = 0.f;
  w.sub = (int)ngb;
  const float one[3] = {1.f, 1.f, 1.f};
  start_index(w, opposite_direction(d), g.sc, one);
  return 0;
}

// Outcome of one visit's opening check (reference loop condition before the
// first iteration): kContinue, kAbsorbed (tau_target not positive), kLeft.
template <class T>
__device__ __forceinline__ int visit_open(const Walker<T> &w, const GridDev &g) {
  if (!(w.tau_done < w.tau_target)) {
    return (w.tau_done >= w.tau_target) ? kAbsorbed
                                        : (inside3(w.idx, g.sc) ? kAbsorbed
                                                                : kLeft);
  }
  return inside3(w.idx, g.sc) ? kContinue : kLeft;
}

// Wave-level sum of a 64-bit counter; lane 0 adds it to *dst.
__device__ __forceinline__ void wave_add(unsigned long long *dst,
                                         unsigned long long v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  if ((threadIdx.x & 63) == 0 && v) atomicAdd(dst, v);
}

// ---------------------------------------------------------------------------
// Kernel: DensitySubGrid::interact on one subgrid for an explicit packet list
// (the deterministic oracle's entry point).  One thread per packet.
// ---------------------------------------------------------------------------
template <int kKind>
__global__ void interact_batch_kernel(GridDev g, const double *n,
                                      const double *xn, double *J, int sub,
                                      int in_dir, PhotonRec *p, size_t np,
                                      int32_t *exit_dir, uint32_t *crossings,
                                      int max_steps) {
#pragma clang fp contract(off)
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np) return;
  PhotonRec ph = p[i];
  Walker<double> w;
  load_packet(w, g, sub, ph.pos, ph.dir, ph.tau, in_dir);
  Mode64T<kKind> m{n, xn, J};
  uint32_t nc = 0;
  int r = kKind == kOpticalDepth ? (inside3(w.idx, g.sc) ? kContinue : kLeft)
                                 : visit_open(w, g);
  while (r == kContinue && (int)nc < max_steps) {
    r = m.step(w, g, ph.sigma, ph.weight);
    ++nc;
  }
  int out;
  if (r == kContinue) {
    out = -2;  // walk bound exceeded
  } else if (r == kAbsorbed) {
    out = TD_INSIDE;
  } else {
    out = exit_direction(w.idx[0], w.idx[1], w.idx[2], g.sc[0], g.sc[1],
                         g.sc[2]);
  }
  // interact / propagate: the remaining optical depth (DensitySubGrid.hpp:
  // 1307, 1456); compute_optical_depth: the accumulated one (1592-1593)
  ph.tau = kKind == kOpticalDepth ? w.tau_target + w.tau_done : w.tau_target - w.tau_done;
  abs_position(w, g, ph.pos);
  p[i] = ph;
  exit_dir[i] = out;
  crossings[i] = nc;
}

// ---------------------------------------------------------------------------
// Kernel: whole-grid transport of an explicit packet list (status 0 absorbed,
// 1 escaped, -1 invalid exit, -2 walk bound).  One thread per packet.
// ---------------------------------------------------------------------------
// Reemission of an absorbed packet (do_reemission and the reemit task,
// testDensitySubGrid_omp.cpp:675-702, 1348-): with probability re.p the
// packet gets the draws of its nre-th absorption and is walked again through
// the same subgrid from where it stopped, entering INSIDE (its absolute
// position, as the reference keeps it between the tasks).  Returns true if
// the packet was reemitted; the new direction also goes to dir_out.
template <class T>
__device__ __forceinline__ bool reemit(Walker<T> &w, const GridDev &g,
                                       const ReemitDraws &re, uint64_t k,
                                       uint32_t nre, double dir_out[3]) {
  double pa[3], dir[3], tau;
  if (!reemit_draw(k, nre, re, dir, &tau)) return false;
  abs_position(w, g, pa);
  load_packet(w, g, w.sub, pa, dir, tau, TD_INSIDE);
  dir_out[0] = dir[0];
  dir_out[1] = dir[1];
  dir_out[2] = dir[2];
  return true;
}

template <class Mode>
__global__ void walk_list_kernel(GridDev g, Mode m, PhotonRec *p,
                                 const uint32_t *start_sub, size_t np,
                                 int32_t *status, uint32_t *crossings,
                                 uint32_t *hops, int max_steps, int max_hops,
                                 ReemitDraws re, Counters *cnt) {
  using T = typename Mode::real;
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  unsigned long long my_cross = 0, my_hops = 0, my_abs = 0, my_esc = 0,
                     my_err = 0, my_re = 0;
  if (i < np) {
    PhotonRec ph = p[i];
    Walker<T> w;
    load_packet(w, g, (int)start_sub[i], ph.pos, ph.dir, ph.tau, TD_INSIDE);
    uint32_t nc = 0, nh = 1, nre = 0;
    int st = -2;
    int steps = 0, leg_hops = 1;
    for (;;) {
      int r = visit_open(w, g);
      while (r == kContinue && steps < max_steps) {
        r = m.step(w, g, (T)ph.sigma, (T)ph.weight);
        ++nc;
        ++steps;
      }
      if (r == kContinue) {
        st = -2;
        break;
      }
      if (r == kAbsorbed) {
        st = 0;
        if (re.p > 0.) {
          if (nre >= kReemitMax) {
            st = -2;
          } else if (reemit(w, g, re, (uint64_t)i, nre, ph.dir)) {
            ++nre;
            ++nh;
            steps = 0;
            leg_hops = 1;
            continue;
          }
        }
        break;
      }
      const int h = hop(w, g);
      if (h != 0) {
        st = (h == 1) ? 1 : -1;
        break;
      }
      steps = 0;
      ++nh;
      if (++leg_hops > max_hops) {
        st = -2;
        break;
      }
    }
    ph.tau = (double)(w.tau_target - w.tau_done);
    abs_position(w, g, ph.pos);
    p[i] = ph;
    status[i] = st;
    crossings[i] = nc;
    hops[i] = nh;
    my_cross = nc;
    my_hops = nh;
    my_abs = (st == 0);
    my_esc = (st == 1);
    my_err = (st < 0);
    my_re = nre;
  }
  wave_add(&cnt->crossings, my_cross);
  wave_add(&cnt->hops, my_hops);
  wave_add(&cnt->absorbed, my_abs);
  wave_add(&cnt->escaped, my_esc);
  wave_add(&cnt->errors, my_err);
  wave_add(&cnt->reemitted, my_re);
}

// ---------------------------------------------------------------------------
// Kernel: emission + whole-grid transport, persistent threads with per-lane
// refill.  Each loop trip every live lane performs one crossing (or one
// subgrid hop); a lane whose packet ended takes the next packet number from
// a wave-aggregated ticket (ballot + popcount + one atomic per wave).
// ---------------------------------------------------------------------------
template <class Mode>
__global__ void __launch_bounds__(256) walk_emit_kernel(
    GridDev g, Mode m, const SourcesDev *__restrict__ S, uint64_t npackets,
    uint32_t iteration, uint32_t key0, uint32_t key1, double sigma,
    int max_steps, int max_hops, ReemitDraws re, Counters *cnt) {
  using T = typename Mode::real;
  const int lane = threadIdx.x & 63;
  const uint64_t lanemask_lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  Walker<T> w;
  bool live = false, exhausted = false;
  int steps = 0, nh = 0;
  uint64_t my_k = 0;  // packet number of the lane's packet
  uint32_t nre = 0;   // its reemissions so far
  unsigned long long my_cross = 0, my_hops = 0, my_abs = 0, my_esc = 0,
                     my_err = 0, my_re = 0;
  for (;;) {
    const bool need = !live && !exhausted;
    const uint64_t want = __ballot(need);
    if (want) {
      const int leader = __ffsll((long long)want) - 1;
      unsigned long long base = 0;
      if (lane == leader) base = atomicAdd(&cnt->next_packet,
                                           (unsigned long long)__popcll(want));
      base = __shfl(base, leader, 64);
      if (need) {
        const uint64_t k = base + (uint64_t)__popcll(want & lanemask_lt);
        if (k < npackets) {
          double pos[3], dir[3], tau;
          const int s = emit_packet(*S, k, iteration, key0, key1, pos, dir,
                                    &tau);
          load_packet(w, g, (int)S->start_sub[s], pos, dir, tau, TD_INSIDE);
          live = true;
          steps = 0;
          nh = 1;
          my_k = k;
          nre = 0;
          ++my_hops;
        } else {
          exhausted = true;
        }
      }
    }
    if (!__any(live)) break;
    if (!live) continue;
    int r = (steps == 0) ? visit_open(w, g) : kContinue;
    if (r == kContinue) {
      if (steps < max_steps) {
        r = m.step(w, g, (T)sigma, (T)1);
        ++my_cross;
        ++steps;
      } else {
        ++my_err;
        live = false;
        continue;
      }
    }
    if (r == kContinue) continue;
    if (r == kAbsorbed) {
      if (re.p > 0.) {
        double dir_unused[3];
        if (nre >= kReemitMax) {
          ++my_err;
          live = false;
          continue;
        }
        if (reemit(w, g, re, my_k, nre, dir_unused)) {
          ++nre;
          ++my_re;
          ++my_hops;
          steps = 0;
          nh = 1;
          continue;
        }
      }
      ++my_abs;
      live = false;
      continue;
    }
    const int h = hop(w, g);
    if (h == 0) {
      steps = 0;
      ++my_hops;
      if (++nh > max_hops) {
        ++my_err;
        live = false;
      }
    } else {
      if (h == 1) ++my_esc; else ++my_err;
      live = false;
    }
  }
  wave_add(&cnt->crossings, my_cross);
  wave_add(&cnt->hops, my_hops);
  wave_add(&cnt->absorbed, my_abs);
  wave_add(&cnt->escaped, my_esc);
  wave_add(&cnt->errors, my_err);
  wave_add(&cnt->reemitted, my_re);
}

// ---------------------------------------------------------------------------
// Kernel: emission only (writes the packets the transport kernels draw).
// ---------------------------------------------------------------------------
__global__ void emit_kernel(const SourcesDev *__restrict__ S, uint64_t first,
                            uint64_t count, uint32_t iteration, uint32_t key0,
                            uint32_t key1, double sigma, PhotonRec *out,
                            uint32_t *start_sub) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  PhotonRec ph;
  const int s = emit_packet(*S, first + i, iteration, key0, key1, ph.pos,
                            ph.dir, &ph.tau);
  ph.sigma = sigma;
  ph.weight = 1.;
  out[i] = ph;
  start_sub[i] = S->start_sub[s];
}

// ---------------------------------------------------------------------------
// Host-side launchers
// ---------------------------------------------------------------------------
hipError_t launch_interact_batch(const GridDev &g, const double *n,
                                 const double *xn, double *J, int sub,
                                 int in_dir, PhotonRec *p, size_t np,
                                 int32_t *exit_dir, uint32_t *crossings,
                                 hipStream_t st, int kind) {
  if (np == 0) return hipSuccess;
  const int max_steps = g.sc[0] + g.sc[1] + g.sc[2] + 8;
  const unsigned blocks = (unsigned)((np + 255) / 256);
  if (kind == kPropagate)
    hipLaunchKernelGGL(interact_batch_kernel<kPropagate>, dim3(blocks), dim3(256), 0, st, g,
                       n, xn, J, sub, in_dir, p, np, exit_dir, crossings, max_steps);
  else if (kind == kOpticalDepth)
    hipLaunchKernelGGL(interact_batch_kernel<kOpticalDepth>, dim3(blocks), dim3(256), 0, st,
                       g, n, xn, J, sub, in_dir, p, np, exit_dir, crossings, max_steps);
  else
    hipLaunchKernelGGL(interact_batch_kernel<kInteract>, dim3(blocks), dim3(256), 0, st, g,
                       n, xn, J, sub, in_dir, p, np, exit_dir, crossings, max_steps);
  return hipGetLastError();
}

template <class Mode>
static hipError_t launch_walk_list_t(const GridDev &g, Mode m, PhotonRec *p,
                                     const uint32_t *start_sub, size_t np,
                                     int32_t *status, uint32_t *crossings,
                                     uint32_t *hops, const ReemitDraws &re,
                                     Counters *cnt, hipStream_t st) {
  if (np == 0) return hipSuccess;
  const int max_steps = g.sc[0] + g.sc[1] + g.sc[2] + 8;
  const int max_hops = 4 * (g.nsub[0] + g.nsub[1] + g.nsub[2]) + 16;
  const unsigned blocks = (unsigned)((np + 255) / 256);
  hipLaunchKernelGGL(walk_list_kernel<Mode>, dim3(blocks), dim3(256), 0, st, g,
                     m, p, start_sub, np, status, crossings, hops, max_steps,
                     max_hops, re, cnt);
  return hipGetLastError();
}

hipError_t launch_walk_list(const GridDev &g, const FieldsDev &f, int prec,
                            PhotonRec *p, const uint32_t *start_sub, size_t np,
                            int32_t *status, uint32_t *crossings,
                            uint32_t *hops, const ReemitDraws &re,
                            Counters *cnt, hipStream_t st) {
  if (prec == 0)
    return launch_walk_list_t(g, Mode64{f.n, f.xn, f.J}, p, start_sub, np,
                              status, crossings, hops, re, cnt, st);
  return launch_walk_list_t(g, Mode32{f.kappa, f.J}, p, start_sub, np, status,
                            crossings, hops, re, cnt, st);
}

hipError_t launch_walk_emit(const GridDev &g, const FieldsDev &f, int prec,
                            const SourcesDev *S, uint64_t npackets,
                            uint32_t iteration, uint64_t seed, double sigma,
                            int blocks, const ReemitDraws &re, Counters *cnt,
                            hipStream_t st) {
  const int max_steps = g.sc[0] + g.sc[1] + g.sc[2] + 8;
  const int max_hops = 4 * (g.nsub[0] + g.nsub[1] + g.nsub[2]) + 16;
  const uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
  if (prec == 0) {
    hipLaunchKernelGGL(walk_emit_kernel<Mode64>, dim3(blocks), dim3(256), 0,
                       st, g, Mode64{f.n, f.xn, f.J}, S, npackets, iteration,
                       k0, k1, sigma, max_steps, max_hops, re, cnt);
  } else {
    hipLaunchKernelGGL(walk_emit_kernel<Mode32>, dim3(blocks), dim3(256), 0,
                       st, g, Mode32{f.kappa, f.J}, S, npackets, iteration,
                       k0, k1, sigma, max_steps, max_hops, re, cnt);
  }
  return hipGetLastError();
}

hipError_t launch_emit(const SourcesDev *S, uint64_t first, uint64_t count,
                       uint32_t iteration, uint64_t seed, double sigma,
                       PhotonRec *out, uint32_t *start_sub, hipStream_t st) {
  if (count == 0) return hipSuccess;
  const unsigned blocks = (unsigned)((count + 255) / 256);
  hipLaunchKernelGGL(emit_kernel, dim3(blocks), dim3(256), 0, st, S, first,
                     count, iteration, (uint32_t)seed, (uint32_t)(seed >> 32),
                     sigma, out, start_sub);
  return hipGetLastError();
}

}  // namespace cmi
