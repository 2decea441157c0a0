// cmi_tiles.hip -- the tiles engine: packets binned per LDS tile, moved
// between tiles in rounds.
//
// The reference moves packets between subgrids in 200-packet PhotonBuffers
// tagged with the target subgrid and the entry TravelDirection
// (PhotonBuffer.hpp:42-65, MemorySpace::add_photons, MemorySpace.hpp:92-125,
// testDensitySubGrid_omp.cpp:1149-1329).  Here the unit of work is an LDS
// tile (a power-of-two block of cells inside a subgrid) and one round moves
// every packet in flight by one tile visit:
//
//   process   one workgroup per (tile, chunk of its packets): the tile's
//             opacities are staged in LDS, every packet is walked until it
//             leaves the tile or is absorbed, the path lengths accumulate in an
//             LDS copy of the tile's J (fixed-point integers, exact in any
//             order) that is flushed to HBM once per visit;
//             survivors are written to the workgroup's own output region
//             (LDS-counter compaction) with their destination tile; the
//             per-face exit counts go to the next round's tile histogram
//             (whose old values are the visit's bases) and every survivor
//             gets its position relative to its destination's segment;
//   plan      one workgroup: exclusive scan of the histogram into tile
//             segments, and the next round's work descriptors (tile, start,
//             count) in chunks, smaller ones for rounds with few packets;
//   place     segment start + relative position of every survivor -> the
//             index list the next round gathers packets through.
//
// Round 0 has no visits before it: for emitted packets its per-tile counts
// follow from the source split, so it is planned before the emission, which
// writes every packet's index-list entry itself; an explicit packet list is
// scattered by destination (per-block LDS counts, per-tile cursors).
//
// Refills are batched per wave: finished lanes wait until a threshold of
// lanes is free (or the wave runs dry) and then write out and reload
// together, so the divergent refill path is paid once per batch rather than
// once per packet.
//
// The last packets of a single transport, once a host check (at the start
// and every kRoundsPerCheck rounds) finds at most one per resident lane, are
// walked to their ends by one walk-out launch instead of further rounds.
#include <float.h>
#include <string.h>

#include "cmi_ctx.hpp"

namespace cmi {

constexpr uint32_t kNone = 0xffffffffu;

// A problem is one grid of subgrids; a problem set is one or more problems
// that share the LDS tile shape and are transported together.  Tiles of all
// problems form one list: problem p owns tiles [tile_base, tile_base +
// ntiles_p), and inside a problem the order is subgrid-major then
// tile-in-subgrid (the order cell_addr() uses), so tile * tile_cells is the
// tile's first cell in the concatenated cell arrays.
struct ProbGeo {
  int N[3];          // cells per axis
  int nts[3];        // tiles per subgrid per axis
  int nsub[3];       // subgrids per axis
  int ntl[3];        // tiles per axis
  int tiles_per_sub;
  int tile_base;     // first tile of this problem in the tile list
  double box[3];     // anchor of the box (m)
  double inv_cell_size[3];
  double cell_size[3];
  double unit_scale[3];  // Delta_0 / Delta_a
};

// Walk state of a straight flight inside a tile, in one 32-bit register:
// the byte offset of the cell's FP32 opacity, 4 (cell + bias), with a bias
// of one axis-0 plane (T1 * T2 cells rounded up to a multiple of 4, so that
// it stays non-negative one step past any face) in the low bits (offmask),
// and above it one step counter per axis (log2(T_a) + 1 bits at bit fs[a]).
// Byte offsets let a visit read the opacity at the offset itself (the
// opacities' biased base is the start of the LDS block).  The counter of
// axis a starts at the cells already behind the packet along its direction
// (l_a for a positive direction, T_a - 1 - l_a for a negative one) and grows
// by one per step along a, so the flight leaves the tile exactly when a
// counter reaches T_a, its top bit (guard).  A step
// along a adds one increment, (1 << fs[a]) plus the signed byte stride along
// a: one select chain, one add and one mask test per crossing move the cell,
// count the step and detect the exit.  The offset field never borrows from or
// carries into the counters (its width covers one step past either face).
// Any tile of at most 8192 cells and 64 cells per axis fits.
struct WalkLayout {
  uint32_t fs[3];       // bit of each axis' counter
  uint32_t guard_a[3];  // top bit of each axis' counter
  uint32_t guard;       // all three
  uint32_t offmask;     // the biased byte offset
  int bias;             // T1 * T2 rounded up to a multiple of 4
  int pad;              // floats staged before the tile's opacities (= bias)
};

struct TileGeo {
  int T[3];          // tile size per axis (cells, powers of two)
  int tshift[3];
  int tile_cells;
  WalkLayout W;
  int ntiles;        // tiles of all problems
  int nprob;
  float reemit_p;    // reemission probability of an absorbed packet (0: off)
  float fix_scale;   // 2^S of the fixed-point deposits (FixAcc)
  double fix_inv;    // 2^-S
  uint32_t reemit_key[2];     // Philox key of the reemission draws
  const ProbGeo *probs;       // [nprob]
  const uint16_t *tile_prob;  // [ntiles] problem of each tile
};

__device__ __forceinline__ int tile_index(const ProbGeo &P, int tx, int ty,
                                          int tz) {
  const int sx = tx / P.nts[0], sy = ty / P.nts[1], sz = tz / P.nts[2];
  const int sub = (sx * P.nsub[1] + sy) * P.nsub[2] + sz;
  const int t = ((tx - sx * P.nts[0]) * P.nts[1] + (ty - sy * P.nts[1])) *
                    P.nts[2] +
                (tz - sz * P.nts[2]);
  return P.tile_base + sub * P.tiles_per_sub + t;
}

__device__ __forceinline__ void tile_coords(const ProbGeo &P, int tile,
                                            int &tx, int &ty, int &tz) {
  const int local = tile - P.tile_base;
  const int sub = local / P.tiles_per_sub;
  const int t = local - sub * P.tiles_per_sub;
  const int sx = sub / (P.nsub[1] * P.nsub[2]);
  const int sy = (sub / P.nsub[2]) % P.nsub[1];
  const int sz = sub % P.nsub[2];
  const int lx = t / (P.nts[1] * P.nts[2]);
  const int ly = (t / P.nts[2]) % P.nts[1];
  const int lz = t % P.nts[2];
  tx = sx * P.nts[0] + lx;
  ty = sy * P.nts[1] + ly;
  tz = sz * P.nts[2] + lz;
}

// Sources of a problem set: problem p emits packets [pk_cum[p-1], pk_cum[p])
// of the call, from its sources src_begin[p] .. src_begin[p+1]-1 (packet
// numbers within the problem split by src_cum, cumulative per problem).
constexpr int kMaxProblems = 256;
constexpr int kMaxSetSources = 4096;
struct SetSources {
  int nprob;
  uint32_t iteration;
  uint64_t seed;
  uint64_t pk_cum[kMaxProblems];
  int src_begin[kMaxProblems + 1];
  uint64_t src_cum[kMaxSetSources];
  double pos[kMaxSetSources][3];
};

// Per-problem ionization constants.
struct IonParams {
  double jfac;         // L / N / V_cell
  double jscale;       // J units -> m^3 (sigma Delta_0 in FP32 mode)
  double kappa_scale;  // sigma Delta_0
};

// FP32 packet record, 32 bytes: position relative to the tile origin in cell
// units, direction scaled to cells per Delta_0, remaining optical depth, and
// the global cell (10 bits per axis).
struct Rec32 {
  float4 a;  // px, py, pz, dx
  float4 b;  // dy, dz, tau, cell bits
};

__device__ __forceinline__ uint32_t pack_cell(int x, int y, int z) {
  return (uint32_t)x | ((uint32_t)y << 10) | ((uint32_t)z << 20);
}
__device__ __forceinline__ void unpack_cell(uint32_t c, int &x, int &y,
                                            int &z) {
  x = (int)(c & 1023u);
  y = (int)((c >> 10) & 1023u);
  z = (int)((c >> 20) & 1023u);
}

struct Desc {
  int tile;
  uint32_t start;
  uint32_t count;
  uint32_t pad;
};

// Per-packet outputs of explicit-list transports (indexed by packet id).
struct PacketOut {
  int32_t *status;
  uint32_t *crossings;
  uint32_t *hops;
  float4 *final_a;  // last record (position in the final tile frame ...)
  float4 *final_b;
  int32_t *final_tile;
};

// ---------------------------------------------------------------------------
// Emission straight into FP32 records (round-0 input).
// ---------------------------------------------------------------------------
// Packet k of problem p draws counters {k_lo, k_hi ^ (p << 20), iteration, 0}
// and {..., 1} (problem 0 draws exactly what emit_packet() draws).
// Returns the problem; *src is the emitting source and *rank the packet's
// number among that source's packets.
__device__ __forceinline__ int set_emit(const SetSources &S, uint64_t kg,
                                        uint32_t iteration, uint32_t key0,
                                        uint32_t key1, double pos[3],
                                        double dir[3], double *tau, int *src,
                                        uint64_t *rank) {
  // the problem: the first p with kg < pk_cum[p] (the last one if none);
  // the source: the first s of the problem with k < src_cum[s] (likewise).
  // Binary searches: a problem set has up to 64 problems of up to 16
  // sources, and every step of a search is a dependent load
  int lo = 0, hi = S.nprob - 1;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (kg < S.pk_cum[mid]) hi = mid; else lo = mid + 1;
  }
  const int p = lo;
  const uint64_t k = kg - (p ? S.pk_cum[p - 1] : 0);
  lo = S.src_begin[p];
  hi = S.src_begin[p + 1] - 1;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (k < S.src_cum[mid]) hi = mid; else lo = mid + 1;
  }
  const int s = lo;
  const u32x4 r0 = philox4x32_10(
      u32x4{(uint32_t)k, (uint32_t)(k >> 32) ^ ((uint32_t)p << 20), iteration, 0u},
      key0, key1);
  const u32x4 r1 = philox4x32_10(
      u32x4{(uint32_t)k, (uint32_t)(k >> 32) ^ ((uint32_t)p << 20), iteration, 1u},
      key0, key1);
  const double cost = 2. * uniform_open(r0.x, r0.y) - 1.;
  const double phi = 2. * 3.14159265358979323846 * uniform_open(r0.z, r0.w);
  const double sint = sqrt(fmax(1. - cost * cost, 0.));
  double sp, cp;
  sincos(phi, &sp, &cp);
  dir[0] = sint * cp;
  dir[1] = sint * sp;
  dir[2] = cost;
  *tau = -log(uniform_open(r1.x, r1.y));
  pos[0] = S.pos[s][0];
  pos[1] = S.pos[s][1];
  pos[2] = S.pos[s][2];
  *src = s;
  *rank = k - (s > S.src_begin[p] ? S.src_cum[s - 1] : 0);
  return p;
}

// Where a source's packets start, the same for all of them: the tile of the
// start cell, the source position in that tile's frame (cells), the packed
// global cell, and the source's offset within the tile's round-0 segment.
struct SrcTile {
  float px, py, pz;
  uint32_t tile;
  uint32_t cell;
  uint32_t off;
  uint32_t pad[2];
};

// Seed and iteration come from S (device memory), so that a captured graph
// of the iteration can be replayed with new values.  Round 0 is planned
// before emission (the per-tile counts follow from the source split), so
// every packet also goes straight into round 0's index list: its source's
// share of the tile's segment starts at src[source].off.
__global__ void tiles_emit_kernel(TileGeo G, const SetSources *__restrict__ S,
                                  uint64_t npackets, Rec32 *__restrict__ out,
                                  const uint32_t *__restrict__ seg,
                                  const SrcTile *__restrict__ src_tile,
                                  uint32_t *__restrict__ idx) {
  const uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < npackets) {
    double pos[3], dir[3], tau;
    int src;
    uint64_t rank;
    const int p = set_emit(*S, k, S->iteration, (uint32_t)S->seed,
                           (uint32_t)(S->seed >> 32), pos, dir, &tau, &src, &rank);
    const ProbGeo &P = G.probs[p];
    const SrcTile st = src_tile[src];
    Rec32 r;
    r.a.x = st.px;
    r.a.y = st.py;
    r.a.z = st.pz;
    r.a.w = (float)(dir[0] * P.unit_scale[0]);
    r.b.x = (float)(dir[1] * P.unit_scale[1]);
    r.b.y = (float)(dir[2] * P.unit_scale[2]);
    r.b.z = (float)tau;
    r.b.w = __uint_as_float(st.cell);
    out[k] = r;
    // in range whenever the source split sums to npackets (as it does by
    // construction); the check keeps a broken split from writing past idx
    const uint64_t at = (uint64_t)seg[st.tile] + st.off + rank;
    if (at < npackets) idx[at] = (uint32_t)k;
  }
}

// Tile of every source's start cell (the cell the emission kernel files its
// packets under) gets that source's packet count; the old count is the
// source's offset within the tile's round-0 segment.  The start cell and
// the source's position in its tile's frame are kept for the emission.
__global__ void tiles_source_hist_kernel(TileGeo G,
                                         const SetSources *__restrict__ S,
                                         uint32_t *__restrict__ hist,
                                         SrcTile *__restrict__ src_tile) {
  // one thread per source of every problem; a source's problem is the last
  // p with src_begin[p] <= s
  const int ns = S->src_begin[S->nprob];
  for (int s = (int)threadIdx.x; s < ns; s += blockDim.x) {
    int lo = 0, hi = S->nprob - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (S->src_begin[mid] <= s) lo = mid; else hi = mid - 1;
    }
    const int p = lo;
    const ProbGeo &P = G.probs[p];
    const int b = S->src_begin[p];
    {
      const uint64_t c = S->src_cum[s] - (s > b ? S->src_cum[s - 1] : 0);
      int g[3];
      double pc[3];
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        pc[a] = (S->pos[s][a] - P.box[a]) * P.inv_cell_size[a];
        const int cc = (int)pc[a];
        g[a] = cc < 0 ? 0 : (cc >= P.N[a] ? P.N[a] - 1 : cc);
      }
      const int tx = g[0] >> G.tshift[0], ty = g[1] >> G.tshift[1],
                tz = g[2] >> G.tshift[2];
      SrcTile st;
      st.tile = (uint32_t)tile_index(P, tx, ty, tz);
      st.px = (float)(pc[0] - (double)(tx << G.tshift[0]));
      st.py = (float)(pc[1] - (double)(ty << G.tshift[1]));
      st.pz = (float)(pc[2] - (double)(tz << G.tshift[2]));
      st.cell = pack_cell(g[0], g[1], g[2]);
      st.off = c ? atomicAdd(&hist[st.tile], (uint32_t)c) : 0u;
      st.pad[0] = st.pad[1] = 0u;
      src_tile[s] = st;
    }
  }
}

// Explicit FP64 packet list -> FP32 records (start cell found like the
// reference: subgrid by truncation, then the cell within it).
__global__ void tiles_load_list_kernel(TileGeo G, GridDev g,
                                       const PhotonRec *__restrict__ p,
                                       const uint32_t *__restrict__ start_sub,
                                       size_t n, Rec32 *__restrict__ out,
                                       uint32_t *__restrict__ dest,
                                       uint32_t *__restrict__ pid,
                                       uint32_t *__restrict__ hist) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const PhotonRec ph = p[i];
  const ProbGeo &P = G.probs[0];
  const int sub = (int)start_sub[i];
  const int sx = sub / (P.nsub[1] * P.nsub[2]);
  const int sy = (sub / P.nsub[2]) % P.nsub[1];
  const int sz = sub % P.nsub[2];
  const int so[3] = {sx * g.sc[0], sy * g.sc[1], sz * g.sc[2]};
  const double *an = g.anchors + 3 * (size_t)sub;
  int gc[3];
  double pc[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const double rel = (ph.pos[a] - an[a]) * g.inv_cell_size[a];
    int c = (int)rel;
    c = c < 0 ? 0 : (c >= g.sc[a] ? g.sc[a] - 1 : c);
    gc[a] = so[a] + c;
    pc[a] = rel + (double)so[a];
  }
  const int tx = gc[0] >> G.tshift[0], ty = gc[1] >> G.tshift[1],
            tz = gc[2] >> G.tshift[2];
  const uint32_t tile = (uint32_t)tile_index(P, tx, ty, tz);
  Rec32 r;
  r.a.x = (float)(pc[0] - (double)(tx << G.tshift[0]));
  r.a.y = (float)(pc[1] - (double)(ty << G.tshift[1]));
  r.a.z = (float)(pc[2] - (double)(tz << G.tshift[2]));
  r.a.w = (float)(ph.dir[0] * P.unit_scale[0]);
  r.b.x = (float)(ph.dir[1] * P.unit_scale[1]);
  r.b.y = (float)(ph.dir[2] * P.unit_scale[2]);
  r.b.z = (float)ph.tau;
  r.b.w = __uint_as_float(pack_cell(gc[0], gc[1], gc[2]));
  out[i] = r;
  dest[i] = tile;
  pid[i] = (uint32_t)i;
  atomicAdd(&hist[tile], 1u);
}

// ---------------------------------------------------------------------------
// Plan: histogram -> tile segments and work descriptors (one workgroup).
// ---------------------------------------------------------------------------
constexpr int kPlanThreads = 1024;

__device__ __forceinline__ uint64_t block_exclusive_scan(uint64_t v,
                                                         uint64_t *sh,
                                                         uint64_t *total) {
  // inclusive scan within the wave
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  uint64_t x = v;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const uint64_t y = __shfl_up(x, off, 64);
    if (lane >= off) x += y;
  }
  if (lane == 63) sh[wave] = x;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint64_t acc = 0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) {
      const uint64_t t = sh[w];
      sh[w] = acc;
      acc += t;
    }
    sh[16] = acc;
  }
  __syncthreads();
  const uint64_t r = sh[wave] + x - v;
  *total = sh[16];
  __syncthreads();
  return r;
}

// A tile with more than kPlanSerialChunks chunks is crowded: the descriptors
// of up to kPlanCrowded crowded tiles are written by the whole workgroup from
// one flattened range (a tile found by binary search over the crowded tiles'
// chunk prefix in LDS); the other tiles' descriptors by their own thread.
constexpr uint32_t kPlanSerialChunks = 2;
constexpr int kPlanCrowded = 1024;

static const uint32_t kChunk = 8192;      // packets per workgroup visit
static const uint32_t kChunkFloor = 256;  // smallest chunk accepted

// total[0]: packets of the round being planned (in: those of the round just
// processed, kept in total[1] for the placement); nonempty (optional) counts
// the plans that found packets.  A round's chunk is `chunk`, or smaller when
// the round holds fewer than min_visits * chunk packets: then the chunk is cut
// (down to kChunkFloor) so that the round still has about min_visits visits.
// Late rounds hold few packets; at full chunks they occupy a fraction of the
// GPU's workgroup slots and the round's time is that of its largest visits,
// which walk several packets per lane one after the other.
__global__ void __launch_bounds__(kPlanThreads) tiles_plan_kernel(
    int ntiles, uint32_t *__restrict__ hist, uint32_t *__restrict__ seg,
    uint32_t *__restrict__ cursor, Desc *__restrict__ desc,
    uint32_t *__restrict__ ndesc, uint32_t *__restrict__ total,
    unsigned *__restrict__ ticket, uint32_t chunk, uint32_t min_visits,
    const uint16_t *__restrict__ tile_prob, int nprob,
    uint32_t *__restrict__ prob_round,
    unsigned long long *__restrict__ prob_time,
    uint32_t *__restrict__ round_ctr, uint32_t *__restrict__ nonempty) {
  __shared__ uint64_t sh[17];
  __shared__ uint32_t s_pc[kMaxProblems];
  __shared__ uint32_t s_round;
  __shared__ uint32_t s_ncrowded;
  __shared__ uint4 s_crowded[kPlanCrowded];  // tile, first descriptor, count, segment start
  __shared__ uint32_t s_cpre[kPlanCrowded];  // exclusive prefix of their chunks
  const bool track = prob_round != nullptr;
  if (threadIdx.x == 0) {
    total[1] = total[0];
    s_ncrowded = 0;
  }
  if (track)
    for (int q = threadIdx.x; q < nprob; q += kPlanThreads) s_pc[q] = 0;
  __syncthreads();
  const int per = (ntiles + kPlanThreads - 1) / kPlanThreads;
  const int t0 = threadIdx.x * per;
  const int t1 = min(ntiles, t0 + per);
  // each thread's tile counts are read once, into registers, when it has at
  // most kPlanPer tiles (ntiles <= 16384): the three passes below then make
  // no dependent global loads (a set of 64 problems has ~11000 tiles)
  constexpr int kPlanPer = 16;
  const bool cached = per <= kPlanPer;
  uint32_t hc[kPlanPer];
  if (cached) {
#pragma unroll
    for (int j = 0; j < kPlanPer; ++j) hc[j] = (t0 + j < t1) ? hist[t0 + j] : 0u;
  }
  uint64_t cnt = 0;
  auto pass1 = [&](int t, uint32_t c) {
    cnt += c;
    if (track && c) atomicAdd(&s_pc[tile_prob[t]], c);
  };
  if (cached) {
#pragma unroll
    for (int j = 0; j < kPlanPer; ++j)
      if (t0 + j < t1) pass1(t0 + j, hc[j]);
  } else {
    for (int t = t0; t < t1; ++t) pass1(t, hist[t]);
  }
  uint64_t tot_c, tot_d;
  uint64_t base_c = block_exclusive_scan(cnt, sh, &tot_c);
  if (min_visits) {
    const uint64_t want = (tot_c + min_visits - 1) / min_visits;
    chunk = (uint32_t)(want < chunk ? (want > kChunkFloor ? want : kChunkFloor) : chunk);
  }
  uint64_t nch = 0;
  if (cached) {
#pragma unroll
    for (int j = 0; j < kPlanPer; ++j)
      if (t0 + j < t1) nch += (hc[j] + chunk - 1) / chunk;
  } else {
    for (int t = t0; t < t1; ++t) nch += (hist[t] + chunk - 1) / chunk;
  }
  uint64_t base_d = block_exclusive_scan(nch, sh, &tot_d);
  auto pass3 = [&](int t, uint32_t c) {
    seg[t] = (uint32_t)base_c;
    cursor[t] = 0;
    hist[t] = 0;
    const uint32_t nch = (c + chunk - 1) / chunk;
    uint32_t slot = kPlanCrowded;
    if (nch > kPlanSerialChunks) slot = atomicAdd(&s_ncrowded, 1u);
    if (slot < kPlanCrowded) {
      s_crowded[slot] = make_uint4((uint32_t)t, (uint32_t)base_d, c, (uint32_t)base_c);
    } else {
      for (uint32_t j = 0; j < nch; ++j) {
        Desc d;
        d.tile = t;
        d.start = (uint32_t)base_c + j * chunk;
        d.count = (c - j * chunk < chunk) ? (c - j * chunk) : chunk;
        d.pad = 0;
        desc[base_d + j] = d;
      }
    }
    base_d += nch;
    base_c += c;
  };
  if (cached) {
#pragma unroll
    for (int j = 0; j < kPlanPer; ++j)
      if (t0 + j < t1) pass3(t0 + j, hc[j]);
  } else {
    for (int t = t0; t < t1; ++t) pass3(t, hist[t]);
  }
  __syncthreads();
  const uint32_t ncrowded = min(s_ncrowded, (uint32_t)kPlanCrowded);
  if (ncrowded) {
    static_assert(kPlanCrowded <= kPlanThreads, "one crowded tile per thread in the scan");
    const uint64_t mine = threadIdx.x < ncrowded
                              ? (s_crowded[threadIdx.x].z + chunk - 1) / chunk : 0;
    uint64_t ncd;
    const uint64_t pre = block_exclusive_scan(mine, sh, &ncd);
    if (threadIdx.x < ncrowded) s_cpre[threadIdx.x] = (uint32_t)pre;
    __syncthreads();
    for (uint32_t k = threadIdx.x; k < (uint32_t)ncd; k += kPlanThreads) {
      // the last crowded tile whose first flattened descriptor is <= k
      uint32_t lo = 0, hi = ncrowded - 1;
      while (lo < hi) {
        const uint32_t mid = (lo + hi + 1) >> 1;
        if (s_cpre[mid] <= k) lo = mid; else hi = mid - 1;
      }
      const uint4 B = s_crowded[lo];
      const uint32_t j = k - s_cpre[lo];
      Desc d;
      d.tile = (int)B.x;
      d.start = B.w + j * chunk;
      d.count = (B.z - j * chunk < chunk) ? (B.z - j * chunk) : chunk;
      d.pad = 0;
      desc[B.y + j] = d;
    }
  }
  if (threadIdx.x == 0) {
    *ndesc = (uint32_t)tot_d;
    total[0] = (uint32_t)tot_c;
    *ticket = 0;
    if (nonempty && tot_c) *nonempty += 1u;
  }
  if (track) {
    // completion of each problem: the first plan that finds it empty
    if (threadIdx.x == 0) {
      s_round = *round_ctr;
      *round_ctr = s_round + 1;
    }
    __syncthreads();
    const unsigned long long now = wall_clock64();
    for (int q = threadIdx.x; q < nprob; q += kPlanThreads) {
      if (s_pc[q] == 0 && prob_round[q] == kNone) {
        prob_round[q] = s_round;
        prob_time[q] = now;
      }
    }
  }
}

// Start of a set iteration: per-problem completion marks cleared, the round
// counter zeroed and the start time taken.
__global__ void tiles_begin_kernel(int nprob, uint32_t *__restrict__ prob_round,
                                   unsigned long long *__restrict__ prob_time,
                                   uint32_t *__restrict__ round_ctr,
                                   unsigned long long *__restrict__ t0) {
  for (int q = threadIdx.x; q < nprob; q += blockDim.x) {
    prob_round[q] = kNone;
    prob_time[q] = 0;
  }
  if (threadIdx.x == 0) {
    *round_ctr = 0;
    *t0 = wall_clock64();
  }
}

// ---------------------------------------------------------------------------
// Scatter: survivors' destination tiles -> index list sorted by tile.
// ---------------------------------------------------------------------------
constexpr int kScatterThreads = 256;
constexpr int kScatterItems = 16;                 // slots per thread
constexpr int kScatterHash = 512;                 // LDS hash entries

// One block takes kScatterThreads * kScatterItems consecutive slots: the
// slots' destination tiles are counted in an LDS hash table (the rank of
// each slot within its (block, tile) group comes from the LDS atomic), one
// global atomic per distinct tile reserves the block's range in that tile's
// segment, and the slots are written to the index list.  Consecutive slots
// come from a few workgroup regions, so a block sees few distinct tiles and
// the per-tile cursors see one atomic per block instead of one per wave.
__global__ void __launch_bounds__(kScatterThreads) tiles_scatter_kernel(
    const uint32_t *__restrict__ nslots_ptr, const uint32_t *__restrict__ dest,
    const uint32_t *__restrict__ seg, uint32_t *__restrict__ cursor,
    uint32_t *__restrict__ idx) {
  __shared__ uint32_t keys[kScatterHash];
  __shared__ uint32_t cnts[kScatterHash];
  __shared__ uint32_t bases[kScatterHash];
  const uint32_t nslots = *nslots_ptr;
  const uint32_t per_block = kScatterThreads * kScatterItems;
  for (uint32_t base = blockIdx.x * per_block; base < nslots;
       base += gridDim.x * per_block) {
    for (int i = threadIdx.x; i < kScatterHash; i += kScatterThreads) {
      keys[i] = kNone;
      cnts[i] = 0;
    }
    __syncthreads();
    uint32_t t[kScatterItems];
    int h[kScatterItems];
    uint32_t rk[kScatterItems];
#pragma unroll
    for (int k = 0; k < kScatterItems; ++k) {
      const uint32_t s = base + k * kScatterThreads + threadIdx.x;
      t[k] = (s < nslots) ? dest[s] : kNone;
      h[k] = -1;
      rk[k] = 0;
      if (t[k] != kNone) {
        uint32_t hh = (t[k] * 2654435761u) >> 23;  // 9 bits
        for (int probe = 0; probe < 32; ++probe) {
          const uint32_t old = atomicCAS(&keys[hh], kNone, t[k]);
          if (old == kNone || old == t[k]) {
            h[k] = (int)hh;
            rk[k] = atomicAdd(&cnts[hh], 1u);
            break;
          }
          hh = (hh + 1) & (kScatterHash - 1);
        }
        if (h[k] < 0) rk[k] = atomicAdd(&cursor[t[k]], 1u);  // table full
      }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < kScatterHash; i += kScatterThreads) {
      const uint32_t c = cnts[i];
      if (c) bases[i] = atomicAdd(&cursor[keys[i]], c);
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kScatterItems; ++k) {
      if (t[k] == kNone) continue;
      const uint32_t s = base + k * kScatterThreads + threadIdx.x;
      const uint32_t b = (h[k] >= 0) ? bases[h[k]] + rk[k] : rk[k];
      idx[seg[t[k]] + b] = s;
    }
    __syncthreads();
  }
}

// Placement of one round's survivors: the process kernel left each survivor
// slot its position relative to its destination tile's segment (the visit's
// base for that face, reserved by the histogram atomic, plus the survivor's
// rank among the visit's exits through that face), so the index list is one
// streaming pass without atomics.
__global__ void tiles_place_kernel(const uint32_t *__restrict__ nslots_ptr,
                                   const uint32_t *__restrict__ dest,
                                   const uint32_t *__restrict__ rel,
                                   const uint32_t *__restrict__ seg,
                                   uint32_t *__restrict__ idx) {
  const uint32_t n = *nslots_ptr;
  for (uint32_t s = blockIdx.x * blockDim.x + threadIdx.x; s < n;
       s += gridDim.x * blockDim.x) {
    const uint32_t t = dest[s];
    if (t != kNone) idx[seg[t] + rel[s]] = s;
  }
}

// ---------------------------------------------------------------------------
// Process: one tile visit per descriptor, FP32 DDA inside the tile.
// ---------------------------------------------------------------------------
template <class Acc>
struct AccAtomic;
template <>
struct AccAtomic<float> {
  __device__ static void add(float *p, float v, float) { atomicAdd(p, v); }
  __device__ static double value(float v, double) { return (double)v; }
};
// Fixed-point deposits: a path length s >= 0 (in Delta_0 units) adds
// fix_quantum(s) = s 2^S rounded to an integer as an unsigned 64-bit
// integer.  S (TileGeo::fix_scale) keeps the longest segment a cell allows,
// its diagonal, within 2^31: S = 30 for cubic cells, a quantum of 9.3e-10,
// far below the FP32 rounding of a segment (tn - t, with t up to the tile's
// size).  A wave's 64 random-address 64-bit integer LDS adds take about 12
// CU cycles on gfx950 against 21 for FP64 adds (and about 190 for FP32
// ones), and integer sums are exact in any order.
using FixAcc = unsigned long long;
__device__ __forceinline__ uint32_t fix_quantum(float s, float scale) {
  return (uint32_t)fmaf(s, scale, 0.5f);
}
template <>
struct AccAtomic<unsigned long long> {
  __device__ static void add(unsigned long long *p, float v, float scale) {
    atomicAdd(p, (unsigned long long)fix_quantum(v, scale));
  }
  __device__ static double value(unsigned long long v, double inv) { return (double)v * inv; }
};

// Walk layout of a T0 x T1 x T2 tile (a constant expression for the
// kernels specialised to one tile shape).
constexpr __host__ __device__ WalkLayout walk_layout(int T0, int T1, int T2) {
  WalkLayout W{};
  const int cells = T0 * T1 * T2;
  W.bias = (T1 * T2 + 3) & ~3;
  W.pad = W.bias;
  // biased byte offsets one step past either face: [0, 4 (cells + 2 bias))
  uint32_t ob = 2;
  while ((1u << ob) < 4u * (uint32_t)(cells + 2 * W.bias)) ++ob;
  const int T[3] = {T0, T1, T2};
  uint32_t w[3] = {1, 1, 1};
  for (int a = 0; a < 3; ++a)
    while ((1 << (w[a] - 1)) < T[a]) ++w[a];
  W.fs[2] = ob;
  W.fs[1] = ob + w[2];
  W.fs[0] = ob + w[2] + w[1];
  for (int a = 0; a < 3; ++a) W.guard_a[a] = (uint32_t)T[a] << W.fs[a];
  W.guard = W.guard_a[0] | W.guard_a[1] | W.guard_a[2];
  W.offmask = (1u << ob) - 1u;
  return W;
}

// Walk state (WalkLayout) of a flight in local cell (l0, l1, l2) moving
// along d, and the increment of a step along each axis.
__device__ __forceinline__ uint32_t walk_state(const WalkLayout &W, int l0, int l1,
                                               int l2, float d0, float d1, float d2,
                                               int T0, int T1, int T2) {
  const uint32_t f0 = (uint32_t)(d0 > 0.f ? l0 : T0 - 1 - l0);
  const uint32_t f1 = (uint32_t)(d1 > 0.f ? l1 : T1 - 1 - l1);
  const uint32_t f2 = (uint32_t)(d2 > 0.f ? l2 : T2 - 1 - l2);
  const uint32_t off = (uint32_t)((l0 * T1 + l1) * T2 + l2 + W.bias) << 2;
  return (f0 << W.fs[0]) | (f1 << W.fs[1]) | (f2 << W.fs[2]) | off;
}
__device__ __forceinline__ uint32_t walk_inc(uint32_t fs, int stride, float d) {
  return (1u << fs) + (uint32_t)(d > 0.f ? stride : -stride);
}
// Step counter of an axis (T_a once the flight has left through it).
__device__ __forceinline__ uint32_t walk_count(uint32_t q, uint32_t fs, int T) {
  return (q >> fs) & (2u * (uint32_t)T - 1u);
}
// Cell index along an axis from its counter (T or -1 past the face).
__device__ __forceinline__ int walk_index(uint32_t f, float d, int T) {
  return d > 0.f ? (int)f : T - 1 - (int)f;
}
// Sum of the three counters: steps taken between two states of one flight
// are the difference of the sums (no counter ever decreases).
__device__ __forceinline__ uint32_t walk_sum(const WalkLayout &W, uint32_t q, int T0,
                                             int T1, int T2) {
  return walk_count(q, W.fs[0], T0) + walk_count(q, W.fs[1], T1) +
         walk_count(q, W.fs[2], T2);
}

// DDA set-up along one axis for a packet at p (cells, tile frame) in cell l
// moving with direction component d: the path parameter of the next face
// ahead (never negative: a position a rounding step past its cell's face
// starts at 0, so that t only ever grows) and the path between faces; an
// axis the packet does not move along never steps.  Computed for both signs
// and selected, without branches.
__device__ __forceinline__ void dda_axis(int l, float p, float d, float &tm,
                                         float &td) {
  const float i = __builtin_amdgcn_rcpf(d);
  const float face = (float)(d > 0.f ? l + 1 : l);
  tm = d != 0.f ? fmaxf((face - p) * i, 0.f) : FLT_MAX;
  td = d != 0.f ? fabsf(i) : FLT_MAX;
}

// kReemit: an absorbed packet is reemitted with probability G.reemit_p
// (do_reemission, testDensitySubGrid_omp.cpp:675-702): a new isotropic
// direction and target optical depth from the absorption point, in the same
// cell.  The draws are Philox4x32-10 keyed by the transport's seed and
// iteration, with the absorption point's coordinate bits, the tile and the
// packet's reemission count as the counter.
template <class Acc, int kThreads, bool kList, bool kReemit = false, int kTile = 0>
__global__ void __launch_bounds__(kThreads) tiles_process_kernel(
    TileGeo G, const float *__restrict__ kappa, double *__restrict__ J,
    const Desc *__restrict__ desc, const uint32_t *__restrict__ ndesc_ptr,
    unsigned *__restrict__ ticket, const Rec32 *__restrict__ Pin,
    const uint32_t *__restrict__ idx, Rec32 *__restrict__ Pout,
    uint32_t *__restrict__ dest_out, uint32_t *__restrict__ hist_next,
    const uint32_t *__restrict__ pid_in, uint32_t *__restrict__ pid_out,
    PacketOut po, Counters *__restrict__ cnt,
    unsigned long long *__restrict__ prob_cross, int refill_batch,
    uint32_t *__restrict__ rel_out) {
  // LDS: W.pad floats of padding, the tile's opacities, its accumulators.
  // The padding and the accumulators behind the opacities keep the read
  // ahead of a cell one step past any face in bounds (its value unused).
  // kTile > 0: the kernel serves kTile^3 tiles only, from a static block
  // whose address the compiler folds into the LDS instructions' offsets;
  // kTile = 0: any shape, dynamic LDS.
  constexpr int kCellsF = kTile * kTile * kTile;
  constexpr WalkLayout kWF = walk_layout(kTile ? kTile : 1, kTile ? kTile : 1, kTile ? kTile : 1);
  constexpr size_t kStatBytes =
      kTile ? ((sizeof(float) * (size_t)(kWF.pad + kCellsF) + 15) & ~(size_t)15) +
                  sizeof(Acc) * (size_t)kCellsF
            : 16;
  __shared__ __align__(16) unsigned char sblk[kStatBytes];
  extern __shared__ __align__(16) unsigned char dyn[];
  unsigned char *const smem = kTile ? sblk : dyn;
  const WalkLayout W = kTile ? kWF : G.W;
  const int tile_cells = kTile ? kCellsF : G.tile_cells;
  // the offset mask with its sign bit known clear, so that the compiler may
  // fold the LDS base into the instruction's offset field
  const uint32_t offmask = W.offmask & 0x7fffffffu;
  const float fscale = G.fix_scale;
  float *kapL = reinterpret_cast<float *>(smem) + W.pad;
  Acc *accL = reinterpret_cast<Acc *>(
      smem + ((sizeof(float) * (size_t)(W.pad + tile_cells) + 15) & ~(size_t)15));
  // addressed by the biased byte offset of the walk state: the opacities'
  // biased base is smem itself (pad == bias), an accumulator sits at
  // sizeof(Acc) / 4 times the offset from its biased base
  const unsigned char *kapB = smem;
  unsigned char *accB = reinterpret_cast<unsigned char *>(accL - W.bias);
  auto kap_at = [&](uint32_t o) { return *reinterpret_cast<const float *>(kapB + o); };
  auto acc_at = [&](uint32_t o) {
    return reinterpret_cast<Acc *>(accB + o * (uint32_t)(sizeof(Acc) / sizeof(float)));
  };
  __shared__ uint32_t s_next, s_surv, s_desc;
  __shared__ uint32_t s_face[6];
  __shared__ uint32_t s_fbase[6];
  __shared__ uint32_t s_nbr[6];  // the tile's face neighbours, by exit face

  const int lane = threadIdx.x & 63;
  const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  const uint32_t ndesc = *ndesc_ptr;
  const int T0 = kTile ? kTile : G.T[0], T1 = kTile ? kTile : G.T[1],
            T2 = kTile ? kTile : G.T[2];
  const int st0 = (T1 * T2) << 2, st1 = T2 << 2;  // byte strides
  // per-lane counters of this launch (32 bits: a lane walks at most a few
  // hundred thousand crossings per launch); widened for the final sums
  uint32_t my_cross = 0, my_hops = 0, my_abs = 0, my_esc = 0, my_err = 0;
  uint32_t my_re = 0;  // reemissions (kReemit only)

  // a workgroup's first visit is the descriptor of its own index; later ones
  // come from the ticket, counted on from the grid size (no start-of-launch
  // queue of every workgroup on the one counter)
  bool first = true;
  for (;;) {
    uint32_t di = blockIdx.x;
    if (!first) {
      if (threadIdx.x == 0) s_desc = gridDim.x + atomicAdd(ticket, 1u);
      __syncthreads();
      di = s_desc;
    }
    first = false;
    if (di >= ndesc) break;
    const Desc D = desc[di];
    const int tile = D.tile;
    const int prob = G.tile_prob[tile];
    const ProbGeo &P = G.probs[prob];
    const uint32_t cross0 = my_cross;
    // stage the tile
    {
      const float4 *src =
          reinterpret_cast<const float4 *>(kappa + (size_t)tile * tile_cells);
      float4 *dst = reinterpret_cast<float4 *>(kapL);
      for (int i = threadIdx.x; i < (tile_cells >> 2); i += kThreads)
        dst[i] = src[i];
      for (int i = threadIdx.x; i < tile_cells; i += kThreads) accL[i] = (Acc)0;
      if (threadIdx.x == 0) {
        s_next = kThreads;  // the first kThreads packets are claimed below
        s_surv = 0;
      }
      if (threadIdx.x < 6) s_face[threadIdx.x] = 0;
    }
    int ttx, tty, ttz;
    tile_coords(P, tile, ttx, tty, ttz);
    // the problem's extent and the 6 face neighbours of the tile in
    // registers: a scalar load inside the crossing loop would make every LDS
    // access wait for it (one shared counter), and the neighbours' indices
    // need integer divisions
    const int PN0 = P.N[0], PN1 = P.N[1], PN2 = P.N[2];
    const int nb_xp = tile_index(P, ttx + 1, tty, ttz), nb_xn = tile_index(P, ttx - 1, tty, ttz);
    const int nb_yp = tile_index(P, ttx, tty + 1, ttz), nb_yn = tile_index(P, ttx, tty - 1, ttz);
    const int nb_zp = tile_index(P, ttx, tty, ttz + 1), nb_zn = tile_index(P, ttx, tty, ttz - 1);
    const int ox = ttx << G.tshift[0], oy = tty << G.tshift[1],
              oz = ttz << G.tshift[2];
    // exit face f = 2 axis + (negative direction): its neighbour tile
    if (threadIdx.x < 6) {
      const int f = threadIdx.x;
      s_nbr[f] = (uint32_t)(f == 0 ? nb_xp : f == 1 ? nb_xn : f == 2 ? nb_yp
                            : f == 3 ? nb_yn : f == 4 ? nb_zp : nb_zn);
    }

    // prefetched next record of this lane
    bool has_nxt = false;
    float4 na = make_float4(0.f, 0.f, 0.f, 0.f), nb = na;
    uint32_t npid = 0;
    if (threadIdx.x < D.count) {
      const uint32_t src = idx ? idx[D.start + threadIdx.x] : D.start + threadIdx.x;
      na = Pin[src].a;
      nb = Pin[src].b;
      if (kList) npid = pid_in[src];
      has_nxt = true;
    }
    __syncthreads();

    // per-lane packet state
    bool live = false, pending = false;
    uint32_t q = 0;                      // walk state (WalkLayout)
    uint32_t cur = 0;                    // its biased byte offset (q & offmask)
    uint32_t inc0 = 0, inc1 = 0, inc2 = 0;  // its increments per axis
    uint32_t q_sum0 = 0;                 // counters' sum at the leg's start
    float p0 = 0.f, p1 = 0.f, p2 = 0.f, d0 = 0.f, d1 = 0.f, d2 = 0.f;
    float tm0 = 0.f, tm1 = 0.f, tm2 = 0.f, td0 = 0.f, td1 = 0.f, td2 = 0.f;
    float t = 0.f, tau = 0.f;
    uint32_t my_pid = 0, vis_acc = 0, n_reemit = 0;
    float kap_c = 0.f;  // opacity of the lane's current cell (read ahead)

    for (;;) {
      // ---- write out finished packets ----
      bool moving = false;
      int wl0 = 0, wl1 = 0, wl2 = 0, exit_axis = 0;
      if (pending) {
        const bool exited = (q & W.guard) != 0u;
        if (!exited) {
          // absorbed: the partial last step, up to the target optical depth
          // (reemitting kernels deposited it at the absorption and left
          // tau = 0)
          const float sa = kap_c > 0.f ? tau * __builtin_amdgcn_rcpf(kap_c) : 0.f;
          AccAtomic<Acc>::add(acc_at(cur), sa, fscale);
          t = t + sa;
        }
        const uint32_t vis =
            vis_acc + (walk_sum(W, q, T0, T1, T2) - q_sum0) + (exited ? 0u : 1u);
        wl0 = walk_index(walk_count(q, W.fs[0], T0), d0, T0);
        wl1 = walk_index(walk_count(q, W.fs[1], T1), d1, T1);
        wl2 = walk_index(walk_count(q, W.fs[2], T2), d2, T2);
        exit_axis = (q & W.guard_a[0]) ? 0 : ((q & W.guard_a[1]) ? 1 : 2);
        const int gx = ox + wl0, gy = oy + wl1, gz = oz + wl2;
        const bool in_box = (unsigned)gx < (unsigned)PN0 && (unsigned)gy < (unsigned)PN1 &&
                            (unsigned)gz < (unsigned)PN2;
        // 0 absorbed, 1 moved to a neighbour tile, 2 escaped
        const int outcome = !exited ? 0 : (in_box ? 1 : 2);
        moving = outcome == 1;
        my_cross += vis;
        my_abs += (outcome == 0);
        my_esc += (outcome == 2);
        if (kList) {
          po.crossings[my_pid] += vis;
          po.hops[my_pid] += 1;
          if (outcome != 1) po.status[my_pid] = outcome == 0 ? 0 : 1;
        }
      }
      const uint64_t mv = __ballot(moving);
      uint32_t sbase = 0;
      if (mv) {
        const int leader = __ffsll((long long)mv) - 1;
        if (lane == leader) sbase = atomicAdd(&s_surv, (uint32_t)__popcll(mv));
        sbase = __shfl(sbase, leader, 64);
      }
      if (pending) {
        // position at the exit (or the absorption), in the tile frame
        const float q0 = fmaf(t, d0, p0), q1 = fmaf(t, d1, p1), q2 = fmaf(t, d2, p2);
        if (moving) {
          // the new tile differs along the exit axis only
          Rec32 r;
          r.a.x = (exit_axis == 0) ? (d0 > 0.f ? 0.f : (float)T0) : q0;
          r.a.y = (exit_axis == 1) ? (d1 > 0.f ? 0.f : (float)T1) : q1;
          r.a.z = (exit_axis == 2) ? (d2 > 0.f ? 0.f : (float)T2) : q2;
          r.a.w = d0;
          r.b.x = d1;
          r.b.y = d2;
          r.b.z = tau;
          r.b.w = __uint_as_float(pack_cell(ox + wl0, oy + wl1, oz + wl2));
          const uint32_t slot = D.start + sbase + (uint32_t)__popcll(mv & lt);
          Pout[slot] = r;
          const float de = exit_axis == 0 ? d0 : (exit_axis == 1 ? d1 : d2);
          const uint32_t f = 2u * (uint32_t)exit_axis + (de > 0.f ? 0u : 1u);
          dest_out[slot] = s_nbr[f];
          if (kList) pid_out[slot] = my_pid;
          // the rank among this visit's exits through this face; the
          // face's base is added once the visit has ended
          rel_out[slot] = (f << 29) | atomicAdd(&s_face[f], 1u);
        } else if (kList) {
          po.final_a[my_pid] = make_float4(q0, q1, q2, d0);
          po.final_b[my_pid] = make_float4(d1, d2, tau, 0.f);
          po.final_tile[my_pid] = tile;
        }
        pending = false;
      }
      // ---- take the prefetched record, prefetch the next one ----
      const bool take = !live && has_nxt;
      if (take) {
        p0 = na.x;
        p1 = na.y;
        p2 = na.z;
        d0 = na.w;
        d1 = nb.x;
        d2 = nb.y;
        tau = nb.z;
        my_pid = npid;
        int gx, gy, gz;
        unpack_cell(__float_as_uint(nb.w), gx, gy, gz);
        const int l0 = gx - ox, l1 = gy - oy, l2 = gz - oz;
        ++my_hops;
        vis_acc = 0;
        n_reemit = 0;
        t = 0.f;
        has_nxt = false;
        if ((unsigned)l0 >= (unsigned)T0 || (unsigned)l1 >= (unsigned)T1 ||
            (unsigned)l2 >= (unsigned)T2) {
          // a record that does not belong to this tile: an error, written
          // out at once (it never walks)
          ++my_err;
          if (kList) {
            po.hops[my_pid] += 1;
            po.status[my_pid] = -2;
            po.final_a[my_pid] = make_float4(p0, p1, p2, d0);
            po.final_b[my_pid] = make_float4(d1, d2, tau, 0.f);
            po.final_tile[my_pid] = tile;
          }
        } else {
          dda_axis(l0, p0, d0, tm0, td0);
          dda_axis(l1, p1, d1, tm1, td1);
          dda_axis(l2, p2, d2, tm2, td2);
          inc0 = walk_inc(W.fs[0], st0, d0);
          inc1 = walk_inc(W.fs[1], st1, d1);
          inc2 = walk_inc(W.fs[2], 4, d2);
          q = walk_state(W, l0, l1, l2, d0, d1, d2, T0, T1, T2);
          q_sum0 = walk_sum(W, q, T0, T1, T2);
          cur = q & offmask;
          kap_c = kap_at(cur);
          live = true;
        }
      }
      const uint64_t tm = __ballot(take);
      if (tm) {
        const int leader = __ffsll((long long)tm) - 1;
        uint32_t kb = 0;
        if (lane == leader) kb = atomicAdd(&s_next, (uint32_t)__popcll(tm));
        kb = __shfl(kb, leader, 64);
        if (take) {
          const uint32_t k = kb + (uint32_t)__popcll(tm & lt);
          if (k < D.count) {
            const uint32_t src = idx ? idx[D.start + k] : D.start + k;
            na = Pin[src].a;
            nb = Pin[src].b;
            if (kList) npid = pid_in[src];
            has_nxt = true;
          }
        }
      }
      // (no lane is pending here: every one was written out above)
      const uint64_t live_mask = __ballot(live);
      const uint64_t nxt_mask = __ballot(has_nxt);
      if ((live_mask | nxt_mask) == 0ull) break;
      // ---- walk: until refill_batch of the wave's lanes have finished
      //      while records are left for them, else until every lane has ----
      const int nlive0 = __popcll(live_mask);
      const int stop_at = nxt_mask ? max(nlive0 - refill_batch, 0) : 0;
      int nlive = nlive0;
      // six steps per pass: the count of finished lanes (a ballot) and the
      // loop's exec-mask bookkeeping are paid once per six crossings; a lane
      // that finishes in a pass sits out its remaining steps (measured on the
      // 256^3 clumpy workload, 1e8 packets: 2 steps 31.4 ms, 4 31.2, 6 31.0,
      // 8 31.0)
      constexpr int kSteps = kReemit ? 1 : 6;
      while (nlive > stop_at) {
        bool fin = false;  // lanes that finished in this pass
#pragma unroll
        for (int u = 0; u < kSteps; ++u) {
        if (live) {
          bool f;
          const float tn = fminf(tm0, fminf(tm1, tm2));
          const float seg = tn - t;
          const float dtau = seg * kap_c;
          if (dtau >= tau) {
            // absorbed in this cell
            bool again = false;
            if (kReemit) {
              const float sa = kap_c > 0.f ? tau * __builtin_amdgcn_rcpf(kap_c) : 0.f;
              AccAtomic<Acc>::add(acc_at(cur), sa, fscale);
              t = t + sa;
              tau = 0.f;
              const float q0 = fmaf(t, d0, p0), q1 = fmaf(t, d1, p1), q2 = fmaf(t, d2, p2);
              const u32x4 r = philox4x32_10(
                  u32x4{__float_as_uint(q0), __float_as_uint(q1), __float_as_uint(q2),
                        (uint32_t)tile ^ (n_reemit << 24)},
                  G.reemit_key[0], G.reemit_key[1]);
              const float k24 = 1.f / 16777216.f;
              if (((float)(r.x >> 8) + 0.5f) * k24 < G.reemit_p) {
                // get_random_direction, testDensitySubGrid_omp.cpp:529-545,
                // and tau = -ln u: FP32 draws on 24-bit uniforms in (0, 1)
                const int l0 = walk_index(walk_count(q, W.fs[0], T0), d0, T0);
                const int l1 = walk_index(walk_count(q, W.fs[1], T1), d1, T1);
                const int l2 = walk_index(walk_count(q, W.fs[2], T2), d2, T2);
                vis_acc += (walk_sum(W, q, T0, T1, T2) - q_sum0) + 1u;
                const float cost = 2.f * (((float)(r.y >> 8) + 0.5f) * k24) - 1.f;
                const float phi = 6.28318530717958647692f * (((float)(r.z >> 8) + 0.5f) * k24);
                const float sint = sqrtf(fmaxf(1.f - cost * cost, 0.f));
                float sp, cp;
                sincosf(phi, &sp, &cp);
                d0 = sint * cp * (float)P.unit_scale[0];
                d1 = sint * sp * (float)P.unit_scale[1];
                d2 = cost * (float)P.unit_scale[2];
                tau = -__logf(((float)(r.w >> 8) + 0.5f) * k24);
                p0 = q0;
                p1 = q1;
                p2 = q2;
                t = 0.f;
                ++n_reemit;
                ++my_re;
                dda_axis(l0, p0, d0, tm0, td0);
                dda_axis(l1, p1, d1, tm1, td1);
                dda_axis(l2, p2, d2, tm2, td2);
                inc0 = walk_inc(W.fs[0], st0, d0);
                inc1 = walk_inc(W.fs[1], st1, d1);
                inc2 = walk_inc(W.fs[2], 4, d2);
                q = walk_state(W, l0, l1, l2, d0, d1, d2, T0, T1, T2);
                q_sum0 = walk_sum(W, q, T0, T1, T2);
                again = true;
              }
            }
            f = !again;
          } else {
            AccAtomic<Acc>::add(acc_at(cur), seg, fscale);
            tau -= dtau;
            t = tn;
            const bool a0 = tm0 == tn;
            const bool a1 = !a0 && (tm1 == tn);
            q += a0 ? inc0 : (a1 ? inc1 : inc2);
            tm0 += a0 ? td0 : 0.f;
            tm1 += a1 ? td1 : 0.f;
            tm2 += (!a0 && !a1) ? td2 : 0.f;
            f = (q & W.guard) != 0u;
            cur = q & offmask;
            kap_c = kap_at(cur);  // the next cell, read ahead
          }
          if (f) {
            live = false;
            pending = true;
            fin = true;
          }
        }
        }
        nlive -= __popcll(__ballot(fin));
      }
    }
    __syncthreads();
    // flush the visit: J deposits, tail of the output region, face counts
    {
      double *Jt = J + (size_t)tile * tile_cells;
      for (int i = threadIdx.x; i < tile_cells; i += kThreads) {
        const Acc v = accL[i];
        if (v != (Acc)0) atomicAdd(&Jt[i], AccAtomic<Acc>::value(v, G.fix_inv));
      }
      const uint32_t ns = s_surv;
      for (uint32_t i = ns + threadIdx.x; i < D.count; i += kThreads)
        dest_out[D.start + i] = kNone;
      if (threadIdx.x < 6) {
        // the next round's count of each neighbour; the old value is this
        // visit's base within the neighbour's segment
        const uint32_t c = s_face[threadIdx.x];
        uint32_t base = 0;
        if (c) {
          const int f = threadIdx.x;
          const int nt = f == 0 ? nb_xp : f == 1 ? nb_xn : f == 2 ? nb_yp
                       : f == 3 ? nb_yn : f == 4 ? nb_zp : nb_zn;
          base = atomicAdd(&hist_next[nt], c);
        }
        s_fbase[threadIdx.x] = base;
      }
      __syncthreads();
      // eight loads in flight per thread, then their rewrites: a loop of
      // one load, wait and store per slot pays a memory latency per slot
      constexpr int kFix = 8;
      for (uint32_t i0 = threadIdx.x; i0 < ns; i0 += kFix * kThreads) {
        uint32_t v[kFix];
#pragma unroll
        for (int k = 0; k < kFix; ++k) {
          const uint32_t i = i0 + (uint32_t)k * kThreads;
          v[k] = i < ns ? rel_out[D.start + i] : 0u;
        }
#pragma unroll
        for (int k = 0; k < kFix; ++k) {
          const uint32_t i = i0 + (uint32_t)k * kThreads;
          if (i < ns) rel_out[D.start + i] = s_fbase[v[k] >> 29] + (v[k] & ((1u << 29) - 1u));
        }
      }
      if (prob_cross) {
        // crossings of this visit, per problem (one atomic per wave)
        unsigned long long v = my_cross - cross0;  // this visit, widened
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
        if (lane == 0 && v) atomicAdd(&prob_cross[prob], v);
      }
    }
    __syncthreads();
  }
  // kernel-wide counters
  unsigned long long w_cross = my_cross, w_hops = my_hops, w_abs = my_abs, w_esc = my_esc,
                     w_err = my_err;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    w_cross += __shfl_xor(w_cross, o, 64);
    w_hops += __shfl_xor(w_hops, o, 64);
    w_abs += __shfl_xor(w_abs, o, 64);
    w_esc += __shfl_xor(w_esc, o, 64);
    w_err += __shfl_xor(w_err, o, 64);
  }
  if (lane == 0) {
    if (w_cross) atomicAdd(&cnt->crossings, w_cross);
    if (w_hops) atomicAdd(&cnt->hops, w_hops);
    if (w_abs) atomicAdd(&cnt->absorbed, w_abs);
    if (w_esc) atomicAdd(&cnt->escaped, w_esc);
    if (w_err) atomicAdd(&cnt->errors, w_err);
  }
  if (kReemit) {
    unsigned long long w_re = my_re;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) w_re += __shfl_xor(w_re, o, 64);
    if (lane == 0 && w_re) atomicAdd(&cnt->reemitted, w_re);
  }
}

// A deposit of the walk-out: the value a visit's accumulator adds for it.
__device__ __forceinline__ double walkout_dep(const TileGeo &G, float s) {
  return G.fix_scale > 0.f ? (double)fix_quantum(s, G.fix_scale) * G.fix_inv : (double)s;
}

// Walk-out of a small round (one problem, no reemission, at most a few
// packets per resident lane): every packet left is walked to its end in this
// one launch, tile after tile, with the opacities read from memory and the
// deposits (quantized as a visit's accumulator quantizes them) added to J by
// FP64 atomics, instead of one process / plan / place round per tile.  Each tile leg repeats the arithmetic of a visit exactly --
// the record a visit writes at the tile's face, the set-up from that record,
// the same steps -- so a packet's cells, crossings and fate do not depend on
// whether it finished here.  A kernel of its own: inside the process kernel
// its registers would lower the visits' occupancy.  One packet per lane, the
// round's slots (total[0] of them, records through idx) striding over the
// grid.
template <int kThreads, bool kList>
__global__ void __launch_bounds__(kThreads) tiles_walkout_kernel(
    TileGeo G, const float *__restrict__ kappa, double *__restrict__ J,
    const uint32_t *__restrict__ total, const Rec32 *__restrict__ Pin,
    const uint32_t *__restrict__ idx, const uint32_t *__restrict__ pid_in,
    PacketOut po, Counters *__restrict__ cnt, int max_tiles) {
  const uint32_t n = total[0];
  const ProbGeo &P = G.probs[0];
  const WalkLayout W = G.W;
  const int T0 = G.T[0], T1 = G.T[1], T2 = G.T[2];
  const int st0 = (T1 * T2) << 2, st1 = T2 << 2;  // byte strides
  const int PN0 = P.N[0], PN1 = P.N[1], PN2 = P.N[2];
  uint32_t my_cross = 0, my_hops = 0, my_abs = 0, my_esc = 0, my_err = 0;
  for (uint32_t j = blockIdx.x * kThreads + threadIdx.x; j < n; j += gridDim.x * kThreads) {
    const uint32_t src = idx ? idx[j] : j;
    const float4 ra = Pin[src].a, rb = Pin[src].b;
    const uint32_t pid = kList ? pid_in[src] : 0u;
    float p0 = ra.x, p1 = ra.y, p2 = ra.z;
    const float d0 = ra.w, d1 = rb.x, d2 = rb.y;
    float tau = rb.z, t = 0.f;
    int gx, gy, gz;
    unpack_cell(__float_as_uint(rb.w), gx, gy, gz);
    int tile = -1;
    uint32_t pc_cross = 0, pc_hops = 1;  // a record outside the box: one hop
    int status = -2;  // a record outside the box, or the tile cap reached
    if ((unsigned)gx < (unsigned)PN0 && (unsigned)gy < (unsigned)PN1 &&
        (unsigned)gz < (unsigned)PN2) {
      pc_hops = 0;
      tile = tile_index(P, gx >> G.tshift[0], gy >> G.tshift[1], gz >> G.tshift[2]);
      for (int h = 0; h < max_tiles; ++h) {
        const int ox = (gx >> G.tshift[0]) << G.tshift[0];
        const int oy = (gy >> G.tshift[1]) << G.tshift[1];
        const int oz = (gz >> G.tshift[2]) << G.tshift[2];
        const int l0 = gx - ox, l1 = gy - oy, l2 = gz - oz;
        ++pc_hops;
        float tm0, tm1, tm2, td0, td1, td2;
        dda_axis(l0, p0, d0, tm0, td0);
        dda_axis(l1, p1, d1, tm1, td1);
        dda_axis(l2, p2, d2, tm2, td2);
        const uint32_t inc0 = walk_inc(W.fs[0], st0, d0), inc1 = walk_inc(W.fs[1], st1, d1),
                       inc2 = walk_inc(W.fs[2], 4, d2);
        uint32_t q = walk_state(W, l0, l1, l2, d0, d1, d2, T0, T1, T2);
        const uint32_t q_sum0 = walk_sum(W, q, T0, T1, T2);
        // indexed by the biased cell of the walk state (its byte offset / 4)
        const float *kt = kappa + (size_t)tile * G.tile_cells - W.bias;
        double *Jt = J + (size_t)tile * G.tile_cells - W.bias;
        float kap_c = kt[(q & W.offmask) >> 2];
        t = 0.f;
        bool exited = false;
        for (;;) {
          const float tn = fminf(tm0, fminf(tm1, tm2));
          const float seg = tn - t;
          const float dtau = seg * kap_c;
          if (dtau >= tau) break;
          atomicAdd(&Jt[(q & W.offmask) >> 2], walkout_dep(G, seg));
          tau -= dtau;
          t = tn;
          const bool a0 = tm0 == tn;
          const bool a1 = !a0 && (tm1 == tn);
          q += a0 ? inc0 : (a1 ? inc1 : inc2);
          tm0 += a0 ? td0 : 0.f;
          tm1 += a1 ? td1 : 0.f;
          tm2 += (!a0 && !a1) ? td2 : 0.f;
          if ((q & W.guard) != 0u) {
            exited = true;
            break;
          }
          kap_c = kt[(q & W.offmask) >> 2];
        }
        pc_cross += (walk_sum(W, q, T0, T1, T2) - q_sum0) + (exited ? 0u : 1u);
        if (!exited) {
          // absorbed: the partial last step, up to the target optical depth
          const float sa = kap_c > 0.f ? tau * __builtin_amdgcn_rcpf(kap_c) : 0.f;
          atomicAdd(&Jt[(q & W.offmask) >> 2], walkout_dep(G, sa));
          t = t + sa;
          status = 0;
          break;
        }
        const int e0 = walk_index(walk_count(q, W.fs[0], T0), d0, T0);
        const int e1 = walk_index(walk_count(q, W.fs[1], T1), d1, T1);
        const int e2 = walk_index(walk_count(q, W.fs[2], T2), d2, T2);
        const int nx = ox + e0, ny = oy + e1, nz = oz + e2;
        if (!((unsigned)nx < (unsigned)PN0 && (unsigned)ny < (unsigned)PN1 &&
              (unsigned)nz < (unsigned)PN2)) {
          status = 1;  // out of the box; the position stays in this tile's frame
          break;
        }
        // the record the next tile's visit would read
        const int exit_axis = (q & W.guard_a[0]) ? 0 : ((q & W.guard_a[1]) ? 1 : 2);
        const float q0 = fmaf(t, d0, p0), q1 = fmaf(t, d1, p1), q2 = fmaf(t, d2, p2);
        p0 = (exit_axis == 0) ? (d0 > 0.f ? 0.f : (float)T0) : q0;
        p1 = (exit_axis == 1) ? (d1 > 0.f ? 0.f : (float)T1) : q1;
        p2 = (exit_axis == 2) ? (d2 > 0.f ? 0.f : (float)T2) : q2;
        t = 0.f;
        gx = nx;
        gy = ny;
        gz = nz;
        tile = tile_index(P, gx >> G.tshift[0], gy >> G.tshift[1], gz >> G.tshift[2]);
      }
    }
    my_cross += pc_cross;
    my_hops += pc_hops;
    my_abs += (status == 0);
    my_esc += (status == 1);
    my_err += (status < 0);
    if (kList) {
      po.crossings[pid] += pc_cross;
      po.hops[pid] += pc_hops;
      po.status[pid] = status;
      // the final position in the last tile's frame (t = 0 unless the
      // packet ended in it)
      po.final_a[pid] = make_float4(fmaf(t, d0, p0), fmaf(t, d1, p1), fmaf(t, d2, p2), d0);
      po.final_b[pid] = make_float4(d1, d2, tau, 0.f);
      po.final_tile[pid] = tile;
    }
  }
  unsigned long long w_cross = my_cross, w_hops = my_hops, w_abs = my_abs, w_esc = my_esc,
                     w_err = my_err;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    w_cross += __shfl_xor(w_cross, o, 64);
    w_hops += __shfl_xor(w_hops, o, 64);
    w_abs += __shfl_xor(w_abs, o, 64);
    w_esc += __shfl_xor(w_esc, o, 64);
    w_err += __shfl_xor(w_err, o, 64);
  }
  if ((threadIdx.x & 63) == 0) {
    if (w_cross) atomicAdd(&cnt->crossings, w_cross);
    if (w_hops) atomicAdd(&cnt->hops, w_hops);
    if (w_abs) atomicAdd(&cnt->absorbed, w_abs);
    if (w_esc) atomicAdd(&cnt->escaped, w_esc);
    if (w_err) atomicAdd(&cnt->errors, w_err);
    if (w_cross) atomicAdd(&cnt->walkout_crossings, w_cross);
    if (w_hops) atomicAdd(&cnt->walkout_hops, w_hops);
  }
}

__global__ void zero_u32_kernel(uint32_t *p, size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    p[i] = 0;
}

__global__ void zero_counters_kernel(Counters *c) {
  if (threadIdx.x < sizeof(Counters) / 8)
    reinterpret_cast<unsigned long long *>(c)[threadIdx.x] = 0;
}

// Final FP32 records of an explicit-list transport -> absolute positions.
__global__ void tiles_finish_list_kernel(TileGeo G, GridDev g, size_t n,
                                         const float4 *__restrict__ fa,
                                         const int32_t *__restrict__ ftile,
                                         const float4 *__restrict__ fb,
                                         PhotonRec *__restrict__ p) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int tile = ftile[i];
  if (tile < 0) return;
  const ProbGeo &P = G.probs[0];
  int tx, ty, tz;
  tile_coords(P, tile, tx, ty, tz);
  const float4 a = fa[i];
  const float4 b = fb[i];
  const int o[3] = {tx << G.tshift[0], ty << G.tshift[1], tz << G.tshift[2]};
  const float q[3] = {a.x, a.y, a.z};
#pragma unroll
  for (int k = 0; k < 3; ++k)
    p[i].pos[k] = P.box[k] + ((double)o[k] + (double)q[k]) * g.cell_size[k];
  p[i].tau = (double)b.z;
}

// ---------------------------------------------------------------------------
// Per-problem ionization update of a problem set (FP64): one block per tile,
// the tile's cells belong to one problem; the L1 change is summed per problem.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) set_ionize_kernel(
    TileGeo G, const IonParams *__restrict__ ip, const double *__restrict__ n,
    double *__restrict__ xn, double *__restrict__ J, float *__restrict__ kappa,
    double alpha, int form, double *__restrict__ l1) {
#pragma clang fp contract(off)
  __shared__ double sh[4];
  for (int tile = blockIdx.x; tile < G.ntiles; tile += gridDim.x) {
    const int p = G.tile_prob[tile];
    const IonParams P = ip[p];
    double acc = 0.;
    const size_t base = (size_t)tile * G.tile_cells;
    for (int k = threadIdx.x; k < G.tile_cells; k += blockDim.x) {
      const size_t i = base + k;
      const double jH = P.jfac * (J[i] * P.jscale);
      const double ntot = n[i];
      double x;
      if (jH > 0.) {
        const double aa = 0.5 * jH / (ntot * alpha);
        if (form == 0) {
          x = 1. + aa * (1. - sqrt(2. / aa + 1.));
        } else {
          const double b = 2. * aa;
          x = 2. / (2. + b + sqrt(b * (b + 4.)));
        }
      } else {
        x = 1.;
      }
      acc += fabs(x - xn[i]);
      xn[i] = x;
      J[i] = 0.;
      kappa[i] = (float)(P.kappa_scale * ntot * x);
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
      double s = 0.;
      for (int w = 0; w < (int)(blockDim.x >> 6); ++w) s += sh[w];
      if (s != 0.) atomicAdd(&l1[p], s);
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------
struct TilesEngine {
  size_t cap = 0;         // packet capacity
  int ntiles = 0;
  int nprob = 0;
  size_t desc_cap = 0;
  Rec32 *P[2] = {nullptr, nullptr};
  uint32_t *dest[2] = {nullptr, nullptr};
  uint32_t *pid[2] = {nullptr, nullptr};
  uint32_t *idx = nullptr;
  uint32_t *rel = nullptr;      // survivors' positions in their next segment
  uint32_t *hist = nullptr, *seg = nullptr, *cursor = nullptr;
  Desc *desc = nullptr;        // work descriptors of the round
  uint32_t *ndesc = nullptr;
  // [3]: packets of the round being planned, packets of the round just
  // processed, number of plans that found packets
  uint32_t *total = nullptr;
  unsigned *ticket = nullptr;
  uint32_t *h_total = nullptr;  // pinned, [3]
  SrcTile *src_tile = nullptr;  // each source's start tile, cell and round-0 offset
  int tile_key[3] = {0, 0, 0};
  std::vector<hipEvent_t> ev;   // start/stop of every process launch
  // problem set
  ProbGeo *d_probs = nullptr;
  uint16_t *d_tile_prob = nullptr;
  SetSources *d_src = nullptr, *h_src = nullptr;  // device / pinned host
  IonParams *d_ion = nullptr, *h_ion = nullptr;
  double *d_l1 = nullptr;
  uint32_t *d_prob_round = nullptr, *d_round_ctr = nullptr;
  unsigned long long *d_prob_time = nullptr, *d_t0 = nullptr;
  unsigned long long *d_prob_cross = nullptr;
  std::vector<ProbGeo> probs;
  // captured iteration graph of a problem set
  hipGraph_t graph = nullptr;
  hipGraphExec_t graph_exec = nullptr;
  uint64_t graph_key = 0;
};

static void free_engine(TilesEngine *e) {
  if (!e) return;
  if (e->graph_exec) hipGraphExecDestroy(e->graph_exec);
  if (e->graph) hipGraphDestroy(e->graph);
  for (int i = 0; i < 2; ++i) {
    hipFree(e->P[i]);
    hipFree(e->dest[i]);
    hipFree(e->pid[i]);
  }
  hipFree(e->idx);
  hipFree(e->rel);
  hipFree(e->hist);
  hipFree(e->seg);
  hipFree(e->cursor);
  hipFree(e->desc);
  hipFree(e->ndesc);
  hipFree(e->total);
  hipFree(e->ticket);
  hipFree(e->src_tile);
  if (e->h_total) hipHostFree(e->h_total);
  hipFree(e->d_probs);
  hipFree(e->d_tile_prob);
  hipFree(e->d_src);
  if (e->h_src) hipHostFree(e->h_src);
  hipFree(e->d_ion);
  if (e->h_ion) hipHostFree(e->h_ion);
  hipFree(e->d_l1);
  hipFree(e->d_prob_round);
  hipFree(e->d_round_ctr);
  hipFree(e->d_prob_time);
  hipFree(e->d_t0);
  hipFree(e->d_prob_cross);
  for (hipEvent_t v : e->ev) hipEventDestroy(v);
  delete e;
}

void tiles_destroy(cmi_ctx *c) {
  free_engine(c->tiles);
  c->tiles = nullptr;
}

// The problem geometry of a single context: one problem, its own subgrids.
static std::vector<ProbGeo> ctx_problems(const cmi_ctx *c) {
  ProbGeo P{};
  for (int a = 0; a < 3; ++a) {
    P.N[a] = c->ncell[a];
    P.nts[a] = c->g.ntile[a];
    P.nsub[a] = c->nsub[a];
    P.ntl[a] = c->ncell[a] / c->tile[a];
    P.box[a] = c->box[a];
    P.inv_cell_size[a] = c->g.inv_cell_size[a];
    P.cell_size[a] = c->g.cell_size[a];
    P.unit_scale[a] = c->g.unit_scale[a];
  }
  P.tiles_per_sub = P.nts[0] * P.nts[1] * P.nts[2];
  P.tile_base = 0;
  return {P};
}

constexpr int kReemitRoundCap = 100000;
constexpr size_t kLdsBytes = 160 * 1024;  // LDS of one CU on gfx950

// The walk-state layout of a tile shape (WalkLayout); bits() of the result
// is at most 32 for every tile set_tile accepts.

static TileGeo make_geo(const cmi_ctx *c, const TilesEngine *e) {
  TileGeo G{};
  for (int a = 0; a < 3; ++a) {
    G.T[a] = c->tile[a];
    G.tshift[a] = c->g.tshift[a];
  }
  G.tile_cells = c->g.tile_cells;
  G.W = walk_layout(G.T[0], G.T[1], G.T[2]);
  G.ntiles = e->ntiles;
  G.nprob = e->nprob;
  G.probs = e->d_probs;
  G.tile_prob = e->d_tile_prob;
  // fixed-point deposits: a segment inside one cell is at most the cell's
  // diagonal in Delta_0 units (plus FP32 rounding), and s 2^S must stay
  // within 2^31
  double diag = 1.0;
  for (const ProbGeo &P : e->probs) {
    double d2 = 0.0;
    for (int a = 0; a < 3; ++a) d2 += 1.0 / (P.unit_scale[a] * P.unit_scale[a]);
    diag = std::max(diag, std::sqrt(d2));
  }
  const int S = std::max(0, std::min(30, (int)std::floor(31.0 - std::log2(diag * 1.01))));
  G.fix_scale = std::ldexp(1.0f, S);
  G.fix_inv = std::ldexp(1.0, -S);
  return G;
}

static int max_rounds_of(const std::vector<ProbGeo> &probs) {
  // a packet's tile index moves monotonically along every axis, so it visits
  // at most ntl_x + ntl_y + ntl_z - 2 tiles
  int m = 0;
  for (const ProbGeo &P : probs) m = std::max(m, P.ntl[0] + P.ntl[1] + P.ntl[2]);
  return m + 2;
}

// (Re)build the engine for a problem set and a packet capacity.
static int ensure_engine(cmi_ctx *c, const std::vector<ProbGeo> &probs,
                         size_t cap, bool with_pid) {
  TilesEngine *e = c->tiles;
  int ntiles = 0;
  for (const ProbGeo &P : probs)
    ntiles += P.tiles_per_sub * P.nsub[0] * P.nsub[1] * P.nsub[2];
  bool same = e && e->cap >= cap && e->ntiles == ntiles &&
              e->nprob == (int)probs.size() && (!with_pid || e->pid[0]) &&
              e->tile_key[0] == c->tile[0] && e->tile_key[1] == c->tile[1] &&
              e->tile_key[2] == c->tile[2];
  if (same) {
    for (size_t q = 0; q < probs.size() && same; ++q)
      same = memcmp(&e->probs[q], &probs[q], sizeof(ProbGeo)) == 0;
  }
  if (same) return CMI_OK;
  free_engine(e);
  c->tiles = e = nullptr;
  if (probs.size() > (size_t)kMaxProblems)
    return fail(c, CMI_ERR_ARG, "too many problems in one set");
  e = new TilesEngine();
  e->cap = cap;
  e->ntiles = ntiles;
  e->nprob = (int)probs.size();
  e->probs = probs;
  e->desc_cap = (size_t)ntiles + cap / kChunkFloor + 16;
  for (int a = 0; a < 3; ++a) e->tile_key[a] = c->tile[a];
  const int np = e->nprob;
  hipError_t err = hipSuccess;
  for (int i = 0; i < 2 && err == hipSuccess; ++i) {
    err = hipMalloc(&e->P[i], cap * sizeof(Rec32));
    if (err == hipSuccess) err = hipMalloc(&e->dest[i], cap * sizeof(uint32_t));
    if (err == hipSuccess && with_pid) err = hipMalloc(&e->pid[i], cap * sizeof(uint32_t));
  }
  if (err == hipSuccess) err = hipMalloc(&e->idx, cap * sizeof(uint32_t));
  if (err == hipSuccess) err = hipMalloc(&e->rel, cap * sizeof(uint32_t));
  if (err == hipSuccess) err = hipMalloc(&e->hist, ntiles * sizeof(uint32_t));
  if (err == hipSuccess) err = hipMalloc(&e->seg, ntiles * sizeof(uint32_t));
  if (err == hipSuccess) err = hipMalloc(&e->cursor, ntiles * sizeof(uint32_t));
  if (err == hipSuccess) err = hipMalloc(&e->desc, e->desc_cap * sizeof(Desc));
  if (err == hipSuccess) err = hipMalloc(&e->ndesc, sizeof(uint32_t));
  if (err == hipSuccess) err = hipMalloc(&e->total, 3 * sizeof(uint32_t));
  if (err == hipSuccess) err = hipMalloc(&e->ticket, sizeof(unsigned));
  if (err == hipSuccess) err = hipHostMalloc(&e->h_total, 3 * sizeof(uint32_t));
  if (err == hipSuccess) err = hipMalloc(&e->src_tile, kMaxSetSources * sizeof(SrcTile));
  if (err == hipSuccess) err = hipMalloc(&e->d_probs, np * sizeof(ProbGeo));
  if (err == hipSuccess) err = hipMalloc(&e->d_tile_prob, ntiles * sizeof(uint16_t));
  if (err == hipSuccess) err = hipMalloc(&e->d_src, sizeof(SetSources));
  if (err == hipSuccess) err = hipHostMalloc(&e->h_src, sizeof(SetSources));
  if (err == hipSuccess) err = hipMalloc(&e->d_ion, np * sizeof(IonParams));
  if (err == hipSuccess) err = hipHostMalloc(&e->h_ion, np * sizeof(IonParams));
  if (err == hipSuccess) err = hipMalloc(&e->d_l1, np * sizeof(double));
  if (err == hipSuccess) err = hipMalloc(&e->d_prob_round, np * sizeof(uint32_t));
  if (err == hipSuccess) err = hipMalloc(&e->d_round_ctr, sizeof(uint32_t));
  if (err == hipSuccess) err = hipMalloc(&e->d_prob_time, np * sizeof(unsigned long long));
  if (err == hipSuccess) err = hipMalloc(&e->d_t0, sizeof(unsigned long long));
  if (err == hipSuccess) err = hipMalloc(&e->d_prob_cross, np * sizeof(unsigned long long));
  if (err != hipSuccess) {
    free_engine(e);
    return fail(c, CMI_ERR_ALLOC, "tiles engine: device allocation failed");
  }
  std::vector<uint16_t> tp(ntiles);
  for (int q = 0; q < np; ++q) {
    const ProbGeo &P = probs[q];
    const int nt = P.tiles_per_sub * P.nsub[0] * P.nsub[1] * P.nsub[2];
    for (int t = 0; t < nt; ++t) tp[P.tile_base + t] = (uint16_t)q;
  }
  err = hipMemcpy(e->d_probs, probs.data(), np * sizeof(ProbGeo), hipMemcpyHostToDevice);
  if (err == hipSuccess)
    err = hipMemcpy(e->d_tile_prob, tp.data(), ntiles * sizeof(uint16_t), hipMemcpyHostToDevice);
  if (err != hipSuccess) {
    free_engine(e);
    return fail(c, CMI_ERR_HIP, "tiles engine: set-up copy failed");
  }
  hipLaunchKernelGGL(zero_u32_kernel, dim3(64), dim3(256), 0, c->stream, e->hist,
                     (size_t)ntiles);
  const int max_rounds = max_rounds_of(probs) + 2;
  e->ev.resize(2 * (size_t)max_rounds, nullptr);
  for (hipEvent_t &v : e->ev) {
    if (hipEventCreate(&v) != hipSuccess) {
      free_engine(e);
      return fail(c, CMI_ERR_ALLOC, "tiles engine: event creation failed");
    }
  }
  c->tiles = e;
  return CMI_OK;
}

// Kernel-launch parameters of the engine.
struct RoundCfg {
  int threads, refill, grid;
  uint32_t chunk, min_visits;
  uint32_t finish_max;  // walk-out threshold: packets left at a host check (0: never)
  bool fp32acc;
  size_t lds;
  unsigned scatter_blocks;
  unsigned place_blocks;
};

// A round with fewer packets than (resident workgroups / kMinVisitsDiv) full
// chunks is cut into about that many visits.  Measured on the 256^3 clumpy
// workload (1e8 packets, 768 resident workgroups): cutting the last ~15
// rounds, whose largest visits then set the round's time, saves 2 % of the
// transport; cutting earlier rounds (4 or 1 visits per workgroup) costs more
// in per-visit latency (descriptor, opacity staging, flush) than it gains.
constexpr uint32_t kMinVisitsDiv = 4;

static RoundCfg round_cfg(const cmi_ctx *c, const TileGeo &G, size_t n0) {
  RoundCfg r;
  r.threads = c->tuning.block_threads == 256 ? 256 : 512;
  r.refill = c->tuning.refill_batch > 0 ? std::min(64, c->tuning.refill_batch) : 32;
  r.chunk = c->tuning.chunk_min > 0
                ? std::max<uint32_t>(kChunkFloor, (uint32_t)c->tuning.chunk_min)
                : kChunk;
  r.fp32acc = c->tuning.lds_fp32 != 0;
  r.lds = ((sizeof(float) * (size_t)(G.W.pad + G.tile_cells) + 15) & ~(size_t)15) +
          (size_t)G.tile_cells * (r.fp32acc ? sizeof(float) : sizeof(double));
  const int wg_per_cu = (int)std::max<size_t>(1, std::min<size_t>(8, (160 * 1024) / (r.lds + 1024)));
  r.grid = c->num_cus * wg_per_cu;
  r.min_visits = std::max<uint32_t>(1, (uint32_t)r.grid / kMinVisitsDiv);
  // walk-out: when a check of the packets left (at the start and every
  // kRoundsPerCheck rounds) finds at most finish_max, one launch walks them
  // all to the end.  Default: one packet per resident lane.  Measured on the
  // 256^3 clumpy workload: 10 % less time for 3e6-packet transports, 4 % for
  // 1e7, none for 1e8 (the walk-out does one global FP64 atomic per
  // crossing, about 10x the cost of a crossing in a visit, so it pays only
  // where it replaces nearly empty rounds).  tuning.finish_packets > 0 sets
  // the threshold, < 0 turns it off.  Not for problem sets or reemission.
  const long long fp = c->tuning.finish_packets;
  r.finish_max = (G.nprob != 1 || G.reemit_p > 0.f || fp < 0)
                     ? 0u
                     : (fp > 0 ? (uint32_t)std::min<long long>(fp, 0x7fffffff)
                               : (uint32_t)((size_t)r.grid * r.threads));
  const size_t per_block = (size_t)kScatterThreads * kScatterItems;
  r.scatter_blocks = (unsigned)std::min<size_t>(8192, (n0 + per_block - 1) / per_block + 1);
  r.place_blocks = (unsigned)std::min<size_t>(8192, (n0 + 4 * 256 - 1) / (4 * 256) + 1);
  return r;
}

static void launch_process(cmi_ctx *c, const TileGeo &G, const RoundCfg &R,
                           bool with_pid, PacketOut po,
                           unsigned long long *prob_cross) {
  TilesEngine *e = c->tiles;
#define CMI_PROC(ACC, TH, LIST, RE, TL)                                            \
  hipLaunchKernelGGL((tiles_process_kernel<ACC, TH, LIST, RE, TL>), dim3(R.grid), dim3(TH), \
                     (TL) ? 0 : R.lds, c->stream, G, c->f.kappa, c->f.J, e->desc,      \
                     e->ndesc, e->ticket, e->P[0], e->idx, e->P[1], e->dest[1], e->hist, \
                     LIST ? e->pid[0] : nullptr, LIST ? e->pid[1] : nullptr, po,       \
                     c->d_cnt, prob_cross, R.refill, e->rel)
  // 16^3 tiles at 512 threads (the tuned shape): the kernel specialised to
  // them, with its LDS block static.  Emission-driven transports without
  // reemission only: the specialisation's constants cost the packet-list and
  // reemitting kernels registers (89 and 99 VGPRs against 75 and 86), and
  // with them a wave per SIMD.
  const bool t16 = R.threads == 512 && G.T[0] == 16 && G.T[1] == 16 && G.T[2] == 16;
  if (G.reemit_p > 0.f) {
    // reemission: emission-driven transports only
    if (R.threads == 256) CMI_PROC(FixAcc, 256, false, true, 0);
    else CMI_PROC(FixAcc, 512, false, true, 0);
    return;
  }
  if (R.fp32acc) {
    if (with_pid) {
      if (R.threads == 256) CMI_PROC(float, 256, true, false, 0);
      else CMI_PROC(float, 512, true, false, 0);
    } else {
      if (R.threads == 256) CMI_PROC(float, 256, false, false, 0);
      else CMI_PROC(float, 512, false, false, 0);
    }
  } else if (with_pid) {
    if (R.threads == 256) CMI_PROC(FixAcc, 256, true, false, 0);
    else CMI_PROC(FixAcc, 512, true, false, 0);
  } else {
    if (R.threads == 256) CMI_PROC(FixAcc, 256, false, false, 0);
    else if (t16) CMI_PROC(FixAcc, 512, false, false, 16);
    else CMI_PROC(FixAcc, 512, false, false, 0);
  }
#undef CMI_PROC
}

// One walk-out launch over the records of the next round (P[0], idx).
static void launch_walkout(cmi_ctx *c, const TileGeo &G, const RoundCfg &R,
                           bool with_pid, PacketOut po) {
  TilesEngine *e = c->tiles;
  const int max_tiles = max_rounds_of(e->probs) + 2;
#define CMI_WALKOUT(TH, LIST)                                                        \
  hipLaunchKernelGGL((tiles_walkout_kernel<TH, LIST>), dim3(R.grid), dim3(TH), 0,      \
                     c->stream, G, c->f.kappa, c->f.J, e->total, e->P[0], e->idx,      \
                     LIST ? e->pid[0] : nullptr, po, c->d_cnt, max_tiles)
  if (with_pid) {
    if (R.threads == 256) CMI_WALKOUT(256, true); else CMI_WALKOUT(512, true);
  } else {
    if (R.threads == 256) CMI_WALKOUT(256, false); else CMI_WALKOUT(512, false);
  }
#undef CMI_WALKOUT
}

// nonempty: count the plans that find packets (total[2]) — the rounds after
// round 0 that had work.
static void launch_plan(cmi_ctx *c, const TileGeo &G, const RoundCfg &R,
                        bool track, bool nonempty = false) {
  TilesEngine *e = c->tiles;
  hipLaunchKernelGGL(tiles_plan_kernel, dim3(1), dim3(kPlanThreads), 0, c->stream,
                     G.ntiles, e->hist, e->seg, e->cursor, e->desc, e->ndesc,
                     e->total, e->ticket, R.chunk, R.min_visits, e->d_tile_prob, e->nprob,
                     track ? e->d_prob_round : nullptr, e->d_prob_time,
                     e->d_round_ctr, nonempty ? e->total + 2 : nullptr);
}

// The host checks the packets left every kRoundsPerCheck rounds: a check
// costs a host round trip (about 30 us), a round launched after the last
// packet has gone about 20 us (its kernels find no work).
constexpr int kRoundsPerCheck = 8;

// Rounds of process / plan / place until no packet is left.  On entry round
// 0 is planned (R sized for its packets) and its index list built, with the
// records in P[0].
static int run_rounds(cmi_ctx *c, const TileGeo &G, const RoundCfg &R, bool with_pid,
                      PacketOut po, size_t n0, uint32_t *rounds_out) {
  TilesEngine *e = c->tiles;
  // straight rays visit a bounded number of tiles; reemitted packets turn,
  // so their rounds are bounded only by the cap
  const int max_rounds = G.reemit_p > 0.f ? kReemitRoundCap : max_rounds_of(e->probs) + 2;
  const int timed = c->kernel_timing ? std::min<int>(max_rounds, (int)(e->ev.size() / 2)) : 0;
  hipLaunchKernelGGL(zero_u32_kernel, dim3(1), dim3(64), 0, c->stream, e->total + 2,
                     (size_t)1);
  int r = 0;
  bool drained = false;
  if (n0 <= R.finish_max) {
    // few enough from the start: one walk-out launch, no rounds
    launch_walkout(c, G, R, with_pid, po);
    CMI_TRY_HIP(c, hipGetLastError());
    drained = true;
  }
  while (r < max_rounds && !drained) {
    const int r_end = std::min(max_rounds, r + kRoundsPerCheck);
    for (; r < r_end; ++r) {
      // round r reads P[0] (through idx) and writes its survivors to P[1];
      // the plan keeps the round's count in total[1] for the placement
      if (r < timed) CMI_TRY_HIP(c, hipEventRecord(e->ev[2 * r], c->stream));
      launch_process(c, G, R, with_pid, po, nullptr);
      if (r < timed) CMI_TRY_HIP(c, hipEventRecord(e->ev[2 * r + 1], c->stream));
      CMI_TRY_HIP(c, hipGetLastError());
      launch_plan(c, G, R, false, true);
      hipLaunchKernelGGL(tiles_place_kernel, dim3(R.place_blocks), dim3(256), 0, c->stream,
                         e->total + 1, e->dest[1], e->rel, e->seg, e->idx);
      CMI_TRY_HIP(c, hipGetLastError());
      std::swap(e->P[0], e->P[1]);
      std::swap(e->dest[0], e->dest[1]);
      if (with_pid) std::swap(e->pid[0], e->pid[1]);
    }
    CMI_TRY_HIP(c, hipMemcpyAsync(e->h_total, e->total, 3 * sizeof(uint32_t),
                                  hipMemcpyDeviceToHost, c->stream));
    CMI_TRY_HIP(c, hipStreamSynchronize(c->stream));
    drained = e->h_total[0] == 0;
    if (!drained && e->h_total[0] <= R.finish_max) {
      // few packets left: one launch walks them to the end (the plan that
      // found them counted this round in total[2])
      launch_walkout(c, G, R, with_pid, po);
      CMI_TRY_HIP(c, hipGetLastError());
      drained = true;
    }
  }
  if (rounds_out) *rounds_out = r == 0 ? 1u : 1u + e->h_total[2];
  if (!drained) return fail(c, CMI_ERR_DIRECTION, "tiles engine: round limit reached");
  // device time of every process launch (the dominant kernel; walk-out
  // launches are not among them)
  const int nt = std::min(r, timed);
  for (int k = 0; k < nt; ++k) {
    float ms = 0.f;
    CMI_TRY_HIP(c, hipEventElapsedTime(&ms, e->ev[2 * k], e->ev[2 * k + 1]));
    c->kernel_time_sum += ms * 1e-3;
  }
  c->kernel_calls += (uint64_t)nt;
  return CMI_OK;
}

static void fill_set_sources_single(cmi_ctx *c, TilesEngine *e, uint64_t seed,
                                    uint32_t iteration) {
  SetSources &S = *e->h_src;
  const SourcesDev &H = *c->h_src;
  S.nprob = 1;
  S.seed = seed;
  S.iteration = iteration;
  S.pk_cum[0] = H.cum[H.nsrc - 1];
  S.src_begin[0] = 0;
  S.src_begin[1] = H.nsrc;
  for (int s = 0; s < H.nsrc; ++s) {
    S.src_cum[s] = H.cum[s];
    for (int a = 0; a < 3; ++a) S.pos[s][a] = H.pos[s][a];
  }
}

static void fill_stats_tiles(cmi_ctx *c, cmi_stats *stats, uint64_t n,
                             uint32_t rounds, double seconds) {
  if (!stats) return;
  stats->reemitted = c->h_cnt->reemitted;
  stats->packets = n;
  stats->crossings = c->h_cnt->crossings;
  stats->hops = c->h_cnt->hops;
  stats->absorbed = c->h_cnt->absorbed;
  stats->escaped = c->h_cnt->escaped;
  stats->errors = c->h_cnt->errors;
  stats->rounds = rounds;
  stats->reserved = 0;
  stats->seconds = seconds;
  stats->walkout_crossings = c->h_cnt->walkout_crossings;
  stats->walkout_hops = c->h_cnt->walkout_hops;
}

int tiles_transport(cmi_ctx *c, uint64_t npackets, uint64_t seed,
                    uint32_t iteration, double, cmi_stats *stats) {
  if (c->precision != CMI_PREC_FP32)
    return fail(c, CMI_ERR_ARG, "the tiles engine runs FP32 contexts");
  if (npackets == 0 || npackets > 0xfffffff0ull)
    return fail(c, CMI_ERR_ARG, "tiles engine: 1 .. 2^32-16 packets per call");
  for (int a = 0; a < 3; ++a)
    if (c->ncell[a] > 1024) return fail(c, CMI_ERR_ARG, "tiles engine: at most 1024 cells per axis");
  int rc = ensure_engine(c, ctx_problems(c), (size_t)npackets, false);
  if (rc != CMI_OK) return rc;
  TilesEngine *e = c->tiles;
  TileGeo G = make_geo(c, e);
  G.reemit_p = (float)c->reemission;
  G.reemit_key[0] = (uint32_t)seed ^ 0xA511E9B3u;
  G.reemit_key[1] = (uint32_t)(seed >> 32) ^ (iteration * 0x9E3779B9u) ^ 0x63D83595u;
  fill_set_sources_single(c, e, seed, iteration);
  CMI_TRY_HIP(c, hipMemcpyAsync(e->d_src, e->h_src, sizeof(SetSources),
                                hipMemcpyHostToDevice, c->stream));
  const RoundCfg R = round_cfg(c, G, (size_t)npackets);
  if (R.lds > kLdsBytes)
    return fail(c, CMI_ERR_ARG, "tiles engine: the tile's opacity and accumulator exceed the LDS");
  hipLaunchKernelGGL(zero_counters_kernel, dim3(1), dim3(64), 0, c->stream, c->d_cnt);
  CMI_TRY_HIP(c, hipEventRecord(c->ev0, c->stream));
  // round 0: per-tile counts from the source split, planned, then the
  // emission fills the records and the index list
  hipLaunchKernelGGL(tiles_source_hist_kernel, dim3(1), dim3(256), 0, c->stream, G,
                     e->d_src, e->hist, e->src_tile);
  launch_plan(c, G, R, false);
  const unsigned eb = (unsigned)((npackets + 255) / 256);
  hipLaunchKernelGGL(tiles_emit_kernel, dim3(eb), dim3(256), 0, c->stream, G, e->d_src,
                     npackets, e->P[0], e->seg, e->src_tile, e->idx);
  CMI_TRY_HIP(c, hipGetLastError());
  PacketOut po{};
  uint32_t rounds = 0;
  rc = run_rounds(c, G, R, false, po, (size_t)npackets, &rounds);
  if (rc != CMI_OK) return rc;
  CMI_TRY_HIP(c, hipEventRecord(c->ev1, c->stream));
  CMI_TRY_HIP(c, hipMemcpyAsync(c->h_cnt, c->d_cnt, sizeof(Counters),
                                hipMemcpyDeviceToHost, c->stream));
  CMI_TRY_HIP(c, hipStreamSynchronize(c->stream));
  float ms = 0.f;
  CMI_TRY_HIP(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
  fill_stats_tiles(c, stats, npackets, rounds, ms * 1e-3);
  if (c->h_cnt->errors)
    return fail(c, CMI_ERR_DIRECTION, "tiles engine: packets stopped on a walk bound");
  return CMI_OK;
}

int tiles_transport_packets(cmi_ctx *c, cmi_photon *p, const uint32_t *start_sub,
                            size_t n, int32_t *status, uint32_t *crossings,
                            uint32_t *hops, cmi_stats *stats) {
  if (c->precision != CMI_PREC_FP32)
    return fail(c, CMI_ERR_ARG, "the tiles engine runs FP32 contexts");
  if (n > 0xfffffff0ull) return fail(c, CMI_ERR_ARG, "too many packets");
  if (c->reemission > 0.)
    return fail(c, CMI_ERR_ARG,
                "reemission on explicit packet lists runs on the walk engine");
  int rc = ensure_engine(c, ctx_problems(c), n, true);
  if (rc != CMI_OK) return rc;
  TilesEngine *e = c->tiles;
  const TileGeo G = make_geo(c, e);
  PhotonRec *dp = nullptr;
  uint32_t *ds = nullptr;
  int32_t *dst = nullptr, *dtile = nullptr;
  uint32_t *dc = nullptr, *dh = nullptr;
  float4 *fa = nullptr, *fb = nullptr;
  hipError_t err = hipMalloc(&dp, n * sizeof(PhotonRec));
  if (err == hipSuccess) err = hipMalloc(&ds, n * sizeof(uint32_t));
  if (err == hipSuccess) err = hipMalloc(&dst, n * sizeof(int32_t));
  if (err == hipSuccess) err = hipMalloc(&dtile, n * sizeof(int32_t));
  if (err == hipSuccess) err = hipMalloc(&dc, n * sizeof(uint32_t));
  if (err == hipSuccess) err = hipMalloc(&dh, n * sizeof(uint32_t));
  if (err == hipSuccess) err = hipMalloc(&fa, n * sizeof(float4));
  if (err == hipSuccess) err = hipMalloc(&fb, n * sizeof(float4));
  auto cleanup = [&]() {
    hipFree(dp); hipFree(ds); hipFree(dst); hipFree(dtile); hipFree(dc);
    hipFree(dh); hipFree(fa); hipFree(fb);
  };
  if (err != hipSuccess) {
    cleanup();
    return fail(c, CMI_ERR_ALLOC, "tiles list transport: allocation failed");
  }
  err = hipMemcpyAsync(dp, p, n * sizeof(PhotonRec), hipMemcpyHostToDevice, c->stream);
  if (err == hipSuccess)
    err = hipMemcpyAsync(ds, start_sub, n * sizeof(uint32_t), hipMemcpyHostToDevice, c->stream);
  if (err == hipSuccess) err = hipMemsetAsync(dst, 0xff, n * sizeof(int32_t), c->stream);
  if (err == hipSuccess) err = hipMemsetAsync(dtile, 0xff, n * sizeof(int32_t), c->stream);
  if (err == hipSuccess) err = hipMemsetAsync(dc, 0, n * sizeof(uint32_t), c->stream);
  if (err == hipSuccess) err = hipMemsetAsync(dh, 0, n * sizeof(uint32_t), c->stream);
  if (err != hipSuccess) {
    cleanup();
    return fail(c, CMI_ERR_HIP, "tiles list transport: copy failed");
  }
  hipLaunchKernelGGL(zero_counters_kernel, dim3(1), dim3(64), 0, c->stream, c->d_cnt);
  err = hipEventRecord(c->ev0, c->stream);
  if (err != hipSuccess) {
    cleanup();
    return fail(c, CMI_ERR_HIP, "tiles list transport: event record failed");
  }
  hipLaunchKernelGGL(tiles_load_list_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256),
                     0, c->stream, G, c->g, dp, ds, n, e->P[0], e->dest[0], e->pid[0],
                     e->hist);
  const RoundCfg R = round_cfg(c, G, n);
  if (R.lds > kLdsBytes) {
    cleanup();
    return fail(c, CMI_ERR_ARG, "tiles engine: the tile's opacity and accumulator exceed the LDS");
  }
  // round 0: planned from the loaded histogram, index list by destination
  launch_plan(c, G, R, false);
  hipLaunchKernelGGL(tiles_scatter_kernel, dim3(R.scatter_blocks), dim3(kScatterThreads),
                     0, c->stream, e->total, e->dest[0], e->seg, e->cursor, e->idx);
  PacketOut po{dst, dc, dh, fa, fb, dtile};
  uint32_t rounds = 0;
  rc = run_rounds(c, G, R, true, po, n, &rounds);
  if (rc != CMI_OK) {
    cleanup();
    return rc;
  }
  hipLaunchKernelGGL(tiles_finish_list_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256),
                     0, c->stream, G, c->g, n, fa, dtile, fb, dp);
  err = hipEventRecord(c->ev1, c->stream);
  if (err == hipSuccess) err = hipMemcpyAsync(p, dp, n * sizeof(PhotonRec), hipMemcpyDeviceToHost, c->stream);
  if (err == hipSuccess && status)
    err = hipMemcpyAsync(status, dst, n * sizeof(int32_t), hipMemcpyDeviceToHost, c->stream);
  if (err == hipSuccess && crossings)
    err = hipMemcpyAsync(crossings, dc, n * sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream);
  if (err == hipSuccess && hops)
    err = hipMemcpyAsync(hops, dh, n * sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream);
  if (err == hipSuccess)
    err = hipMemcpyAsync(c->h_cnt, c->d_cnt, sizeof(Counters), hipMemcpyDeviceToHost, c->stream);
  if (err == hipSuccess) err = hipStreamSynchronize(c->stream);
  float ms = 0.f;
  if (err == hipSuccess) err = hipEventElapsedTime(&ms, c->ev0, c->ev1);
  cleanup();
  if (err != hipSuccess) return fail(c, CMI_ERR_HIP, "tiles list transport failed");
  fill_stats_tiles(c, stats, n, rounds, ms * 1e-3);
  if (c->h_cnt->errors)
    return fail(c, CMI_ERR_DIRECTION, "tiles engine: packets stopped on a walk bound");
  return CMI_OK;
}

}  // namespace cmi

#include "cmi_set.inc"
