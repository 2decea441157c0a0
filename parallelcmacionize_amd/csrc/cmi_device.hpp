// cmi_device.hpp -- device-side building blocks shared by the transport and
// ionization kernels: the travel-direction algebra, the grid geometry and cell
// addressing, the counter-based RNG and packet emission, and the per-crossing
// steps in FP64 (reference arithmetic) and FP32 (cell units).
//
// Written for gfx950 (CDNA4, wave64).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace cmi {

// TravelDirection numbering of TravelDirections.hpp:46-75 (the ABI keeps it).
enum : int {
  TD_INSIDE = 0,
  TD_NUMBER = 27,
};

// Side of the cube named by each TravelDirection, as {-1,0,1} per axis
// (P = +1, N = -1, axis not named = 0).  The 27 rows are in enum order:
// INSIDE, CORNER PPP..NNN, EDGE_X PP..NN, EDGE_Y PP..NN, EDGE_Z PP..NN,
// FACE X_P, X_N, Y_P, Y_N, Z_P, Z_N.  Encoded as (ox+1) + 3(oy+1) + 9(oz+1).
__host__ __device__ constexpr int td_code(int ox, int oy, int oz) {
  return (ox + 1) + 3 * (oy + 1) + 9 * (oz + 1);
}

struct DirTables {
  signed char offset[TD_NUMBER][3];
  signed char from_code[27];  // code -> direction
};

__host__ __device__ constexpr DirTables make_dir_tables() {
  DirTables t{{{0, 0, 0},   {1, 1, 1},   {1, 1, -1},  {1, -1, 1},  {1, -1, -1},
               {-1, 1, 1},  {-1, 1, -1}, {-1, -1, 1}, {-1, -1, -1},
               {0, 1, 1},   {0, 1, -1},  {0, -1, 1},  {0, -1, -1},
               {1, 0, 1},   {1, 0, -1},  {-1, 0, 1},  {-1, 0, -1},
               {1, 1, 0},   {1, -1, 0},  {-1, 1, 0},  {-1, -1, 0},
               {1, 0, 0},   {-1, 0, 0},  {0, 1, 0},   {0, -1, 0},
               {0, 0, 1},   {0, 0, -1}},
              {}};
  for (int d = 0; d < TD_NUMBER; ++d) {
    t.from_code[td_code(t.offset[d][0], t.offset[d][1], t.offset[d][2])] =
        (signed char)d;
  }
  return t;
}

__constant__ inline DirTables c_dir = make_dir_tables();

// Exit direction of an index triple that may lie one cell outside the
// subgrid on each axis (DensitySubGrid.hpp:483-506: the high test is
// "index / n > 0", here "index >= n"; identical on [-1, n] and on every
// out-of-range value the walk can produce).  Returns -1 for no valid side.
__device__ __forceinline__ int exit_direction(int ix, int iy, int iz, int nx,
                                              int ny, int nz) {
  const int ox = (ix >= nx) - (ix < 0);
  const int oy = (iy >= ny) - (iy < 0);
  const int oz = (iz >= nz) - (iz < 0);
  return c_dir.from_code[td_code(ox, oy, oz)];
}

__device__ __forceinline__ int opposite_direction(int d) {
  return c_dir.from_code[td_code(-c_dir.offset[d][0], -c_dir.offset[d][1],
                                 -c_dir.offset[d][2])];
}

// Grid of subgrids as the device sees it.  The cell order in device memory
// is tile-major: subgrid, then tile within the subgrid (tx, ty, tz), then
// the cell within the tile (x slowest).  With tile == subgrid this is the
// reference's own order (DensitySubGrid.hpp:207-210).
struct GridDev {
  int nsub[3];       // subgrids per axis
  int sc[3];         // cells per subgrid per axis
  int tshift[3];     // log2 of the tile size per axis
  int ntile[3];      // tiles per subgrid per axis
  int tile_cells;    // cells per tile
  int cps;           // cells per subgrid
  int nsubgrid;
  int pad_;
  double cell_size[3];      // subgrid _cell_size (m)
  double inv_cell_size[3];  // subgrid _inv_cell_size (1/m)
  double unit_scale[3];     // FP32 mode: Delta_0 / Delta_a
  const double *anchors;    // [nsubgrid][3], subgrid _anchor (m)
  const uint32_t *ngbs;     // [nsubgrid][27], NEIGHBOUR_OUTSIDE = ~0u
};

__device__ __forceinline__ size_t cell_addr(const GridDev &g, int sub, int ix,
                                            int iy, int iz) {
  const int tx = ix >> g.tshift[0], ty = iy >> g.tshift[1],
            tz = iz >> g.tshift[2];
  const int lx = ix - (tx << g.tshift[0]), ly = iy - (ty << g.tshift[1]),
            lz = iz - (tz << g.tshift[2]);
  const int t = (tx * g.ntile[1] + ty) * g.ntile[2] + tz;
  const int l = (((lx << g.tshift[1]) + ly) << g.tshift[2]) + lz;
  return (size_t)sub * (size_t)g.cps + (size_t)t * (size_t)g.tile_cells +
         (size_t)l;
}

// ---------------------------------------------------------------------------
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as
// easy as 1, 2, 3", SC'11).
// ---------------------------------------------------------------------------
struct u32x4 {
  uint32_t x, y, z, w;
};

__host__ __device__ __forceinline__ u32x4 philox4x32_10(u32x4 c, uint32_t k0,
                                                        uint32_t k1) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    if (r > 0) {
      k0 += 0x9E3779B9u;
      k1 += 0xBB67AE85u;
    }
    const uint64_t p0 = (uint64_t)0xD2511F53u * c.x;
    const uint64_t p1 = (uint64_t)0xCD9E8D57u * c.z;
    c = u32x4{(uint32_t)(p1 >> 32) ^ c.y ^ k0, (uint32_t)p1,
              (uint32_t)(p0 >> 32) ^ c.w ^ k1, (uint32_t)p0};
  }
  return c;
}

// 53 random bits -> (0, 1); never 0 or 1.
__host__ __device__ __forceinline__ double uniform_open(uint32_t hi,
                                                        uint32_t lo) {
  const uint64_t m = ((uint64_t)(hi >> 5) << 26) | (uint64_t)(lo >> 6);
  return ((double)m + 0.5) * (1.0 / 9007199254740992.0);
}

// Sources for one transport call: positions, start subgrid, and cumulative
// packet counts (packet k belongs to the first source with k < cum[s]).
constexpr int kMaxSources = 64;
struct SourcesDev {
  int nsrc;
  uint32_t start_sub[kMaxSources];
  uint64_t cum[kMaxSources];
  double pos[kMaxSources][3];
};

// testDensitySubGrid_omp.cpp:529-597: isotropic direction and tau = -ln u,
// drawn from counters {k, iteration, 0} and {k, iteration, 1}.
__device__ __forceinline__ int emit_packet(const SourcesDev &S, uint64_t k,
                                           uint32_t iteration, uint32_t key0,
                                           uint32_t key1, double pos[3],
                                           double dir[3], double *tau) {
  int s = 0;
  while (s + 1 < S.nsrc && k >= S.cum[s]) ++s;
  const u32x4 r0 = philox4x32_10(
      u32x4{(uint32_t)k, (uint32_t)(k >> 32), iteration, 0u}, key0, key1);
  const u32x4 r1 = philox4x32_10(
      u32x4{(uint32_t)k, (uint32_t)(k >> 32), iteration, 1u}, key0, key1);
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
  return s;
}

// Counters a transport writes back (device-side, one slot each).
struct Counters {
  unsigned long long crossings;
  unsigned long long hops;
  unsigned long long absorbed;
  unsigned long long escaped;
  unsigned long long errors;  // invalid exit mask or exceeded walk bound
  unsigned long long next_packet;  // persistent-refill ticket
  unsigned long long reemitted;    // reemission events
  unsigned long long walkout_crossings;  // of crossings: those of walk-out launches
  unsigned long long walkout_hops;       // of hops: tile legs of walk-out launches
};

// ---------------------------------------------------------------------------
// Reemission draws (do_reemission, testDensitySubGrid_omp.cpp:675-702: a
// uniform below the probability, then get_random_direction 529-545 and
// tau = -ln u) for the r-th absorption of packet k.  Only operations IEEE 754
// rounds exactly (+ - * / sqrt and compares), so that a CPU restatement gives
// the same bits:
//   - the direction by Marsaglia's rejection pair (Ann. Math. Stat. 43 (1972)
//     645): a, b uniform in (-1, 1) with q = a^2 + b^2 < 1, then
//     (2a sqrt(1 - q), 2b sqrt(1 - q), 1 - 2q) is uniform on the sphere;
//   - the optical depth by von Neumann's comparison method (NBS Appl. Math.
//     Ser. 12 (1951) 36; Knuth, TAOCP 2, 3.4.1): trial j draws U0 > U1 > ...
//     until a draw exceeds its predecessor; an odd run length accepts
//     j + U0, which is Exp(1) like -ln u.
// Uniforms are taken in order, two per Philox block, from counters
// {k_lo, k_hi, iteration, 2 + 256 r + b}, b = 0, 1, ... (emission uses 0, 1).
// ---------------------------------------------------------------------------
constexpr int kReemitBlocks = 256;       // Philox blocks per absorption, at most
constexpr int kReemitDirTries = 32;      // direction pairs, at most
constexpr uint32_t kReemitMax = 65536;   // reemissions per packet, at most

struct ReemitDraws {
  double p;  // probability (0: off)
  uint32_t key0, key1, iteration, pad;
};

__host__ __device__ inline bool reemit_draw(uint64_t k, uint32_t r,
                                            const ReemitDraws &R, double dir[3],
                                            double *tau) {
#pragma clang fp contract(off)
  u32x4 ctr{(uint32_t)k, (uint32_t)(k >> 32), R.iteration, 2u + 256u * r};
  u32x4 buf{0u, 0u, 0u, 0u};
  int avail = 0, blocks = 0;
  auto uniform = [&]() -> double {
    if (avail == 0) {
      buf = philox4x32_10(ctr, R.key0, R.key1);
      ctr.w += 1u;
      blocks += 1;
      avail = 2;
    }
    const bool first_pair = avail == 2;
    avail -= 1;
    return first_pair ? uniform_open(buf.x, buf.y) : uniform_open(buf.z, buf.w);
  };
  if (!(uniform() < R.p)) return false;
  dir[0] = 0.;
  dir[1] = 0.;
  dir[2] = 1.;
  for (int t = 0; t < kReemitDirTries; ++t) {
    const double a = 2. * uniform() - 1.;
    const double b = 2. * uniform() - 1.;
    const double q = a * a + b * b;
    if (q < 1. && q > 0.) {
      const double f = 2. * sqrt(1. - q);
      dir[0] = a * f;
      dir[1] = b * f;
      dir[2] = 1. - 2. * q;
      break;
    }
  }
  double trials = 0.;
  *tau = 0.5;
  while (blocks < kReemitBlocks) {
    const double u0 = uniform();
    double prev = u0;
    int run = 1;
    for (;;) {
      const double v = uniform();
      if (!(v < prev)) break;
      prev = v;
      ++run;
      if (blocks >= kReemitBlocks) break;  // bound (never reached)
    }
    if (run & 1) {
      *tau = trials + u0;
      break;
    }
    trials += 1.;
  }
  return true;
}

}  // namespace cmi
