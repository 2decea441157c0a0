/*
 * ref_harness.cpp -- a thin C-ABI harness around the reference's own hot-path
 * headers (DensitySubGrid.hpp, TravelDirections.hpp, Photon.hpp,
 * PhotonBuffer.hpp, MemorySpace.hpp, Lock.hpp), compiled by oracle/Makefile
 * from the sources where they lie under the reference checkout into
 * oracle/_ref/.
 *
 * TEST INFRASTRUCTURE ONLY (the checker, and the timed CPU baseline).  None of
 * the reference's source is copied here: this file only #includes it and
 * drives it.
 *
 * Crossing counts: the reference keeps no crossing counter.  Every loop
 * iteration of DensitySubGrid::interact (DensitySubGrid.hpp:1221-1305)
 * deposits weight * lmin * sigma into exactly one cell, a straight ray visits
 * a cell of a subgrid at most once, and the weight takes no part in the walk.
 * So the harness counts crossings by walking a copy of the packet with
 * weight NaN over a zeroed scratch J and counting the cells that became NaN.
 */
#define NO_HYDRO
#include "DensitySubGrid.hpp"
#include "MemoryMap.hpp"
#include "MemorySpace.hpp"
#include "PhotonBuffer.hpp"
#include "TravelDirections.hpp"

#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <mutex>
#include <omp.h>
#include <vector>

int MPI_rank = 0, MPI_size = 1;

namespace {

struct HPhoton { /* field order of Photon.hpp:52-65 minus _type */
  double pos[3], dir[3], tau, sigma, weight;
};

void to_photon(const HPhoton &h, Photon &p) {
  p.set_position(h.pos[0], h.pos[1], h.pos[2]);
  p.set_direction(h.dir[0], h.dir[1], h.dir[2]);
  p.set_target_optical_depth(h.tau);
  p.set_photoionization_cross_section(h.sigma);
  p.set_weight(h.weight);
  p.set_type(PHOTONPACKETTYPE_DIRECT);
}

void from_photon(const Photon &p, HPhoton &h) {
  for (int a = 0; a < 3; ++a) {
    h.pos[a] = p.get_position()[a];
    h.dir[a] = p.get_direction()[a];
  }
  h.tau = p.get_target_optical_depth();
  h.sigma = p.get_photoionization_cross_section();
  h.weight = p.get_weight();
}

/* The grid of subgrids as testDensitySubGrid_omp.cpp:1805-1877 builds it. */
struct RefGrid {
  std::vector<DensitySubGrid *> sub;
  int nsub[3];
  size_t cps;
  RefGrid(const double *box, const int *ncell, const int *ns) {
    for (int a = 0; a < 3; ++a) nsub[a] = ns[a];
    const int tot = ns[0] * ns[1] * ns[2];
    const double side[3] = {box[3] / ns[0], box[4] / ns[1], box[5] / ns[2]};
    const int sc[3] = {ncell[0] / ns[0], ncell[1] / ns[1], ncell[2] / ns[2]};
    cps = (size_t)sc[0] * sc[1] * sc[2];
    sub.resize(tot);
    for (int index = 0; index < tot; ++index) {
      const int ix = index / (ns[1] * ns[2]);
      const int iy = (index - ix * ns[1] * ns[2]) / ns[2];
      const int iz = index - ix * ns[1] * ns[2] - iy * ns[2];
      const double sb[6] = {box[0] + ix * side[0], box[1] + iy * side[1],
                            box[2] + iz * side[2], side[0], side[1], side[2]};
      sub[index] = new DensitySubGrid(sb, sc);
      DensitySubGrid &g = *sub[index];
      for (int i = 0; i < TRAVELDIRECTION_NUMBER; ++i) {
        g.set_neighbour(i, NEIGHBOUR_OUTSIDE);
        g.set_active_buffer(i, NEIGHBOUR_OUTSIDE);
      }
      for (int nix = -1; nix < 2; ++nix)
        for (int niy = -1; niy < 2; ++niy)
          for (int niz = -1; niz < 2; ++niz) {
            const int cix = ix + nix, ciy = iy + niy, ciz = iz + niz;
            if (cix >= 0 && cix < ns[0] && ciy >= 0 && ciy < ns[1] &&
                ciz >= 0 && ciz < ns[2]) {
              const int ti[3] = {nix * sc[0], niy * sc[1], niz * sc[2]};
              g.set_neighbour(g.get_output_direction(ti),
                              cix * ns[1] * ns[2] + ciy * ns[2] + ciz);
            }
          }
    }
  }
  ~RefGrid() {
    for (auto *s : sub) delete s;
  }
  void set_fields(const double *n, const double *xn) {
    for (size_t k = 0; k < sub.size(); ++k)
      for (size_t i = 0; i < cps; ++i) {
        sub[k]->set_number_density(i, n[k * cps + i]);
        sub[k]->set_neutral_fraction(i, xn[k * cps + i]);
        sub[k]->set_intensity_integral(i, 0.);
      }
  }
  /* cell (flat) index order of DensitySubGrid.hpp:207-210 */
  void get_J(double *J, int ncx, int ncy, int ncz) const {
    for (size_t k = 0; k < sub.size(); ++k)
      for (int ix = 0; ix < ncx; ++ix)
        for (int iy = 0; iy < ncy; ++iy)
          for (int iz = 0; iz < ncz; ++iz)
            J[k * cps + ((size_t)ix * ncy + iy) * ncz + iz] =
                sub[k]->get_intensity_integral(ix, iy, iz);
  }
};

/* Count cells of subgrid s that are NaN in the shadow copy, and clear them. */
uint32_t count_and_clear_nan(DensitySubGrid &shadow, int nc[3]) {
  uint32_t c = 0;
  for (int ix = 0; ix < nc[0]; ++ix)
    for (int iy = 0; iy < nc[1]; ++iy)
      for (int iz = 0; iz < nc[2]; ++iz) {
        if (std::isnan(shadow.get_intensity_integral(ix, iy, iz))) {
          ++c;
          shadow.set_intensity_integral((ix * nc[1] + iy) * nc[2] + iz, 0.);
        }
      }
  return c;
}

size_t grid_cells(const int ncell[3]) {
  return (size_t)ncell[0] * ncell[1] * ncell[2];
}

} // namespace

extern "C" {

/* DensitySubGrid::get_output_direction (DensitySubGrid.hpp:483-506) */
int ref_get_output_direction(const int ncell[3], const int three_index[3]) {
  const double box[6] = {0., 0., 0., 1., 1., 1.};
  DensitySubGrid g(box, ncell);
  return g.get_output_direction(three_index);
}

/* TravelDirections::get_output_direction(mask) (TravelDirections.hpp:315) */
int ref_direction_from_mask(int mask) {
  return TravelDirections::get_output_direction(mask);
}

/* TravelDirections::output_to_input_direction (TravelDirections.hpp:89) */
int ref_output_to_input_direction(int d) {
  return TravelDirections::output_to_input_direction(d);
}

int ref_is_compatible_output_direction(const double dir[3], int d) {
  return TravelDirections::is_compatible_output_direction(dir, d) ? 1 : 0;
}

int ref_is_compatible_input_direction(const double dir[3], int d) {
  return TravelDirections::is_compatible_input_direction(dir, d) ? 1 : 0;
}

/* DensitySubGrid::get_start_index (DensitySubGrid.hpp:428-469) */
int ref_get_start_index(const double box[6], const int ncell[3],
                        const double position[3], int input_direction,
                        int three_index[3]) {
  DensitySubGrid g(box, ncell);
  return g.get_start_index(position, input_direction, three_index);
}

/* DensitySubGrid::interact (DensitySubGrid.hpp:1180-1323) on one subgrid:
 * packets are walked one after the other, all entering through
 * input_direction.  J (reference cell order) receives the deposits. */
void ref_interact_list(const double box[6], const int ncell[3],
                       const double *n, const double *xn, double *J,
                       double *photons, size_t np, int input_direction,
                       int32_t *exit_dir, uint32_t *crossings) {
  DensitySubGrid g(box, ncell), shadow(box, ncell);
  const size_t nc = (size_t)ncell[0] * ncell[1] * ncell[2];
  for (size_t i = 0; i < nc; ++i) {
    g.set_number_density(i, n[i]);
    g.set_neutral_fraction(i, xn[i]);
    g.set_intensity_integral(i, 0.);
    shadow.set_number_density(i, n[i]);
    shadow.set_neutral_fraction(i, xn[i]);
    shadow.set_intensity_integral(i, 0.);
  }
  int nc3[3] = {ncell[0], ncell[1], ncell[2]};
  HPhoton *hp = reinterpret_cast<HPhoton *>(photons);
  for (size_t k = 0; k < np; ++k) {
    Photon p, q;
    to_photon(hp[k], p);
    to_photon(hp[k], q);
    q.set_weight(NAN);
    exit_dir[k] = g.interact(p, input_direction);
    shadow.interact(q, input_direction);
    crossings[k] = count_and_clear_nan(shadow, nc3);
    from_photon(p, hp[k]);
  }
  for (int ix = 0; ix < ncell[0]; ++ix)
    for (int iy = 0; iy < ncell[1]; ++iy)
      for (int iz = 0; iz < ncell[2]; ++iz)
        J[((size_t)ix * ncell[1] + iy) * ncell[2] + iz] =
            g.get_intensity_integral(ix, iy, iz);
}

/* DensitySubGrid::propagate (kind 1) or compute_optical_depth (kind 2) for
 * a packet list.  Crossings from a shadow interact of the same packet (NaN
 * weight): propagate walks exactly as interact; compute_optical_depth walks
 * to the edge, as interact does with an infinite target optical depth. */
extern "C" void ref_walk_list(const double box[6], const int ncell[3],
                              const double *n, const double *xn, double *photons,
                              size_t np, int input_direction, int kind,
                              int32_t *exit_dir, uint32_t *crossings) {
  DensitySubGrid g(box, ncell), shadow(box, ncell);
  const size_t nc = (size_t)ncell[0] * ncell[1] * ncell[2];
  for (size_t i = 0; i < nc; ++i) {
    g.set_number_density(i, n[i]);
    g.set_neutral_fraction(i, xn[i]);
    g.set_intensity_integral(i, 0.);
    shadow.set_number_density(i, n[i]);
    shadow.set_neutral_fraction(i, xn[i]);
    shadow.set_intensity_integral(i, 0.);
  }
  int nc3[3] = {ncell[0], ncell[1], ncell[2]};
  HPhoton *hp = reinterpret_cast<HPhoton *>(photons);
  for (size_t k = 0; k < np; ++k) {
    Photon p, q;
    to_photon(hp[k], p);
    to_photon(hp[k], q);
    q.set_weight(NAN);
    if (kind == 2) q.set_target_optical_depth(INFINITY);
    exit_dir[k] = kind == 2 ? g.compute_optical_depth(p, input_direction)
                            : g.propagate(p, input_direction);
    shadow.interact(q, input_direction);
    crossings[k] = count_and_clear_nan(shadow, nc3);
    from_photon(p, hp[k]);
  }
}

/* Whole-grid transport of an explicit packet list, one packet at a time,
 * with the routing of testDensitySubGrid_omp.cpp:607-666 / 1149-1329
 * (INSIDE = absorbed, NEIGHBOUR_OUTSIDE = escaped, otherwise enter the
 * neighbour with output_to_input_direction).  status 0 absorbed, 1 escaped. */
void ref_transport_list(const double box[6], const int ncell[3],
                        const int nsub[3], const double *n, const double *xn,
                        double *J, double *photons, const uint32_t *start_sub,
                        size_t np, int32_t *status, uint32_t *crossings,
                        uint32_t *hops) {
  RefGrid grid(box, ncell, nsub), shadow(box, ncell, nsub);
  grid.set_fields(n, xn);
  shadow.set_fields(n, xn);
  int sc[3] = {ncell[0] / nsub[0], ncell[1] / nsub[1], ncell[2] / nsub[2]};
  HPhoton *hp = reinterpret_cast<HPhoton *>(photons);
  for (size_t k = 0; k < np; ++k) {
    Photon p, q;
    to_photon(hp[k], p);
    to_photon(hp[k], q);
    q.set_weight(NAN);
    unsigned int s = start_sub[k];
    int in_dir = TRAVELDIRECTION_INSIDE;
    uint32_t nc = 0, nh = 0;
    int32_t st = -1;
    for (;;) {
      ++nh;
      const int out = grid.sub[s]->interact(p, in_dir);
      shadow.sub[s]->interact(q, in_dir);
      nc += count_and_clear_nan(*shadow.sub[s], sc);
      if (out == TRAVELDIRECTION_INSIDE) {
        st = 0;
        break;
      }
      const unsigned int ngb = grid.sub[s]->get_neighbour(out);
      if (ngb == NEIGHBOUR_OUTSIDE) {
        st = 1;
        break;
      }
      in_dir = TravelDirections::output_to_input_direction(out);
      s = ngb;
    }
    from_photon(p, hp[k]);
    status[k] = st;
    crossings[k] = nc;
    hops[k] = nh;
  }
  grid.get_J(J, sc[0], sc[1], sc[2]);
}

/* DensitySubGrid::compute_neutral_fraction (DensitySubGrid.hpp:1628-1658),
 * alpha = 4e-19 as hard-coded there. */
void ref_compute_neutral_fraction(const double box[6], const int ncell[3],
                                  const double *n, const double *J,
                                  double luminosity, unsigned int num_photon,
                                  double *xn_out) {
  DensitySubGrid g(box, ncell);
  const size_t nc = (size_t)ncell[0] * ncell[1] * ncell[2];
  for (size_t i = 0; i < nc; ++i) {
    g.set_number_density(i, n[i]);
    g.set_intensity_integral(i, J[i]);
  }
  g.compute_neutral_fraction(luminosity, num_photon);
  const double *x = g.get_neutral_fraction();
  for (size_t i = 0; i < nc; ++i) xn_out[i] = x[i];
}

/* Timed CPU baseline: parallel transport of an explicit packet list with the
 * reference's own DensitySubGrid::interact and PhotonBuffer routing
 * (do_photon_traversal into 27 output buffers, output_to_input_direction,
 * get_neighbour; testDensitySubGrid_omp.cpp:607-666, 1149-1329).  Threads
 * take 200-packet source buffers from a shared counter; each thread owns a
 * private copy of the grid, which is the reference's subgrid-copy scheme
 * (testDensitySubGrid_omp.cpp:720-, 2116-2171) taken to one copy of every
 * subgrid per thread: no subgrid locks, no contention.  The copies' J are
 * summed at the end (update_intensities, DensitySubGrid.hpp:1096).  Returns
 * the wall time in seconds of the transport (the copies are set up before
 * the clock starts); J is returned in reference order. */
double ref_transport_tasks(const double box[6], const int ncell[3],
                           const int nsub[3], const double *n,
                           const double *xn, double *J, const double *photons,
                           const uint32_t *start_sub, size_t np, int nthreads,
                           size_t /*pool_size*/) {
  std::vector<RefGrid *> grids(nthreads, nullptr);
#pragma omp parallel num_threads(nthreads)
  {
    const int t = omp_get_thread_num();
    grids[t] = new RefGrid(box, ncell, nsub);
    grids[t]->set_fields(n, xn);
  }
  int sc[3] = {ncell[0] / nsub[0], ncell[1] / nsub[1], ncell[2] / nsub[2]};
  const HPhoton *hp = reinterpret_cast<const HPhoton *>(photons);
  std::atomic<size_t> next(0);
  const auto t0 = std::chrono::steady_clock::now();
#pragma omp parallel num_threads(nthreads)
  {
    RefGrid &grid = *grids[omp_get_thread_num()];
    std::vector<PhotonBuffer *> work, spare;
    PhotonBuffer local[TRAVELDIRECTION_NUMBER];
    auto get_buffer = [&]() {
      if (spare.empty()) return new PhotonBuffer();
      PhotonBuffer *b = spare.back();
      spare.pop_back();
      b->reset();
      return b;
    };
    for (;;) {
      const size_t k0 = next.fetch_add(PHOTONBUFFER_SIZE);
      if (k0 >= np) break;
      const size_t k1 = std::min(np, k0 + PHOTONBUFFER_SIZE);
      // source buffers: consecutive packets that share a start subgrid
      for (size_t k = k0; k < k1;) {
        PhotonBuffer *b = get_buffer();
        b->set_subgrid_index(start_sub[k]);
        b->set_direction(TRAVELDIRECTION_INSIDE);
        const uint32_t sg = start_sub[k];
        while (k < k1 && start_sub[k] == sg) {
          to_photon(hp[k], (*b)[b->get_next_free_photon()]);
          ++k;
        }
        work.push_back(b);
      }
      while (!work.empty()) {
        PhotonBuffer *in = work.back();
        work.pop_back();
        DensitySubGrid &g = *grid.sub[in->get_subgrid_index()];
        for (int d = 0; d < TRAVELDIRECTION_NUMBER; ++d) local[d].reset();
        for (unsigned int i = 0; i < in->size(); ++i) {
          Photon &p = (*in)[i];
          const int out = g.interact(p, in->get_direction());
          if (out != TRAVELDIRECTION_INSIDE &&
              g.get_neighbour(out) != NEIGHBOUR_OUTSIDE) {
            local[out][local[out].get_next_free_photon()] = p;
          }
        }
        spare.push_back(in);
        for (int d = 1; d < TRAVELDIRECTION_NUMBER; ++d) {
          if (local[d].size() == 0) continue;
          PhotonBuffer *b = get_buffer();
          *b = local[d];
          b->set_subgrid_index(g.get_neighbour(d));
          b->set_direction(TravelDirections::output_to_input_direction(d));
          work.push_back(b);
        }
      }
    }
    for (PhotonBuffer *b : spare) delete b;
  }
  const auto t1 = std::chrono::steady_clock::now();
  const size_t nc = grid_cells(ncell);
  std::vector<double> Jt(nc);
  for (size_t i = 0; i < nc; ++i) J[i] = 0.;
  for (int t = 0; t < nthreads; ++t) {
    grids[t]->get_J(Jt.data(), sc[0], sc[1], sc[2]);
    for (size_t i = 0; i < nc; ++i) J[i] += Jt[i];
    delete grids[t];
  }
  return std::chrono::duration<double>(t1 - t0).count();
}

} /* extern "C" */

/* The reference's memory-mapped result file (output_neutral_fractions,
 * testDensitySubGrid_omp.cpp:394-440): one block of get_output_size() bytes
 * per subgrid, written by DensitySubGrid::output_intensities
 * (DensitySubGrid.hpp:1738-1796). */
extern "C" size_t ref_output_dat(const double box[6], const int ncell[3],
                                 const int nsub[3], const double *n,
                                 const double *xn, const char *path) {
  RefGrid g(box, ncell, nsub);
  g.set_fields(n, xn);
  const size_t block = g.sub[0]->get_output_size();
  {
    MemoryMap file(path, g.sub.size() * block);
    for (size_t k = 0; k < g.sub.size(); ++k)
      g.sub[k]->output_intensities(k * block, file);
  }
  return g.sub.size() * block;
}
