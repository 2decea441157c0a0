// cmi.hpp -- C++ host interface over the C ABI (include/cmi.h), shaped like
// the reference's hot-path classes so that code written against
// DensitySubGrid / Photon / PhotonBuffer / TravelDirections / MemorySpace
// keeps its call sites:
//
//   Photon               Photon.hpp:49-248 (same accessors; layout of
//                        cmi_photon, no packet type)
//   PhotonBuffer         PhotonBuffer.hpp:53-209 (200-packet bin with target
//                        subgrid and entry TravelDirection)
//   TravelDirections     TravelDirections.hpp:89 / 315
//   MemorySpace          MemorySpace.hpp:62-125 (pool of PhotonBuffers with
//                        the add_photons spill)
//   DensitySubGridGrid   the grid of DensitySubGrids a driver builds
//                        (testDensitySubGrid_omp.cpp:1805-1877); owns one
//                        cmi_ctx
//   DensitySubGrid       one subgrid: interact (per packet and per buffer),
//                        the cell setters/getters, neighbours, active
//                        buffers and the copy reduction; a view into a
//                        DensitySubGridGrid, or stand-alone (the reference's
//                        constructor and copy constructor) with a
//                        one-subgrid grid of its own
//
// Errors: the reference aborts through cmac_error (Error.hpp:104-113); this
// interface throws cmi::Error carrying the ABI's error code and message
// instead, so a caller can recover.
#pragma once

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/cmi.h"

namespace cmi {

class Error : public std::runtime_error {
 public:
  Error(int code, const std::string &msg)
      : std::runtime_error(msg + " (code " + std::to_string(code) + ")"),
        code_(code) {}
  int code() const { return code_; }

 private:
  int code_;
};

// TravelDirection values (TravelDirections.hpp:46-75).
enum TravelDirection {
  TRAVELDIRECTION_INSIDE = 0,
  TRAVELDIRECTION_CORNER_PPP, TRAVELDIRECTION_CORNER_PPN,
  TRAVELDIRECTION_CORNER_PNP, TRAVELDIRECTION_CORNER_PNN,
  TRAVELDIRECTION_CORNER_NPP, TRAVELDIRECTION_CORNER_NPN,
  TRAVELDIRECTION_CORNER_NNP, TRAVELDIRECTION_CORNER_NNN,
  TRAVELDIRECTION_EDGE_X_PP, TRAVELDIRECTION_EDGE_X_PN,
  TRAVELDIRECTION_EDGE_X_NP, TRAVELDIRECTION_EDGE_X_NN,
  TRAVELDIRECTION_EDGE_Y_PP, TRAVELDIRECTION_EDGE_Y_PN,
  TRAVELDIRECTION_EDGE_Y_NP, TRAVELDIRECTION_EDGE_Y_NN,
  TRAVELDIRECTION_EDGE_Z_PP, TRAVELDIRECTION_EDGE_Z_PN,
  TRAVELDIRECTION_EDGE_Z_NP, TRAVELDIRECTION_EDGE_Z_NN,
  TRAVELDIRECTION_FACE_X_P, TRAVELDIRECTION_FACE_X_N,
  TRAVELDIRECTION_FACE_Y_P, TRAVELDIRECTION_FACE_Y_N,
  TRAVELDIRECTION_FACE_Z_P, TRAVELDIRECTION_FACE_Z_N,
  TRAVELDIRECTION_NUMBER
};

static const unsigned int NEIGHBOUR_OUTSIDE = CMI_NEIGHBOUR_OUTSIDE;

namespace TravelDirections {
// TravelDirections.hpp:89
inline int output_to_input_direction(int output_direction) {
  const int d = cmi_output_to_input_direction(output_direction);
  if (d < 0) throw Error(CMI_ERR_ARG, "Unknown output direction");
  return d;
}
// TravelDirections.hpp:315 (-1 for an invalid mask, like the reference)
inline int get_output_direction(int mask) { return cmi_direction_from_mask(mask); }
}  // namespace TravelDirections

// Photon.hpp:49-248
class Photon {
 public:
  const double *get_position() const { return p_.pos; }
  void set_position(double x, double y, double z) {
    p_.pos[0] = x;
    p_.pos[1] = y;
    p_.pos[2] = z;
  }
  const double *get_direction() const { return p_.dir; }
  double *get_direction() { return p_.dir; }
  void set_direction(double x, double y, double z) {
    p_.dir[0] = x;
    p_.dir[1] = y;
    p_.dir[2] = z;
  }
  double get_target_optical_depth() const { return p_.tau; }
  void set_target_optical_depth(double t) { p_.tau = t; }
  double get_photoionization_cross_section() const { return p_.sigma; }
  void set_photoionization_cross_section(double s) { p_.sigma = s; }
  double get_weight() const { return p_.weight; }
  void set_weight(double w) { p_.weight = w; }
  cmi_photon &raw() { return p_; }
  const cmi_photon &raw() const { return p_; }

 private:
  cmi_photon p_{};
};
static_assert(sizeof(Photon) == sizeof(cmi_photon), "Photon must wrap cmi_photon");

// PhotonBuffer.hpp:42-209
static const unsigned int PHOTONBUFFER_SIZE = 200u;

class PhotonBuffer {
 public:
  unsigned int size() const { return size_; }
  void reset() { size_ = 0; }
  const Photon &operator[](unsigned int i) const { return photons_[i]; }
  Photon &operator[](unsigned int i) { return photons_[i]; }
  unsigned int get_next_free_photon() {
    if (size_ >= PHOTONBUFFER_SIZE) throw Error(CMI_ERR_OVERFLOW, "PhotonBuffer overflow");
    return size_++;
  }
  void grow(unsigned int size) {
    if (size > PHOTONBUFFER_SIZE) throw Error(CMI_ERR_OVERFLOW, "PhotonBuffer overflow");
    size_ = size;
  }
  void set_subgrid_index(unsigned int s) { subgrid_ = s; }
  unsigned int get_subgrid_index() const { return subgrid_; }
  void set_direction(int d) { direction_ = d; }
  int get_direction() const { return direction_; }
  Photon *data() { return photons_; }

 private:
  unsigned int subgrid_ = 0;
  int direction_ = TRAVELDIRECTION_INSIDE;
  unsigned int size_ = 0;
  Photon photons_[PHOTONBUFFER_SIZE];
};

// MemorySpace.hpp:62-125: a pool of buffers addressed by index.
class MemorySpace {
 public:
  explicit MemorySpace(size_t size) : buffers_(size), used_(size, false) {}
  PhotonBuffer &operator[](size_t i) { return buffers_[i]; }
  size_t get_free_buffer() {
    for (size_t k = 0; k < used_.size(); ++k) {
      const size_t i = (next_ + k) % used_.size();
      if (!used_[i]) {
        used_[i] = true;
        next_ = i + 1;
        buffers_[i].reset();
        return i;
      }
    }
    throw Error(CMI_ERR_OVERFLOW, "No more free elements in memory space!");
  }
  void free_buffer(size_t i) {
    buffers_[i].reset();
    used_[i] = false;
  }
  // Append buffer to the pool buffer `index`; when it fills up, the rest
  // spills into a fresh buffer with the same target, whose index is returned
  // (the reference returns the new active index the same way).
  size_t add_photons(size_t index, PhotonBuffer &buffer) {
    PhotonBuffer &dst = buffers_[index];
    unsigned int k = 0;
    while (dst.size() < PHOTONBUFFER_SIZE && k < buffer.size())
      dst[dst.get_next_free_photon()] = buffer[k++];
    size_t out = index;
    if (dst.size() == PHOTONBUFFER_SIZE) {
      out = get_free_buffer();
      PhotonBuffer &nb = buffers_[out];
      nb.set_subgrid_index(dst.get_subgrid_index());
      nb.set_direction(dst.get_direction());
      while (k < buffer.size()) nb[nb.get_next_free_photon()] = buffer[k++];
    }
    return out;
  }

 private:
  std::vector<PhotonBuffer> buffers_;
  std::vector<bool> used_;
  size_t next_ = 0;
};

class DensitySubGridGrid;

// One subgrid of a DensitySubGridGrid.
class DensitySubGrid {
 public:
  // subgrid `index` of grid g (the grid hands these out)
  DensitySubGrid(DensitySubGridGrid *g, unsigned int index) : g_(g), index_(index) {}
  // DensitySubGrid.hpp:516: a stand-alone subgrid of ncell cells in box
  // (anchor box[0..2], sides box[3..5]); n = 1e8, x_n = 1e-6, J = 0.  It
  // owns a one-subgrid grid on the GPU.
  inline DensitySubGrid(const double *box, const int *ncell,
                        int precision = CMI_PREC_FP64);
  // DensitySubGrid.hpp:618: a copy of a subgrid (the reference's subgrid
  // copies): the original's geometry, n and x_n, J = 0, in a one-subgrid
  // grid of its own.  Deposits into a copy stay in the copy until
  // update_intensities adds them to the original.
  inline DensitySubGrid(const DensitySubGrid &original);
  inline DensitySubGrid(DensitySubGrid &&other) noexcept;
  DensitySubGrid &operator=(const DensitySubGrid &) = delete;
  inline ~DensitySubGrid();
  // DensitySubGrid::interact (DensitySubGrid.hpp:1180): one packet.
  inline int interact(Photon &photon, int input_direction);
  // The packet loop of do_photon_traversal (testDensitySubGrid_omp.cpp:
  // 619-665) for one input buffer: every packet walks this subgrid, its exit
  // direction is written to exit_dirs (size() entries).
  inline void interact(PhotonBuffer &buffer, int *exit_dirs,
                       uint32_t *crossings = nullptr);
  // DensitySubGrid::propagate (DensitySubGrid.hpp:1333): the walk without
  // deposits.
  inline int propagate(Photon &photon, int input_direction);
  // DensitySubGrid::compute_optical_depth (DensitySubGrid.hpp:1482): walk to
  // the subgrid's edge, adding the path's optical depth to the packet's.
  inline int compute_optical_depth(Photon &photon, int input_direction);
  // DensitySubGrid.hpp:820, 832: the neighbour table as the driver set it.
  // A grid's transport routes by the table its geometry defines; shoot()
  // checks that every neighbour a driver set agrees with it (CMI_ERR_ARG
  // otherwise).  The per-subgrid walks do not read it: a driver that routes
  // on the host reads its own values back.
  inline unsigned int get_neighbour(int direction) const;
  inline void set_neighbour(int direction, unsigned int ngb);
  // DensitySubGrid.hpp:878-892: the index of the subgrid's active buffer in
  // each direction, for drivers that route buffers on the host
  inline unsigned int get_active_buffer(int direction) const;
  inline void set_active_buffer(int direction, unsigned int index);
  // DensitySubGrid.hpp:1083-1089: x_n from the original, J reset (called on
  // a copy after the original's ionization update)
  inline void update_neutral_fractions(const DensitySubGrid &original);
  // DensitySubGrid.hpp:1096-1108: add a copy's J to this subgrid's
  inline void update_intensities(const DensitySubGrid &copy);
  // DensitySubGrid.hpp:483-506 (public in the reference for unit tests)
  inline int get_output_direction(const int three_index[3]) const;
  // DensitySubGrid.hpp:428-469
  inline int get_start_index(const double position[3], int input_direction,
                             int three_index[3]) const;
  inline size_t get_number_of_cells() const;
  inline void set_number_density(unsigned int i, double v);
  inline void set_neutral_fraction(unsigned int i, double v);
  inline void set_intensity_integral(unsigned int i, double v);
  inline double get_intensity_integral(int ix, int iy, int iz) const;
  inline const double *get_neutral_fraction() const;
  unsigned int index() const { return index_; }
  DensitySubGridGrid &grid() { return *g_; }

 private:
  DensitySubGridGrid *g_;
  unsigned int index_;
  std::unique_ptr<DensitySubGridGrid> own_;  // stand-alone subgrids and copies
};

// The grid of subgrids (testDensitySubGrid_omp.cpp:1805-1877) on one GPU.
class DensitySubGridGrid {
 public:
  DensitySubGridGrid(const double box[6], const int ncell[3], const int nsub[3],
                     int precision = CMI_PREC_FP64)
      : precision_(precision) {
    check(cmi_create(box, ncell, nsub, precision, &ctx_), "cmi_create");
    uint64_t ns = 0, nc = 0;
    cmi_get_sizes(ctx_, &ns, &nc);
    nsubgrid_ = ns;
    ncell_ = nc;
    cps_ = nc / ns;
    for (int a = 0; a < 6; ++a) box_[a] = box[a];
    for (int a = 0; a < 3; ++a) {
      sc_[a] = ncell[a] / nsub[a];
      nsub_[a] = nsub[a];
      // DensitySubGrid.hpp:519-520 with the subgrid box of
      // testDensitySubGrid_omp.cpp:1806-1808
      side_[a] = box[3 + a] / nsub[a];
      cell_size_[a] = side_[a] / sc_[a];
      inv_cell_size_[a] = sc_[a] / side_[a];
    }
    n_.assign(nc, 1.e8);
    xn_.assign(nc, 1.e-6);
    J_.assign(nc, 0.);
    ngb_geometry_.resize(27 * ns);
    for (uint64_t s = 0; s < ns; ++s)
      check(cmi_get_neighbours(ctx_, (uint32_t)s, &ngb_geometry_[27 * s]),
            "cmi_get_neighbours");
    ngb_declared_ = ngb_geometry_;
    active_.assign(27 * ns, NEIGHBOUR_OUTSIDE);
    subgrids_.reserve(ns);
    for (uint64_t s = 0; s < ns; ++s) subgrids_.emplace_back(this, (unsigned)s);
  }
  ~DensitySubGridGrid() { cmi_destroy(ctx_); }
  DensitySubGridGrid(const DensitySubGridGrid &) = delete;
  DensitySubGridGrid &operator=(const DensitySubGridGrid &) = delete;

  size_t size() const { return nsubgrid_; }
  DensitySubGrid &operator[](size_t i) { return subgrids_[i]; }
  cmi_ctx *context() { return ctx_; }

  void set_engine(int engine, const cmi_tuning *t = nullptr) {
    check(cmi_set_engine(ctx_, engine, t), "cmi_set_engine");
  }
  void set_sources(const std::vector<double> &pos, const std::vector<double> &L) {
    check(cmi_set_sources(ctx_, pos.data(), L.data(), (int)L.size()), "cmi_set_sources");
  }
  // One photon-shooting phase (the worker loop of
  // testDensitySubGrid_omp.cpp:2274-2401) with the GPU's own routing.
  cmi_stats shoot(uint64_t npackets, uint64_t seed, uint32_t iteration,
                  double sigma = 6.3e-22) {
    check_neighbours();
    sync_in();
    cmi_stats st{};
    check(cmi_transport(ctx_, npackets, seed, iteration, sigma, &st), "cmi_transport");
    host_stale_ = true;
    return st;
  }
  // DensitySubGrid::compute_neutral_fraction (DensitySubGrid.hpp:1628) on
  // every subgrid; alpha defaults to the reference's 4e-19.
  double compute_neutral_fraction(double luminosity, uint64_t num_photon,
                                  double alphaH = 4.e-19, int form = 0) {
    sync_in();
    double l1 = 0.;
    check(cmi_update_ionization(ctx_, luminosity, num_photon, alphaH, form, &l1),
          "cmi_update_ionization");
    host_stale_ = true;
    return l1;
  }

  // reemission_probability of the reference's parameter file (its
  // do_reemission, testDensitySubGrid_omp.cpp:675-702): shoot() on both
  // engines, explicit lists on the walk engine
  void set_reemission_probability(double p) {
    check(cmi_set_reemission(ctx_, p), "cmi_set_reemission");
  }
  // key of the reemission draws of explicit-list transports (no counterpart
  // in the reference, whose threads own RANLUX generators)
  void set_reemission_draws(uint64_t seed, uint32_t iteration) {
    check(cmi_set_reemission_draws(ctx_, seed, iteration),
          "cmi_set_reemission_draws");
  }
  // output_neutral_fractions (testDensitySubGrid_omp.cpp:394-440) and
  // output_neutral_fractions_binary (507-521): the reference's result files
  void output_neutral_fractions(const char *dat_path = "neutral_fractions.dat") {
    sync_in();
    check(cmi_write_results(ctx_, dat_path, nullptr), "cmi_write_results");
  }
  void output_neutral_fractions_binary(const char *bin_path = "neutral_fractions.bin") {
    sync_in();
    check(cmi_write_results(ctx_, nullptr, bin_path), "cmi_write_results");
  }

  // --- used by the DensitySubGrid views ---
  void sync_in() {
    if (dirty_fields_) {
      check(cmi_set_fields(ctx_, n_.data(), xn_.data()), "cmi_set_fields");
      dirty_fields_ = false;
    }
    if (dirty_J_) {
      check(cmi_set_intensity(ctx_, J_.data()), "cmi_set_intensity");
      dirty_J_ = false;
    }
  }
  void sync_out() const {
    if (host_stale_) {
      auto *self = const_cast<DensitySubGridGrid *>(this);
      self->check(cmi_get_fields(ctx_, self->J_.data(), self->xn_.data()),
                  "cmi_get_fields");
      self->host_stale_ = false;
    }
  }
  void check(int rc, const char *what) {
    if (rc != CMI_OK) throw Error(rc, std::string(what) + ": " + cmi_last_error(ctx_));
  }
  // The neighbours a driver set (set_neighbour) against the ones the
  // transport routes by (the grid's geometry, testDensitySubGrid_omp.cpp:
  // 1836-1874).
  void check_neighbours() const {
    if (!ngb_dirty_) return;
    for (size_t k = 0; k < ngb_declared_.size(); ++k) {
      if (ngb_declared_[k] != ngb_geometry_[k])
        throw Error(CMI_ERR_ARG,
                    "set_neighbour: subgrid " + std::to_string(k / 27) +
                        ", direction " + std::to_string(k % 27) + ": " +
                        std::to_string(ngb_declared_[k]) + " differs from the " +
                        "grid's neighbour " + std::to_string(ngb_geometry_[k]));
    }
    ngb_dirty_ = false;
  }

 private:
  friend class DensitySubGrid;
  cmi_ctx *ctx_ = nullptr;
  int precision_;
  size_t nsubgrid_ = 0, ncell_ = 0, cps_ = 0;
  double box_[6];
  int sc_[3], nsub_[3];
  double side_[3], cell_size_[3], inv_cell_size_[3];
  std::vector<DensitySubGrid> subgrids_;
  std::vector<double> n_, xn_, J_;
  std::vector<uint32_t> ngb_geometry_, ngb_declared_, active_;
  mutable bool ngb_dirty_ = false;
  bool dirty_fields_ = true, dirty_J_ = false, host_stale_ = false;
};

inline DensitySubGrid::DensitySubGrid(const double *box, const int *ncell,
                                      int precision)
    : g_(nullptr), index_(0) {
  const int one[3] = {1, 1, 1};
  own_.reset(new DensitySubGridGrid(box, ncell, one, precision));
  g_ = own_.get();
}

inline DensitySubGrid::DensitySubGrid(const DensitySubGrid &original)
    : g_(nullptr), index_(0) {
  DensitySubGridGrid &og = *original.g_;
  og.sync_out();
  // the original's subgrid box, as testDensitySubGrid_omp.cpp:1799-1808
  // computes it (anchor = box anchor + index * side)
  const size_t s = original.index_;
  const int ix = (int)(s / ((size_t)og.nsub_[1] * og.nsub_[2]));
  const int iy = (int)((s / og.nsub_[2]) % og.nsub_[1]);
  const int iz = (int)(s % og.nsub_[2]);
  const double sub[6] = {og.box_[0] + ix * og.side_[0], og.box_[1] + iy * og.side_[1],
                         og.box_[2] + iz * og.side_[2], og.side_[0], og.side_[1],
                         og.side_[2]};
  const int one[3] = {1, 1, 1};
  own_.reset(new DensitySubGridGrid(sub, og.sc_, one, og.precision_));
  g_ = own_.get();
  const size_t cps = og.cps_;
  std::memcpy(g_->n_.data(), &og.n_[s * cps], cps * sizeof(double));
  std::memcpy(g_->xn_.data(), &og.xn_[s * cps], cps * sizeof(double));
  g_->dirty_fields_ = true;
}

inline DensitySubGrid::DensitySubGrid(DensitySubGrid &&other) noexcept
    : g_(other.g_), index_(other.index_), own_(std::move(other.own_)) {}

inline DensitySubGrid::~DensitySubGrid() = default;

inline int DensitySubGrid::interact(Photon &photon, int input_direction) {
  g_->sync_out();
  g_->sync_in();
  int32_t out = 0;
  uint32_t nc = 0;
  g_->check(cmi_interact_batch(g_->ctx_, index_, input_direction, &photon.raw(), 1,
                               &out, &nc),
            "cmi_interact_batch");
  g_->host_stale_ = true;
  return out;
}

inline void DensitySubGrid::interact(PhotonBuffer &buffer, int *exit_dirs,
                                     uint32_t *crossings) {
  g_->sync_out();
  g_->sync_in();
  std::vector<uint32_t> nc(buffer.size());
  g_->check(cmi_interact_batch(g_->ctx_, index_, buffer.get_direction(),
                               &buffer.data()->raw(), buffer.size(),
                               reinterpret_cast<int32_t *>(exit_dirs), nc.data()),
            "cmi_interact_batch");
  if (crossings) std::memcpy(crossings, nc.data(), nc.size() * sizeof(uint32_t));
  g_->host_stale_ = true;
}

inline int DensitySubGrid::propagate(Photon &photon, int input_direction) {
  g_->sync_out();
  g_->sync_in();
  int32_t out = 0;
  uint32_t nc = 0;
  g_->check(cmi_propagate_batch(g_->ctx_, index_, input_direction, &photon.raw(), 1,
                                &out, &nc),
            "cmi_propagate_batch");
  return out;
}

inline int DensitySubGrid::compute_optical_depth(Photon &photon, int input_direction) {
  g_->sync_out();
  g_->sync_in();
  int32_t out = 0;
  uint32_t nc = 0;
  g_->check(cmi_optical_depth_batch(g_->ctx_, index_, input_direction, &photon.raw(), 1,
                                    &out, &nc),
            "cmi_optical_depth_batch");
  return out;
}

inline unsigned int DensitySubGrid::get_neighbour(int direction) const {
  if (direction < 0 || direction >= 27) throw Error(CMI_ERR_ARG, "output_direction");
  return g_->ngb_declared_[27 * (size_t)index_ + direction];
}

inline void DensitySubGrid::set_neighbour(int direction, unsigned int ngb) {
  if (direction < 0 || direction >= 27) throw Error(CMI_ERR_ARG, "output_direction");
  g_->ngb_declared_[27 * (size_t)index_ + direction] = ngb;
  // a stand-alone subgrid or copy is routed by its driver: nothing to check
  if (!own_) g_->ngb_dirty_ = true;
}

inline unsigned int DensitySubGrid::get_active_buffer(int direction) const {
  if (direction < 0 || direction >= 27) throw Error(CMI_ERR_ARG, "direction");
  return g_->active_[27 * (size_t)index_ + direction];
}

inline void DensitySubGrid::set_active_buffer(int direction, unsigned int index) {
  if (direction < 0 || direction >= 27) throw Error(CMI_ERR_ARG, "direction");
  g_->active_[27 * (size_t)index_ + direction] = index;
}

inline void DensitySubGrid::update_neutral_fractions(const DensitySubGrid &original) {
  if (original.g_->cps_ != g_->cps_)
    throw Error(CMI_ERR_ARG, "update_neutral_fractions: cell counts differ");
  original.g_->sync_out();
  g_->sync_out();
  const size_t cps = g_->cps_;
  std::memcpy(&g_->xn_[index_ * cps], &original.g_->xn_[original.index_ * cps],
              cps * sizeof(double));
  std::fill(g_->J_.begin() + index_ * cps, g_->J_.begin() + (index_ + 1) * cps, 0.);
  g_->dirty_fields_ = true;
  g_->dirty_J_ = true;
}

inline void DensitySubGrid::update_intensities(const DensitySubGrid &copy) {
  if (copy.g_->cps_ != g_->cps_)
    throw Error(CMI_ERR_ARG, "update_intensities: cell counts differ");
  copy.g_->sync_out();
  g_->sync_out();
  const size_t cps = g_->cps_;
  const double *src = &copy.g_->J_[copy.index_ * cps];
  double *dst = &g_->J_[index_ * cps];
  for (size_t i = 0; i < cps; ++i) dst[i] += src[i];
  g_->dirty_J_ = true;
}

inline size_t DensitySubGrid::get_number_of_cells() const { return g_->cps_; }

inline int DensitySubGrid::get_output_direction(const int idx[3]) const {
  int mask = 0;
  for (int a = 0; a < 3; ++a) {
    const int lo = idx[a] < 0;
    const int hi = (idx[a] / g_->sc_[a]) > 0;
    mask |= (hi << (5 - 2 * a)) | (lo << (4 - 2 * a));
  }
  const int d = TravelDirections::get_output_direction(mask);
  if (d < 0) throw Error(CMI_ERR_DIRECTION, "Unknown outgoing check mask");
  return d;
}

inline int DensitySubGrid::get_start_index(const double position[3],
                                           int input_direction,
                                           int idx[3]) const {
  if (input_direction < 0 || input_direction >= TRAVELDIRECTION_NUMBER)
    throw Error(CMI_ERR_ARG, "Unknown incoming direction");
  // the side an entry direction names: the N<->P swap of the exit side of
  // the same name, read off the exit mask table
  int side[3] = {0, 0, 0};
  for (int mask = 0; mask < 64; ++mask) {
    if (TravelDirections::get_output_direction(mask) == input_direction) {
      for (int a = 0; a < 3; ++a)
        side[a] = ((mask >> (5 - 2 * a)) & 1) - ((mask >> (4 - 2 * a)) & 1);
      break;
    }
  }
  for (int a = 0; a < 3; ++a) {
    idx[a] = side[a] == 0 ? (int)(position[a] * g_->inv_cell_size_[a])
                          : (side[a] < 0 ? 0 : g_->sc_[a] - 1);
  }
  return (idx[0] * g_->sc_[1] + idx[1]) * g_->sc_[2] + idx[2];
}

inline void DensitySubGrid::set_number_density(unsigned int i, double v) {
  g_->sync_out();
  g_->n_[index_ * g_->cps_ + i] = v;
  g_->dirty_fields_ = true;
}

inline void DensitySubGrid::set_neutral_fraction(unsigned int i, double v) {
  g_->sync_out();
  g_->xn_[index_ * g_->cps_ + i] = v;
  g_->dirty_fields_ = true;
}

inline void DensitySubGrid::set_intensity_integral(unsigned int i, double v) {
  g_->sync_out();
  g_->J_[index_ * g_->cps_ + i] = v;
  g_->dirty_J_ = true;
}

inline double DensitySubGrid::get_intensity_integral(int ix, int iy, int iz) const {
  g_->sync_out();
  return g_->J_[index_ * g_->cps_ + ((size_t)ix * g_->sc_[1] + iy) * g_->sc_[2] + iz];
}

inline const double *DensitySubGrid::get_neutral_fraction() const {
  g_->sync_out();
  return g_->xn_.data() + index_ * g_->cps_;
}

}  // namespace cmi
