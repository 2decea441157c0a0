/*
 * cmi_oracle.h -- CPU oracle for the photon-packet transport and ionization
 * update of ParallelCMacIonize.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: it is loaded by
 * tests/, by __graft_entry__.smoke() and by bench.py's cpu_baseline leg, and by
 * nothing else.  The product path (parallelcmacionize_amd/, libcmi.so) never
 * links, loads or calls it.
 *
 * Everything here is a plain-C, all-FP64 restatement of the reference
 * algorithm; every function cites the reference file:line it follows.
 * Compile with -ffp-contract=off -fno-fast-math so that the arithmetic is the
 * IEEE sequence the reference source spells out.
 */
#ifndef CMI_ORACLE_H
#define CMI_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* TravelDirection numbering, identical to TravelDirections.hpp:46-75. */
enum {
  ORC_INSIDE = 0,
  ORC_CORNER_PPP, ORC_CORNER_PPN, ORC_CORNER_PNP, ORC_CORNER_PNN,
  ORC_CORNER_NPP, ORC_CORNER_NPN, ORC_CORNER_NNP, ORC_CORNER_NNN,
  ORC_EDGE_X_PP, ORC_EDGE_X_PN, ORC_EDGE_X_NP, ORC_EDGE_X_NN,
  ORC_EDGE_Y_PP, ORC_EDGE_Y_PN, ORC_EDGE_Y_NP, ORC_EDGE_Y_NN,
  ORC_EDGE_Z_PP, ORC_EDGE_Z_PN, ORC_EDGE_Z_NP, ORC_EDGE_Z_NN,
  ORC_FACE_X_P, ORC_FACE_X_N, ORC_FACE_Y_P, ORC_FACE_Y_N,
  ORC_FACE_Z_P, ORC_FACE_Z_N,
  ORC_NDIR
};

#define ORC_NEIGHBOUR_OUTSIDE 0xffffffffu /* DensitySubGrid.hpp:53 */

/* Packet record in the field order of Photon.hpp:52-65 (the _type member of
 * the reference is not carried: the transport only handles direct packets). */
typedef struct {
  double pos[3];
  double dir[3];
  double tau;    /* target optical depth (remaining) */
  double sigma;  /* photoionization cross section (m^2) */
  double weight;
} orc_photon;

/* One subgrid: geometry of DensitySubGrid.hpp:516-527; the cell arrays live
 * in the owning grid (subgrid-major, cell order ix*ny*nz + iy*nz + iz as in
 * DensitySubGrid.hpp:207-210). */
typedef struct {
  double anchor[3];
  double cell_size[3];
  double inv_cell_size[3];
  int ncell[4]; /* nx, ny, nz, ny*nz */
  uint32_t ngbs[ORC_NDIR];
  double *n;  /* number density (m^-3) */
  double *xn; /* neutral fraction */
  double *J;  /* intensity integral (m^3) */
} orc_subgrid;

typedef struct {
  double box[6];
  int ncell[3];
  int nsub[3];
  int nsubgrid;
  size_t cells_per_subgrid;
  size_t ncell_total;
  orc_subgrid *sub;
  double *n, *xn, *J; /* subgrid-major arrays of ncell_total doubles */
} orc_grid;

/* ---- travel-direction algebra (TravelDirections.hpp:89-403) ---- */
void orc_direction_offset(int direction, int offset[3]);
int orc_direction_from_offset(const int offset[3]);
int orc_output_to_input_direction(int output_direction);
int orc_direction_from_mask(int mask);
int orc_is_compatible_output_direction(const double dir[3], int direction);
int orc_is_compatible_input_direction(const double dir[3], int direction);

/* ---- subgrid index helpers (DensitySubGrid.hpp:207-506) ---- */
int orc_get_output_direction(const orc_subgrid *s, const int three_index[3]);
int orc_get_start_index(const orc_subgrid *s, const double position[3],
                        int input_direction, int three_index[3]);
void orc_subgrid_init(orc_subgrid *s, const double box[6], const int ncell[3]);

/* ---- traversal (DensitySubGrid.hpp:1180-1323) ---- */
int orc_interact(const orc_subgrid *s, orc_photon *photon, int input_direction,
                 uint32_t *crossings);
enum { ORC_WALK_INTERACT = 0, ORC_WALK_PROPAGATE = 1, ORC_WALK_OPTICAL_DEPTH = 2 };
int orc_walk(const orc_subgrid *s, orc_photon *photon, int input_direction,
             int kind, uint32_t *crossings);

/* ---- grid of subgrids (testDensitySubGrid_omp.cpp:1805-1877) ---- */
int orc_grid_create(orc_grid *g, const double box[6], const int ncell[3],
                    const int nsub[3]);
void orc_grid_destroy(orc_grid *g);
uint32_t orc_grid_source_subgrid(const orc_grid *g, const double pos[3]);

/* Serial transport of an explicit packet list through the whole grid.
 * status: 0 = absorbed, 1 = left the box.  Per packet: crossings, hops
 * (subgrid visits).  Packets are updated in place. */
void orc_transport_list(orc_grid *g, orc_photon *p, const uint32_t *start_sub,
                        size_t n, int32_t *status, uint32_t *crossings,
                        uint32_t *hops, int nthreads);

/* The same with reemission (do_reemission, testDensitySubGrid_omp.cpp:
 * 675-702): an absorbed packet is reemitted with the given probability from
 * its absorption point, in the same subgrid, entering INSIDE.  Packet i of
 * the list draws as packet number first + i of (seed, iteration), see
 * orc_reemit_draw.  status -2: ORC_REEMIT_MAX reemissions reached.
 * reemissions (may be NULL) receives each packet's count. */
#define ORC_REEMIT_MAX 65536
void orc_transport_list_reemit(orc_grid *g, orc_photon *p,
                               const uint32_t *start_sub, size_t n,
                               double probability, uint64_t seed,
                               uint32_t iteration, uint64_t first,
                               int32_t *status, uint32_t *crossings,
                               uint32_t *hops, uint32_t *reemissions,
                               int nthreads);

/* Draws of packet k's r-th absorption: 1 (and a new direction and target
 * optical depth) if it is reemitted, 0 if not.  At most ORC_REEMIT_BLOCKS
 * Philox blocks per absorption and ORC_REEMIT_DIR_TRIES direction pairs. */
#define ORC_REEMIT_BLOCKS 256
#define ORC_REEMIT_DIR_TRIES 32
int orc_reemit_draw(uint64_t k, uint32_t r, double probability, uint64_t seed,
                    uint32_t iteration, double dir[3], double *tau);

/* ---- ionization (DensitySubGrid.hpp:1628-1658) ---- */
/* form 0: the reference expression 1 + aa(1 - sqrt(2/aa + 1));
 * form 1: the cancellation-free root 2 / (2 + b + sqrt(b(b + 4))), b = 2 aa. */
void orc_compute_neutral_fraction(orc_grid *g, double luminosity,
                                  double num_photon, double alphaH, int form);

/* ---- counter-based RNG and emission ---- */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2],
                       uint32_t out[4]);
double orc_uniform_open(uint32_t hi, uint32_t lo);
void orc_source_counts(const double *L, int nsrc, uint64_t npackets,
                       uint64_t seed, uint32_t iteration, uint64_t *counts);
void orc_emit(const orc_grid *g, const double *src_pos, const double *src_L,
              int nsrc, uint64_t npackets, uint64_t first, uint64_t count,
              uint64_t seed, uint32_t iteration, double sigma,
              orc_photon *out, uint32_t *start_sub);
/* Problem q of a problem set: source split keyed by seed ^ q, counter high
 * word XOR q << 20; problem 0 is orc_emit. */
void orc_emit_problem(const orc_grid *g, const double *src_pos,
                      const double *src_L, int nsrc, uint64_t npackets,
                      uint64_t first, uint64_t count, uint64_t seed,
                      uint32_t iteration, double sigma, uint32_t problem,
                      orc_photon *out, uint32_t *start_sub);

/* One plain iteration (emission + transport + ionization); returns the total
 * number of cell crossings and writes the mean |dx_n| to *l1. */
uint64_t orc_iteration(orc_grid *g, const double *src_pos, const double *src_L,
                       int nsrc, uint64_t npackets, uint64_t seed,
                       uint32_t iteration, double sigma, double alphaH,
                       int form, int nthreads, double *l1);

#ifdef __cplusplus
}
#endif

#endif
