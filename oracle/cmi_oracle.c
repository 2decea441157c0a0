/*
 * cmi_oracle.c -- CPU oracle (all FP64) for the photon-packet transport and
 * ionization update of ParallelCMacIonize.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP path, never the product.
 * Loaded only by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg.
 *
 * Pinning (what each part is checked against; see tests/test_oracle.py):
 *   - direction algebra, get_start_index, get_output_direction: the 54 cases
 *     of the reference's own known-answer test
 *     (testDensitySubGridDirections.cpp:105-167), the N<->P involution, and
 *     bit-for-bit agreement with oracle/_ref (the reference headers compiled
 *     by oracle/Makefile) on every mask and entry direction;
 *   - orc_interact / orc_transport_list: bit-for-bit agreement (crossing
 *     counts, exit directions, positions, remaining tau, per-cell sums) with
 *     oracle/_ref on seeded packet lists, committed golden fixtures written by
 *     the reference build, closed forms (chord length in an empty box,
 *     tau/(sigma n x) absorption length in a uniform box, crossing count of a
 *     straight ray) and conservation (sum of deposits = sigma w path length);
 *   - orc_compute_neutral_fraction: bit-for-bit against oracle/_ref and the
 *     golden fixtures (form 0); the ionization-balance residual, jH = 0 -> 1
 *     and the aa -> infinity limit (both forms);
 *   - Philox4x32-10: the published known-answer vectors of Salmon et al.
 *     (SC'11, Random123 kat_vectors);
 *   - orc_emit / orc_iteration: statistical properties (isotropy, exponential
 *     tau, photon conservation) and the analytic Stromgren radius.  The
 *     reference's own emission uses RANLUX (RandomGenerator.hpp), so stream
 *     parity with it is neither possible nor claimed.
 */
#include "cmi_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------------ */
/* Travel-direction algebra.                                                 */
/*                                                                           */
/* Each TravelDirection is a side of the cube: an offset in {-1,0,1}^3 (P =  */
/* +1, N = -1, an axis not named = 0).  The reference spells the same        */
/* relations out as switch statements (TravelDirections.hpp:89-403); the     */
/* table below is the restatement: output_to_input negates the offset, the   */
/* exit mask sets the offset from the out-of-range bits, and entry clamping  */
/* uses the offset of the entry side.                                        */
/* ------------------------------------------------------------------------ */
static const signed char kOffset[ORC_NDIR][3] = {
    {0, 0, 0},                                                    /* INSIDE */
    {1, 1, 1},   {1, 1, -1},   {1, -1, 1},   {1, -1, -1},         /* CORNER P** */
    {-1, 1, 1},  {-1, 1, -1},  {-1, -1, 1},  {-1, -1, -1},        /* CORNER N** */
    {0, 1, 1},   {0, 1, -1},   {0, -1, 1},   {0, -1, -1},         /* EDGE_X */
    {1, 0, 1},   {1, 0, -1},   {-1, 0, 1},   {-1, 0, -1},         /* EDGE_Y */
    {1, 1, 0},   {1, -1, 0},   {-1, 1, 0},   {-1, -1, 0},         /* EDGE_Z */
    {1, 0, 0},   {-1, 0, 0},   {0, 1, 0},    {0, -1, 0},          /* FACE X,Y */
    {0, 0, 1},   {0, 0, -1}};                                     /* FACE Z */

void orc_direction_offset(int direction, int offset[3]) {
  offset[0] = kOffset[direction][0];
  offset[1] = kOffset[direction][1];
  offset[2] = kOffset[direction][2];
}

int orc_direction_from_offset(const int offset[3]) {
  for (int d = 0; d < ORC_NDIR; ++d) {
    if (kOffset[d][0] == offset[0] && kOffset[d][1] == offset[1] &&
        kOffset[d][2] == offset[2]) {
      return d;
    }
  }
  return -1;
}

/* TravelDirections.hpp:89-152: swap N and P. */
int orc_output_to_input_direction(int output_direction) {
  if (output_direction < 0 || output_direction >= ORC_NDIR) return -1;
  const int o[3] = {-kOffset[output_direction][0],
                    -kOffset[output_direction][1],
                    -kOffset[output_direction][2]};
  return orc_direction_from_offset(o);
}

/* TravelDirections.hpp:315-403: mask bit 5..0 = x_high x_low y_high y_low
 * z_high z_low; -1 if an axis is both high and low. */
int orc_direction_from_mask(int mask) {
  if (mask < 0 || mask > 63) return -1;
  int o[3];
  for (int a = 0; a < 3; ++a) {
    const int hi = (mask >> (5 - 2 * a)) & 1;
    const int lo = (mask >> (4 - 2 * a)) & 1;
    if (hi && lo) return -1;
    o[a] = hi - lo;
  }
  return orc_direction_from_offset(o);
}

/* TravelDirections.hpp:163-225: every axis named by the exit side must have a
 * direction component of that sign. */
int orc_is_compatible_output_direction(const double dir[3], int direction) {
  for (int a = 0; a < 3; ++a) {
    const int o = kOffset[direction][a];
    if (o > 0 && !(dir[a] > 0.)) return 0;
    if (o < 0 && !(dir[a] < 0.)) return 0;
  }
  return 1;
}

/* TravelDirections.hpp:236-298: entering through side o needs sign -o. */
int orc_is_compatible_input_direction(const double dir[3], int direction) {
  for (int a = 0; a < 3; ++a) {
    const int o = kOffset[direction][a];
    if (o > 0 && !(dir[a] < 0.)) return 0;
    if (o < 0 && !(dir[a] > 0.)) return 0;
  }
  return 1;
}

/* ------------------------------------------------------------------------ */
/* Subgrid helpers                                                           */
/* ------------------------------------------------------------------------ */

/* DensitySubGrid.hpp:516-527 (geometry part of the constructor). */
void orc_subgrid_init(orc_subgrid *s, const double box[6], const int ncell[3]) {
  for (int a = 0; a < 3; ++a) {
    s->anchor[a] = box[a];
    s->cell_size[a] = box[3 + a] / ncell[a];
    s->inv_cell_size[a] = ncell[a] / box[3 + a];
    s->ncell[a] = ncell[a];
  }
  s->ncell[3] = ncell[1] * ncell[2];
  for (int d = 0; d < ORC_NDIR; ++d) s->ngbs[d] = ORC_NEIGHBOUR_OUTSIDE;
  s->n = s->xn = s->J = NULL;
}

/* DensitySubGrid.hpp:483-506: out-of-range mask -> exit direction.  The
 * reference tests "index / n > 0" for the high side, which for indices in
 * [-1, n] is "index == n"; the division is kept. */
int orc_get_output_direction(const orc_subgrid *s, const int idx[3]) {
  int mask = 0;
  for (int a = 0; a < 3; ++a) {
    const int lo = idx[a] < 0;
    const int hi = (idx[a] / s->ncell[a]) > 0;
    mask |= (hi << (5 - 2 * a)) | (lo << (4 - 2 * a));
  }
  return orc_direction_from_mask(mask);
}

/* DensitySubGrid.hpp:247-469: along an axis that the entry side names, the
 * index is clamped to 0 (N) or n-1 (P); along a free axis it is the truncated
 * position * (1 / cell size). */
int orc_get_start_index(const orc_subgrid *s, const double position[3],
                        int input_direction, int idx[3]) {
  for (int a = 0; a < 3; ++a) {
    const int o = kOffset[input_direction][a];
    if (o == 0) {
      idx[a] = (int)(position[a] * s->inv_cell_size[a]);
    } else if (o < 0) {
      idx[a] = 0;
    } else {
      idx[a] = s->ncell[a] - 1;
    }
  }
  return idx[0] * s->ncell[3] + idx[1] * s->ncell[2] + idx[2];
}

/* std::min(a, b) of <algorithm>: (b < a) ? b : a. */
static inline double std_min(double a, double b) { return (b < a) ? b : a; }

/* DensitySubGrid.hpp:1180-1323: walk one packet through one subgrid.  The
 * arithmetic is the reference's, step for step (inverse direction, anchor
 * subtraction, face distances, lmin, tau, the partial last step, the index
 * update on l == lmin, the deposit lmin*sigma*weight, the face snap of the
 * position). */
int orc_interact(const orc_subgrid *s, orc_photon *photon, int input_direction,
                 uint32_t *crossings) {
  return orc_walk(s, photon, input_direction, ORC_WALK_INTERACT, crossings);
}

/* The reference's three walk variants share the loop above:
 * ORC_WALK_INTERACT (interact, 1180-1323), ORC_WALK_PROPAGATE (propagate,
 * 1333-1472: no deposit) and ORC_WALK_OPTICAL_DEPTH (compute_optical_depth,
 * 1482-1604: no deposit, the loop runs while the index is inside, and the
 * path's optical depth is added to the packet's tau, 1592-1593). */
int orc_walk(const orc_subgrid *s, orc_photon *photon, int input_direction,
             int kind, uint32_t *crossings) {
  const double *direction = photon->dir;
  const double inverse_direction[3] = {1. / direction[0], 1. / direction[1],
                                       1. / direction[2]};
  double position[3] = {photon->pos[0] - s->anchor[0],
                        photon->pos[1] - s->anchor[1],
                        photon->pos[2] - s->anchor[2]};
  double tau_done = 0.;
  const double tau_target = photon->tau;
  const double cross_section = photon->sigma;
  const double photon_weight = photon->weight;
  int three_index[3];
  int active_cell = orc_get_start_index(s, position, input_direction,
                                        three_index);
  uint32_t ncross = 0;
  const int od = kind == ORC_WALK_OPTICAL_DEPTH;
  while ((od || tau_done < tau_target) && three_index[0] < s->ncell[0] &&
         three_index[0] >= 0 && three_index[1] < s->ncell[1] &&
         three_index[1] >= 0 && three_index[2] < s->ncell[2] &&
         three_index[2] >= 0) {
    const double cell_low[3] = {three_index[0] * s->cell_size[0],
                                three_index[1] * s->cell_size[1],
                                three_index[2] * s->cell_size[2]};
    const double cell_high[3] = {(three_index[0] + 1.) * s->cell_size[0],
                                 (three_index[1] + 1.) * s->cell_size[1],
                                 (three_index[2] + 1.) * s->cell_size[2]};
    double l[3];
    for (int idim = 0; idim < 3; ++idim) {
      if (direction[idim] > 0.) {
        l[idim] = (cell_high[idim] - position[idim]) * inverse_direction[idim];
      } else if (direction[idim] < 0.) {
        l[idim] = (cell_low[idim] - position[idim]) * inverse_direction[idim];
      } else {
        l[idim] = DBL_MAX;
      }
    }
    double lmin = std_min(l[0], std_min(l[1], l[2]));
    double lminsigma = lmin * cross_section;
    const double tau =
        lminsigma * s->n[active_cell] * s->xn[active_cell];
    tau_done += tau;
    if (!od && tau_done >= tau_target) {
      const double correction = (tau_done - tau_target) / tau;
      lmin *= (1. - correction);
      lminsigma = lmin * cross_section;
    } else {
      for (int idim = 0; idim < 3; ++idim) {
        if (l[idim] == lmin) {
          three_index[idim] += (direction[idim] > 0.) ? 1 : -1;
        }
      }
    }
    if (kind == ORC_WALK_INTERACT) s->J[active_cell] += lminsigma * photon_weight;
    for (int idim = 0; idim < 3; ++idim) {
      position[idim] =
          (l[idim] == lmin)
              ? ((direction[idim] > 0.) ? cell_high[idim] : cell_low[idim])
              : position[idim] + lmin * direction[idim];
    }
    active_cell = three_index[0] * s->ncell[3] + three_index[1] * s->ncell[2] +
                  three_index[2];
    ++ncross;
  }
  photon->tau = od ? tau_target + tau_done : tau_target - tau_done;
  for (int a = 0; a < 3; ++a) photon->pos[a] = position[a] + s->anchor[a];
  if (crossings) *crossings += ncross;
  if (!od && tau_done >= tau_target) return ORC_INSIDE;
  return orc_get_output_direction(s, three_index);
}

/* ------------------------------------------------------------------------ */
/* Grid of subgrids                                                          */
/* ------------------------------------------------------------------------ */

/* testDensitySubGrid_omp.cpp:1805-1877: subgrid boxes, and the 27-entry
 * neighbour table (slot of direction get_output_direction(offset * ncell),
 * slot 0 = the subgrid itself). Fields start at n = 1e8, x_n = 1e-6, J = 0
 * (DensitySubGrid.hpp:544-551). */
int orc_grid_create(orc_grid *g, const double box[6], const int ncell[3],
                    const int nsub[3]) {
  memset(g, 0, sizeof(*g));
  for (int a = 0; a < 3; ++a) {
    if (nsub[a] <= 0 || ncell[a] <= 0 || ncell[a] % nsub[a] != 0) return -1;
    g->ncell[a] = ncell[a];
    g->nsub[a] = nsub[a];
  }
  memcpy(g->box, box, 6 * sizeof(double));
  g->nsubgrid = nsub[0] * nsub[1] * nsub[2];
  const int sub_ncell[3] = {ncell[0] / nsub[0], ncell[1] / nsub[1],
                            ncell[2] / nsub[2]};
  const double sub_side[3] = {box[3] / nsub[0], box[4] / nsub[1],
                              box[5] / nsub[2]};
  g->cells_per_subgrid = (size_t)sub_ncell[0] * sub_ncell[1] * sub_ncell[2];
  g->ncell_total = g->cells_per_subgrid * g->nsubgrid;
  g->sub = (orc_subgrid *)calloc(g->nsubgrid, sizeof(orc_subgrid));
  g->n = (double *)malloc(g->ncell_total * sizeof(double));
  g->xn = (double *)malloc(g->ncell_total * sizeof(double));
  g->J = (double *)calloc(g->ncell_total, sizeof(double));
  if (!g->sub || !g->n || !g->xn || !g->J) return -2;
  for (size_t i = 0; i < g->ncell_total; ++i) {
    g->n[i] = 1.e8;
    g->xn[i] = 1.e-6;
  }
  for (int index = 0; index < g->nsubgrid; ++index) {
    const int ix = index / (nsub[1] * nsub[2]);
    const int iy = (index - ix * nsub[1] * nsub[2]) / nsub[2];
    const int iz = index - ix * nsub[1] * nsub[2] - iy * nsub[2];
    const double subbox[6] = {box[0] + ix * sub_side[0],
                              box[1] + iy * sub_side[1],
                              box[2] + iz * sub_side[2],
                              sub_side[0], sub_side[1], sub_side[2]};
    orc_subgrid *s = &g->sub[index];
    orc_subgrid_init(s, subbox, sub_ncell);
    s->n = g->n + (size_t)index * g->cells_per_subgrid;
    s->xn = g->xn + (size_t)index * g->cells_per_subgrid;
    s->J = g->J + (size_t)index * g->cells_per_subgrid;
    for (int nix = -1; nix < 2; ++nix) {
      for (int niy = -1; niy < 2; ++niy) {
        for (int niz = -1; niz < 2; ++niz) {
          const int cix = ix + nix, ciy = iy + niy, ciz = iz + niz;
          if (cix >= 0 && cix < nsub[0] && ciy >= 0 && ciy < nsub[1] &&
              ciz >= 0 && ciz < nsub[2]) {
            const int ti[3] = {nix * sub_ncell[0], niy * sub_ncell[1],
                               niz * sub_ncell[2]};
            const int ngbi = orc_get_output_direction(s, ti);
            s->ngbs[ngbi] = (uint32_t)(cix * nsub[1] * nsub[2] +
                                       ciy * nsub[2] + ciz);
          }
        }
      }
    }
  }
  return 0;
}

void orc_grid_destroy(orc_grid *g) {
  free(g->sub);
  free(g->n);
  free(g->xn);
  free(g->J);
  memset(g, 0, sizeof(*g));
}

/* testDensitySubGrid_omp.cpp:1956-1969: subgrid that holds a source,
 * (unsigned)((x - anchor) / side * nsub) per axis. */
uint32_t orc_grid_source_subgrid(const orc_grid *g, const double pos[3]) {
  unsigned int si[3];
  for (int a = 0; a < 3; ++a) {
    si[a] = (unsigned int)((pos[a] - g->box[a]) / g->box[3 + a] * g->nsub[a]);
  }
  return si[0] * g->nsub[1] * g->nsub[2] + si[1] * g->nsub[2] + si[2];
}

/* Routing of testDensitySubGrid_omp.cpp:607-666 and 1149-1329 for one packet
 * at a time: interact; INSIDE = absorbed (reemission probability 0 disables
 * the INSIDE buffer, 1188-1191); an exit towards NEIGHBOUR_OUTSIDE leaves the
 * box; otherwise the packet enters the neighbour with the N<->P swapped
 * direction.  Packets are independent while x_n is frozen, so the order in
 * which the reference drains its 200-packet buffers only changes the order of
 * the J sums, not any packet's path. */
/* Reemission of a list transport: probability, Philox key and iteration;
 * packet i of the list draws as packet number first + i. */
typedef struct {
  double prob;
  uint64_t seed;
  uint32_t iteration;
  uint64_t first;
} reemit_cfg;

static void transport_one(orc_grid *g, orc_photon *p, uint32_t sub,
                          int32_t *status, uint32_t *crossings,
                          uint32_t *hops, const reemit_cfg *re, uint64_t k,
                          uint32_t *reemissions) {
  int in_dir = ORC_INSIDE;
  uint32_t nc = 0, nh = 0, nre = 0;
  for (;;) {
    orc_subgrid *s = &g->sub[sub];
    ++nh;
    const int out = orc_interact(s, p, in_dir, &nc);
    if (out == ORC_INSIDE) {
      /* do_reemission (testDensitySubGrid_omp.cpp:675-702): the absorbed
       * packet sits in its subgrid's INSIDE buffer (1188-1191) and is
       * reemitted with the given probability from where it stopped (its
       * position stays), then traversed again in the same subgrid with
       * input direction INSIDE (execute_photon_reemit_task, 1348-) */
      if (re && re->prob > 0. && nre < ORC_REEMIT_MAX &&
          orc_reemit_draw(k, nre, re->prob, re->seed, re->iteration, p->dir,
                          &p->tau)) {
        ++nre;
        in_dir = ORC_INSIDE;
        continue;
      }
      *status = (re && re->prob > 0. && nre >= ORC_REEMIT_MAX) ? -2 : 0;
      break;
    }
    if (out < 0) {
      *status = -1;
      break;
    }
    const uint32_t ngb = s->ngbs[out];
    if (ngb == ORC_NEIGHBOUR_OUTSIDE) {
      *status = 1;
      break;
    }
    in_dir = orc_output_to_input_direction(out);
    sub = ngb;
  }
  if (crossings) *crossings = nc;
  if (hops) *hops = nh;
  if (reemissions) *reemissions = nre;
}

static void transport_list_impl(orc_grid *g, orc_photon *p,
                                const uint32_t *start_sub, size_t n,
                                int32_t *status, uint32_t *crossings,
                                uint32_t *hops, uint32_t *reemissions,
                                const reemit_cfg *re, int nthreads);

void orc_transport_list(orc_grid *g, orc_photon *p, const uint32_t *start_sub,
                        size_t n, int32_t *status, uint32_t *crossings,
                        uint32_t *hops, int nthreads) {
  transport_list_impl(g, p, start_sub, n, status, crossings, hops, NULL, NULL,
                      nthreads);
}

void orc_transport_list_reemit(orc_grid *g, orc_photon *p,
                               const uint32_t *start_sub, size_t n,
                               double probability, uint64_t seed,
                               uint32_t iteration, uint64_t first,
                               int32_t *status, uint32_t *crossings,
                               uint32_t *hops, uint32_t *reemissions,
                               int nthreads) {
  const reemit_cfg re = {probability, seed, iteration, first};
  transport_list_impl(g, p, start_sub, n, status, crossings, hops, reemissions,
                      &re, nthreads);
}

static void transport_list_impl(orc_grid *g, orc_photon *p,
                                const uint32_t *start_sub, size_t n,
                                int32_t *status, uint32_t *crossings,
                                uint32_t *hops, uint32_t *reemissions,
                                const reemit_cfg *re, int nthreads) {
  const uint64_t first = re ? re->first : 0;
  if (nthreads <= 1) {
    for (size_t i = 0; i < n; ++i) {
      int32_t st;
      uint32_t c, h, r;
      transport_one(g, &p[i], start_sub[i], &st, &c, &h, re, first + i, &r);
      if (status) status[i] = st;
      if (crossings) crossings[i] = c;
      if (hops) hops[i] = h;
      if (reemissions) reemissions[i] = r;
    }
    return;
  }
  /* Parallel variant: every thread accumulates into its own J copy, which
   * are summed at the end (the reference gets exclusive J access from its
   * subgrid locks, Task.hpp:247-269; a private copy is the CPU analogue). */
#ifdef _OPENMP
  const size_t nc = g->ncell_total;
  double *J0 = g->J;
  double **priv = (double **)calloc((size_t)nthreads, sizeof(double *));
#pragma omp parallel num_threads(nthreads)
  {
    const int t = omp_get_thread_num();
    priv[t] = (double *)calloc(nc, sizeof(double));
    orc_grid local = *g;
    orc_subgrid *subs =
        (orc_subgrid *)malloc((size_t)g->nsubgrid * sizeof(orc_subgrid));
    memcpy(subs, g->sub, (size_t)g->nsubgrid * sizeof(orc_subgrid));
    for (int k = 0; k < g->nsubgrid; ++k) {
      subs[k].J = priv[t] + (size_t)k * g->cells_per_subgrid;
    }
    local.sub = subs;
    local.J = priv[t];
#pragma omp for schedule(dynamic, 256)
    for (long long i = 0; i < (long long)n; ++i) {
      int32_t st;
      uint32_t c, h, r;
      transport_one(&local, &p[i], start_sub[i], &st, &c, &h, re,
                    first + (uint64_t)i, &r);
      if (status) status[i] = st;
      if (crossings) crossings[i] = c;
      if (hops) hops[i] = h;
      if (reemissions) reemissions[i] = r;
    }
    free(subs);
  }
  for (int t = 0; t < nthreads; ++t) {
    for (size_t c = 0; c < nc; ++c) J0[c] += priv[t][c];
    free(priv[t]);
  }
  free(priv);
#else
  transport_list_impl(g, p, start_sub, n, status, crossings, hops, reemissions,
                      re, 1);
#endif
}

/* ------------------------------------------------------------------------ */
/* Ionization balance                                                        */
/* ------------------------------------------------------------------------ */

/* DensitySubGrid.hpp:1628-1658, with alpha a parameter (the reference fixes
 * it to 4e-19 at 1637). */
void orc_compute_neutral_fraction(orc_grid *g, double luminosity,
                                  double num_photon, double alphaH, int form) {
  for (int k = 0; k < g->nsubgrid; ++k) {
    orc_subgrid *s = &g->sub[k];
    const double jfac = luminosity / num_photon * s->inv_cell_size[0] *
                        s->inv_cell_size[1] * s->inv_cell_size[2];
    const int ncell_tot = s->ncell[0] * s->ncell[3];
    for (int i = 0; i < ncell_tot; ++i) {
      const double jH = jfac * s->J[i];
      if (jH > 0.) {
        const double ntot = s->n[i];
        const double aa = 0.5 * jH / (ntot * alphaH);
        if (form == 0) {
          const double bb = 2. / aa;
          const double cc = sqrt(bb + 1.);
          s->xn[i] = 1. + aa * (1. - cc);
        } else {
          const double b = 2. * aa;
          s->xn[i] = 2. / (2. + b + sqrt(b * (b + 4.)));
        }
      } else {
        s->xn[i] = 1.;
      }
      s->J[i] = 0.;
    }
  }
}

/* ------------------------------------------------------------------------ */
/* Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11) and emission            */
/* ------------------------------------------------------------------------ */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2],
                       uint32_t out[4]) {
  uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
  uint32_t k0 = key[0], k1 = key[1];
  for (int r = 0; r < 10; ++r) {
    if (r > 0) {
      k0 += 0x9E3779B9u;
      k1 += 0xBB67AE85u;
    }
    const uint64_t p0 = (uint64_t)0xD2511F53u * c0;
    const uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
    const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
    const uint32_t n1 = (uint32_t)p1;
    const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
    const uint32_t n3 = (uint32_t)p0;
    c0 = n0;
    c1 = n1;
    c2 = n2;
    c3 = n3;
  }
  out[0] = c0;
  out[1] = c1;
  out[2] = c2;
  out[3] = c3;
}

/* 53 random bits -> (0, 1), never exactly 0 or 1. */
double orc_uniform_open(uint32_t hi, uint32_t lo) {
  const uint64_t m = ((uint64_t)(hi >> 5) << 26) | (uint64_t)(lo >> 6);
  return ((double)m + 0.5) * (1.0 / 9007199254740992.0);
}

/* Reemission draws of packet k's r-th absorption (do_reemission,
 * testDensitySubGrid_omp.cpp:675-702: a uniform below the probability, then
 * get_random_direction 529-545 and tau = -ln u).  The draws use only
 * operations IEEE 754 rounds exactly (+ - * / sqrt and compares), so that
 * this restatement and the GPU walk give the same bits:
 *   - the direction by Marsaglia's rejection pair (Ann. Math. Stat. 43
 *     (1972) 645): a, b uniform in (-1, 1) with s = a^2 + b^2 < 1, then
 *     (2a sqrt(1 - s), 2b sqrt(1 - s), 1 - 2s) is uniform on the sphere;
 *   - the optical depth by von Neumann's comparison method (NBS Appl. Math.
 *     Ser. 12 (1951) 36; Knuth, TAOCP 2, 3.4.1): trial k draws U0 > U1 > ...
 *     until a draw exceeds its predecessor; an odd run length accepts
 *     k + U0, which is then Exp(1), like -ln u.
 * The uniforms come in order from Philox counters {k_lo, k_hi, iteration,
 * 2 + 256 r + j}, j = 0, 1, ..., two per counter (the emission uses 0 and
 * 1).  Returns 1 and writes dir and tau when the packet is reemitted. */
typedef struct {
  uint32_t key[2], ctr[4], buf[4];
  int avail, blocks;
} reemit_stream;

static double rs_uniform(reemit_stream *s) {
  if (s->avail == 0) {
    orc_philox4x32_10(s->ctr, s->key, s->buf);
    s->ctr[3] += 1u;
    s->blocks += 1;
    s->avail = 2;
  }
  const int at = 2 - s->avail;
  s->avail -= 1;
  return orc_uniform_open(s->buf[2 * at], s->buf[2 * at + 1]);
}

int orc_reemit_draw(uint64_t k, uint32_t r, double probability, uint64_t seed,
                    uint32_t iteration, double dir[3], double *tau) {
  reemit_stream s;
  s.key[0] = (uint32_t)seed;
  s.key[1] = (uint32_t)(seed >> 32);
  s.ctr[0] = (uint32_t)k;
  s.ctr[1] = (uint32_t)(k >> 32);
  s.ctr[2] = iteration;
  s.ctr[3] = 2u + 256u * r;
  s.avail = 0;
  s.blocks = 0;
  if (!(rs_uniform(&s) < probability)) return 0;
  dir[0] = 0.;
  dir[1] = 0.;
  dir[2] = 1.;
  for (int t = 0; t < ORC_REEMIT_DIR_TRIES; ++t) {
    const double a = 2. * rs_uniform(&s) - 1.;
    const double b = 2. * rs_uniform(&s) - 1.;
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
  while (s.blocks < ORC_REEMIT_BLOCKS) {
    const double u0 = rs_uniform(&s);
    double prev = u0;
    int run = 1;
    for (;;) {
      const double v = rs_uniform(&s);
      if (!(v < prev)) break;
      prev = v;
      ++run;
      if (s.blocks >= ORC_REEMIT_BLOCKS) break; /* bound (never reached) */
    }
    if (run & 1) {
      *tau = trials + u0;
      break;
    }
    trials += 1.;
  }
  return 1;
}

/* testDensitySubGrid_omp.cpp:2199-2208 and 2240-2273 (with distribute(),
 * 1650-1677): every source gets floor(L_i / L_tot * N) packets; the
 * remainder goes, one each, to the sources with the smallest random keys.
 * The keys come from Philox (counter {i, 0, iteration, 0x5EED}) instead of
 * the reference's RANLUX stream.  Beyond the reference: a rounded share that
 * overshoots is taken back from the largest share, and a remainder of more
 * than one packet per source goes round the key order again, so that the
 * split always sums to N. */
void orc_source_counts(const double *L, int nsrc, uint64_t npackets,
                       uint64_t seed, uint32_t iteration, uint64_t *counts) {
  double Ltot = 0.;
  for (int i = 0; i < nsrc; ++i) Ltot += L[i];
  uint64_t added = 0;
  for (int i = 0; i < nsrc; ++i) {
    counts[i] = (uint64_t)(L[i] / Ltot * (double)npackets);
    added += counts[i];
  }
  while (added > npackets) {
    int imax = 0;
    for (int i = 1; i < nsrc; ++i)
      if (counts[i] > counts[imax]) imax = i;
    --counts[imax];
    --added;
  }
  uint64_t rem = npackets - added;
  if (rem == 0) return;
  double *keys = (double *)malloc((size_t)nsrc * sizeof(double));
  int *order = (int *)malloc((size_t)nsrc * sizeof(int));
  const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
  for (int i = 0; i < nsrc; ++i) {
    const uint32_t ctr[4] = {(uint32_t)i, 0u, iteration, 0x5EEDu};
    uint32_t o[4];
    orc_philox4x32_10(ctr, key, o);
    keys[i] = orc_uniform_open(o[0], o[1]);
    order[i] = i;
  }
  /* insertion sort by key (nsrc is small), ties by index */
  for (int i = 1; i < nsrc; ++i) {
    int j = i;
    while (j > 0 && (keys[order[j]] < keys[order[j - 1]])) {
      const int t = order[j];
      order[j] = order[j - 1];
      order[j - 1] = t;
      --j;
    }
  }
  for (uint64_t r = 0; r < rem; ++r) counts[order[r % (uint64_t)nsrc]]++;
  free(keys);
  free(order);
}

/* testDensitySubGrid_omp.cpp:529-597: isotropic direction (cos theta =
 * 2u - 1, phi = 2 pi u, sin theta = sqrt(max(1 - cos^2, 0))), tau = -ln u,
 * sigma fixed, weight 1, position = source, entry INSIDE.  Packet k of the
 * iteration draws counter {k_lo, k_hi, iteration, 0} (direction) and
 * {k_lo, k_hi, iteration, 1} (tau); packets are numbered source by source. */
void orc_emit(const orc_grid *g, const double *src_pos, const double *src_L,
              int nsrc, uint64_t npackets, uint64_t first, uint64_t count,
              uint64_t seed, uint32_t iteration, double sigma, orc_photon *out,
              uint32_t *start_sub) {
  orc_emit_problem(g, src_pos, src_L, nsrc, npackets, first, count, seed,
                   iteration, sigma, 0, out, start_sub);
}

/* The emission of problem q of a problem set (no counterpart in the
 * reference, which runs one grid per process): the source split keyed by
 * seed ^ q, and the high counter word of packet k XOR q << 20 under the
 * set's key; problem 0 is orc_emit. */
void orc_emit_problem(const orc_grid *g, const double *src_pos,
                      const double *src_L, int nsrc, uint64_t npackets,
                      uint64_t first, uint64_t count, uint64_t seed,
                      uint32_t iteration, double sigma, uint32_t problem,
                      orc_photon *out, uint32_t *start_sub) {
  uint64_t *counts = (uint64_t *)malloc((size_t)nsrc * sizeof(uint64_t));
  orc_source_counts(src_L, nsrc, npackets, seed ^ (uint64_t)problem, iteration,
                    counts);
  const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
  for (uint64_t j = 0; j < count; ++j) {
    const uint64_t k = first + j;
    int src = 0;
    uint64_t acc = counts[0];
    while (src + 1 < nsrc && k >= acc) {
      ++src;
      acc += counts[src];
    }
    const uint32_t hi = (uint32_t)(k >> 32) ^ (problem << 20);
    const uint32_t c0[4] = {(uint32_t)k, hi, iteration, 0u};
    const uint32_t c1[4] = {(uint32_t)k, hi, iteration, 1u};
    uint32_t r0[4], r1[4];
    orc_philox4x32_10(c0, key, r0);
    orc_philox4x32_10(c1, key, r1);
    const double cost = 2. * orc_uniform_open(r0[0], r0[1]) - 1.;
    const double phi = 2. * M_PI * orc_uniform_open(r0[2], r0[3]);
    const double sint = sqrt(fmax(1. - cost * cost, 0.));
    orc_photon *p = &out[j];
    p->pos[0] = src_pos[3 * src + 0];
    p->pos[1] = src_pos[3 * src + 1];
    p->pos[2] = src_pos[3 * src + 2];
    p->dir[0] = sint * cos(phi);
    p->dir[1] = sint * sin(phi);
    p->dir[2] = cost;
    p->tau = -log(orc_uniform_open(r1[0], r1[1]));
    p->sigma = sigma;
    p->weight = 1.;
    if (start_sub) start_sub[j] = orc_grid_source_subgrid(g, p->pos);
  }
  free(counts);
}

uint64_t orc_iteration(orc_grid *g, const double *src_pos, const double *src_L,
                       int nsrc, uint64_t npackets, uint64_t seed,
                       uint32_t iteration, double sigma, double alphaH,
                       int form, int nthreads, double *l1) {
  const size_t batch = 1u << 18;
  orc_photon *p = (orc_photon *)malloc(batch * sizeof(orc_photon));
  uint32_t *ss = (uint32_t *)malloc(batch * sizeof(uint32_t));
  uint32_t *cr = (uint32_t *)malloc(batch * sizeof(uint32_t));
  uint64_t total = 0;
  memset(g->J, 0, g->ncell_total * sizeof(double));
  for (uint64_t first = 0; first < npackets; first += batch) {
    const uint64_t cnt = (npackets - first < batch) ? npackets - first : batch;
    orc_emit(g, src_pos, src_L, nsrc, npackets, first, cnt, seed, iteration,
             sigma, p, ss);
    orc_transport_list(g, p, ss, (size_t)cnt, NULL, cr, NULL, nthreads);
    for (uint64_t j = 0; j < cnt; ++j) total += cr[j];
  }
  double Ltot = 0.;
  for (int i = 0; i < nsrc; ++i) Ltot += src_L[i];
  double *old = NULL;
  if (l1) {
    old = (double *)malloc(g->ncell_total * sizeof(double));
    memcpy(old, g->xn, g->ncell_total * sizeof(double));
  }
  orc_compute_neutral_fraction(g, Ltot, (double)npackets, alphaH, form);
  if (l1) {
    double s = 0.;
    for (size_t i = 0; i < g->ncell_total; ++i) s += fabs(g->xn[i] - old[i]);
    *l1 = s / (double)g->ncell_total;
    free(old);
  }
  free(p);
  free(ss);
  free(cr);
  return total;
}
