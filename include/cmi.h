/*
 * cmi.h -- C ABI of the MI355X photon-packet transport and ionization engine.
 *
 * Drop-in for the hot path of ParallelCMacIonize: DensitySubGrid cell walking
 * with optical-depth deposition into the intensity integrals, PhotonBuffer /
 * TravelDirections routing of packets between subgrids, and the per-cell
 * ionization-equilibrium update.  The reference exposes this path as inline
 * C++ class methods; each entry point below names the method(s) it replaces.
 *
 * Conventions
 *  - Every call returns CMI_OK (0) or a negative CMI_ERR_* code; nothing
 *    aborts and no C++ exception crosses the ABI.  cmi_last_error() gives the
 *    message of the last failure on a context.
 *  - A context owns all device memory and one HIP stream on one GPU and is
 *    used from one host thread at a time.  Several contexts may coexist.
 *  - Host arrays are caller-owned and copied in and out.  Cell arrays use the
 *    reference layout: subgrid-major (subgrid index ix*ny*nz + iy*nz + iz of
 *    testDensitySubGrid_omp.cpp:1818-1822), and inside a subgrid the cell
 *    order ix*ny*nz + iy*nz + iz (DensitySubGrid.hpp:207-210).
 *  - TravelDirection values are those of TravelDirections.hpp:46-75.
 */
#ifndef CMI_H
#define CMI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- error codes ---- */
#define CMI_OK 0
#define CMI_ERR_ARG (-1)       /* invalid argument */
#define CMI_ERR_HIP (-2)       /* HIP runtime error */
#define CMI_ERR_ALLOC (-3)     /* device or host allocation failed */
#define CMI_ERR_OVERFLOW (-4)  /* packet pool / bin capacity exceeded */
#define CMI_ERR_DIRECTION (-5) /* invalid exit mask or walk bound exceeded */
#define CMI_ERR_STATE (-6)     /* call out of order (e.g. no sources set) */
#define CMI_ERR_NODEVICE (-7)  /* no usable GPU */

/* ---- arithmetic modes ---- */
/* FP64 throughout, reference arithmetic: bit-identical walks to
 * DensitySubGrid::interact. */
#define CMI_PREC_FP64 0
/* FP32 geometry and optical depth (cell units, kappa = sigma n x_n stored as
 * FP32); path lengths summed per tile visit in 64-bit fixed point (2^-30 of
 * a cell for cubic cells, exact integer sums) and added into the FP64 J. */
#define CMI_PREC_FP32 1

/* ---- transport engines ---- */
#define CMI_ENGINE_WALK 0  /* one persistent thread per packet, global J */
#define CMI_ENGINE_TILES 1 /* packets binned per LDS tile, rounds of hops */

/* TravelDirection values (TravelDirections.hpp:46-75) used by the ABI. */
#define CMI_TRAVELDIRECTION_INSIDE 0
#define CMI_TRAVELDIRECTION_NUMBER 27
#define CMI_NEIGHBOUR_OUTSIDE 0xffffffffu /* DensitySubGrid.hpp:53 */

/* Photon packet (Photon.hpp:52-65, without the packet type). 72 bytes. */
typedef struct {
  double pos[3];   /* position (m) */
  double dir[3];   /* unit direction */
  double tau;      /* remaining target optical depth */
  double sigma;    /* photoionization cross section (m^2) */
  double weight;   /* packet weight */
} cmi_photon;

/* Per-transport statistics (the reference keeps no such counters). */
typedef struct {
  uint64_t packets;   /* packets transported */
  uint64_t crossings; /* cell crossings = deposits into J */
  uint64_t hops;      /* subgrid (walk) or tile (tiles) visits */
  uint64_t absorbed;  /* packets that reached their target optical depth */
  uint64_t escaped;   /* packets that left the box */
  uint64_t errors;    /* packets stopped on an invalid exit or a walk bound */
  uint32_t rounds;    /* hop rounds (tiles engine), 0 for walk */
  uint32_t reserved;
  double seconds;     /* device time of the transport (HIP events) */
  uint64_t reemitted; /* reemission events (an absorbed packet sent on) */
  uint64_t walkout_crossings; /* tiles: the crossings (of those above) walked
                                 in a walk-out launch */
  uint64_t walkout_hops;      /* tiles: the tile visits (of hops) walked in a
                                 walk-out launch */
} cmi_stats;

/* Tuning of the engines (0 = default for any field). */
typedef struct {
  int tile[3];       /* LDS tile in cells (powers of two up to 64 dividing
                        the subgrid, at most 8192 cells) */
  int block_threads; /* tiles: threads per workgroup (256 or 512) */
  int refill_batch;  /* tiles: finished lanes a wave gathers before it writes
                        them out and reloads together (1..64) */
  int chunk_min;     /* tiles: packets of one tile one workgroup visit
                        takes (a tile's packets are split in such chunks;
                        >= 256, default 8192); a round with fewer packets
                        than a quarter of the resident workgroups' worth of
                        such chunks uses smaller ones (down to 256) */
  int walk_blocks;   /* walk: persistent grid size */
  int lds_fp32;      /* tiles: 1 = FP32 LDS accumulators for J (flushed into
                        FP64; slower on gfx950), 0 = 64-bit fixed point */
  int finish_packets;  /* tiles: when the packets left (counted at the start
                          and every 8 rounds) are at most this many, one
                          walk-out launch walks them all to the end
                          (opacities read from memory, the visits'
                          quantized deposits added by FP64 atomics; the
                          same cells, crossings and fates as the rounds).  0 = the default threshold, < 0 =
                          never.  Single contexts without reemission only. */
} cmi_tuning;

typedef struct cmi_ctx cmi_ctx;

/* Create a grid of nsub[0]*nsub[1]*nsub[2] subgrids covering box (anchor
 * box[0..2], sides box[3..5], m) with ncell cells in total per axis; fields
 * start at n = 1e8 m^-3, x_n = 1e-6, J = 0 (DensitySubGrid.hpp:544-551).
 * Replaces the subgrid set-up of testDensitySubGrid_omp.cpp:1805-1877
 * (DensitySubGrid::DensitySubGrid, DensitySubGrid.hpp:516, and
 * set_neighbour, DensitySubGrid.hpp:831). */
int cmi_create(const double box[6], const int ncell[3], const int nsub[3],
               int precision, cmi_ctx **out);

/* Free everything the context owns (DensitySubGrid::~DensitySubGrid,
 * DensitySubGrid.hpp:758). */
void cmi_destroy(cmi_ctx *ctx);

/* Message of the last failed call on ctx (never NULL). */
const char *cmi_last_error(const cmi_ctx *ctx);

/* Select the transport engine and its tuning (tuning may be NULL). */
int cmi_set_engine(cmi_ctx *ctx, int engine, const cmi_tuning *tuning);

/* Reemission (testDensitySubGrid_omp.cpp:675-702 do_reemission, the INSIDE
 * buffer enabled at 1188-1191, reemission_probability of the parameter file):
 * an absorbed packet is reemitted from its absorption point with this
 * probability, with a new isotropic direction and target optical depth.
 * 0 (the default) disables it, as in the reference.  It applies to
 * cmi_transport on both engines and to cmi_transport_packets (and its
 * _noalloc variant) on the walk engine; cmi_transport_packets on the tiles
 * engine refuses a non-zero probability (CMI_ERR_ARG).  The per-subgrid walks
 * (cmi_interact_batch, cmi_propagate_batch, cmi_optical_depth_batch) are
 * DensitySubGrid::interact: an absorbed packet comes back with
 * TRAVELDIRECTION_INSIDE and reemitting it is the caller's step, as
 * do_reemission is the reference driver's.
 * Draws: the walk engine gives the r-th absorption of packet number k of
 * cmi_transport(seed, iteration) the Philox counters {k, iteration,
 * 2 + 256 r + j} under key seed, with draws exact in IEEE arithmetic (a
 * uniform < p, Marsaglia's pair for the direction, von Neumann's comparison
 * method for the optical depth); the tiles engine draws FP32 values keyed by
 * the absorption point. */
int cmi_set_reemission(cmi_ctx *ctx, double probability);

/* Key of the reemission draws of explicit-list transports on the walk engine
 * (default 0, 0): packet i of the list draws as packet number i of
 * cmi_transport(seed, iteration).  No counterpart in the reference, whose
 * threads draw from their own RANLUX generators. */
int cmi_set_reemission_draws(cmi_ctx *ctx, uint64_t seed, uint32_t iteration);

/* Copy number density and neutral fraction in (either may be NULL to keep
 * the current values).  Replaces DensitySubGrid::set_number_density and
 * set_neutral_fraction (DensitySubGrid.hpp:1116-1131). */
int cmi_set_fields(cmi_ctx *ctx, const double *nH, const double *xn);

/* Copy intensity integrals in (DensitySubGrid::set_intensity_integral,
 * DensitySubGrid.hpp:1139).  Values are in the reference's units (m^3). */
int cmi_set_intensity(cmi_ctx *ctx, const double *J);

/* Copy intensity integrals (m^3) and neutral fractions out (either may be
 * NULL).  Replaces get_intensity_integral and get_neutral_fraction
 * (DensitySubGrid.hpp:1614, 790). */
int cmi_get_fields(cmi_ctx *ctx, double *J, double *xn);

/* Zero the intensity integrals (the J reset of
 * DensitySubGrid::compute_neutral_fraction, DensitySubGrid.hpp:1656). */
int cmi_reset_intensity(cmi_ctx *ctx);

/* DensitySubGrid::interact (DensitySubGrid.hpp:1180-1323) for n packets that
 * all enter subgrid `subgrid` through TravelDirection entry_dir: packets are
 * updated in place (position, remaining tau), exit_dir receives the returned
 * TravelDirection (0 = absorbed) and crossings the number of cells crossed.
 * Deposits are added to J.  FP64 contexts only. */
int cmi_interact_batch(cmi_ctx *ctx, uint32_t subgrid, int entry_dir,
                       cmi_photon *p, size_t n, int32_t *exit_dir,
                       uint32_t *crossings);

/* DensitySubGrid::propagate (DensitySubGrid.hpp:1333-1472): the same walk
 * without deposits into J.  Same arguments as cmi_interact_batch. */
int cmi_propagate_batch(cmi_ctx *ctx, uint32_t subgrid, int entry_dir,
                        cmi_photon *p, size_t n, int32_t *exit_dir,
                        uint32_t *crossings);

/* DensitySubGrid::compute_optical_depth (DensitySubGrid.hpp:1482-1604): walk
 * to the subgrid's edge without deposits and without stopping; the optical
 * depth of the path is added to the packet's tau. */
int cmi_optical_depth_batch(cmi_ctx *ctx, uint32_t subgrid, int entry_dir,
                            cmi_photon *p, size_t n, int32_t *exit_dir,
                            uint32_t *crossings);

/* Whole-grid transport of an explicit packet list: each packet starts INSIDE
 * start_sub[i] and is routed from subgrid to subgrid (the loop of
 * testDensitySubGrid_omp.cpp:607-666 and 1149-1329 with
 * TravelDirections::output_to_input_direction, TravelDirections.hpp:89)
 * until absorbed (status 0) or out of the box (status 1); negative status =
 * invalid exit or walk bound.  Deposits are added to J.  Any of status,
 * crossings, hops may be NULL. */
int cmi_transport_packets(cmi_ctx *ctx, cmi_photon *p,
                          const uint32_t *start_sub, size_t n, int32_t *status,
                          uint32_t *crossings, uint32_t *hops,
                          cmi_stats *stats);

/* The same transport as cmi_transport_packets on the walk engine (whatever
 * cmi_set_engine selected), for callers that cannot allocate device memory:
 * code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc.  No counterpart in the reference.  Host arrays as in
 * cmi_transport_packets; they are staged through a fixed device array the
 * library holds on each GPU until the process ends, so a call allocates
 * nothing.  At most CMI_NOALLOC_MAX_PACKETS packets per call (CMI_ERR_ARG
 * beyond: packets are independent, split longer lists).  Every call on one
 * GPU shares that array: calls that overlap on one GPU race. */
#define CMI_NOALLOC_MAX_PACKETS 65536
int cmi_transport_packets_noalloc(cmi_ctx *ctx, cmi_photon *p,
                                  const uint32_t *start_sub, size_t n,
                                  int32_t *status, uint32_t *crossings,
                                  uint32_t *hops, cmi_stats *stats);

/* Point sources (Source.hpp:57): positions (3 per source, m) and ionizing
 * luminosities (s^-1).  Up to 64 sources. */
int cmi_set_sources(cmi_ctx *ctx, const double *pos, const double *L,
                    int nsrc);

/* Packets the transport draws for (npackets, seed, iteration), numbers
 * first .. first+count-1: isotropic directions, tau = -ln u, the given cross
 * section, weight 1 (fill_buffer, testDensitySubGrid_omp.cpp:556-597, with a
 * Philox4x32-10 counter-based generator instead of RANLUX).  start_sub may be
 * NULL. */
int cmi_emit(cmi_ctx *ctx, uint64_t npackets, uint64_t first, uint64_t count,
             uint64_t seed, uint32_t iteration, double sigma, cmi_photon *out,
             uint32_t *start_sub);

/* One photon-shooting phase: emit npackets over the sources (split as
 * testDensitySubGrid_omp.cpp:2199-2208 / 2240-2273) and transport them all,
 * depositing into J (the worker loop of testDensitySubGrid_omp.cpp:2274-2401).
 */
int cmi_transport(cmi_ctx *ctx, uint64_t npackets, uint64_t seed,
                  uint32_t iteration, double sigma, cmi_stats *stats);

/* DensitySubGrid::compute_neutral_fraction (DensitySubGrid.hpp:1628-1658) on
 * every cell, with the recombination rate alphaH a parameter (the reference
 * fixes 4e-19).  form 0 = the reference expression, form 1 = the
 * cancellation-free root.  J is reset.  l1_change (may be NULL) receives the
 * mean |x_n(new) - x_n(old)| over all cells. */
int cmi_update_ionization(cmi_ctx *ctx, double luminosity, uint64_t npackets,
                          double alphaH, int form, double *l1_change);

/* Total luminosity of the sources set by cmi_set_sources. */
double cmi_total_luminosity(const cmi_ctx *ctx);

/* TravelDirections::output_to_input_direction (TravelDirections.hpp:89) and
 * TravelDirections::get_output_direction(mask) (TravelDirections.hpp:315):
 * host-side, no device needed; -1 on invalid input. */
int cmi_output_to_input_direction(int output_direction);
int cmi_direction_from_mask(int mask);

/* Neighbour table of subgrid `subgrid`: 27 entries indexed by
 * TravelDirection, CMI_NEIGHBOUR_OUTSIDE where none
 * (DensitySubGrid::get_neighbour, DensitySubGrid.hpp:820). */
int cmi_get_neighbours(const cmi_ctx *ctx, uint32_t subgrid,
                       uint32_t ngbs[27]);

/* Number of subgrids and cells (DensitySubGrid::get_number_of_cells,
 * DensitySubGrid.hpp:780, summed over subgrids). */
int cmi_get_sizes(const cmi_ctx *ctx, uint64_t *nsubgrid, uint64_t *ncell);

/* Device timing (HIP events on the context's stream) of the dominant
 * transport kernel, for the roofline figure: its average launch duration (s)
 * over the launches since the last reset, and the number of launches.  The
 * dominant kernel is walk_emit_kernel (one launch per cmi_transport) for the
 * walk engine and tiles_process_kernel (one launch per round) for the tiles
 * engine. */
int cmi_kernel_timing(cmi_ctx *ctx, double *avg_seconds, uint64_t *calls,
                      int reset);

/* Switch the tiles engine's per-launch timing of tiles_process_kernel on
 * (default) or off.  The two events per round cost about 6 us each on the
 * stream; runs that do not need cmi_kernel_timing (convergence runs, the
 * warm start's cheap passes) switch them off.  No counterpart in the
 * reference, which has no device timing. */
int cmi_set_kernel_timing(cmi_ctx *ctx, int on);

/* Synchronise the context's stream. */
int cmi_synchronize(cmi_ctx *ctx);

/* ---- result writers (the reference's on-disk formats) ----
 * neutral_fractions.dat: output_neutral_fractions, testDensitySubGrid_omp.cpp:
 * 394-440, with DensitySubGrid::output_intensities, DensitySubGrid.hpp:
 * 1738-1796 -- per subgrid, per cell (z fastest): centre x, y, z, x_n, n_H as
 * f64, then 5 f32 hydro primitives (zero: no hydrodynamics), 60 B per cell.
 * Arrays in the reference layout; no GPU needed. */
int cmi_write_neutral_fractions_dat(const char *path, const double box[6],
                                    const int ncell[3], const int nsub[3],
                                    const double *nH, const double *xn);
/* neutral_fractions.bin: output_neutral_fractions_binary,
 * testDensitySubGrid_omp.cpp:507-521 -- x_n as raw f64, reference layout */
int cmi_write_neutral_fractions_bin(const char *path, uint64_t ncell_total,
                                    const double *xn);
/* both files from a context's current fields (either path may be NULL) */
int cmi_write_results(cmi_ctx *ctx, const char *dat_path, const char *bin_path);

/* ---- problem sets ----
 * Many independent grids transported together (no counterpart in the
 * reference, which runs one grid per process): every problem has its own box,
 * cubic-cell grid, density, sources and packet budget; all share the LDS tile
 * shape (tuning->tile[0], default 16, which must divide every problem's cell
 * counts), run in FP32 on the tiles engine, and one iteration of all of them
 * -- emission, transport, ionization update -- is one replay of a captured
 * hipGraph.  Cell arrays of a problem use the plain lattice order
 * x*ny*nz + y*nz + z.  Up to 256 problems and 4096 sources in a set. */
typedef struct cmi_batch cmi_batch;
int cmi_batch_create(int nprob, const double *boxes, const int *ncells,
                     const cmi_tuning *tuning, cmi_batch **out);
void cmi_batch_destroy(cmi_batch *b);
const char *cmi_batch_last_error(const cmi_batch *b);
int cmi_batch_set_fields(cmi_batch *b, int problem, const double *nH,
                         const double *xn);
int cmi_batch_get_fields(cmi_batch *b, int problem, double *J, double *xn);
int cmi_batch_set_sources(cmi_batch *b, int problem, const double *pos,
                          const double *L, int nsrc);
/* Change the launch knobs of a set (threads, refill batch, chunk, LDS
 * accumulator); the tile edge is fixed at creation and must be 0 or equal. */
int cmi_batch_set_tuning(cmi_batch *b, const cmi_tuning *tuning);
/* One iteration of every problem: npackets[p] packets for problem p, then the
 * FP64 ionization update (alphaH, form as cmi_update_ionization).  Per
 * problem (arrays of nprob, any may be NULL): mean |dx_n|, the time from the
 * start of the iteration until the problem's last packet finished (device
 * clock), and its cell crossings. */
int cmi_batch_iteration(cmi_batch *b, const uint64_t *npackets, uint64_t seed,
                        uint32_t iteration, double alphaH, int form,
                        double *l1_change, double *problem_seconds,
                        uint64_t *problem_crossings, cmi_stats *stats);

/* ---- learned warm start ----
 * A 4-layer 3-D convolutional network (3x3x3 kernels, 32 channels, ReLU;
 * no counterpart in the reference) from the features of the input mode (log10
 * n_H and the first-pass photoionization rate or the cheap iterate's x_n, see
 * cmi_net_set_input_mode) to the logit of the converged neutral fraction,
 * trained and run on BF16 MFMA with FP32 accumulation.  The plain iteration
 * still decides convergence; the network only supplies the starting x_n.  Feature arrays are [..][D][H][W][2] (FP32), target and logit
 * arrays [..][D][H][W]; the last extent W must be a multiple of 32.  Master
 * parameters are FP32, per layer l: W[27][NOUT][32] (tap = 9 kd + 3 kh + kw,
 * output channel, input channel; zero beyond the real 2 or 32 input and 32 or
 * 1 output channels) then b[NOUT], NOUT = 32 (16 for the last layer). */
typedef struct cmi_net cmi_net;
int cmi_net_create(uint64_t seed, cmi_net **out);
void cmi_net_destroy(cmi_net *net);
const char *cmi_net_last_error(const cmi_net *net);
size_t cmi_net_param_count(const cmi_net *net);
int cmi_net_get_params(cmi_net *net, float *params);
int cmi_net_set_params(cmi_net *net, const float *params);
/* features are normalised as (f - mean) * scale before the first layer; the
 * last layer's output is the logit divided by logit_scale (training targets
 * are divided by it, predictions multiplied) */
int cmi_net_set_normalisation(cmi_net *net, const float mean[2],
                              const float scale[2], float logit_scale);
/* dilation of each layer's 3x3x3 taps (default 1, 1, 1, 1; the taps of
 * layer l sit at offsets -d_l, 0, +d_l per axis, zero padded) */
int cmi_net_set_dilation(cmi_net *net, const int dilation[4]);
int cmi_net_get_dilation(const cmi_net *net, int dilation[4]);
/* Input mode of the warm start and of cmi_net_context_features: 0 (default)
 * feature 1 is log10 jH of the context's last transport and the output is
 * the logit of x_n; 1 feature 1 is the logit of the context's current x_n
 * (e.g. after a few cheap iterations) and the output is added to it. */
int cmi_net_set_input_mode(cmi_net *net, int mode);
int cmi_net_get_input_mode(const cmi_net *net);
/* Forward convolutions through the LDS-tiled kernel where the volume divides
 * into its blocks (default 1), or always through the per-run kernel (0).
 * Both give identical outputs; the switch is there to check that. */
int cmi_net_set_tiled(cmi_net *net, int on);
/* weight of the squared logit error where the prediction is more neutral
 * than the target (default 1: plain mean squared error) */
int cmi_net_set_loss_weight(cmi_net *net, float w_too_neutral);
/* append training samples (kept on the device) */
int cmi_net_add_samples(cmi_net *net, const float *features,
                        const float *target_logits, int nsamples,
                        const int dim[3]);
int cmi_net_sample_count(const cmi_net *net);
/* Adam steps on random crops of the samples (mean squared error on logits);
 * losses (may be NULL) receives one value per step */
int cmi_net_train(cmi_net *net, int steps, int batch, const int crop[3],
                  float lr, uint64_t seed, float *losses);
/* loss and FP32 parameter gradient of one batch, no update */
int cmi_net_loss_and_grad(cmi_net *net, const float *features,
                          const float *target_logits, int batch,
                          const int dim[3], float *loss, float *grad);
int cmi_net_predict(cmi_net *net, const float *features, const int dim[3],
                    float *logits);
/* Warm start of a context, all on the device.  Input mode 0: the context's J
 * must hold the deposits of a first transport of `npackets` packets at total
 * luminosity `luminosity` (per cell jH = L / N / V_cell * J, the jfac of
 * cmi_update_ionization); features log10 n_H and log10(jH + 1e-12 s^-1); the
 * network's output is the logit of x_n.  Input mode 1: the context's x_n is
 * the state after a few cheap iterations; features log10 n_H and the logit
 * of that x_n (clamped to [1e-7, 1 - 1/150]); the output is added to that
 * logit.  Either way x_n = 1 / (1 + exp(-logit)), clamped to [1e-7, 1], the
 * start rule applies, J is zeroed and the FP32 opacity refreshed.  ncell[2]
 * of the context must be a multiple of 16. */
int cmi_net_warm_start(cmi_net *net, cmi_ctx *ctx, double luminosity,
                       uint64_t npackets);
/* The warm start's rule for cells the network calls neutral: a cell whose
 * predicted logit exceeds logit_cut starts at fallback_xn instead (the plain
 * start, 1e-6, lets the iteration form such fronts from the ionized side,
 * where they converge fast).  Default: no cut. */
int cmi_net_set_start_rule(cmi_net *net, float logit_cut, double fallback_xn);
int cmi_net_get_start_rule(const cmi_net *net, float *logit_cut,
                           double *fallback_xn);
/* Coarse start of the warm start (no reference counterpart: the reference
 * always starts from x_n = 1e-6, DensitySubGrid.hpp:549).  `coarse` is a
 * context of the same box whose ncell is an integer factor f >= 2 smaller
 * per axis, on the same device.  cmi_restrict_density: coarse n_H = the
 * mean of each f^3 block of the fine n_H, coarse x_n = 1e-6, J = 0, FP32
 * opacity refreshed.  cmi_prolong_neutral_fraction: fine x_n = the coarse
 * x_n's logit (clamped to [1e-7, 1 - 1/150]) interpolated trilinearly at
 * each fine cell centre, J = 0, FP32 opacity refreshed.  Each waits for the
 * source context's stream and launches on the destination's. */
int cmi_restrict_density(cmi_ctx *fine, cmi_ctx *coarse);
int cmi_prolong_neutral_fraction(cmi_ctx *coarse, cmi_ctx *fine);
/* The same features of a context (by the input mode), as [D][H][W][2] FP32
 * on the host (for building training sets from contexts). */
int cmi_net_context_features(cmi_net *net, cmi_ctx *ctx, double luminosity,
                             uint64_t npackets, float *features);

#ifdef __cplusplus
}
#endif

#endif /* CMI_H */
