// cmi_internal.hpp -- types shared between the C-ABI layer and the kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "cmi_device.hpp"

namespace cmi {

// Device copy of cmi_photon (identical layout, 72 bytes).
struct PhotonRec {
  double pos[3];
  double dir[3];
  double tau;
  double sigma;
  double weight;
};
static_assert(sizeof(PhotonRec) == 72, "PhotonRec must match cmi_photon");

// Cell fields in device memory (tile-major cell order, see GridDev).
struct FieldsDev {
  double *n;      // number density (m^-3)
  double *xn;     // neutral fraction
  double *J;      // intensity integral (FP64 mode: m^3; FP32 mode: Delta_0)
  float *kappa;   // FP32 mode: sigma n x_n Delta_0
};

hipError_t launch_interact_batch(const GridDev &g, const double *n,
                                 const double *xn, double *J, int sub,
                                 int in_dir, PhotonRec *p, size_t np,
                                 int32_t *exit_dir, uint32_t *crossings,
                                 hipStream_t st, int kind = 0);
hipError_t launch_walk_list(const GridDev &g, const FieldsDev &f, int prec,
                            PhotonRec *p, const uint32_t *start_sub, size_t np,
                            int32_t *status, uint32_t *crossings,
                            uint32_t *hops, const ReemitDraws &re,
                            Counters *cnt, hipStream_t st);
hipError_t launch_walk_emit(const GridDev &g, const FieldsDev &f, int prec,
                            const SourcesDev *S, uint64_t npackets,
                            uint32_t iteration, uint64_t seed, double sigma,
                            int blocks, const ReemitDraws &re, Counters *cnt,
                            hipStream_t st);
hipError_t launch_emit(const SourcesDev *S, uint64_t first, uint64_t count,
                       uint32_t iteration, uint64_t seed, double sigma,
                       PhotonRec *out, uint32_t *start_sub, hipStream_t st);

// Ionization update over all cells; per-block partial sums of |dx| go to
// partial[0 .. blocks-1].
hipError_t launch_ionize(const FieldsDev &f, size_t ncell, double jfac,
                         double jscale, double alpha, int form,
                         double kappa_scale, int write_kappa, double *partial,
                         int blocks, hipStream_t st);
// kappa = kappa_scale * n * x_n for every cell.
hipError_t launch_kappa(const FieldsDev &f, size_t ncell, double kappa_scale,
                        hipStream_t st);
// p[i] = value for i < n.
hipError_t launch_fill(double *p, size_t n, double value, hipStream_t st);

}  // namespace cmi
