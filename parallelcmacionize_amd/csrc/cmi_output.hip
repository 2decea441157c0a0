// cmi_output.hip -- result writers in the reference's on-disk formats, so
// that its analysis scripts read this library's output unchanged.
//
//   neutral_fractions.dat  output_neutral_fractions (testDensitySubGrid_omp.cpp:
//                          394-440) with DensitySubGrid::output_intensities
//                          (DensitySubGrid.hpp:1738-1796): one block per
//                          subgrid in subgrid order, per cell (z fastest) the
//                          cell centre x, y, z, x_n and n_H as FP64, then the
//                          five FP32 hydro primitives, zero in a build without
//                          hydrodynamics (as the reference's NO_HYDRO build
//                          writes them); 60 bytes per cell.
//   neutral_fractions.bin  output_neutral_fractions_binary (testDensitySubGrid_
//                          omp.cpp:507-521): x_n as raw FP64 in the
//                          reference's subgrid-major order.
//
// Host code only; the array entry points need no GPU.
#include <stdio.h>
#include <string.h>

#include <vector>

#include "cmi_ctx.hpp"

namespace {

constexpr size_t kDatCell = 5 * sizeof(double) + 5 * sizeof(float);

int write_all(FILE *f, const void *p, size_t n) {
  return fwrite(p, 1, n, f) == n ? CMI_OK : CMI_ERR_ARG;
}

}  // namespace

namespace cmi {
int get_density_ref(cmi_ctx *c, double *nH);  // cmi_api.hip
}

extern "C" {

int cmi_write_neutral_fractions_dat(const char *path, const double box[6],
                                    const int ncell[3], const int nsub[3],
                                    const double *nH, const double *xn) {
  if (!path || !box || !ncell || !nsub || !nH || !xn) return CMI_ERR_ARG;
  int sc[3];
  double side[3], cell[3];
  for (int a = 0; a < 3; ++a) {
    if (nsub[a] <= 0 || ncell[a] <= 0 || ncell[a] % nsub[a] != 0 || !(box[3 + a] > 0.))
      return CMI_ERR_ARG;
    sc[a] = ncell[a] / nsub[a];
    // subgrid box and cell size as the driver and the subgrid constructor
    // form them (testDensitySubGrid_omp.cpp:1805-1831, DensitySubGrid.hpp:518)
    side[a] = box[3 + a] / nsub[a];
    cell[a] = side[a] / sc[a];
  }
  FILE *f = fopen(path, "wb");
  if (!f) return CMI_ERR_ARG;
  const size_t cps = (size_t)sc[0] * sc[1] * sc[2];
  std::vector<unsigned char> buf(cps * kDatCell);
  size_t r = 0;
  int rc = CMI_OK;
  for (int ix = 0; ix < nsub[0] && rc == CMI_OK; ++ix)
    for (int iy = 0; iy < nsub[1] && rc == CMI_OK; ++iy)
      for (int iz = 0; iz < nsub[2] && rc == CMI_OK; ++iz) {
        const double anchor[3] = {box[0] + ix * side[0], box[1] + iy * side[1],
                                  box[2] + iz * side[2]};
        unsigned char *o = buf.data();
        for (int cx = 0; cx < sc[0]; ++cx) {
          const double px = anchor[0] + (cx + 0.5) * cell[0];
          for (int cy = 0; cy < sc[1]; ++cy) {
            const double py = anchor[1] + (cy + 0.5) * cell[1];
            for (int cz = 0; cz < sc[2]; ++cz, ++r) {
              const double pz = anchor[2] + (cz + 0.5) * cell[2];
              const double d[5] = {px, py, pz, xn[r], nH[r]};
              const float h[5] = {0.f, 0.f, 0.f, 0.f, 0.f};
              memcpy(o, d, sizeof(d));
              memcpy(o + sizeof(d), h, sizeof(h));
              o += kDatCell;
            }
          }
        }
        rc = write_all(f, buf.data(), buf.size());
      }
  if (fclose(f) != 0) rc = CMI_ERR_ARG;
  return rc;
}

int cmi_write_neutral_fractions_bin(const char *path, uint64_t ncell_total,
                                    const double *xn) {
  if (!path || !xn) return CMI_ERR_ARG;
  FILE *f = fopen(path, "wb");
  if (!f) return CMI_ERR_ARG;
  int rc = write_all(f, xn, ncell_total * sizeof(double));
  if (fclose(f) != 0) rc = CMI_ERR_ARG;
  return rc;
}

int cmi_write_results(cmi_ctx *c, const char *dat_path, const char *bin_path) {
  if (!c) return CMI_ERR_ARG;
  std::vector<double> xn(c->ncell_total), n(c->ncell_total);
  int rc = cmi_get_fields(c, nullptr, xn.data());
  if (rc != CMI_OK) return rc;
  if (dat_path) {
    rc = cmi::get_density_ref(c, n.data());
    if (rc != CMI_OK) return rc;
    rc = cmi_write_neutral_fractions_dat(dat_path, c->box, c->ncell, c->nsub, n.data(),
                                         xn.data());
    if (rc != CMI_OK) return cmi::fail(c, rc, std::string("cannot write ") + dat_path);
  }
  if (bin_path) {
    rc = cmi_write_neutral_fractions_bin(bin_path, c->ncell_total, xn.data());
    if (rc != CMI_OK) return cmi::fail(c, rc, std::string("cannot write ") + bin_path);
  }
  return CMI_OK;
}

}  // extern "C"
