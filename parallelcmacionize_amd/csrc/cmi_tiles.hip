// cmi_tiles.hip -- tiles engine (placeholder until the LDS-tile kernels land).
#include "cmi_ctx.hpp"

namespace cmi {

int tiles_transport_packets(cmi_ctx *c, cmi_photon *, const uint32_t *,
                            size_t, int32_t *, uint32_t *, uint32_t *,
                            cmi_stats *) {
  return fail(c, CMI_ERR_STATE, "tiles engine not available");
}

int tiles_transport(cmi_ctx *c, uint64_t, uint64_t, uint32_t, double,
                    cmi_stats *) {
  return fail(c, CMI_ERR_STATE, "tiles engine not available");
}

void tiles_destroy(cmi_ctx *) {}

}  // namespace cmi
