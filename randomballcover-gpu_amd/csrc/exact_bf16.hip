// BF16 MFMA stages of the exact query.
#include "exact_query.hpp"

namespace rbc {

int bf16_qreps(rbc_exact *, const float *, size_t, uint32_t, float *, hipStream_t) {
  set_error("BF16 filter not built yet");
  return RBC_ERR_ARG;
}
int bf16_probe(rbc_exact *, const float *, size_t, const uint32_t *, const ScanItem *, uint32_t,
               uint32_t, float *, uint32_t *, hipStream_t) {
  set_error("BF16 filter not built yet");
  return RBC_ERR_ARG;
}
int bf16_bounds(rbc_exact *, const float *, size_t, uint32_t, uint32_t, uint32_t, const float *,
                const float *, const uint32_t *, float *, double *, uint32_t *, hipStream_t) {
  set_error("BF16 filter not built yet");
  return RBC_ERR_ARG;
}
int bf16_scan(rbc_exact *, const float *, size_t, const uint32_t *, const ScanItem *, uint32_t,
              const float *, float *, uint32_t *, uint32_t *, uint32_t, hipStream_t) {
  set_error("BF16 filter not built yet");
  return RBC_ERR_ARG;
}
double bf16_plan_err(const rbc_exact *) { return 0.0; }

}  // namespace rbc
