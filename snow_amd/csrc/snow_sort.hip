// snow_sort.hip - stable (key, slot) radix sort of the particle cell keys.
// Uses rocPRIM's device radix sort (LSD, stable, deterministic), limited to the key bits
// the batch actually uses. Kept in its own translation unit: its templates are heavy.
#include <hip/hip_runtime.h>
#include <rocprim/device/device_radix_sort.hpp>
#include <stdint.h>

namespace snow {

hipError_t radix_sort_pairs(void* temp, size_t* temp_bytes, const uint32_t* keys_in, uint32_t* keys_out,
                            const uint32_t* vals_in, uint32_t* vals_out, uint32_t n, int end_bit,
                            hipStream_t stream) {
  return rocprim::radix_sort_pairs(temp, *temp_bytes, keys_in, keys_out, vals_in, vals_out, size_t(n), 0u,
                                   unsigned(end_bit), stream, false);
}

}  // namespace snow
