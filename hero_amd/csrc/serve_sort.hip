// Radix sort of the serving path's seeds by (subject << 1 | strand), on rocPRIM. Kept in its own
// translation unit: the rocPRIM headers dominate compile time.
#include <hip/hip_runtime.h>

#include <rocprim/device/device_radix_sort.hpp>

#include "kernels.h"

namespace hero {

hipError_t sort_pairs_u64(const uint64_t *keys_in, uint64_t *keys_out, const uint64_t *vals_in, uint64_t *vals_out,
		uint32_t n, int key_bits, void *tmp, size_t &tmp_bytes, hipStream_t st) {
	return rocprim::radix_sort_pairs(tmp, tmp_bytes, keys_in, keys_out, vals_in, vals_out, n, 0u, (unsigned) key_bits, st);
}

} // namespace hero
