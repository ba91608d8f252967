// iq_sort.hip -- stable ascending sort of (FP64 distance, candidate index)
// pairs.  Distances are non-negative (or +inf), so their IEEE bit patterns
// order like the values and a stable LSD radix sort over the bits keeps equal
// distances in index order (Julia's Perm ordering, used by partialsortperm
// in relaxation.jl:12,27).
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>
#include <stdint.h>

size_t iq_sort_temp_bytes(int n) {
  size_t bytes = 0;
  (void)hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, (const unsigned long long*)nullptr,
                                     (unsigned long long*)nullptr, (const int32_t*)nullptr,
                                     (int32_t*)nullptr, n, 0, 63, (hipStream_t)0);
  return bytes;
}

hipError_t iq_sort_pairs(void* tmp, size_t tmp_bytes, const double* keys_in, double* keys_out,
                         const int32_t* vals_in, int32_t* vals_out, int n, hipStream_t s) {
  return hipcub::DeviceRadixSort::SortPairs(tmp, tmp_bytes, (const unsigned long long*)keys_in,
                                            (unsigned long long*)keys_out, vals_in, vals_out, n, 0, 63, s);
}

// Segmented form: segment i covers [begin[i], end[i]) of keys/values (one
// map of one query each), all sorted by one launch sequence.
size_t iq_segsort_temp_bytes(int n, int nseg) {
  size_t bytes = 0;
  (void)hipcub::DeviceSegmentedRadixSort::SortPairs(nullptr, bytes, (const unsigned long long*)nullptr,
                                                    (unsigned long long*)nullptr, (const int32_t*)nullptr,
                                                    (int32_t*)nullptr, n, nseg, (const int64_t*)nullptr,
                                                    (const int64_t*)nullptr, 0, 63, (hipStream_t)0);
  return bytes;
}

hipError_t iq_segsort_pairs(void* tmp, size_t tmp_bytes, const double* keys_in, double* keys_out,
                            const int32_t* vals_in, int32_t* vals_out, int n, int nseg, const int64_t* begin,
                            const int64_t* end, hipStream_t s) {
  return hipcub::DeviceSegmentedRadixSort::SortPairs(tmp, tmp_bytes, (const unsigned long long*)keys_in,
                                                     (unsigned long long*)keys_out, vals_in, vals_out, n, nseg,
                                                     begin, end, 0, 63, s);
}
