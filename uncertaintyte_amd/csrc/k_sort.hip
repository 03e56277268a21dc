// k_sort.hip -- stable grouping of observations by camera.
//
// The reference accepts any row order (its sparse products do not care), so
// camera a's observations need not be contiguous (04_Daliborka.jacob is
// point-major).  A stable LSD radix sort of (camera, observation) pairs gives
// every camera its observations in row order, the order the camera-block
// assembly and the Schur sweep accumulate in.  Dropped observations (fixed
// points) get the key num_cams and sort to the end.
#include <cstring>

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_run_length_encode.hpp>
#include <rocprim/device/device_scan.hpp>

#include "common.h"

namespace unc {

__global__ void k_sort_keys(int nobs, int nc, const int *obs_cam, int *keys, int *vals) {
  int o = blockIdx.x * blockDim.x + threadIdx.x;
  if (o >= nobs) return;
  int c = obs_cam[o];
  keys[o] = c >= 0 ? c : nc;
  vals[o] = o;
}

static int key_bits(int nc) {
  int b = 1;
  while ((1 << b) <= nc) ++b;
  return b;
}

size_t camera_sort_temp_bytes(int nobs, int nc) {
  size_t bytes = 0;
  if (rocprim::radix_sort_pairs(nullptr, bytes, (int *)nullptr, (int *)nullptr, (int *)nullptr,
                                (int *)nullptr, (unsigned int)nobs, 0, key_bits(nc)) != hipSuccess)
    return 0;
  return bytes;
}

// cam_off[c] = first position of key c in the sorted keys (keys in 0..nc,
// nc marking dropped observations); no atomics, gaps filled for empty cameras
__global__ void k_segment_offsets(const int *keys, int n, int nc, int *off) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i > n) return;
  int prev = (i == 0) ? -1 : min(keys[i - 1], nc);
  int cur = (i == n) ? nc : min(keys[i], nc);
  for (int c = prev + 1; c <= cur; ++c) off[c] = i;
}

hipError_t sort_obs_by_camera(int nobs, int nc, const int *obs_cam, int *keys_in, int *keys_out,
                              int *vals_in, int *cam_list, int *cam_off, void *temp,
                              size_t temp_bytes, hipStream_t st) {
  if (nobs <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_sort_keys, dim3(cdiv(nobs, 256)), dim3(256), 0, st, nobs, nc, obs_cam,
                     keys_in, vals_in);
  size_t tb = temp_bytes;
  hipError_t e = rocprim::radix_sort_pairs(temp, tb, keys_in, keys_out, vals_in, cam_list,
                                           (unsigned int)nobs, 0, key_bits(nc), st);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_segment_offsets, dim3(cdiv(nobs + 1, 256)), dim3(256), 0, st, keys_out, nobs,
                     nc, cam_off);
  return hipGetLastError();
}

// ---- pair grouping for the Schur complement (k_schur.hip)
// stable sort of (block key, packed pair); then runs of equal keys; exclusive
// scans of int counts.  Temp sizes are for n elements.
size_t pair_sort_temp_bytes(int n, int key_bits) {
  size_t b = 0;
  if (rocprim::radix_sort_pairs(nullptr, b, (unsigned int *)nullptr, (unsigned int *)nullptr,
                                (unsigned long long *)nullptr, (unsigned long long *)nullptr,
                                (unsigned int)n, 0, key_bits) != hipSuccess)
    return 0;
  return b;
}

hipError_t pair_sort(const unsigned int *keys_in, unsigned int *keys_out,
                     const unsigned long long *vals_in, unsigned long long *vals_out, int n,
                     int key_bits, void *temp, size_t temp_bytes, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  size_t tb = temp_bytes;
  return rocprim::radix_sort_pairs(temp, tb, keys_in, keys_out, vals_in, vals_out, (unsigned int)n,
                                   0, key_bits, st);
}

size_t runs_temp_bytes(int n) {
  size_t b = 0;
  if (rocprim::run_length_encode(nullptr, b, (const unsigned int *)nullptr, (unsigned int)n,
                                 (unsigned int *)nullptr, (int *)nullptr, (int *)nullptr) !=
      hipSuccess)
    return 0;
  return b;
}

hipError_t runs(const unsigned int *keys, int n, unsigned int *unique, int *counts, int *nruns,
                void *temp, size_t temp_bytes, hipStream_t st) {
  size_t tb = temp_bytes;
  return rocprim::run_length_encode(temp, tb, keys, (unsigned int)n, unique, counts, nruns, st);
}

size_t scan_temp_bytes(int n) {
  size_t b = 0;
  if (rocprim::exclusive_scan(nullptr, b, (const int *)nullptr, (int *)nullptr, 0, (size_t)n,
                              rocprim::plus<int>()) != hipSuccess)
    return 0;
  return b;
}

hipError_t exclusive_scan_int(const int *in, int *out, int n, void *temp, size_t temp_bytes,
                              hipStream_t st) {
  if (n <= 0) return hipSuccess;
  size_t tb = temp_bytes;
  return rocprim::exclusive_scan(temp, tb, in, out, 0, (size_t)n, rocprim::plus<int>(), st);
}

}  // namespace unc
