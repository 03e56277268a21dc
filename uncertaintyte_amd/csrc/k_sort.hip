// k_sort.hip -- grouping of observations by camera.
//
// The reference accepts any row order (its sparse products do not care), so
// camera a's observations need not be contiguous (04_Daliborka.jacob is
// point-major).  A stable LSD radix sort by camera of the point-grouped list
// gives every camera its observations in point order, the order the
// camera-block assembly and the Schur complement accumulate in.  Observations
// dropped by fixPts are not in the point lists; the tail of the key array gets
// the key num_cams and sorts to the end.
#include <cstring>

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include "common.h"

namespace unc {

// key of the t-th point-grouped observation (its camera), value the observation
__global__ void k_sort_keys(Structure s, int *pt_cams, int *keys, int *vals) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= s.nobs) return;
  if (t < s.pt_off[s.npc]) {
    const int o = s.pt_list[t], c = s.obs_cam[o];
    pt_cams[t] = c;
    keys[t] = c;
    vals[t] = o;
  } else {
    keys[t] = s.nc;
    vals[t] = -1;
  }
}

__global__ void k_cam_pts(Structure s, int *cam_pts) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < s.cam_off[s.nc]) cam_pts[t] = s.obs_pt[s.cam_list[t]];
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

hipError_t sort_obs_by_camera(const Structure &s, int *pt_cams, int *cam_pts, int *keys_in,
                              int *keys_out, int *vals_in, void *temp, size_t temp_bytes,
                              hipStream_t st) {
  const int nobs = s.nobs, nc = s.nc;
  if (nobs <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_sort_keys, dim3(cdiv(nobs, 256)), dim3(256), 0, st, s, pt_cams, keys_in,
                     vals_in);
  size_t tb = temp_bytes;
  hipError_t e = rocprim::radix_sort_pairs(temp, tb, keys_in, keys_out, vals_in, s.cam_list,
                                           (unsigned int)nobs, 0, key_bits(nc), st);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_segment_offsets, dim3(cdiv(nobs + 1, 256)), dim3(256), 0, st, keys_out, nobs,
                     nc, s.cam_off);
  hipLaunchKernelGGL(k_cam_pts, dim3(cdiv(nobs, 256)), dim3(256), 0, st, s, cam_pts);
  return hipGetLastError();
}

// ---- exclusive scans of int counts (temp sizes are for n elements)
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
