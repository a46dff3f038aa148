// kernels_pre.hip -- CloudPreProcessor's filter commands on the device
// (cloud-pre-processor.cc:30-44): the voxel-grid downsampling of
// VoxelGridCmd (voxel-grid-cmd.cc:11-24, PCL VoxelGrid with every field
// averaged) and the z pass-through of PassThroughGndFilterCmd
// (pass-through-gnd-filter-cmd.cc:14-26, PCL PassThrough).  Points are
// float4 (x, y, z, intensity).
//
// Voxel grid: the bounding box of the finite points (block min/max, then
// one block), the voxel index of each point (PCL VoxelGrid 1.8's
// arithmetic: floor of the float product with the inverse leaf, minus the
// box's lower voxel), a stable radix sort of (index, point), runs of equal
// indices, and one centroid per run summed in float in the sorted order
// (each voxel's points in their input order).  The output is ordered by
// voxel index, as the reference's.
//
// Also here: the gather of a batched call's input clouds from the caller's
// device buffers into a context's cloud slots (one launch for the batch).
#include <hipcub/hipcub.hpp>

#include "kernels.hpp"

namespace phz {

namespace {

__device__ __forceinline__ bool finite3(float4 p) {
  return isfinite(p.x) && isfinite(p.y) && isfinite(p.z);
}

// per-block min/max of the finite points' x, y, z: part[b] = (min xyz, max xyz)
__global__ void k_minmax_blocks(const float4* __restrict__ pts, int64_t P,
                                float* __restrict__ part) {
  __shared__ float sm[6][256];
  float mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < P;
       i += (int64_t)gridDim.x * blockDim.x) {
    const float4 p = pts[i];
    if (!finite3(p)) continue;
    mn[0] = fminf(mn[0], p.x);
    mn[1] = fminf(mn[1], p.y);
    mn[2] = fminf(mn[2], p.z);
    mx[0] = fmaxf(mx[0], p.x);
    mx[1] = fmaxf(mx[1], p.y);
    mx[2] = fmaxf(mx[2], p.z);
  }
  for (int d = 0; d < 3; ++d) {
    sm[d][threadIdx.x] = mn[d];
    sm[3 + d][threadIdx.x] = mx[d];
  }
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w)
      for (int d = 0; d < 3; ++d) {
        sm[d][threadIdx.x] = fminf(sm[d][threadIdx.x], sm[d][threadIdx.x + w]);
        sm[3 + d][threadIdx.x] = fmaxf(sm[3 + d][threadIdx.x], sm[3 + d][threadIdx.x + w]);
      }
    __syncthreads();
  }
  if (threadIdx.x == 0)
    for (int d = 0; d < 6; ++d) part[blockIdx.x * 6 + d] = sm[d][0];
}

__global__ void k_minmax_final(const float* __restrict__ part, int nb, float* __restrict__ out) {
  if (threadIdx.x != 0) return;
  float r[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
  for (int b = 0; b < nb; ++b)
    for (int d = 0; d < 3; ++d) {
      r[d] = fminf(r[d], part[b * 6 + d]);
      r[3 + d] = fmaxf(r[3 + d], part[b * 6 + 3 + d]);
    }
  for (int d = 0; d < 6; ++d) out[d] = r[d];
}

// voxel index of each point (voxel_grid.hpp: ijk = (int)(floor(p * inv) -
// (float)min_b), idx = ijk . (1, div_x, div_x div_y)); non-finite points
// take the largest key (dropped)
__global__ void k_voxel_grid_keys(const float4* __restrict__ pts, int64_t P, float3 inv,
                                  int3 min_b, int3 mul, uint32_t* __restrict__ keys,
                                  uint32_t* __restrict__ vals) {
#pragma clang fp contract(off)
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= P) return;
  const float4 p = pts[i];
  uint32_t k = 0xffffffffu;
  if (finite3(p)) {
    const int i0 = (int)(floorf(p.x * inv.x) - (float)min_b.x);
    const int i1 = (int)(floorf(p.y * inv.y) - (float)min_b.y);
    const int i2 = (int)(floorf(p.z * inv.z) - (float)min_b.z);
    k = (uint32_t)(i0 * mul.x + i1 * mul.y + i2 * mul.z);
  }
  keys[i] = k;
  vals[i] = (uint32_t)i;
}

// one centroid per run: the run's points summed in order, divided by the count
__global__ void k_voxel_grid_centroids(const float4* __restrict__ pts,
                                       const uint32_t* __restrict__ order,
                                       const uint32_t* __restrict__ counts,
                                       const uint32_t* __restrict__ offsets,
                                       const int* __restrict__ nruns, float4* __restrict__ out) {
#pragma clang fp contract(off)
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= *nruns) return;
  const uint32_t o = offsets[r], c = counts[r];
  float sx = 0.f, sy = 0.f, sz = 0.f, si = 0.f;
  for (uint32_t k = 0; k < c; ++k) {
    const float4 p = pts[order[o + k]];
    sx += p.x;
    sy += p.y;
    sz += p.z;
    si += p.w;
  }
  const float n = (float)c;
  out[r] = make_float4(sx / n, sy / n, sz / n, si / n);
}

// PCL PassThrough (passthrough.hpp, applyFilterIndices): non-finite points
// and non-finite field values are dropped; kept are the points whose field
// (0 x, 1 y, 2 z, 3 intensity) lies in [lo, hi], or outside it with negative
__global__ void k_pass_through_flags(const float4* __restrict__ pts, int64_t P, int field, float lo,
                                     float hi, int negative, unsigned char* __restrict__ keep) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= P) return;
  const float4 p = pts[i];
  const float v = field == 0 ? p.x : (field == 1 ? p.y : (field == 2 ? p.z : p.w));
  bool k = finite3(p) && isfinite(v);
  if (k) k = negative ? !(v >= lo && v <= hi) : !(v < lo || v > hi);
  keep[i] = k ? 1 : 0;
}

}  // namespace

int launch_bbox(const float4* pts, int64_t P, float* d_part, float* d_out, hipStream_t s) {
  constexpr int kBlocks = 256;
  hipLaunchKernelGGL(k_minmax_blocks, dim3(kBlocks), dim3(256), 0, s, pts, P, d_part);
  hipLaunchKernelGGL(k_minmax_final, dim3(1), dim3(64), 0, s, d_part, kBlocks, d_out);
  return kBlocks;
}

void launch_voxel_grid_keys(const float4* pts, int64_t P, float3 inv, int3 min_b, int3 mul,
                            uint32_t* keys, uint32_t* vals, hipStream_t s) {
  if (P <= 0) return;
  hipLaunchKernelGGL(k_voxel_grid_keys, dim3((unsigned)((P + 255) / 256)), dim3(256), 0, s, pts,
                     P, inv, min_b, mul, keys, vals);
}

hipError_t voxel_grid_sort_runs(void* temp, size_t* temp_bytes, const uint32_t* keys_in,
                                uint32_t* keys_out, const uint32_t* vals_in, uint32_t* vals_out,
                                int n, uint32_t* unique, uint32_t* counts, uint32_t* offsets,
                                int* nruns, hipStream_t s) {
  // three library passes share one temporary buffer: its size is the
  // largest of the three (temp == null: size query only)
  size_t a = 0, b = 0, c = 0;
  hipError_t e;
  if ((e = hipcub::DeviceRadixSort::SortPairs(nullptr, a, keys_in, keys_out, vals_in, vals_out, n,
                                               0, 32, s)) != hipSuccess)
    return e;
  if ((e = hipcub::DeviceRunLengthEncode::Encode(nullptr, b, keys_out, unique, counts, nruns, n,
                                                  s)) != hipSuccess)
    return e;
  if ((e = hipcub::DeviceScan::ExclusiveSum(nullptr, c, counts, offsets, n, s)) != hipSuccess)
    return e;
  const size_t need = std::max(a, std::max(b, c));
  if (!temp) {
    *temp_bytes = need;
    return hipSuccess;
  }
  if ((e = hipcub::DeviceRadixSort::SortPairs(temp, a, keys_in, keys_out, vals_in, vals_out, n, 0,
                                               32, s)) != hipSuccess)
    return e;
  if ((e = hipcub::DeviceRunLengthEncode::Encode(temp, b, keys_out, unique, counts, nruns, n, s)) !=
      hipSuccess)
    return e;
  return hipcub::DeviceScan::ExclusiveSum(temp, c, counts, offsets, n, s);
}

void launch_voxel_grid_centroids(const float4* pts, const uint32_t* order, const uint32_t* counts,
                                 const uint32_t* offsets, const int* nruns, int max_runs,
                                 float4* out, hipStream_t s) {
  if (max_runs <= 0) return;
  hipLaunchKernelGGL(k_voxel_grid_centroids, dim3((max_runs + 255) / 256), dim3(256), 0, s, pts,
                     order, counts, offsets, nruns, out);
}

namespace {
// The input clouds of a batched call, gathered from the caller's buffers
// into the context's slots: one workgroup of 256 threads per 1024 points of
// cloud blockIdx.y; each thread issues its four loads before its stores
// (indices clamped into the cloud, stores predicated).
__global__ void __launch_bounds__(256) k_gather_clouds(const GatherEntry* __restrict__ tab,
                                                       float4* __restrict__ pts,
                                                       float2* __restrict__ extra, int64_t cap,
                                                       int* __restrict__ npts) {
  const int slot = blockIdx.y;
  const GatherEntry e = tab[slot];
  if (blockIdx.x == 0 && threadIdx.x == 0) npts[slot] = (int)e.n;
  float4* dp = pts + (int64_t)slot * cap;
  const int64_t stride = (int64_t)gridDim.x * 1024;
  for (int64_t b = (int64_t)blockIdx.x * 1024; b < e.n; b += stride) {
    float4 v[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const int64_t i = b + threadIdx.x + 256 * k;
      v[k] = e.pts[i < e.n ? i : e.n - 1];
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const int64_t i = b + threadIdx.x + 256 * k;
      if (i < e.n) dp[i] = v[k];
    }
    if (extra) {
      float2* dx = extra + (int64_t)slot * cap;
      float2 x[4];
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const int64_t i = b + threadIdx.x + 256 * k;
        x[k] = e.extra ? e.extra[i < e.n ? i : e.n - 1] : make_float2(0.f, 0.f);
      }
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const int64_t i = b + threadIdx.x + 256 * k;
        if (i < e.n) dx[i] = x[k];
      }
    }
  }
}
}  // namespace

void launch_gather_clouds(const GatherEntry* tab, int ncl, int64_t max_n, float4* pts,
                          float2* extra, int64_t cap, int* npts, hipStream_t s) {
  if (ncl <= 0) return;
  const int64_t per = (max_n + 1023) / 1024;
  const unsigned bx = (unsigned)(per < 1 ? 1 : (per > 256 ? 256 : per));
  hipLaunchKernelGGL(k_gather_clouds, dim3(bx, ncl), dim3(256), 0, s, tab, pts, extra, cap, npts);
}

hipError_t pass_through_select(void* temp, size_t* temp_bytes, const float4* pts, int64_t P,
                               int field, float lo, float hi, int negative, unsigned char* keep,
                               float4* out, int* nsel, hipStream_t s) {
  if (!temp)
    return hipcub::DeviceSelect::Flagged(nullptr, *temp_bytes, pts, keep, out, nsel, (int)P, s);
  if (P > 0)
    hipLaunchKernelGGL(k_pass_through_flags, dim3((unsigned)((P + 255) / 256)), dim3(256), 0, s,
                       pts, P, field, lo, hi, negative, keep);
  return hipcub::DeviceSelect::Flagged(temp, *temp_bytes, pts, keep, out, nsel, (int)P, s);
}

}  // namespace phz
