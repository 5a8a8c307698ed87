// Device-side helpers for gfx950 (wave64): wavefront reductions and scans.
#pragma once

#include <hip/hip_runtime.h>

namespace hn {

constexpr int kWave = 64;

template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
  for (int off = kWave / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
  return v;
}

// Maximum that ignores NaN inputs the way std::max(best, v) does when best
// starts at a number: (m < v) ? v : m.
template <typename T>
__device__ __forceinline__ T nan_ignoring_max(T m, T v) {
  return (m < v) ? v : m;
}

template <typename T>
__device__ __forceinline__ T wave_max(T v) {
#pragma unroll
  for (int off = kWave / 2; off > 0; off >>= 1) v = nan_ignoring_max(v, __shfl_xor(v, off, kWave));
  return v;
}

template <typename T>
__device__ __forceinline__ T wave_min(T v) {
#pragma unroll
  for (int off = kWave / 2; off > 0; off >>= 1) {
    T o = __shfl_xor(v, off, kWave);
    v = (o < v) ? o : v;
  }
  return v;
}

// Inclusive segmented scan across the wave: lanes holding equal keys form a
// segment; keys must be non-decreasing in lane order.
template <typename T>
__device__ __forceinline__ T wave_segscan(int key, T v) {
  const int lane = threadIdx.x & (kWave - 1);
#pragma unroll
  for (int off = 1; off < kWave; off <<= 1) {
    T ov = __shfl_up(v, off, kWave);
    int ok = __shfl_up(key, off, kWave);
    if (lane >= off && ok == key) v += ov;
  }
  return v;
}

// Block-wide sum for blockDim.x == NT (multiple of 64); result valid in every
// thread.  `red` must hold NT/64 elements.
template <typename T, int NT>
__device__ __forceinline__ T block_sum(T v, T* red) {
  v = wave_sum(v);
  const int w = threadIdx.x / kWave;
  if ((threadIdx.x & (kWave - 1)) == 0) red[w] = v;
  __syncthreads();
  T s = 0;
#pragma unroll
  for (int i = 0; i < NT / kWave; ++i) s += red[i];
  __syncthreads();
  return s;
}

template <typename T, int NT>
__device__ __forceinline__ T block_max(T v, T* red) {
  v = wave_max(v);
  const int w = threadIdx.x / kWave;
  if ((threadIdx.x & (kWave - 1)) == 0) red[w] = v;
  __syncthreads();
  T s = red[0];
#pragma unroll
  for (int i = 1; i < NT / kWave; ++i) s = nan_ignoring_max(s, red[i]);
  __syncthreads();
  return s;
}

template <typename T, int NT>
__device__ __forceinline__ T block_min(T v, T* red) {
  v = wave_min(v);
  const int w = threadIdx.x / kWave;
  if ((threadIdx.x & (kWave - 1)) == 0) red[w] = v;
  __syncthreads();
  T s = red[0];
#pragma unroll
  for (int i = 1; i < NT / kWave; ++i) s = (red[i] < s) ? red[i] : s;
  __syncthreads();
  return s;
}

}  // namespace hn
