// Validation metrics of the mask-propagation harness (warping.py:245-270),
// which scores propagated masks with monai's compute_dice and
// compute_hausdorff_distance on one binary channel.  A voxel is set when its
// value is > 0.5 (the harness thresholds the propagated masks at 0.5).
//
//   counts     |A|, |B|, |A and B| as exact 64-bit integers (Dice is formed
//              from them on the host).
//   edges      set voxels with a 6-neighbour that is unset or outside the
//              volume: mask XOR binary_erosion(mask) with SciPy's default
//              cross, which is what monai's get_mask_edges computes.
//   Hausdorff  per direction A -> B, the largest squared Euclidean distance
//              (voxel units, an exact integer) from an edge voxel of A to the
//              nearest edge voxel of B.  monai takes that distance from a
//              Euclidean distance transform of B's edges, i.e. the same value.
//              The workspace's first two 64-bit words receive the edge counts
//              of A and B (0 edges: the distance is undefined).
#include <algorithm>
#include <mutex>

#include "svs_common.hpp"
#include "svs_device.hpp"
#include "svsflow.h"

namespace svs {

constexpr int kMetricBlock = 256;

// Zeroes the counters a launch accumulates into (a fill kernel rather than a
// memset, so the sequence can be captured in a graph).
__global__ void k_zero_u64(unsigned long long* p, int n) {
  if ((int)threadIdx.x < n) p[threadIdx.x] = 0ull;
}

__global__ __launch_bounds__(kMetricBlock) void k_mask_counts(const float* __restrict__ a,
                                                              const float* __restrict__ b, int64_t n,
                                                              unsigned long long* __restrict__ counts) {
  __shared__ unsigned int part[3][kMetricBlock / 64];
  unsigned int na = 0, nb = 0, nab = 0;
  for (int64_t i = (int64_t)blockIdx.x * kMetricBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kMetricBlock) {
    const bool x = a[i] > 0.5f, y = b[i] > 0.5f;
    na += x;
    nb += y;
    nab += (x && y);
  }
  // Wavefront sums (64 lanes), then one atomic per block and counter.
  for (int off = 32; off > 0; off >>= 1) {
    na += __shfl_down(na, off);
    nb += __shfl_down(nb, off);
    nab += __shfl_down(nab, off);
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) { part[0][wave] = na; part[1][wave] = nb; part[2][wave] = nab; }
  __syncthreads();
  if (threadIdx.x < 3) {
    unsigned long long s = 0;
    for (int w = 0; w < kMetricBlock / 64; ++w) s += part[threadIdx.x][w];
    atomicAdd(&counts[threadIdx.x], s);
  }
}

__device__ __forceinline__ bool mask_at(const float* m, int z, int y, int x, int NZ, int NY, int NX) {
  return z >= 0 && z < NZ && y >= 0 && y < NY && x >= 0 && x < NX && m[((int64_t)z * NY + y) * NX + x] > 0.5f;
}

// Appends the linear index of every edge voxel to `list`; `count` receives
// the number (the order depends on scheduling; the distances do not).
__global__ __launch_bounds__(kMetricBlock) void k_mask_edges(const float* __restrict__ m, int NZ, int NY, int NX,
                                                             unsigned int* __restrict__ list,
                                                             unsigned long long* __restrict__ count) {
  const int64_t i = (int64_t)blockIdx.x * kMetricBlock + threadIdx.x;
  if (i >= (int64_t)NZ * NY * NX) return;
  int z, y, x;
  unflatten(i, NY, NX, z, y, x);
  if (!(m[i] > 0.5f)) return;
  const bool inner = mask_at(m, z - 1, y, x, NZ, NY, NX) && mask_at(m, z + 1, y, x, NZ, NY, NX) &&
                     mask_at(m, z, y - 1, x, NZ, NY, NX) && mask_at(m, z, y + 1, x, NZ, NY, NX) &&
                     mask_at(m, z, y, x - 1, NZ, NY, NX) && mask_at(m, z, y, x + 1, NZ, NY, NX);
  if (inner) return;
  const unsigned long long slot = atomicAdd(count, 1ull);
  list[slot] = (unsigned int)i;
}

// One thread per edge voxel of A: the squared distance to the nearest edge
// voxel of B, with B's list streamed through LDS in tiles; the block maximum
// goes to *worst with an integer atomic max.
__global__ __launch_bounds__(kMetricBlock) void k_directed_hausdorff2(
    const unsigned int* __restrict__ ea, const unsigned long long* __restrict__ na_p,
    const unsigned int* __restrict__ eb, const unsigned long long* __restrict__ nb_p, int NY, int NX,
    unsigned long long* __restrict__ worst) {
  __shared__ int tz[kMetricBlock], ty[kMetricBlock], tx[kMetricBlock];
  const unsigned long long na = *na_p, nb = *nb_p;
  // Work-groups past A's edges leave before any barrier (uniform per group).
  if ((unsigned long long)blockIdx.x * kMetricBlock >= na) return;
  const unsigned long long i = (unsigned long long)blockIdx.x * kMetricBlock + threadIdx.x;
  const bool active = i < na;
  int az = 0, ay = 0, ax = 0;
  if (active) {
    int z, y, x;
    unflatten((int64_t)ea[i], NY, NX, z, y, x);
    az = z; ay = y; ax = x;
  }
  unsigned long long best = ~0ull;
  for (unsigned long long base = 0; base < nb; base += kMetricBlock) {
    __syncthreads();
    const unsigned long long j = base + threadIdx.x;
    if (j < nb) {
      int z, y, x;
      unflatten((int64_t)eb[j], NY, NX, z, y, x);
      tz[threadIdx.x] = z; ty[threadIdx.x] = y; tx[threadIdx.x] = x;
    }
    __syncthreads();
    const int m = (int)((nb - base) < (unsigned long long)kMetricBlock ? (nb - base) : kMetricBlock);
    if (active)
      for (int k = 0; k < m; ++k) {
        const long long dz = az - tz[k], dy = ay - ty[k], dx = ax - tx[k];
        const unsigned long long d = (unsigned long long)(dz * dz + dy * dy + dx * dx);
        best = d < best ? d : best;
      }
  }
  if (active && nb > 0) atomicMax(worst, best);
}

// Scratch of svs_mask_hausdorff2_noalloc: the two edge counts (64 bytes)
// and two edge lists of up to kNoAllocMaxVoxels entries, 16 MiB + 64 bytes,
// held by every GPU the library runs on until the process ends.  Host code
// never passes this symbol to a kernel: its device address comes from
// hipGetSymbolAddress, once per device (noalloc_scratch).
constexpr int64_t kNoAllocMaxVoxels = (int64_t)1 << 21;  // 128^3
constexpr int kMaxDevices = 64;
__device__ __align__(256) unsigned char scratch[64 + 2 * 4 * kNoAllocMaxVoxels];

static std::mutex g_scratch_mu;
static void* g_scratch_ptr[kMaxDevices] = {};

// The device address of `scratch` on the current device, looked up on the
// first call there.
static int noalloc_scratch(void** out) {
  int dev = -1;
  SVS_HIP(hipGetDevice(&dev));
  SVS_REQUIRE(dev >= 0 && dev < kMaxDevices, "device %d beyond the %d the library tracks", dev, kMaxDevices);
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    SVS_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    SVS_REQUIRE(p, "no device address for the scratch array on device %d", dev);
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return 0;
}

// Both directed distances; cnt[0..1] and out[0..1] are zeroed first.
static int enqueue_hausdorff(const float* a, const float* b, int NZ, int NY, int NX, unsigned long long* cnt,
                             unsigned int* la, unsigned int* lb, unsigned long long* out, hipStream_t st) {
  const int64_t n = (int64_t)NZ * NY * NX;
  k_zero_u64<<<1, 64, 0, st>>>(cnt, 2);
  k_zero_u64<<<1, 64, 0, st>>>(out, 2);
  const int grid = grid_for(n, kMetricBlock);
  k_mask_edges<<<grid, kMetricBlock, 0, st>>>(a, NZ, NY, NX, la, cnt);
  k_mask_edges<<<grid, kMetricBlock, 0, st>>>(b, NZ, NY, NX, lb, cnt + 1);
  // Edge counts stay on the device: the launches cover every voxel and
  // threads past the count return.
  k_directed_hausdorff2<<<grid, kMetricBlock, 0, st>>>(la, cnt, lb, cnt + 1, NY, NX, out);
  k_directed_hausdorff2<<<grid, kMetricBlock, 0, st>>>(lb, cnt + 1, la, cnt, NY, NX, out + 1);
  SVS_LAUNCHED();
  return SVS_STATUS_OK;
}

}  // namespace svs

using namespace svs;

extern "C" {

int svs_mask_counts(const float* a, const float* b, int NZ, int NY, int NX, unsigned long long* counts,
                    void* stream) {
  SVS_REQUIRE(a && b && counts, "mask counts: null pointer");
  SVS_REQUIRE_DIMS(NZ, NY, NX);
  hipStream_t st = (hipStream_t)stream;
  const int64_t n = (int64_t)NZ * NY * NX;
  k_zero_u64<<<1, 64, 0, st>>>(counts, 3);
  const int grid = (int)std::min<int64_t>(grid_for(n, kMetricBlock), 4096);
  k_mask_counts<<<grid, kMetricBlock, 0, st>>>(a, b, n, counts);
  SVS_LAUNCHED();
  return SVS_STATUS_OK;
}

size_t svs_mask_hausdorff_workspace_bytes(int NZ, int NY, int NX) {
  if (NZ <= 0 || NY <= 0 || NX <= 0) return 0;
  const int64_t n = (int64_t)NZ * NY * NX;
  return 64 + 2 * 4 * (size_t)n;
}

int svs_mask_hausdorff2(const float* a, const float* b, int NZ, int NY, int NX, void* workspace, size_t ws_bytes,
                        unsigned long long* out, void* stream) {
  SVS_REQUIRE(a && b && workspace && out, "mask Hausdorff: null pointer");
  SVS_REQUIRE_DIMS(NZ, NY, NX);
  SVS_REQUIRE(ws_bytes >= svs_mask_hausdorff_workspace_bytes(NZ, NY, NX),
              "mask Hausdorff: workspace of %zu bytes, need %zu", ws_bytes,
              svs_mask_hausdorff_workspace_bytes(NZ, NY, NX));
  SVS_REQUIRE(((uintptr_t)workspace & 7) == 0, "mask Hausdorff: workspace must be 8-byte aligned");
  const int64_t n = (int64_t)NZ * NY * NX;
  unsigned long long* cnt = (unsigned long long*)workspace;     // [0] |edges A|, [1] |edges B|
  unsigned int* la = (unsigned int*)((char*)workspace + 64);
  unsigned int* lb = la + n;
  return enqueue_hausdorff(a, b, NZ, NY, NX, cnt, la, lb, out, (hipStream_t)stream);
}

int svs_mask_hausdorff2_noalloc(const float* a, const float* b, int NZ, int NY, int NX, unsigned long long* out,
                                unsigned long long* edge_counts, void* stream) {
  SVS_REQUIRE(a && b && out, "mask Hausdorff (no-alloc): null pointer");
  SVS_REQUIRE_DIMS(NZ, NY, NX);
  const int64_t n = (int64_t)NZ * NY * NX;
  SVS_REQUIRE(n <= kNoAllocMaxVoxels,
              "mask Hausdorff (no-alloc): %dx%dx%d is %lld voxels, beyond the %lld its fixed scratch holds; "
              "use svs_mask_hausdorff2 with a workspace", NZ, NY, NX, (long long)n, (long long)kNoAllocMaxVoxels);
  void* ws = nullptr;
  if (int rc = noalloc_scratch(&ws)) return rc;
  unsigned long long* cnt = edge_counts ? edge_counts : (unsigned long long*)ws;
  unsigned int* la = (unsigned int*)((char*)ws + 64);
  unsigned int* lb = la + n;
  return enqueue_hausdorff(a, b, NZ, NY, NX, cnt, la, lb, out, (hipStream_t)stream);
}

}  // extern "C"
