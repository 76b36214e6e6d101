// Flow-consistent spatial augmentation (svs/modules/transforms/spatial.py):
// a RandomRotate followed by any number of RandomFlips, applied to every key
// of a training sample (images, masks, flows) in one launch.
//
// Output voxel o of the flipped result is the rotated volume at q = flip(o)
// (a flip is a pure index map).  The rotated volume at q is F.grid_sample of
// the source at G(q) = identity_grid(q) @ R (spatial.py:56, :77-83); a flow
// key then rotates the sampled vector by R (spatial.py:115-128) and negates
// the flipped components (spatial.py:158-168: axis Z/Y/X negates component
// 2/1/0).  Without a rotation the launch is the flips alone.
//
// Rounding follows the reference on the host CPU:
//   * G comes from the caller (the reference's own torch.matmul), or is
//     formed here from the identity grid's axes with the 3-term rounding the
//     host's BLAS uses for that product (grid_rounding);
//   * the vector rotation uses the host BLAS rounding of torch.einsum
//     (vector_rounding).  Roundings: 0 fma(a2,b2,fma(a1,b1,fma(a0,b0,+0)))
//     (MKL's sgemm on Intel hosts), 1 (a0*b0 + a1*b1) + a2*b2, 2 the same
//     summed from +0;
//   * the sampling is torch's CPU grid_sampler_3d expression by expression:
//     unnormalise, pad, floor / round-half-even, corner weights multiplied
//     x * y * z, corner values accumulated from +0 in (z, y, x) order with
//     out-of-volume corners skipped.
//
// Rotation kernel: a work-group owns an 8 x 8 x 8 tile of rotated-volume
// positions q and streams about 8 channels.  It computes the source
// coordinates, weights and bounding box of the tile once (a rigid motion of
// an 8^3 cube spans at most 14^3 source voxels), stages that box of each
// channel in LDS with coalesced loads (two channels ahead in flight while
// the current one is sampled), with zeros where the box leaves the volume,
// and gathers the 8 trilinear corners from LDS without bounds tests.
// Staging loads and output stores are buffer instructions on per-channel
// descriptors with offsets computed once per tile: an offset past the buffer
// (apron elements, voxels outside the output) loads 0 or drops the store.  A
// tile whose box does not fit (resampling to a coarser grid, reflection
// folds) gathers straight from global memory instead.  Work-groups run in plain block
// order (tile fastest, then channel group): giving each XCD a contiguous run
// of tiles measured 6 % slower.
#include <algorithm>
#include <climits>

#include "svs_common.hpp"
#include "svs_device.hpp"
#include "svsflow.h"

namespace svs {

constexpr int kAugMaxVolumes = 8;
constexpr int kFlipTX = 16, kFlipTY = 16;
constexpr int kRotEdge = 8;
constexpr int kRotThreads = 256;
constexpr int kRotVox = kRotEdge * kRotEdge * kRotEdge / kRotThreads;   // voxels per thread
constexpr int kStage = 12;                                               // staged floats per thread
constexpr int kBoxMax = kStage * kRotThreads;                            // 3072 >= 14^3
constexpr int kLdsStride = 3;                                            // dwords per staged element
constexpr int kRotChannelsPerGroup = 8;                                  // channels a work-group streams

struct AugVolume {
  const float* src;
  float* dst;
  int batch, channels, mode, vector;
};

struct AugArgs {
  AugVolume vol[kAugMaxVolumes];
  int unit_start[kAugMaxVolumes + 1];   // rotation: units (a channel or a vector triple); flips: (key, batch)
  int nvol;
  int IZ, IY, IX, OZ, OY, OX;
  int padding, align, flips, grid_rounding, vector_rounding;
  float R[9];
  const float* grid;   // [OZ, OY, OX, 3] or null
  const float* axes;   // lx[OX], ly[OY], lz[OZ] when grid is null
  int tiles_x, tiles_y, ntiles, units_per_group, total;
};

// grid_sampler_unnormalize / clip_coordinates / reflect_coordinates /
// safe_downgrade_to_int_range of torch's grid sampler.
__device__ __forceinline__ float gs_unnormalize(float c, int size, int align) {
  return align ? ((c + 1.f) / 2.f) * (float)(size - 1) : ((c + 1.f) * (float)size - 1.f) / 2.f;
}
// std::min(size - 1, std::max(c, 0)) with its comparisons (NaN -> size - 1).
__device__ __forceinline__ float gs_clip(float c, int size) {
  const float lim = (float)(size - 1);
  const float m = (c < 0.f) ? 0.f : c;
  return (m < lim) ? m : lim;
}
__device__ float gs_reflect(float in, int twice_low, int twice_high) {
  if (twice_low == twice_high) return 0.f;
  const float lo = (float)twice_low / 2.f, span = (float)(twice_high - twice_low) / 2.f;
  in = fabsf(in - lo);
  const float extra = fmodf(in, span);
  const int flips = (int)floorf(in / span);
  return (flips % 2 == 0) ? extra + lo : span - extra + lo;
}
__device__ __forceinline__ float gs_source(float c, int size, int padding, int align) {
  c = gs_unnormalize(c, size, align);
  if (padding == 1) {
    c = gs_clip(c, size);
  } else if (padding == 2) {
    c = align ? gs_reflect(c, 0, 2 * (size - 1)) : gs_reflect(c, -1, 2 * size - 1);
    c = gs_clip(c, size);
  }
  // Non-finite or out-of-int-range coordinates become -100 (outside).
  return (c > 2147483646.f || c < -2147483648.f || !isfinite(c)) ? -100.f : c;
}

// a0*b0 + a1*b1 + a2*b2 in the host BLAS's rounding (see the file comment).
__device__ __forceinline__ float dot3(int rounding, float a0, float b0, float a1, float b1, float a2, float b2) {
  if (rounding == 0) return fmaf(a2, b2, fmaf(a1, b1, fmaf(a0, b0, 0.f)));
  if (rounding == 1) return (a0 * b0 + a1 * b1) + a2 * b2;
  return ((0.f + a0 * b0) + a1 * b1) + a2 * b2;
}

__device__ __forceinline__ int volume_of(const AugArgs& a, int unit) {
  int v = 0;
  while (v + 1 < a.nvol && unit >= a.unit_start[v + 1]) ++v;
  return v;
}

// ---------------------------------------------------------------- flips only
// One thread per voxel of one (key, batch entry) item; a tile is 16 x 16
// voxels of one z plane.
__global__ __launch_bounds__(kFlipTX* kFlipTY) void k_spatial_flip(const AugArgs a) {
  const int tile = (int)blockIdx.x;
  const int item = tile / a.ntiles;
  int r = tile - item * a.ntiles;
  const int plane = a.tiles_x * a.tiles_y;
  const int oz = r / plane;
  r -= oz * plane;
  const int ty = r / a.tiles_x, tx = r - ty * a.tiles_x;
  const int ox = tx * kFlipTX + (int)(threadIdx.x % kFlipTX);
  const int oy = ty * kFlipTY + (int)(threadIdx.x / kFlipTX);
  if (ox >= a.OX || oy >= a.OY) return;
  const int v = volume_of(a, item);
  const AugVolume V = a.vol[v];
  const int b = item - a.unit_start[v];
  const size_t N = (size_t)a.OZ * a.OY * a.OX;
  const float* src = V.src + (size_t)b * V.channels * N;
  float* dst = V.dst + (size_t)b * V.channels * N + ((size_t)oz * a.OY + oy) * a.OX + ox;
  const int qx = (a.flips & 1) ? a.OX - 1 - ox : ox;
  const int qy = (a.flips & 2) ? a.OY - 1 - oy : oy;
  const int qz = (a.flips & 4) ? a.OZ - 1 - oz : oz;
  const size_t q = ((size_t)qz * a.OY + qy) * a.OX + qx;
  if (V.vector) {
    const float s0 = src[q], s1 = src[N + q], s2 = src[2 * N + q];
    dst[0] = (a.flips & 1) ? -s0 : s0;
    dst[N] = (a.flips & 2) ? -s1 : s1;
    dst[2 * N] = (a.flips & 4) ? -s2 : s2;
  } else {
    for (int c = 0; c < V.channels; ++c) dst[(size_t)c * N] = src[(size_t)c * N + q];
  }
}

// ---------------------------------------------------------------- rotation
// The unit (a channel, or the 3 channels of a flow vector) at `unit`.
struct Chan {
  const float* src;  // this component's source channel
  float* dst;        // output channel (scalar) or the vector's component 0
  int ncomp, mode, vector;
};

__device__ __forceinline__ Chan chan_at(const AugArgs& a, int unit, int comp, size_t Nin, size_t Nout) {
  const int v = volume_of(a, unit);
  const AugVolume V = a.vol[v];
  const int r = unit - a.unit_start[v];
  Chan c;
  c.mode = V.mode;
  c.vector = V.vector;
  if (V.vector) {
    c.ncomp = 3;
    c.src = V.src + ((size_t)r * 3 + comp) * Nin;
    c.dst = V.dst + (size_t)r * 3 * Nout;
  } else {
    c.ncomp = 1;
    c.src = V.src + (size_t)r * Nin;   // r = b * channels + channel
    c.dst = V.dst + (size_t)r * Nout;
  }
  return c;
}

__device__ __forceinline__ void store_vector(const AugArgs& a, svs_rsrc dst, uint32_t off, uint32_t comp_bytes,
                                             float v0, float v1, float v2) {
  float r0 = dot3(a.vector_rounding, a.R[0], v0, a.R[1], v1, a.R[2], v2);
  float r1 = dot3(a.vector_rounding, a.R[3], v0, a.R[4], v1, a.R[5], v2);
  float r2 = dot3(a.vector_rounding, a.R[6], v0, a.R[7], v1, a.R[8], v2);
  if (a.flips & 1) r0 = -r0;
  if (a.flips & 2) r1 = -r1;
  if (a.flips & 4) r2 = -r2;
  bst(dst, off, 0, r0);
  bst(dst, off, comp_bytes, r1);
  bst(dst, off, 2 * comp_bytes, r2);
}

// ---- staged path.  LDS holds two staging buffers interleaved with a
// stride of 3 dwords: element e of buffer b is dword 3e + b, so a corner pair
// (e, e + 1) of buffer b is one ds_read2_b32 at byte 12e with dword offsets b
// and b + 3 (a stride coprime with the bank count, so random pairs spread over
// all banks).  Two channels are in flight in registers: while channel k is
// sampled from LDS, channel k + 1 (loaded a step earlier) waits in registers
// and channel k + 2 is being loaded.
struct VoxL {
  float w[8];        // trilinear corner weights, corner k = (dz << 2) | (dy << 1) | dx
  uint32_t pair[4];  // byte address 12 e of the corner pairs q = (dz << 1) | dy (dx = 0, 1)
  uint32_t near;     // byte address 12 e of the nearest neighbour
  uint32_t out;      // byte offset of flip(q) in an output channel, kOutside when q is not inside
};

// All corner reads of the thread's voxels are issued before any product, so
// a channel costs one LDS round trip per wave.
template <int B>
__device__ __forceinline__ void sample_box2(const float* stage, const VoxL (&vx)[kRotVox], int mode,
                                            float (&val)[kRotVox]) {
  const char* sb = reinterpret_cast<const char*>(stage);
  if (mode == 1) {
#pragma unroll
    for (int j = 0; j < kRotVox; ++j) val[j] = reinterpret_cast<const float*>(sb + vx[j].near)[B];
    return;
  }
  float v[kRotVox][8];
#pragma unroll
  for (int j = 0; j < kRotVox; ++j)
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const float* p = reinterpret_cast<const float*>(sb + vx[j].pair[q]);
      v[j][2 * q] = p[B];
      v[j][2 * q + 1] = p[B + kLdsStride];
    }
#pragma unroll
  for (int j = 0; j < kRotVox; ++j) {
    float acc = 0.f;
#pragma unroll
    for (int k = 0; k < 8; ++k) acc = acc + v[j][k] * vx[j].w[k];
    val[j] = acc;
  }
}

// Position (unit, component) in the work-group's channel sequence.
struct ChanPos {
  int u, comp;
  bool valid;
  Chan c;
};

__device__ __forceinline__ ChanPos chan_next(const AugArgs& a, const ChanPos& p, int u1, size_t Nin, size_t Nout) {
  ChanPos n;
  n.u = p.u;
  n.comp = p.comp + 1;
  if (n.comp >= p.c.ncomp) {
    n.u = p.u + 1;
    n.comp = 0;
  }
  n.valid = p.valid && n.u < u1;
  n.c = n.valid ? chan_at(a, n.u, n.comp, Nin, Nout) : p.c;
  return n;
}

struct StagedLoop {
  ChanPos cur, nxt, nxt2;   // channels k, k + 1, k + 2
  float vec[kRotVox][3];
  float sv[2][kStage];      // register stages: [k & 1] = channel k + 2, [(k + 1) & 1] = channel k + 1
};

__device__ __forceinline__ void stage_load(const Chan& c, size_t Nin, const uint32_t (&goff)[kStage],
                                           float (&sv)[kStage]) {
  const svs_rsrc r = make_rsrc(c.src, (uint32_t)(Nin * 4));
#pragma unroll
  for (int s = 0; s < kStage; ++s) sv[s] = bld(r, goff[s], 0);
}

template <int B>
__device__ __forceinline__ void stage_store(float* stage, const float (&sv)[kStage]) {
  char* sb = reinterpret_cast<char*>(stage) + 4 * kLdsStride * threadIdx.x + 4 * B;
#pragma unroll
  for (int s = 0; s < kStage; ++s) *reinterpret_cast<float*>(sb + 4 * kLdsStride * kRotThreads * s) = sv[s];
}

// Samples channel k (staged in LDS buffer B), starts the loads of channel
// k + 2, stages channel k + 1 in buffer B ^ 1.  Returns false after the last
// channel.
template <int B>
__device__ __forceinline__ bool staged_step(const AugArgs& a, StagedLoop& st, const VoxL (&vx)[kRotVox],
                                            const uint32_t (&goff)[kStage], int u1, float* stage, size_t Nin,
                                            size_t Nout) {
  if (st.nxt2.valid) stage_load(st.nxt2.c, Nin, goff, st.sv[B]);
  const Chan& c = st.cur.c;
  const int comp = st.cur.comp;
  const uint32_t comp_bytes = (uint32_t)(Nout * 4);
  const svs_rsrc dst = make_rsrc(c.dst, (c.vector ? 3u : 1u) * comp_bytes);
  float val[kRotVox];
  sample_box2<B>(stage, vx, c.mode, val);
#pragma unroll
  for (int j = 0; j < kRotVox; ++j) {
    // Every component is rewritten on every path (a select, no branch).
    st.vec[j][0] = comp == 0 ? val[j] : st.vec[j][0];
    st.vec[j][1] = comp == 1 ? val[j] : st.vec[j][1];
    st.vec[j][2] = comp == 2 ? val[j] : st.vec[j][2];
    if (!c.vector) bst(dst, vx[j].out, 0, val[j]);
  }
  if (c.vector && comp == 2) {
#pragma unroll
    for (int j = 0; j < kRotVox; ++j)
      store_vector(a, dst, vx[j].out, comp_bytes, st.vec[j][0], st.vec[j][1], st.vec[j][2]);
  }
  if (!st.nxt.valid) return false;
  stage_store<B ^ 1>(stage, st.sv[B ^ 1]);
  __syncthreads();
  st.cur = st.nxt;
  st.nxt = st.nxt2;
  st.nxt2 = chan_next(a, st.nxt2, u1, Nin, Nout);
  return true;
}

// ---- global path (boxes too large to stage): corners straight from memory,
// out-of-volume corners skipped.
struct VoxG {
  float w[8];
  int base;          // index of corner 0 in the channel
  unsigned ok;       // bit k: corner k inside the source volume
  int nidx;          // nearest-neighbour index, valid when nok
  bool nok;
  uint32_t out;
};

__device__ __forceinline__ float sample_global(const float* p, const VoxG& x, int mode, int dy, int dz) {
  if (mode == 1) {
    const float v = p[x.nok ? x.nidx : 0];
    return x.nok ? v : 0.f;
  }
  float acc = 0.f;
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    const bool ok = (x.ok >> k) & 1;
    const int idx = ok ? x.base + (k & 1) + ((k >> 1) & 1) * dy + (k >> 2) * dz : 0;
    const float v = p[idx];
    acc = ok ? acc + v * x.w[k] : acc;
  }
  return acc;
}

__device__ void global_channels(const AugArgs& a, const VoxG (&vx)[kRotVox], int u0, int u1) {
  const size_t Nin = (size_t)a.IZ * a.IY * a.IX, Nout = (size_t)a.OZ * a.OY * a.OX;
  const uint32_t comp_bytes = (uint32_t)(Nout * 4);
  const int dy = a.IX, dz = a.IX * a.IY;
  for (int u = u0; u < u1; ++u) {
    const Chan c0 = chan_at(a, u, 0, Nin, Nout);
    const svs_rsrc dst = make_rsrc(c0.dst, (c0.vector ? 3u : 1u) * comp_bytes);
    float vec[kRotVox][3] = {};
    for (int comp = 0; comp < c0.ncomp; ++comp) {
      const float* src = c0.src + (size_t)comp * Nin;
#pragma unroll
      for (int j = 0; j < kRotVox; ++j) {
        const float val = sample_global(src, vx[j], c0.mode, dy, dz);
        vec[j][0] = comp == 0 ? val : vec[j][0];
        vec[j][1] = comp == 1 ? val : vec[j][1];
        vec[j][2] = comp == 2 ? val : vec[j][2];
        if (!c0.vector) bst(dst, vx[j].out, 0, val);
      }
    }
    if (c0.vector) {
#pragma unroll
      for (int j = 0; j < kRotVox; ++j)
        store_vector(a, dst, vx[j].out, comp_bytes, vec[j][0], vec[j][1], vec[j][2]);
    }
  }
}

__global__ __launch_bounds__(kRotThreads) void k_spatial_rotate(const AugArgs a) {
  __shared__ float stage[kLdsStride * kBoxMax];
  __shared__ int part[kRotThreads / 64][6];
  const int wg = (int)blockIdx.x;
  if (wg >= a.total) return;   // uniform per work-group, before any barrier
  const int group = wg / a.ntiles, tile = wg - group * a.ntiles;
  const int u0 = group * a.units_per_group;
  const int u1 = min(u0 + a.units_per_group, a.unit_start[a.nvol]);
  const int plane = a.tiles_x * a.tiles_y;
  const int tz = tile / plane, txy = tile - tz * plane;
  const int ty = txy / a.tiles_x, tx = txy - ty * a.tiles_x;
  const int tid = (int)threadIdx.x;
  const size_t Nin = (size_t)a.IZ * a.IY * a.IX, Nout = (size_t)a.OZ * a.OY * a.OX;

  // Source coordinates of this thread's voxels.  Floors are clamped to
  // [-2, I] and nearest indices to [-1, I]: both keep every corner's in/out
  // status, and the bounding box of the clamped corners (the box staged, with
  // zeros outside the volume) stays small.
  float cx[kRotVox], cy[kRotVox], cz[kRotVox];
  bool in[kRotVox];
  uint32_t out[kRotVox];
  int lo[3] = {INT_MAX, INT_MAX, INT_MAX}, hi[3] = {INT_MIN, INT_MIN, INT_MIN};
#pragma unroll
  for (int j = 0; j < kRotVox; ++j) {
    const int l = tid + j * kRotThreads;
    const int qx = tx * kRotEdge + (l & 7), qy = ty * kRotEdge + ((l >> 3) & 7), qz = tz * kRotEdge + (l >> 6);
    in[j] = qx < a.OX && qy < a.OY && qz < a.OZ;
    const int sx = in[j] ? qx : 0, sy = in[j] ? qy : 0, sz = in[j] ? qz : 0;
    float gx, gy, gz;
    if (a.grid) {
      const float* g = a.grid + (((size_t)sz * a.OY + sy) * a.OX + sx) * 3;
      gx = g[0];
      gy = g[1];
      gz = g[2];
    } else {
      const float lx = a.axes[sx], ly = a.axes[a.OX + sy], lz = a.axes[a.OX + a.OY + sz];
      gx = dot3(a.grid_rounding, lx, a.R[0], ly, a.R[3], lz, a.R[6]);
      gy = dot3(a.grid_rounding, lx, a.R[1], ly, a.R[4], lz, a.R[7]);
      gz = dot3(a.grid_rounding, lx, a.R[2], ly, a.R[5], lz, a.R[8]);
    }
    cx[j] = gs_source(gx, a.IX, a.padding, a.align);
    cy[j] = gs_source(gy, a.IY, a.padding, a.align);
    cz[j] = gs_source(gz, a.IZ, a.padding, a.align);
    const int ox = (a.flips & 1) ? a.OX - 1 - sx : sx;
    const int oy = (a.flips & 2) ? a.OY - 1 - sy : sy;
    const int oz = (a.flips & 4) ? a.OZ - 1 - sz : sz;
    out[j] = in[j] ? (uint32_t)((oz * a.OY + oy) * a.OX + ox) * 4u : kOutside;
    if (in[j]) {
      const int x0 = (int)fminf(fmaxf(floorf(cx[j]), -2.f), (float)a.IX);
      const int y0 = (int)fminf(fmaxf(floorf(cy[j]), -2.f), (float)a.IY);
      const int z0 = (int)fminf(fmaxf(floorf(cz[j]), -2.f), (float)a.IZ);
      lo[0] = min(lo[0], x0); hi[0] = max(hi[0], x0 + 1);
      lo[1] = min(lo[1], y0); hi[1] = max(hi[1], y0 + 1);
      lo[2] = min(lo[2], z0); hi[2] = max(hi[2], z0 + 1);
    }
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1)
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      lo[d] = min(lo[d], __shfl_xor(lo[d], off));
      hi[d] = max(hi[d], __shfl_xor(hi[d], off));
    }
  if ((tid & 63) == 0) {
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      part[tid >> 6][d] = lo[d];
      part[tid >> 6][3 + d] = hi[d];
    }
  }
  __syncthreads();
  // The tile has at least one voxel inside the output, so every box edge is
  // at least 2 long.
  int b0[3], bn[3];
  int64_t nbox64 = 1;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    int l = part[0][d], h = part[0][3 + d];
#pragma unroll
    for (int w = 1; w < kRotThreads / 64; ++w) {
      l = min(l, part[w][d]);
      h = max(h, part[w][3 + d]);
    }
    b0[d] = l;
    bn[d] = h - l + 1;
    nbox64 *= bn[d];
  }

  if (nbox64 <= kBoxMax) {
    const int nbox = (int)nbox64;
    const int dy = bn[0], dz = bn[0] * bn[1];
    VoxL vx[kRotVox];
#pragma unroll
    for (int j = 0; j < kRotVox; ++j) {
      VoxL& x = vx[j];
      const float fx = floorf(cx[j]), fy = floorf(cy[j]), fz = floorf(cz[j]);
      const float wx[2] = {(fx + 1.f) - cx[j], cx[j] - fx};
      const float wy[2] = {(fy + 1.f) - cy[j], cy[j] - fy};
      const float wz[2] = {(fz + 1.f) - cz[j], cz[j] - fz};
#pragma unroll
      for (int c = 0; c < 8; ++c) x.w[c] = wx[c & 1] * wy[(c >> 1) & 1] * wz[c >> 2];
      const int x0 = (int)fminf(fmaxf(fx, -2.f), (float)a.IX) - b0[0];
      const int y0 = (int)fminf(fmaxf(fy, -2.f), (float)a.IY) - b0[1];
      const int z0 = (int)fminf(fmaxf(fz, -2.f), (float)a.IZ) - b0[2];
      const int xn = (int)fminf(fmaxf(rintf(cx[j]), -1.f), (float)a.IX) - b0[0];
      const int yn = (int)fminf(fmaxf(rintf(cy[j]), -1.f), (float)a.IY) - b0[1];
      const int zn = (int)fminf(fmaxf(rintf(cz[j]), -1.f), (float)a.IZ) - b0[2];
      // Voxels outside the output read pairs from element 0 (the box edges
      // are at least 2 long, so they stay in the box).
      const int base = in[j] ? z0 * dz + y0 * dy + x0 : 0;
#pragma unroll
      for (int q = 0; q < 4; ++q) x.pair[q] = 4u * kLdsStride * (uint32_t)(base + (q & 1) * dy + (q >> 1) * dz);
      x.near = in[j] ? 4u * kLdsStride * (uint32_t)(zn * dz + yn * dy + xn) : 0u;
      x.out = out[j];
    }
    // Source byte offsets of this thread's share of the box (kOutside where
    // the box leaves the volume: those elements load as 0), the same for every
    // channel.  Element e = s * 256 + tid is (ex, ey, ez), carried without
    // divisions as e steps by the work-group size.
    uint32_t goff[kStage];
    {
      const int bxy = bn[0] * bn[1];
      int ez = tid / bxy;
      int ey = (tid - ez * bxy) / bn[0];
      int ex = tid - ez * bxy - ey * bn[0];
      const int sz = kRotThreads / bxy;
      const int sy = (kRotThreads - sz * bxy) / bn[0];
      const int sx = kRotThreads - sz * bxy - sy * bn[0];
#pragma unroll
      for (int s = 0; s < kStage; ++s) {
        const int e = s * kRotThreads + tid;
        const int gx = b0[0] + ex, gy = b0[1] + ey, gz = b0[2] + ez;
        const bool inside = gx >= 0 && gx < a.IX && gy >= 0 && gy < a.IY && gz >= 0 && gz < a.IZ;
        goff[s] = (e < nbox && inside) ? (uint32_t)((gz * a.IY + gy) * a.IX + gx) * 4u : kOutside;
        ex += sx;
        ey += sy;
        ez += sz;
        if (ex >= bn[0]) {
          ex -= bn[0];
          ++ey;
        }
        if (ey >= bn[1]) {
          ey -= bn[1];
          ++ez;
        }
      }
    }
    StagedLoop st;
    st.cur.u = u0;
    st.cur.comp = 0;
    st.cur.valid = true;
    st.cur.c = chan_at(a, u0, 0, Nin, Nout);
    st.nxt = chan_next(a, st.cur, u1, Nin, Nout);
    st.nxt2 = chan_next(a, st.nxt, u1, Nin, Nout);
#pragma unroll
    for (int j = 0; j < kRotVox; ++j) st.vec[j][0] = st.vec[j][1] = st.vec[j][2] = 0.f;
    stage_load(st.cur.c, Nin, goff, st.sv[0]);
    if (st.nxt.valid) stage_load(st.nxt.c, Nin, goff, st.sv[1]);
    stage_store<0>(stage, st.sv[0]);
    __syncthreads();
    while (staged_step<0>(a, st, vx, goff, u1, stage, Nin, Nout) &&
           staged_step<1>(a, st, vx, goff, u1, stage, Nin, Nout)) {
    }
  } else {
    VoxG vx[kRotVox];
    const int dy = a.IX, dz = a.IX * a.IY;
#pragma unroll
    for (int j = 0; j < kRotVox; ++j) {
      VoxG& x = vx[j];
      const float fx = floorf(cx[j]), fy = floorf(cy[j]), fz = floorf(cz[j]);
      const int x0 = (int)fminf(fmaxf(fx, -2.f), (float)a.IX);
      const int y0 = (int)fminf(fmaxf(fy, -2.f), (float)a.IY);
      const int z0 = (int)fminf(fmaxf(fz, -2.f), (float)a.IZ);
      const float wx[2] = {(fx + 1.f) - cx[j], cx[j] - fx};
      const float wy[2] = {(fy + 1.f) - cy[j], cy[j] - fy};
      const float wz[2] = {(fz + 1.f) - cz[j], cz[j] - fz};
      x.ok = 0;
#pragma unroll
      for (int c = 0; c < 8; ++c) {
        const int ddx = c & 1, ddy = (c >> 1) & 1, ddz = c >> 2;
        x.w[c] = wx[ddx] * wy[ddy] * wz[ddz];
        const int xx = x0 + ddx, yy = y0 + ddy, zz = z0 + ddz;
        const bool ok = in[j] && xx >= 0 && xx < a.IX && yy >= 0 && yy < a.IY && zz >= 0 && zz < a.IZ;
        x.ok |= (unsigned)ok << c;
      }
      x.base = z0 * dz + y0 * dy + x0;
      const int xn = (int)fminf(fmaxf(rintf(cx[j]), -1.f), (float)a.IX);
      const int yn = (int)fminf(fmaxf(rintf(cy[j]), -1.f), (float)a.IY);
      const int zn = (int)fminf(fmaxf(rintf(cz[j]), -1.f), (float)a.IZ);
      x.nok = in[j] && xn >= 0 && xn < a.IX && yn >= 0 && yn < a.IY && zn >= 0 && zn < a.IZ;
      x.nidx = x.nok ? zn * dz + yn * dy + xn : 0;
      x.out = out[j];
    }
    global_channels(a, vx, u0, u1);
  }
}

static bool ranges_overlap(const void* a, size_t na, const void* b, size_t nb) {
  const uintptr_t a0 = (uintptr_t)a, b0 = (uintptr_t)b;
  return a0 < b0 + nb && b0 < a0 + na;
}

}  // namespace svs

using namespace svs;

extern "C" int svs_spatial_augment(const svs_aug_volume* volumes, int nvol, int IZ, int IY, int IX, int OZ, int OY,
                                   int OX, const float* R, const float* grid, const float* axes, int padding,
                                   int align_corners, int flips, int grid_rounding, int vector_rounding,
                                   void* stream) {
  SVS_REQUIRE(volumes, "spatial augment: null volume list");
  SVS_REQUIRE(nvol >= 1 && nvol <= kAugMaxVolumes, "spatial augment: %d volumes (1..%d per launch)", nvol,
              kAugMaxVolumes);
  SVS_REQUIRE_DIMS(IZ, IY, IX);
  SVS_REQUIRE_DIMS(OZ, OY, OX);
  SVS_REQUIRE(padding >= 0 && padding <= 2, "spatial augment: padding %d (0 zeros, 1 border, 2 reflection)",
              padding);
  SVS_REQUIRE(flips >= 0 && flips <= 7, "spatial augment: flip mask %d (bits X=1, Y=2, Z=4)", flips);
  SVS_REQUIRE(grid_rounding >= 0 && grid_rounding <= 2 && vector_rounding >= 0 && vector_rounding <= 2,
              "spatial augment: rounding modes %d, %d (0 fma chain, 1 plain, 2 plain from +0)", grid_rounding,
              vector_rounding);
  if (R) {
    SVS_REQUIRE(grid || axes, "spatial augment: a rotation needs the grid or the identity grid's axes");
    // Byte offsets within a source channel and within an output vector stay
    // below 2^31 (offsets from 2^31 up mark elements outside the buffers).
    SVS_REQUIRE((int64_t)IZ * IY * IX * 4 < ((int64_t)1 << 31) && (int64_t)OZ * OY * OX * 12 < ((int64_t)1 << 31),
                "spatial augment: volumes above 2^29 source / 2^27 output voxels are not supported");
  } else {
    SVS_REQUIRE(OZ == IZ && OY == IY && OX == IX,
                "spatial augment: without a rotation the output size must equal the input size");
  }
  AugArgs a{};
  a.nvol = nvol;
  a.IZ = IZ; a.IY = IY; a.IX = IX;
  a.OZ = OZ; a.OY = OY; a.OX = OX;
  a.padding = padding;
  a.align = align_corners != 0;
  a.flips = flips;
  a.grid_rounding = grid_rounding;
  a.vector_rounding = vector_rounding;
  if (R) std::copy(R, R + 9, a.R);
  a.grid = grid;
  a.axes = axes;
  const size_t Nin = (size_t)IZ * IY * IX, Nout = (size_t)OZ * OY * OX;
  int64_t units = 0, channels = 0;
  for (int v = 0; v < nvol; ++v) {
    const svs_aug_volume& s = volumes[v];
    SVS_REQUIRE(s.src && s.dst, "spatial augment: volume %d has a null pointer", v);
    SVS_REQUIRE(s.batch >= 1 && s.channels >= 1, "spatial augment: volume %d has batch %d, channels %d", v,
                s.batch, s.channels);
    SVS_REQUIRE(s.mode == 0 || s.mode == 1, "spatial augment: volume %d mode %d (0 bilinear, 1 nearest)", v,
                s.mode);
    SVS_REQUIRE(!s.vector || s.channels == 3, "spatial augment: vector volume %d has %d channels (3 expected)", v,
                s.channels);
    const size_t ndst = (size_t)s.batch * s.channels * Nout * sizeof(float);
    for (int w = 0; w < nvol; ++w) {
      const size_t nw = (size_t)volumes[w].batch * volumes[w].channels * Nin * sizeof(float);
      SVS_REQUIRE(!ranges_overlap(s.dst, ndst, volumes[w].src, nw),
                  "spatial augment: output %d overlaps input %d (the pass gathers from other voxels)", v, w);
      if (w < v) {
        const size_t ndw = (size_t)volumes[w].batch * volumes[w].channels * Nout * sizeof(float);
        SVS_REQUIRE(!ranges_overlap(s.dst, ndst, volumes[w].dst, ndw), "spatial augment: outputs %d and %d overlap",
                    w, v);
      }
    }
    a.vol[v] = AugVolume{s.src, s.dst, s.batch, s.channels, s.mode, s.vector};
    a.unit_start[v] = (int)units;
    units += R ? (int64_t)s.batch * (s.vector ? 1 : s.channels) : (int64_t)s.batch;
    channels += (int64_t)s.batch * s.channels;
  }
  SVS_REQUIRE(units < ((int64_t)1 << 30), "spatial augment: too many channels");
  a.unit_start[nvol] = (int)units;
  hipStream_t st = (hipStream_t)stream;
  if (!R) {
    a.tiles_x = (OX + kFlipTX - 1) / kFlipTX;
    a.tiles_y = (OY + kFlipTY - 1) / kFlipTY;
    const int64_t per_item = (int64_t)a.tiles_x * a.tiles_y * OZ;
    const int64_t total = per_item * units;
    SVS_REQUIRE(total < ((int64_t)1 << 30), "spatial augment: %lld tiles exceed the launch limit", (long long)total);
    a.ntiles = (int)per_item;
    a.total = (int)total;
    k_spatial_flip<<<(unsigned)total, kFlipTX * kFlipTY, 0, st>>>(a);
  } else {
    a.tiles_x = (OX + kRotEdge - 1) / kRotEdge;
    a.tiles_y = (OY + kRotEdge - 1) / kRotEdge;
    const int64_t ntiles = (int64_t)a.tiles_x * a.tiles_y * ((OZ + kRotEdge - 1) / kRotEdge);
    // About 8 channels per work-group: enough to spread the per-tile set-up,
    // few enough that work-groups on neighbouring tiles stay within a few
    // channels of each other (measured on 80^3: 4, 8, 16, 32 and all 241
    // channels per work-group take 0.90, 0.85, 0.87, 0.98 and 1.01 ms).
    const int64_t upg = std::max<int64_t>(1, (units * kRotChannelsPerGroup + channels / 2) / channels);
    const int64_t groups = (units + upg - 1) / upg;
    const int64_t total = ntiles * groups;
    SVS_REQUIRE(total < ((int64_t)1 << 30), "spatial augment: %lld work-groups exceed the launch limit",
                (long long)total);
    a.ntiles = (int)ntiles;
    a.units_per_group = (int)upg;
    a.total = (int)total;
    k_spatial_rotate<<<(unsigned)total, kRotThreads, 0, st>>>(a);
  }
  SVS_LAUNCHED();
  return SVS_STATUS_OK;
}
