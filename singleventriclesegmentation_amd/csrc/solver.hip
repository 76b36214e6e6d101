// The whole TV-L1 solve on the device: pyramid, per-warp linearisation and
// the fused primal-dual sweep, for one frame pair or a batch of frame pairs
// of uneven sizes (tvl13d.py:69-255 is the algorithm restated here).
//
// Every kernel reads its per-instance geometry and buffer addresses from a
// device-side descriptor table (LevelDesc).  A captured hipGraph therefore
// holds only table addresses and per-iteration scalars; binding a plan to new
// inputs or outputs rewrites the table and keeps the graph.
//
// State layout in HBM, per instance and pyramid level (N voxels, x fastest):
//   I0, I1     float[N]            restricted (level 0: blurred) frames
//   J          float4[N]           (I1, dI1/dx, dI1/dy, dI1/dz)  gather source
//   SN         float4[N]           (s, n_x, n_y, n_z) anisotropic weight+normal
//   RG         float4[N]           (rho_c, g_x, g_y, g_z) from the warp
//   u[2]       float[3][N]  SoA    flow, ping-pong
//   p[2]       float[9][N]  SoA    dual, p[3c+d] = row c, derivative d
//   z[2]       float[3][N]  SoA    over-relaxed point (extended model only)
// J, SN and RG are shared by the levels of an instance (levels run in turn).
#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "svs_common.hpp"
#include "svs_device.hpp"
#include "svs_math.h"
#include "svsflow.h"

namespace svs {

int launch_fill(float* p, int64_t n, float v, hipStream_t s);

struct alignas(16) LevelDesc {
  int NZ, NY, NX, pad0;
  float hZ, hY, hX, pad1;
  float LZ, LY, LX, pad2;
  float upZ, upY, upX, pad3;  // coarse -> this level flow scale L_fine / L_coarse
  int64_t N;
  float* I0;
  float* I1;
  float4* J;
  float4* SN;
  float4* RG;
  float* u[2];
  float* p[2];
  float* z[2];  // extended model only (see k_sweep's EXT): the dual step's point, ping-pong
  // Level 0 only: caller-bound inputs and outputs.
  const float* src0;
  const float* src1;
  const float* u_in;
  const float* p_in;
  float* u_out;
  float* p_out;
};

constexpr int kBlock = 256;
constexpr int kXcds = 8;  // MI355X: 8 XCDs, work-groups dealt round-robin

// Taps in flight per tricubic gather (the LOOP level of sample_cubic_impl)
// for the per-warp linearisation, the restriction and the prolongation.
// Level 1 (one z-plane of taps at a time) takes the linearisation from 233
// to 97 VGPRs (2 -> 4 waves per SIMD) and the prolongation from 234 to 121;
// the 68 x 80^3 batch solves in 614 ms instead of 627 ms (level 2: 617 ms).
#ifndef SVS_LIN_LOOP
#define SVS_LIN_LOOP 1
#endif
#ifndef SVS_RESTRICT_LOOP
#define SVS_RESTRICT_LOOP 1
#endif
#ifndef SVS_PROLONG_LOOP
#define SVS_PROLONG_LOOP 1
#endif

// Maps a block of a flattened multi-instance launch to (instance, voxel).
__device__ __forceinline__ int find_instance(const int* __restrict__ block_start, int ninst, int blk) {
  int lo = 0, hi = ninst - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (block_start[mid] <= blk) lo = mid; else hi = mid - 1;
  }
  return lo;
}

#define SVS_BATCH_PROLOGUE(descs)                                                  \
  const int inst_ = find_instance(block_start, ninst, (int)blockIdx.x);            \
  const LevelDesc& d = descs[inst_];                                               \
  const int64_t i = (int64_t)(blockIdx.x - block_start[inst_]) * kBlock + threadIdx.x; \
  if (i >= d.N) return;                                                            \
  int z, y, x;                                                                     \
  unflatten(i, d.NY, d.NX, z, y, x)

// ---------------------------------------------------------- level set-up
// Gaussian pre-blur of both frames of every instance (flow_utils.py:94-112).
__global__ __launch_bounds__(kBlock) void k_blur_pair(const LevelDesc* __restrict__ descs, const int* __restrict__ block_start, int ninst,
                            const float* __restrict__ w, int K) {
  SVS_BATCH_PROLOGUE(descs);
  const int P = K / 2;
  float acc0 = 0.0f, acc1 = 0.0f;
  for (int a = 0; a < K; ++a) {
    const int zz = z + a - P;
    if (zz < 0 || zz >= d.NZ) continue;
    for (int b = 0; b < K; ++b) {
      const int yy = y + b - P;
      if (yy < 0 || yy >= d.NY) continue;
      const int64_t row = ((int64_t)zz * d.NY + yy) * d.NX;
      const float* wr = w + (a * K + b) * K;
      for (int c = 0; c < K; ++c) {
        const int xx = x + c - P;
        if (xx < 0 || xx >= d.NX) continue;
        acc0 = acc0 + wr[c] * d.src0[row + xx];
        acc1 = acc1 + wr[c] * d.src1[row + xx];
      }
    }
  }
  d.I0[i] = acc0;
  d.I1[i] = acc1;
}

struct PlaneFetch {
  const float* v;
  int NY, NX;
  __device__ void operator()(int z, int y, int x, float* o) const { o[0] = v[((int64_t)z * NY + y) * NX + x]; }
};

// Pyramid restriction fine -> coarse by resampling (tvl13d.py:108-115,
// prolongationOps.cu:533-559); both frames share the sample geometry.
__global__ __launch_bounds__(kBlock) void k_restrict(const LevelDesc* __restrict__ fine, const LevelDesc* __restrict__ coarse,
                           const int* __restrict__ block_start, int ninst, int ip, int bc) {
  SVS_BATCH_PROLOGUE(coarse);
  const LevelDesc& f = fine[inst_];
  const float cxp = (float)x * d.hX;
  const float cx = cxp * f.LX / d.LX;
  const float cyp = (float)y * d.hY;
  const float cy = cyp * f.LY / d.LY;
  const float czp = (float)z * d.hZ;
  const float cz = czp * f.LZ / d.LZ;
  float v0[1], v1[1];
  sample<1, PlaneFetch, SVS_RESTRICT_LOOP>(ip, PlaneFetch{f.I0, f.NY, f.NX}, f.NZ, f.NY, f.NX, f.hZ, f.hY, f.hX, bc,
                                           cz, cy, cx, v0);
  sample<1, PlaneFetch, SVS_RESTRICT_LOOP>(ip, PlaneFetch{f.I1, f.NY, f.NX}, f.NZ, f.NY, f.NX, f.hZ, f.hY, f.hX, bc,
                                           cz, cy, cx, v1);
  d.I0[i] = v0[0];
  d.I1[i] = v1[0];
}

// Per-level constants (tvl13d.py:179-192): G = grad I1 (Nabla3D_CD.forward)
// packed with I1 as the gather source J, and the anisotropic weight and unit
// normal (anisotropicDifferentialOps.cu:90-141) as SN.
__global__ __launch_bounds__(kBlock) void k_level_setup(const LevelDesc* __restrict__ descs, const int* __restrict__ block_start, int ninst,
                              int aniso, float alpha, float beta, int bc) {
  SVS_BATCH_PROLOGUE(descs);
  const float* I1 = d.I1;
  const int64_t sy = d.NX, sz = (int64_t)d.NY * d.NX;
  const float b0 = I1[i];
  const float gx = cd_fwd(x > 0 ? I1[i - 1] : 0.0f, b0, x < d.NX - 1 ? I1[i + 1] : 0.0f, x, d.NX, d.hX, bc);
  const float gy = cd_fwd(y > 0 ? I1[i - sy] : 0.0f, b0, y < d.NY - 1 ? I1[i + sy] : 0.0f, y, d.NY, d.hY, bc);
  const float gz = cd_fwd(z > 0 ? I1[i - sz] : 0.0f, b0, z < d.NZ - 1 ? I1[i + sz] : 0.0f, z, d.NZ, d.hZ, bc);
  d.J[i] = make_float4(b0, gx, gy, gz);
  if (aniso) {
    const float g[3] = {gx, gy, gz};
    float ns = 0.0f;
    for (int k = 0; k < 3; ++k) ns += g[k] * g[k];
    const float nrm = sqrtf(ns);
    float4 sn = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (nrm > 0.0f) sn = make_float4(svs_expf(-alpha * svs_powf(nrm, beta)), gx / nrm, gy / nrm, gz / nrm);
    d.SN[i] = sn;
  }
}

// Zero start (tvl13d.py:116-118) or caller-provided u, p (used when the
// pyramid has a single level, tvl13d.py:104-105) in SoA layout.
__global__ __launch_bounds__(kBlock) void k_init_state(const LevelDesc* __restrict__ descs, const int* __restrict__ block_start, int ninst,
                             int from_input) {
  SVS_BATCH_PROLOGUE(descs);
  (void)z; (void)y; (void)x;
  const int64_t N = d.N;
  for (int c = 0; c < 3; ++c) d.u[0][c * N + i] = from_input ? d.u_in[3 * i + c] : 0.0f;
  for (int k = 0; k < 9; ++k) d.p[0][k * N + i] = from_input ? d.p_in[9 * i + k] : 0.0f;
}

// The linearisation's taps of J through a buffer descriptor: 32-bit byte
// offsets (16 N < 2^32 by the plan's size limit), whose plane and row terms
// the compiler keeps per z-plane and per row, so a tap costs one add instead
// of a 64-bit index and address computation.
struct Float4BufFetch {
  svs_rsrc r;
  uint32_t sy, sz;  // bytes per row and per plane
  __device__ void operator()(int z, int y, int x, float* o) const {
    const svs_f4 t = bld4(r, (uint32_t)z * sz + (uint32_t)y * sy + (uint32_t)x * 16u, 0);
    o[0] = t.x; o[1] = t.y; o[2] = t.z; o[3] = t.w;
  }
};

// Linearisation of the data term at the current flow (tvl13d.py:198-201):
// I1w = warp(I1, u), g = warp(grad I1, u), rho_c = I1w - sum(u * g) - I0.
// The four fields share the 64-tap geometry; each is evaluated exactly as the
// scalar / vector-field warps do it.
// Five waves per SIMD (85 VGPRs, no spills; six waves at 80 VGPRs are no
// faster).  The arithmetic bounds it: the 84 Catmull-Rom evaluations of a
// voxel (21 per field, 21 VALU instructions each) are about three quarters
// of its ~2,300 VALU instructions per wave.
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(5))) void k_linearise(const LevelDesc* __restrict__ descs, const int* __restrict__ block_start, int ninst,
                            int parity, int ip, int bc) {
  SVS_BATCH_PROLOGUE(descs);
  const int64_t N = d.N;
  const float* u = d.u[parity];
  const float u0 = u[i], u1 = u[N + i], u2 = u[2 * N + i];
  const float cx = (float)x * d.hX + u0;
  const float cy = (float)y * d.hY + u1;
  const float cz = (float)z * d.hZ + u2;
  float v[4];
  const uint32_t sy = (uint32_t)d.NX * 16u;
  const Float4BufFetch J{make_rsrc(d.J, (uint32_t)N * 16u), sy, (uint32_t)d.NY * sy};
  sample<4, Float4BufFetch, SVS_LIN_LOOP>(ip, J, d.NZ, d.NY, d.NX, d.hZ, d.hY, d.hX, bc, cz, cy, cx, v);
  const float s = (u0 * v[1] + u2 * v[3]) + u1 * v[2];  // torch's sum over dim 3: (a + c) + b
  const float rc = (v[0] - s) - d.I0[i];
  d.RG[i] = make_float4(rc, v[1], v[2], v[3]);
}

struct SoAFetch3 {
  const float* v;
  int64_t N;
  int NY, NX;
  __device__ void operator()(int z, int y, int x, float* o) const {
    const int64_t k = ((int64_t)z * NY + y) * NX + x;
    o[0] = v[k]; o[1] = v[N + k]; o[2] = v[2 * N + k];
  }
};

// Coarse -> fine transfer (tvl13d.py:145-158): u by vector-field resampling
// then component c scaled by L_c,fine / L_c,coarse; p by matrix-field
// resampling, unscaled.
__global__ __launch_bounds__(kBlock) void k_prolong_state(const LevelDesc* __restrict__ coarse,
                                                          const LevelDesc* __restrict__ fine,
                                                          const int* __restrict__ block_start, int ninst,
                                                          int parity_c, int ip, int bc) {
  SVS_BATCH_PROLOGUE(fine);
  const LevelDesc& c = coarse[inst_];
  const float cxp = (float)x * d.hX;
  const float cx = cxp * c.LX / d.LX;
  const float cyp = (float)y * d.hY;
  const float cy = cyp * c.LY / d.LY;
  const float czp = (float)z * d.hZ;
  const float cz = czp * c.LZ / d.LZ;
  const int64_t Nf = d.N;
  float v[3];
  sample<3, SoAFetch3, SVS_PROLONG_LOOP>(ip, SoAFetch3{c.u[parity_c], c.N, c.NY, c.NX}, c.NZ, c.NY, c.NX, c.hZ, c.hY,
                                         c.hX, bc, cz, cy, cx, v);
  d.u[0][i] = v[0] * d.upX;
  d.u[0][Nf + i] = v[1] * d.upY;
  d.u[0][2 * Nf + i] = v[2] * d.upZ;
  for (int r = 0; r < 3; ++r) {
    sample<3, SoAFetch3, SVS_PROLONG_LOOP>(ip, SoAFetch3{c.p[parity_c] + 3 * r * c.N, c.N, c.NY, c.NX}, c.NZ, c.NY,
                                           c.NX, c.hZ, c.hY, c.hX, bc,
              cz, cy, cx, v);
    for (int k = 0; k < 3; ++k) d.p[0][(3 * r + k) * Nf + i] = v[k];
  }
}

// Tricubic coarse -> fine transfer in three separable passes.  The per-voxel
// gather above evaluates, for every fine voxel, an x-pass on 16 coarse rows,
// a y-pass per coarse plane and a z-pass.  The x-pass value of coarse row
// (z', y') at fine x depends on nothing else, nor does the y-pass value of
// coarse plane z' at fine (y, x), so computing each once and storing it gives
// the same evaluations on the same values in the same order: the same bits,
// for 12 instead of 64 taps per field and fine voxel.  cr_local is the
// evaluation sample_cubic uses for any lane whose segment rounds up, and for
// the others it gives the same bits as the select-free form.
//   x-pass -> T1[f][z'][y'][x]  in the instance's J/SN/RG block (free here)
//   y-pass -> T2[f][z'][y][x]   in the fine level's second dual buffer
//   z-pass -> u, p of the fine level, u scaled per component
constexpr int kProlongFields = 12;   // u (3 components) and p (9)

__device__ __forceinline__ const float* state_field(const LevelDesc& c, int parity, int f) {
  return f < 3 ? c.u[parity] + (int64_t)f * c.N : c.p[parity] + (int64_t)(f - 3) * c.N;
}

__global__ __launch_bounds__(kBlock) void k_prolong_x(const LevelDesc* __restrict__ coarse,
                                                      const LevelDesc* __restrict__ fine,
                                                      const int* __restrict__ block_start, int ninst, int parity_c,
                                                      int bc) {
  const int inst = find_instance(block_start, ninst, (int)blockIdx.x);
  const LevelDesc& c = coarse[inst];
  const LevelDesc& d = fine[inst];
  const int64_t P1 = (int64_t)c.NZ * c.NY * d.NX;
  const int64_t i = (int64_t)(blockIdx.x - block_start[inst]) * kBlock + threadIdx.x;
  if (i >= P1) return;
  const int x = (int)(i % d.NX);
  const int64_t row = (i / d.NX) * c.NX;   // coarse row (z', y')
  const float cxp = (float)x * d.hX;
  const float cx = cxp * c.LX / d.LX;
  const Axis ax = make_axis(cx, c.hX);
  int xs[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) xs[k] = bidx(ax.f - 1 + k, c.NX, bc);
  const float lx = ax.w + 1;
  float* T1 = reinterpret_cast<float*>(d.J);
#pragma unroll 4
  for (int f = 0; f < kProlongFields; ++f) {
    const float* s = state_field(c, parity_c, f) + row;
    T1[f * P1 + i] = cr_local(s[xs[0]], s[xs[1]], s[xs[2]], s[xs[3]], lx);
  }
}

__global__ __launch_bounds__(kBlock) void k_prolong_y(const LevelDesc* __restrict__ coarse,
                                                      const LevelDesc* __restrict__ fine,
                                                      const int* __restrict__ block_start, int ninst, int bc) {
  const int inst = find_instance(block_start, ninst, (int)blockIdx.x);
  const LevelDesc& c = coarse[inst];
  const LevelDesc& d = fine[inst];
  const int64_t P1 = (int64_t)c.NZ * c.NY * d.NX;
  const int64_t P2 = (int64_t)c.NZ * d.NY * d.NX;
  const int64_t i = (int64_t)(blockIdx.x - block_start[inst]) * kBlock + threadIdx.x;
  if (i >= P2) return;
  const int x = (int)(i % d.NX);
  const int64_t r = i / d.NX;
  const int y = (int)(r % d.NY);
  const int zc = (int)(r / d.NY);
  const float cyp = (float)y * d.hY;
  const float cy = cyp * c.LY / d.LY;
  const Axis ay = make_axis(cy, c.hY);
  int64_t t[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) t[k] = ((int64_t)zc * c.NY + bidx(ay.f - 1 + k, c.NY, bc)) * d.NX + x;
  const float ly = ay.w + 1;
  const float* T1 = reinterpret_cast<const float*>(d.J);
  float* T2 = d.p[1];
#pragma unroll 4
  for (int f = 0; f < kProlongFields; ++f) {
    const float* s = T1 + f * P1;
    T2[f * P2 + i] = cr_local(s[t[0]], s[t[1]], s[t[2]], s[t[3]], ly);
  }
}

__global__ __launch_bounds__(kBlock) void k_prolong_z(const LevelDesc* __restrict__ coarse,
                                                      const LevelDesc* __restrict__ fine,
                                                      const int* __restrict__ block_start, int ninst, int bc) {
  SVS_BATCH_PROLOGUE(fine);
  (void)y; (void)x;
  const LevelDesc& c = coarse[inst_];
  const int64_t Nf = d.N;
  const int64_t P2 = (int64_t)c.NZ * d.NY * d.NX;
  const int64_t plane = (int64_t)d.NY * d.NX;
  const int64_t yx = i - (int64_t)z * plane;
  const float czp = (float)z * d.hZ;
  const float cz = czp * c.LZ / d.LZ;
  const Axis az = make_axis(cz, c.hZ);
  int64_t t[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) t[k] = (int64_t)bidx(az.f - 1 + k, c.NZ, bc) * plane + yx;
  const float lz = az.w + 1;
  const float* T2 = d.p[1];
  const float up[3] = {d.upX, d.upY, d.upZ};
#pragma unroll
  for (int f = 0; f < 3; ++f) {
    const float* s = T2 + f * P2;
    d.u[0][f * Nf + i] = cr_local(s[t[0]], s[t[1]], s[t[2]], s[t[3]], lz) * up[f];
  }
#pragma unroll 3
  for (int f = 3; f < kProlongFields; ++f) {
    const float* s = T2 + f * P2;
    d.p[0][(f - 3) * Nf + i] = cr_local(s[t[0]], s[t[1]], s[t[2]], s[t[3]], lz);
  }
}

// SoA state -> the caller's [Z,Y,X,3] flow and [Z,Y,X,3,3] dual.
__global__ __launch_bounds__(kBlock) void k_store_output(const LevelDesc* __restrict__ descs, const int* __restrict__ block_start, int ninst,
                               int parity) {
  SVS_BATCH_PROLOGUE(descs);
  (void)z; (void)y; (void)x;
  const int64_t N = d.N;
  if (d.u_out)
    for (int c = 0; c < 3; ++c) d.u_out[3 * i + c] = d.u[parity][c * N + i];
  if (d.p_out)
    for (int k = 0; k < 9; ++k) d.p_out[9 * i + k] = d.p[parity][k * N + i];
}

// ------------------------------------------------------ fused PD sweep
// One primal-dual iteration (tvl13d.py:213-253) in a single pass:
//   p' = proj_{|.|<=w_tv}(p + sigma * D grad u)                (dual ascent + prox)
//   u' = prox_L1(u - tau * grad^T (D p'))  with rho at u      (primal descent + prox)
// (the reference sets z = u after every step, so the dual uses grad u).
//
// A work-group owns an (x, y) tile of BX x BY columns including a one-voxel
// ring and marches along z over [z0, z1).  At plane k it computes p' and
// r = D p' on the whole extended tile (the ring is recomputed by the
// neighbouring tile as well), then the primal update of plane k-1 from r of
// planes k-2..k.  LDS holds, double-buffered, u of plane k with a second ring
// (for the x/y derivatives of the ring) and the x/y derivative columns of r;
// a thread's u of plane k-1 is read back from its own slot of the other U
// buffer, u of planes k and k+1 and the z-column of r live in its registers.
// One barrier per plane.
//
// A work-group's tile: instance, origin, z-range and its own shape, so that
// one launch can give every instance of an uneven batch full-row tiles.
struct alignas(16) SweepTile {
  int inst;
  int xy;     // x0 | (y0 << 16): first interior voxel
  int zr;     // z0 | (z1 << 16)
  int shape;  // bx | (by << 16): work-group extent of this tile, ring included
};

// UNIT_H: every instance of the launch has h = 1 on all three axes (always
// so for numDofs meshes, h = (N-1)/(N-1)).  The divisions by h then drop out
// at compile time; x / 1.0f == x, so the results are the same bits.
// STATS: the work-group also sums, over the voxels it updates, the four
// per-iteration statistics terms (squared change of u, squared change of p,
// |rho(u^n)|, and the l2 norms of the rows of D grad u^n) and writes them to
// part[blockIdx.x] (see k_stats_reduce); u and p are computed as without.
// EXT, the extended model (neither part is in the reference): the dual rows
// use the Huber-TV step with factor hub (prox_dual_row) and the energy the
// Huber function with eps; the dual step differentiates the over-relaxed
// point z^n = u^n + theta (u^n - u^{n-1}) instead of u^n.  z lives in its own
// ping-pong field: the sweep stages z (u itself when zfirst, a warp's first
// iteration) through LDS for the gradient, reads this column's u^n for the
// primal step one plane ahead, and writes z' = u' + theta (u' - u^n) beside u'.
// Without EXT, hub = 1, eps = 0 and z = u are compile-time facts, so the
// reference's iteration costs nothing extra.
template <int BC, bool ANISO, bool UNIT_H, bool STATS, bool EXT>
__global__ __launch_bounds__(1024) void k_sweep(const LevelDesc* __restrict__ descs,
                                                const SweepTile* __restrict__ tiles, int parity, float sigma,
                                                float tau, float factor, float wtv, float hub, float eps,
                                                float theta, int zfirst, float4* __restrict__ part) {
  extern __shared__ float lds[];
  const SweepTile t = tiles[blockIdx.x];
  const LevelDesc& d = descs[t.inst];
  const int NZ = d.NZ, NY = d.NY, NX = d.NX;
  const int64_t N = d.N;
  const float hZ = UNIT_H ? 1.0f : d.hZ, hY = UNIT_H ? 1.0f : d.hY, hX = UNIT_H ? 1.0f : d.hX;
  // The work-group is one-dimensional and sized for the largest tile of the
  // launch; threads past this tile's BX * BY only take part in the barriers.
  const int BX = t.shape & 0xffff, BY = t.shape >> 16;
  const int tid = threadIdx.x;
  const bool active = tid < BX * BY;
  const int ex = tid % BX, ey = tid / BX;
  const int x = (t.xy & 0xffff) - 1 + ex, y = (t.xy >> 16) - 1 + ey;
  const int z0 = t.zr & 0xffff, z1 = t.zr >> 16;
  const bool valid = active && x >= 0 && x < NX && y >= 0 && y < NY;
  const bool interior = valid && ex >= 1 && ex <= BX - 2 && ey >= 1 && ey <= BY - 2;

  const int UW = BX + 2, UH = BY + 2, UPL = UW * UH;
  float* U = lds;                 // [2][3][UPL]
  float* R = lds + 6 * UPL;       // [2][6][BX * BY]
  const int RPL = BX * BY;
  const int uo = (ey + 1) * UW + (ex + 1);
  const int ro = tid;

  // Buffer addressing: this thread's in-plane byte offset in a VGPR, the
  // field and plane offsets in SGPRs (the host keeps 36*N < 2^32).
  const uint32_t F4 = (uint32_t)N * 4u;              // bytes of one scalar field
  const uint32_t P4 = (uint32_t)(NY * NX) * 4u;      // bytes of one z-plane of it
  const uint32_t col = valid ? (uint32_t)(y * NX + x) : 0u;
  const uint32_t vo = col * 4u, vo4 = col * 16u;
  const svs_rsrc uin = make_rsrc(d.u[parity], 3u * F4), uout = make_rsrc(d.u[parity ^ 1], 3u * F4);
  // The field the dual step differentiates (and stages with its ring).
  const svs_rsrc zin = (EXT && !zfirst) ? make_rsrc(d.z[parity], 3u * F4) : uin;
  const svs_rsrc pin = make_rsrc(d.p[parity], 9u * F4), pout = make_rsrc(d.p[parity ^ 1], 9u * F4);
  const svs_rsrc SNr = make_rsrc(d.SN, 4u * F4), RGr = make_rsrc(d.RG, 4u * F4);

  // This thread's position on the outer ring of the u tile (one per thread;
  // the host guarantees 2*UW + 2*BY <= BX*BY).
  const int nring = 2 * UW + 2 * BY;
  // Only waves holding ring lanes issue the ring loads, behind a wave-uniform
  // branch (a divergent one made the compiler merge the loads' wave offsets
  // into VGPRs and read them back lane by lane; loads issued by every wave
  // cost the sweep 3 % in address processing).  Within those waves, lanes off
  // the ring, or with a ring position outside the volume, load from past the
  // buffer and get 0 without a memory access.
  const bool ring = tid < nring;
  const bool ring_wave = __builtin_amdgcn_readfirstlane(tid >> 6) * 64 < nring;
  int ring_lds = 0;
  uint32_t ring_vo = kOutside;
  if (ring) {
    int ux, uy;
    if (tid < 2 * UW) {
      uy = tid < UW ? 0 : UH - 1;
      ux = tid < UW ? tid : tid - UW;
    } else {
      const int rr = tid - 2 * UW;
      uy = 1 + (rr >> 1);
      ux = (rr & 1) ? UW - 1 : 0;
    }
    const int gx = (t.xy & 0xffff) - 2 + ux, gy = (t.xy >> 16) - 2 + uy;
    if (gx >= 0 && gx < NX && gy >= 0 && gy < NY) ring_vo = (uint32_t)(gy * NX + gx) * 4u;
    ring_lds = uy * UW + ux;
  }

  const int kstart = z0 > 0 ? z0 - 1 : 0;
  const int kend = z1 < NZ ? z1 : NZ;  // the step k == NZ only finishes plane NZ-1

  // Registers: own-column u (z in the extended model) of planes k-1..k+1 (k+2
  // in flight; k-1 is refreshed from LDS each step), the dual and anisotropy of plane k (k+1 in
  // flight), rho_c/g of plane k-1 (k in flight), the ring value of plane k
  // (k+1 in flight) and the z-derivative column of r = D p' for planes k-2..k.
  float um[3] = {0.0f, 0.0f, 0.0f}, uc[3] = {0.0f, 0.0f, 0.0f}, up[3] = {0.0f, 0.0f, 0.0f};
  float pc[9], snc[4] = {0.0f, 0.0f, 0.0f, 0.0f}, rgc[4] = {0.0f, 0.0f, 0.0f, 0.0f}, ringc[3] = {0.0f, 0.0f, 0.0f};
  for (int e = 0; e < 9; ++e) pc[e] = 0.0f;
  if (valid) {
    if (kstart >= 1)
      for (int c = 0; c < 3; ++c) um[c] = bld(zin, vo, c * F4 + (kstart - 1) * P4);
    for (int c = 0; c < 3; ++c) uc[c] = bld(zin, vo, c * F4 + kstart * P4);
    if (kstart + 1 < NZ)
      for (int c = 0; c < 3; ++c) up[c] = bld(zin, vo, c * F4 + (kstart + 1) * P4);
    for (int e = 0; e < 9; ++e) pc[e] = bld(pin, vo, e * F4 + kstart * P4);
    if (ANISO) {
      const svs_f4 sn = bld4(SNr, vo4, 4u * kstart * P4);
      snc[0] = sn.x; snc[1] = sn.y; snc[2] = sn.z; snc[3] = sn.w;
    }
  }
  if (ring_wave)
    for (int c = 0; c < 3; ++c) ringc[c] = bld(zin, ring_vo, c * F4 + kstart * P4);
  float r2a[3] = {0.0f, 0.0f, 0.0f}, r2b[3] = {0.0f, 0.0f, 0.0f}, r2c[3];
  // EXT: this column's u^n of plane k-1 for the primal step (the staged field
  // is z), loaded one step ahead.  The first step's primal plane is never in
  // the chunk, so it starts unset.
  float uoc[3] = {0.0f, 0.0f, 0.0f};
  float st_du = 0.0f, st_dp = 0.0f, st_rho = 0.0f, st_tv = 0.0f;  // STATS only
  // Each step reads this thread's u of the previous plane back from its own
  // slot of the other LDS buffer (only this thread writes that slot), so no
  // register carries it across a step; seed that slot for the first step.
  if (active) {
    float* Ub = U + ((kstart + 1) & 1) * 3 * UPL;
    for (int c = 0; c < 3; ++c) Ub[c * UPL + uo] = um[c];
  }

  for (int k = kstart; k <= kend; ++k) {
    const int buf = k & 1;
    if (active) {
      const float* Ubm = U + (buf ^ 1) * 3 * UPL;
      for (int c = 0; c < 3; ++c) um[c] = Ubm[c * UPL + uo];
    }
    // ---- prefetch what step k+1 needs; consumed after this step's barrier.
    // Plane indices are clamped into the volume and lanes outside it read
    // in-plane offset 0, so every load is in bounds and needs no branch: a
    // clamped or outside value only reaches steps that are skipped (k >= NZ,
    // outside z0..z1) or derivative branches the boundary rule does not select
    // (the last plane's upper neighbour, a face's outer neighbour).
    // They are issued after the dual step, which needs the most registers
    // (p and SN after it has consumed plane k's copies, RG after the primal
    // step), so nothing is in flight while it runs; the primal step covers
    // their latency.
    const uint32_t k1 = (uint32_t)min(k + 1, NZ - 1), k2 = (uint32_t)min(k + 2, NZ - 1);
    float un[3], pn[9], snn[4] = {0.0f, 0.0f, 0.0f, 0.0f}, rgn[4] = {0.0f, 0.0f, 0.0f, 0.0f}, ringn[3];

    // ---- u of plane k into LDS
    if (k < NZ && active) {
      float* Ub = U + buf * 3 * UPL;
      for (int c = 0; c < 3; ++c) Ub[c * UPL + uo] = uc[c];
      if (ring)
        for (int c = 0; c < 3; ++c) Ub[c * UPL + ring_lds] = ringc[c];
    }
    __syncthreads();

    // ---- dual step at plane k on the extended tile
    float* Rb = R + buf * 6 * RPL;
    if (valid && k < NZ) {
      const float* Ub = U + buf * 3 * UPL;
      float s = 0.0f, n[3] = {0.0f, 0.0f, 0.0f}, t1[3] = {0.0f, 0.0f, 0.0f};
      if (ANISO) {
        s = snc[0]; n[0] = snc[1]; n[1] = snc[2]; n[2] = snc[3];
        tangent1(n, t1);
      }
      const bool store_p = interior && k >= z0 && k < z1;
      for (int c = 0; c < 3; ++c) {
        const float* Uc = Ub + c * UPL;
        const float b0 = Uc[uo];
        float q[3];
        q[0] = cd_fwd(Uc[uo - 1], b0, Uc[uo + 1], x, NX, hX, BC);
        q[1] = cd_fwd(Uc[uo - UW], b0, Uc[uo + UW], y, NY, hY, BC);
        q[2] = cd_fwd(um[c], b0, up[c], k, NZ, hZ, BC);
        float dq[3];
        if (ANISO) aniso_row(s, n, t1, q, dq);
        else { dq[0] = q[0]; dq[1] = q[1]; dq[2] = q[2]; }
        float pv[3];
        for (int e = 0; e < 3; ++e) {
          const float inc = sigma * dq[e];
          pv[e] = pc[3 * c + e] + inc;
        }
        prox_dual_row(pv, wtv, EXT ? hub : 1.0f);
        if (wtv == 0.0f) { pv[0] = 0.0f; pv[1] = 0.0f; pv[2] = 0.0f; }  // tvl13d.py:223-224
        if (STATS && store_p) {
          st_tv += huber_of_norm(__builtin_amdgcn_sqrtf(dq[0] * dq[0] + dq[1] * dq[1] + dq[2] * dq[2]),
                                 EXT ? eps : 0.0f);
          for (int e = 0; e < 3; ++e) {
            const float dd = pv[e] - pc[3 * c + e];
            st_dp += dd * dd;
          }
        }
        if (store_p)
          for (int e = 0; e < 3; ++e) bst(pout, vo, (3 * c + e) * F4 + k * P4, pv[e]);
        float rr[3];
        if (ANISO) aniso_row(s, n, t1, pv, rr);
        else { rr[0] = pv[0]; rr[1] = pv[1]; rr[2] = pv[2]; }
        Rb[(2 * c) * RPL + ro] = rr[0];
        Rb[(2 * c + 1) * RPL + ro] = rr[1];
        r2c[c] = rr[2];
      }
    } else {
      for (int c = 0; c < 3; ++c) {
        if (active) {
          Rb[(2 * c) * RPL + ro] = 0.0f;
          Rb[(2 * c + 1) * RPL + ro] = 0.0f;
        }
        r2c[c] = 0.0f;
      }
    }

    for (int c = 0; c < 3; ++c) un[c] = bld(zin, vo, c * F4 + k2 * P4);
    for (int c = 0; c < 3; ++c) ringn[c] = 0.0f;
    if (ring_wave)
      for (int c = 0; c < 3; ++c) ringn[c] = bld(zin, ring_vo, c * F4 + k1 * P4);
    for (int e = 0; e < 9; ++e) pn[e] = bld(pin, vo, e * F4 + k1 * P4);
    float uon[3];
    if (EXT)
      for (int c = 0; c < 3; ++c) uon[c] = bld(uin, vo, c * F4 + (uint32_t)min(k, NZ - 1) * P4);

    // ---- primal step at plane k-1 on the tile interior
    const int km = k - 1;
    if (interior && km >= z0 && km < z1) {
      const float* Rp = R + (km & 1) * 6 * RPL;
      const float g[3] = {rgc[1], rgc[2], rgc[3]};
      const float* uold = EXT ? uoc : um;  // u^n of this column at plane k-1
      float v[3];
      for (int c = 0; c < 3; ++c) {
        const float* Rx = Rp + (2 * c) * RPL;
        const float* Ry = Rp + (2 * c + 1) * RPL;
        const float ax = cd_adj(Rx[ro - 1], Rx[ro], Rx[ro + 1], x, NX, BC);
        const float ay = cd_adj(Ry[ro - BX], Ry[ro], Ry[ro + BX], y, NY, BC);
        const float az = cd_adj(r2a[c], r2b[c], r2c[c], km, NZ, BC);
        const float div = ax / hX + ay / hY + az / hZ;
        const float step = tau * div;
        v[c] = uold[c] - step;
      }
      const float sm = (uold[0] * g[0] + uold[2] * g[2]) + uold[1] * g[1];  // (a + c) + b, as torch's sum
      const float rho = rgc[0] + sm;
      prox_primal_point(v, rho, g, factor);
      if (STATS) {
        st_rho += fabsf(rho);
        for (int c = 0; c < 3; ++c) {
          const float dd = v[c] - uold[c];
          st_du += dd * dd;
        }
      }
      for (int c = 0; c < 3; ++c) bst(uout, vo, c * F4 + km * P4, v[c]);
      if (EXT) {
        const svs_rsrc zout = make_rsrc(d.z[parity ^ 1], 3u * F4);
        for (int c = 0; c < 3; ++c) {
          const float dd = v[c] - uold[c];
          const float ee = theta * dd;
          bst(zout, vo, c * F4 + km * P4, v[c] + ee);
        }
      }
    }
    if (ANISO) {
      const svs_f4 sn = bld4(SNr, vo4, 4u * k1 * P4);
      snn[0] = sn.x; snn[1] = sn.y; snn[2] = sn.z; snn[3] = sn.w;
    }
    if (interior) {
      const svs_f4 rg = bld4(RGr, vo4, 4u * (uint32_t)min(k, NZ - 1) * P4);
      rgn[0] = rg.x; rgn[1] = rg.y; rgn[2] = rg.z; rgn[3] = rg.w;
    }

    for (int c = 0; c < 3; ++c) {
      r2a[c] = r2b[c];
      r2b[c] = r2c[c];
      uc[c] = up[c];
      up[c] = un[c];
      ringc[c] = ringn[c];
      if (EXT) uoc[c] = uon[c];
    }
    for (int e = 0; e < 9; ++e) pc[e] = pn[e];
    for (int e = 0; e < 4; ++e) { snc[e] = snn[e]; rgc[e] = rgn[e]; }
  }
  if constexpr (STATS) {
    // Every thread of the work-group gets here (none returns early).
    __shared__ float4 red[16];
    float4 w = make_float4(wave_sum(st_du), wave_sum(st_dp), wave_sum(st_rho), wave_sum(st_tv));
    if ((tid & 63) == 0) red[tid >> 6] = w;
    __syncthreads();
    if (tid == 0) {
      const int nw = ((int)blockDim.x + 63) >> 6;
      for (int j = 1; j < nw; ++j) {
        w.x += red[j].x; w.y += red[j].y; w.z += red[j].z; w.w += red[j].w;
      }
      part[blockIdx.x] = w;
    }
  }
}

// ------------------------------------------- resident multi-iteration sweep
// All K primal-dual iterations of one warp in a single launch, for pyramid
// levels whose instances hold at most kResidentPerThread * 1024 voxels: one
// work-group per instance keeps the instance's whole state on chip for the
// warp.  A thread owns up to V voxels (v = tid + j * blockDim.x) and holds
// their u, p, (s, n) and (rho_c, g) in registers; LDS holds the two fields
// neighbours read: u (for grad u) and r = D p' (for the adjoint), 48 bytes
// per voxel.  Per iteration:
//   1. own u -> LDS;                                           barrier
//   2. dual step from the neighbours' u; own r -> LDS;         barrier
//   3. primal step from the neighbours' r (registers only).
// Step 3 reads only r and the next step 1 writes only u; the next step 2
// writes r after the barrier that every thread reaches after its step 3.
// Each voxel is updated with the same operations, in the same order, as the
// marching sweep (k_sweep), so the results are the same bits; the launch
// replaces K launches whose cost at these sizes is latency.
// sched holds (sigma, tau, tau * lambda, Huber factor) per iteration; the state is read
// from buffer `parity` and left in buffer parity ^ (K & 1), where K single
// sweeps would leave it.  With two voxels per thread, (s, n) and (rho_c, g)
// are re-read from global memory (cache-resident after the first iteration)
// instead of held in registers, which keeps the kernel within 128 VGPRs
// without spills.
constexpr int kResidentPerThread = 2;

// part (null: off): per iteration it and wave, the four statistics sums of
// the wave's voxels at part[(it * gridDim.x + blockIdx.x) * waves + wave] (see
// k_sweep's STATS and k_stats_reduce).
// theta (not in the reference, 0 there): the dual step differentiates the
// over-relaxed point z = u + theta (u - u_prev), kept in registers beside u and
// equal to u at the launch's first iteration (a warp's start); theta = 0 keeps
// z = u without arithmetic.
template <int BC, bool ANISO, bool UNIT_H, int V>
__global__ __launch_bounds__(1024) void k_sweep_resident(const LevelDesc* __restrict__ descs,
                                                         const float* __restrict__ sched, int first, int K,
                                                         int sched_len, int parity, float wtv, int wtv_zero,
                                                         float eps, float theta, float4* __restrict__ part) {
  extern __shared__ float lds[];
  const LevelDesc& d = descs[blockIdx.x];
  const int NZ = d.NZ, NY = d.NY, NX = d.NX;
  const int N = (int)d.N;
  const int PL = NY * NX;
  const float hZ = UNIT_H ? 1.0f : d.hZ, hY = UNIT_H ? 1.0f : d.hY, hX = UNIT_H ? 1.0f : d.hX;
  float* U = lds;          // [3][N]
  float* R = lds + 3 * N;  // [9][N]: r[3c + e] = (D p')[c][e]
  const float* uin = d.u[parity];
  const float* pin = d.p[parity];

  constexpr bool KEEP = V == 1;  // (s, n) and (rho_c, g) held in registers
  float u[V][3], zr[V][3], p[V][9], sn[KEEP ? V : 1][4], rg[KEEP ? V : 1][4];
  int vz[V], vy[V], vx[V];
  bool own[V];
#pragma unroll
  for (int j = 0; j < V; ++j) {
    const int v = (int)threadIdx.x + j * (int)blockDim.x;
    own[j] = v < N;
    const int vv = own[j] ? v : 0;
    vx[j] = vv % NX;
    vy[j] = (vv / NX) % NY;
    vz[j] = vv / PL;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      u[j][c] = own[j] ? uin[(int64_t)c * N + vv] : 0.0f;
      zr[j][c] = u[j][c];
    }
#pragma unroll
    for (int e = 0; e < 9; ++e) p[j][e] = own[j] ? pin[(int64_t)e * N + vv] : 0.0f;
    if (KEEP) {
      const float4 a = own[j] && ANISO ? d.SN[vv] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
      const float4 b = own[j] ? d.RG[vv] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
      sn[0][0] = a.x; sn[0][1] = a.y; sn[0][2] = a.z; sn[0][3] = a.w;
      rg[0][0] = b.x; rg[0][1] = b.y; rg[0][2] = b.z; rg[0][3] = b.w;
    }
  }

  for (int it = 0; it < K; ++it) {
    const int n = (first + it) % sched_len;
    const float sigma = sched[4 * n], tau = sched[4 * n + 1], factor = sched[4 * n + 2], hub = sched[4 * n + 3];
    float st_du = 0.0f, st_dp = 0.0f, st_rho = 0.0f, st_tv = 0.0f;
#pragma unroll
    for (int j = 0; j < V; ++j)
      if (own[j]) {
        const int v = (int)threadIdx.x + j * (int)blockDim.x;
#pragma unroll
        for (int c = 0; c < 3; ++c) U[c * N + v] = zr[j][c];
      }
    __syncthreads();
    // ---- dual step
#pragma unroll
    for (int j = 0; j < V; ++j)
      if (own[j]) {
        const int v = (int)threadIdx.x + j * (int)blockDim.x;
        const int x = vx[j], y = vy[j], z = vz[j];
        // Neighbour indices clamped into the volume: a clamped value only
        // reaches derivative branches the boundary rule does not select.
        const int xm = x > 0 ? v - 1 : v, xp = x < NX - 1 ? v + 1 : v;
        const int ym = y > 0 ? v - NX : v, yp = y < NY - 1 ? v + NX : v;
        const int zm = z > 0 ? v - PL : v, zp = z < NZ - 1 ? v + PL : v;
        float s = 0.0f, nn[3] = {0.0f, 0.0f, 0.0f}, t1[3] = {0.0f, 0.0f, 0.0f};
        if (ANISO) {
          if (KEEP) {
            s = sn[0][0]; nn[0] = sn[0][1]; nn[1] = sn[0][2]; nn[2] = sn[0][3];
          } else {
            const float4 a = d.SN[v];
            s = a.x; nn[0] = a.y; nn[1] = a.z; nn[2] = a.w;
          }
          tangent1(nn, t1);
        }
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          const float* Uc = U + c * N;
          const float b0 = Uc[v];
          float q[3];
          q[0] = cd_fwd(Uc[xm], b0, Uc[xp], x, NX, hX, BC);
          q[1] = cd_fwd(Uc[ym], b0, Uc[yp], y, NY, hY, BC);
          q[2] = cd_fwd(Uc[zm], b0, Uc[zp], z, NZ, hZ, BC);
          float dq[3];
          if (ANISO) aniso_row(s, nn, t1, q, dq);
          else { dq[0] = q[0]; dq[1] = q[1]; dq[2] = q[2]; }
          float pv[3];
#pragma unroll
          for (int e = 0; e < 3; ++e) {
            const float inc = sigma * dq[e];
            pv[e] = p[j][3 * c + e] + inc;
          }
          prox_dual_row(pv, wtv, hub);
          if (wtv_zero) { pv[0] = 0.0f; pv[1] = 0.0f; pv[2] = 0.0f; }
          if (part) {
            st_tv += huber_of_norm(__builtin_amdgcn_sqrtf(dq[0] * dq[0] + dq[1] * dq[1] + dq[2] * dq[2]), eps);
#pragma unroll
            for (int e = 0; e < 3; ++e) {
              const float dd = pv[e] - p[j][3 * c + e];
              st_dp += dd * dd;
            }
          }
#pragma unroll
          for (int e = 0; e < 3; ++e) p[j][3 * c + e] = pv[e];
          float rr[3];
          if (ANISO) aniso_row(s, nn, t1, pv, rr);
          else { rr[0] = pv[0]; rr[1] = pv[1]; rr[2] = pv[2]; }
#pragma unroll
          for (int e = 0; e < 3; ++e) R[(3 * c + e) * N + v] = rr[e];
        }
      }
    __syncthreads();
    // ---- primal step
#pragma unroll
    for (int j = 0; j < V; ++j)
      if (own[j]) {
        const int v = (int)threadIdx.x + j * (int)blockDim.x;
        const int x = vx[j], y = vy[j], z = vz[j];
        const int xm = x > 0 ? v - 1 : v, xp = x < NX - 1 ? v + 1 : v;
        const int ym = y > 0 ? v - NX : v, yp = y < NY - 1 ? v + NX : v;
        const int zm = z > 0 ? v - PL : v, zp = z < NZ - 1 ? v + PL : v;
        float rgv[4];
        if (KEEP) {
          rgv[0] = rg[0][0]; rgv[1] = rg[0][1]; rgv[2] = rg[0][2]; rgv[3] = rg[0][3];
        } else {
          const float4 b = d.RG[v];
          rgv[0] = b.x; rgv[1] = b.y; rgv[2] = b.z; rgv[3] = b.w;
        }
        const float g[3] = {rgv[1], rgv[2], rgv[3]};
        float nv[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          const float* Rx = R + (3 * c) * N;
          const float* Ry = R + (3 * c + 1) * N;
          const float* Rz = R + (3 * c + 2) * N;
          const float ax = cd_adj(Rx[xm], Rx[v], Rx[xp], x, NX, BC);
          const float ay = cd_adj(Ry[ym], Ry[v], Ry[yp], y, NY, BC);
          const float az = cd_adj(Rz[zm], Rz[v], Rz[zp], z, NZ, BC);
          const float div = ax / hX + ay / hY + az / hZ;
          const float step = tau * div;
          nv[c] = u[j][c] - step;
        }
        const float sm = (u[j][0] * g[0] + u[j][2] * g[2]) + u[j][1] * g[1];  // (a + c) + b, as torch's sum
        const float rho = rgv[0] + sm;
        prox_primal_point(nv, rho, g, factor);
        if (part) {
          st_rho += fabsf(rho);
#pragma unroll
          for (int c = 0; c < 3; ++c) {
            const float dd = nv[c] - u[j][c];
            st_du += dd * dd;
          }
        }
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          if (theta != 0.0f) {
            const float dd = nv[c] - u[j][c];
            const float ee = theta * dd;
            zr[j][c] = nv[c] + ee;
          } else {
            zr[j][c] = nv[c];
          }
          u[j][c] = nv[c];
        }
      }
    if (part) {  // uniform: every thread of the work-group takes it
      const float4 w = make_float4(wave_sum(st_du), wave_sum(st_dp), wave_sum(st_rho), wave_sum(st_tv));
      const int waves = (int)blockDim.x >> 6;
      if ((threadIdx.x & 63) == 0)
        part[((int64_t)it * gridDim.x + blockIdx.x) * waves + (threadIdx.x >> 6)] = w;
    }
  }
  float* uout = d.u[parity ^ (K & 1)];
  float* pout = d.p[parity ^ (K & 1)];
#pragma unroll
  for (int j = 0; j < V; ++j)
    if (own[j]) {
      const int v = (int)threadIdx.x + j * (int)blockDim.x;
#pragma unroll
      for (int c = 0; c < 3; ++c) uout[(int64_t)c * N + v] = u[j][c];
#pragma unroll
      for (int e = 0; e < 9; ++e) pout[(int64_t)e * N + v] = p[j][e];
    }
}

// --------------------------------------------------- iteration statistics
// Per-iteration statistics of every instance from the sweeps' partial sums,
// part[(n * ngroups + g) * wpg + w] = (sum (u' - u)^2, sum (p' - p)^2,
// sum |rho(u)|, sum_c ||(D grad u)_c||) of work-group g's wave w (wpg = 1 for
// the marching sweep, which sums its waves itself) at iteration n of a warp.
// Work-group (n, k) of K * ninst adds instance k's entries in a fixed order, in
// double, and writes out[(n * ninst + k) * 3 + j]: j = 0 ||u^{n+1} - u^n||,
// j = 1 ||p^{n+1} - p^n|| (the residual norms of tvl13d.py:222, 237, 253),
// j = 2 the energy lambda sum|rho(u^n)| + lambda_tv sum_c ||D grad u^n_c|| (the
// caller passes lam = NaN when the dual step differentiates an over-relaxed
// point instead of u^n: the energy is then not available).
// Instance k's work-groups are csr_idx[csr_off[k] .. csr_off[k + 1]), or work-
// group k alone when csr_off is null (the resident sweep).
__global__ __launch_bounds__(64) void k_stats_reduce(const float4* __restrict__ part, const int* __restrict__ csr_off,
                                                     const int* __restrict__ csr_idx, int ngroups, int wpg,
                                                     int ninst, double lam, double wtv, float* __restrict__ out) {
  const int n = (int)blockIdx.x / ninst, k = (int)blockIdx.x % ninst;
  const int g0 = csr_off ? csr_off[k] : k, g1 = csr_off ? csr_off[k + 1] : k + 1;
  const int units = (g1 - g0) * wpg;
  double a = 0.0, b = 0.0, c = 0.0, e = 0.0;
  for (int q = (int)threadIdx.x; q < units; q += 64) {
    const int g = csr_off ? csr_idx[g0 + q / wpg] : k;
    const float4 v = part[((int64_t)n * ngroups + g) * wpg + q % wpg];
    a += v.x; b += v.y; c += v.z; e += v.w;
  }
  a = wave_sum(a); b = wave_sum(b); c = wave_sum(c); e = wave_sum(e);
  if (threadIdx.x == 0) {
    float* o = out + ((int64_t)n * ninst + k) * 3;
    o[0] = (float)sqrt(a);
    o[1] = (float)sqrt(b);
    o[2] = (float)(lam * c + wtv * e);
  }
}

// ============================================================ host plan
struct LevelPlan {
  std::vector<LevelDesc> host;   // per instance
  LevelDesc* dev = nullptr;      // table in the workspace
  int* block_start = nullptr;    // per instance, + total
  int nblocks = 0;
  // Separable prolongation into this level from the next coarser one:
  // block tables of the x-pass (coarse z, coarse y, fine x) and the y-pass
  // (coarse z, fine y, fine x) domains.
  int* px_start = nullptr;
  int* py_start = nullptr;
  int px_blocks = 0, py_blocks = 0;
  SweepTile* tiles = nullptr;
  int* tile_csr = nullptr;       // per instance, its tiles' launch indices: [ninst + 1] offsets, then indices
  int ntiles = 0;
  int tile_capacity = 0;
  int bx = 0, by = 0, cz = 0;    // shape policy (see inst_tile)
  int threads = 0;               // work-group size: the largest tile's bx * by
  size_t lds = 0;                // dynamic LDS of the largest tile
  bool unit_h = false;           // h == 1 on every axis of every instance
  // Resident sweep (k_sweep_resident): voxels per thread (0: the marching
  // sweep runs this level), its work-group size and dynamic LDS.
  int resident = 0;
  int res_capable = 0;           // voxels per thread the level would need (0: too large)
  int res_threads = 0;
  size_t res_lds = 0;
};

}  // namespace svs

struct svs_plan {
  svs_tvl1_params P;
  int ninst = 0;
  int S = 0;
  std::vector<int> dims;                  // 3 per instance (NZ, NY, NX)
  std::vector<svs::LevelPlan> levels;
  std::vector<float> sig, tau, factor;    // per iteration, rounded to float
  std::vector<float> hub;                 // per iteration: 1 + sigma eps / weight_tv (1: the reference's TV)
  std::vector<float> blur_w;              // K^3 weights (host copy)
  float* blur_dev = nullptr;
  char* ws = nullptr;
  size_t ws_bytes = 0;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  hipStream_t capture_stream = nullptr;
  bool tables_dirty = true;
  // Completion of the plan's last enqueued work (run, sweeps, autotuning):
  // work enqueued later, on any stream, waits for it, because it reuses the
  // workspace and the device tables.
  hipEvent_t last_work = nullptr;
  // Re-bound descriptor tables travel through a ring of pinned staging slots,
  // copied in stream order; a slot is rewritten only after the copy that
  // last read it has executed (its event).
  static constexpr int kStagingSlots = 4;
  char* staging = nullptr;               // kStagingSlots * table_bytes, pinned
  size_t table_bytes = 0;                // all levels' LevelDesc tables
  hipEvent_t staged[kStagingSlots] = {};
  int next_slot = 0;
  float* sched_dev = nullptr;            // (sigma, tau, tau * lambda, Huber factor) per iteration
  // Per-iteration statistics (svs_tvl1_plan_set_stats): the caller's output
  // and the plan's own buffer of the sweeps' partial sums for one warp.
  float* stats_out = nullptr;
  float4* stats_part = nullptr;
  size_t stats_part_bytes = 0;
};

namespace svs {

static size_t align_up(size_t v, size_t a) { return (v + a - 1) / a * a; }

// Domains of the separable prolongation's x-pass and y-pass.
static int64_t prolong_x_domain(const LevelDesc& c, const LevelDesc& f) { return (int64_t)c.NZ * c.NY * f.NX; }
static int64_t prolong_y_domain(const LevelDesc& c, const LevelDesc& f) { return (int64_t)c.NZ * f.NY * f.NX; }

// numDofs meshes use L = N - 1 (flow_utils.py:48-71); h = L / (N - 1) in float.
static void level_geometry(const svs_tvl1_params& P, int NZ, int NY, int NX, LevelDesc& d) {
  d.NZ = NZ; d.NY = NY; d.NX = NX;
  if (P.length_fixed) {
    d.LZ = P.LZ; d.LY = P.LY; d.LX = P.LX;
  } else {
    d.LZ = (float)(NZ - 1); d.LY = (float)(NY - 1); d.LX = (float)(NX - 1);
  }
  d.hZ = d.LZ / (float)(NZ - 1); d.hY = d.LY / (float)(NY - 1); d.hX = d.LX / (float)(NX - 1);
  d.N = (int64_t)NZ * NY * NX;
}

static size_t sweep_lds_bytes(int bx, int by) {
  return sizeof(float) * (6 * (size_t)(bx + 2) * (by + 2) + 12 * (size_t)bx * by);
}

// A level's sweep shape (bx, by, cz) is a policy: tiles are at most bx - 2
// voxels wide and use at most bx * by threads; each instance gets its own tile
// from it (inst_tile), so volumes of different widths in one launch all march
// full (or evenly split) rows.  For NX == bx - 2 the tile is exactly bx x by.
struct SweepShape {
  int bx, by, cz;
};

static void inst_tile(const LevelDesc& d, int bx, int by, int& tbx, int& tby) {
  const int txmax = bx - 2;
  const int ntx = (d.NX + txmax - 1) / txmax;
  const int tx = (d.NX + ntx - 1) / ntx;  // rows split into equal pieces
  tbx = tx + 2;
  const int tymax = std::min((bx * by) / tbx - 2, d.NY);
  if (tymax < 1) {  // not even one row fits the budget: shape_ok rejects it
    tby = tymax + 2;
    return;
  }
  const int nty = (d.NY + tymax - 1) / tymax;
  tby = (d.NY + nty - 1) / nty + 2;       // and the rows into equal bands
}

static int count_tiles(const LevelPlan& L, int bx, int by, int cz) {
  int n = 0;
  for (auto& d : L.host) {
    int tbx, tby;
    inst_tile(d, bx, by, tbx, tby);
    n += ((d.NX + tbx - 3) / (tbx - 2)) * ((d.NY + tby - 3) / (tby - 2)) * ((d.NZ + cz - 1) / cz);
  }
  return n;
}

// Every instance's tile has a row or more, one thread per outer-ring
// position and fits the LDS limit.
static bool shape_ok(const LevelPlan& L, int bx, int by, int cz) {
  if (bx < 3 || by < 3 || cz < 1 || bx * by > 1024) return false;
  for (auto& d : L.host) {
    int tbx, tby;
    inst_tile(d, bx, by, tbx, tby);
    if (tby < 3 || 2 * (tbx + 2) + 2 * tby > tbx * tby || sweep_lds_bytes(tbx, tby) > 96 * 1024) return false;
  }
  return true;
}

// z-chunk that gives the launch about `target` work-groups.
static int default_cz(const LevelPlan& L, int bx, int by, int target, int minz) {
  int maxNZ = 0;
  for (auto& d : L.host) maxNZ = std::max(maxNZ, d.NZ);
  int64_t cols = 0;
  for (auto& d : L.host) {
    int tbx, tby;
    inst_tile(d, bx, by, tbx, tby);
    cols += (int64_t)((d.NX + tbx - 3) / (tbx - 2)) * ((d.NY + tby - 3) / (tby - 2));
  }
  int cz = maxNZ;
  while (cz > minz && cols * ((maxNZ + cz - 1) / cz) < target) cz = (cz + 1) / 2;
  return cz;
}

// Autotuner candidates: (tile width, tile height) pairs with a z-chunk aimed at
// ~2048 work-groups (stage 1), then other z-chunks for the winner (stage 2).
static std::vector<SweepShape> tile_candidates(const LevelPlan& L) {
  int maxNX = 0, maxNY = 0;
  for (auto& d : L.host) { maxNX = std::max(maxNX, d.NX); maxNY = std::max(maxNY, d.NY); }
  std::vector<int> txs = {maxNX, (maxNX + 1) / 2, (maxNX + 2) / 3, (maxNX + 3) / 4, 16, 8};
  // Uneven batches: every instance's own width and its half as well, so that a
  // policy can give full rows to one size while splitting wider ones.
  for (auto& d : L.host)
    for (int tx : {d.NX, (d.NX + 1) / 2})
      if (std::find(txs.begin(), txs.end(), tx) == txs.end()) txs.push_back(tx);
  std::vector<int> tys = {2, 3, 4, 5, 6, 7, 8, 9, 10, 12, 14, 20};
  std::vector<SweepShape> out;
  for (int tx : txs) {
    if (tx < 4) continue;
    for (int ty : tys) {
      if (ty > maxNY && ty != 2) continue;
      const int bx = tx + 2, by = ty + 2;
      bool dup = false;
      for (auto& c : out) dup |= (c.bx == bx && c.by == by);
      if (dup || !shape_ok(L, bx, by, 1)) continue;
      out.push_back({bx, by, default_cz(L, bx, by, 2048, 2)});
    }
  }
  return out;
}

static std::vector<int> cz_candidates(const LevelPlan& L) {
  int maxNZ = 0;
  for (auto& d : L.host) maxNZ = std::max(maxNZ, d.NZ);
  // Short chunks for small levels, and maxNZ split into 1..6 equal chunks.
  std::vector<int> out;
  auto add = [&](int cz) {
    if (cz >= 1 && cz <= maxNZ && std::find(out.begin(), out.end(), cz) == out.end()) out.push_back(cz);
  };
  for (int cz : {1, 2, 4, 8}) add(cz);
  for (int parts = 1; parts <= 6; ++parts) add((maxNZ + parts - 1) / parts);
  return out;
}

// Process-wide memory of tuned shapes, keyed by the level geometry and the
// kernel variant, so later plans of the same geometry start tuned.
static std::mutex g_tune_mu;
static std::map<std::string, SweepShape> g_tuned;
static std::map<std::string, bool> g_tuned_resident;  // levels that qualify: resident faster?

static std::string tune_key(const svs_tvl1_params& P, const LevelPlan& L) {
  std::map<std::array<int, 3>, int> hist;
  for (auto& d : L.host) hist[{d.NZ, d.NY, d.NX}]++;
  std::string k = std::to_string(P.boundary) + "/" + std::to_string(P.use_aniso) + ":";
  for (auto& e : hist)
    k += std::to_string(e.first[0]) + "x" + std::to_string(e.first[1]) + "x" + std::to_string(e.first[2]) + "*" +
         std::to_string(e.second) + ",";
  return k;
}

// Default sweep shape policy of a level, by its voxel count (all instances).
// Measured on MI355X with the autotuner: batches of large volumes run best
// with full-row tiles of ~900 threads marching ~40 planes (68 x 80^3:
// 82 x 11, 40 planes); mid-size levels with rows up to 64 wide and ~600
// work-groups (68 x 40^3: 42 x 10, 20 planes); a single volume of 10^5..10^6
// voxels with half rows marching 2 planes (80^3: 42 x 11); small levels are
// latency-bound and want many small work-groups marching one plane.
static void choose_sweep_shape(const svs_tvl1_params& P, LevelPlan& L, int ninst) {
  (void)ninst;
  int maxNX = 0;
  int64_t vox = 0;
  for (auto& d : L.host) {
    maxNX = std::max(maxNX, d.NX);
    vox += d.N;
  }
  int bx = P.sweep_bx, by = P.sweep_by, cz = P.sweep_cz;
  if (bx <= 0 && by <= 0 && cz <= 0) {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    auto it = g_tuned.find(tune_key(P, L));
    if (it != g_tuned.end()) {
      L.bx = it->second.bx; L.by = it->second.by; L.cz = it->second.cz;
      return;
    }
  }
  int maxtx, budget, minz, target;
  if (vox >= ((int64_t)8 << 20)) { maxtx = 128; budget = 910; minz = 40; target = 1024; }
  else if (vox >= ((int64_t)1 << 20)) { maxtx = 64; budget = 448; minz = 4; target = 600; }
  else if (vox >= ((int64_t)128 << 10)) { maxtx = 40; budget = 462; minz = 2; target = 2048; }
  else if (vox >= ((int64_t)16 << 10)) { maxtx = 14; budget = 192; minz = 1; target = 2048; }
  else { maxtx = 8; budget = 64; minz = 1; target = 2048; }
  if (bx <= 0) {
    const int ntx = (maxNX + maxtx - 1) / maxtx;
    bx = (maxNX + ntx - 1) / ntx + 2;
  }
  if (by <= 0) by = std::max(3, budget / bx);
  if (!shape_ok(L, bx, by, 1)) {
    // A narrow tile may need more rows for one thread per outer-ring position.
    int b = by;
    while (b * bx <= 1024 && !shape_ok(L, bx, b, 1)) ++b;
    if (shape_ok(L, bx, b, 1)) by = b;
    else { bx = 12; by = 12; }
  }
  if (cz <= 0) cz = default_cz(L, bx, by, target, minz);
  L.bx = bx; L.by = by; L.cz = cz;
}

}  // namespace svs

using namespace svs;

// The extended model (Huber-TV, over-relaxation; no reference counterpart):
// its sweeps keep the over-relaxed point z in fields of its own.
static bool extended_model(const svs_tvl1_params& P) { return P.huber_eps > 0.0 || P.relax_theta != 0.0; }

// Workspace layout, computed identically by the sizing and the plan builder.
static int plan_layout(svs_plan* pl, char* base, size_t* total_out) {
  const svs_tvl1_params& P = pl->P;
  const int S = pl->S;
  size_t off = 0;
  int ntake = 0;
  // Buffers start 512 B x (count mod 8) past the 256 B boundary, so that the
  // fields of one volume do not all share one offset modulo 16 KiB (at sizes
  // like 80^3 every field is a multiple of 16 KiB long): a copy of the
  // sweep's 32 streams moves 4.88 TB/s with all of them aligned alike and
  // 5.00 TB/s with the groups (u, p, SN, RG of both parities) staggered.
  auto take = [&](size_t bytes) -> char* {
    off = align_up(off, 256) + 512 * (size_t)(ntake++ % 8);
    char* p = base ? base + off : nullptr;
    off += bytes;
    return p;
  };
  pl->levels.assign(S, LevelPlan());
  for (int s = 0; s < S; ++s) pl->levels[s].host.resize(pl->ninst);
  for (int k = 0; k < pl->ninst; ++k) {
    int nz = pl->dims[3 * k], ny = pl->dims[3 * k + 1], nx = pl->dims[3 * k + 2];
    const int64_t N0 = (int64_t)nz * ny * nx;
    // One block, so that the separable prolongation can use all of it for its
    // x-pass (<= 12 * 0.625^2 floats per fine voxel; free while it runs). SN
    // and RG sit 512 B and 1024 B past N0 and 2 N0 float4s: staggered like
    // the buffers.
    float4* J = (float4*)take(48 * N0 + 1024);
    float4* SN = J + N0 + 32;
    float4* RG = J + 2 * N0 + 64;
    for (int s = 0; s < S; ++s) {
      if (s > 0) { nz = (nz + 1) / 2; ny = (ny + 1) / 2; nx = (nx + 1) / 2; }
      LevelDesc& d = pl->levels[s].host[k];
      std::memset(&d, 0, sizeof(d));
      level_geometry(P, nz, ny, nx, d);
      const int64_t N = d.N;
      d.J = J; d.SN = SN; d.RG = RG;
      if (s > 0 || P.use_blur) {
        d.I0 = (float*)take(4 * N);
        d.I1 = (float*)take(4 * N);
      }
      for (int b = 0; b < 2; ++b) {
        d.u[b] = (float*)take(12 * N);
        d.p[b] = (float*)take(36 * N);
        if (extended_model(P)) d.z[b] = (float*)take(12 * N);
      }
    }
  }
  // Device tables.
  for (int s = 0; s < S; ++s) {
    LevelPlan& L = pl->levels[s];
    choose_sweep_shape(P, L, pl->ninst);
    L.dev = (LevelDesc*)take(sizeof(LevelDesc) * pl->ninst);
    L.block_start = (int*)take(sizeof(int) * (pl->ninst + 1));
    L.px_start = (int*)take(sizeof(int) * (pl->ninst + 1));
    L.py_start = (int*)take(sizeof(int) * (pl->ninst + 1));
    L.ntiles = count_tiles(L, L.bx, L.by, L.cz);
    // Room for any shape the autotuner may try.
    int cap = L.ntiles;
    for (const SweepShape& c : tile_candidates(L)) {
      cap = std::max(cap, count_tiles(L, c.bx, c.by, c.cz));
      for (int cz : cz_candidates(L)) cap = std::max(cap, count_tiles(L, c.bx, c.by, cz));
    }
    L.tile_capacity = cap;
    L.tiles = (SweepTile*)take(sizeof(SweepTile) * cap);
    L.tile_csr = (int*)take(sizeof(int) * ((size_t)pl->ninst + 1 + cap));
    int nb = 0;
    bool unit = true;
    for (auto& d : L.host) {
      nb += grid_for(d.N, kBlock);
      unit = unit && d.hZ == 1.0f && d.hY == 1.0f && d.hX == 1.0f;
    }
    L.nblocks = nb;
    L.unit_h = unit;
    {
      int64_t maxN = 0;
      for (auto& d : L.host) maxN = std::max<int64_t>(maxN, d.N);
      L.res_capable = 0;
      for (int v = 1; v <= kResidentPerThread && !L.res_capable; ++v)
        if (maxN <= (int64_t)1024 * v) L.res_capable = v;
      if (L.res_capable) {
        const int per = (int)((maxN + L.res_capable - 1) / L.res_capable);
        L.res_threads = std::min(1024, (per + 63) / 64 * 64);
        L.res_lds = sizeof(float) * 12 * (size_t)maxN;
      }
      L.resident = L.res_capable;  // automatic: resident whenever the level qualifies ...
      if (L.res_capable) {         // ... unless the autotuner measured the marching sweep faster
        std::lock_guard<std::mutex> lk(g_tune_mu);
        auto it = g_tuned_resident.find(tune_key(P, L));
        if (it != g_tuned_resident.end() && !it->second) L.resident = 0;
      }
    }
    L.px_blocks = L.py_blocks = 0;
    if (s + 1 < S)
      for (int k = 0; k < pl->ninst; ++k) {
        L.px_blocks += grid_for(prolong_x_domain(pl->levels[s + 1].host[k], L.host[k]), kBlock);
        L.py_blocks += grid_for(prolong_y_domain(pl->levels[s + 1].host[k], L.host[k]), kBlock);
      }
  }
  if (P.use_blur) pl->blur_dev = (float*)take(sizeof(float) * pl->blur_w.size());
  pl->sched_dev = (float*)take(sizeof(float) * 4 * (size_t)std::max(1, P.max_outer_iterations));
  *total_out = align_up(off, 256);
  return 0;
}

static int validate_params(const svs_tvl1_params* P) {
  SVS_REQUIRE(P, "null params");
  SVS_REQUIRE(P->num_scales >= 1 && P->num_scales <= 16, "num_scales must be in [1,16]");
  SVS_REQUIRE(P->max_warps >= 0 && P->max_outer_iterations >= 0, "negative warp/iteration count");
  SVS_REQUIRE(valid_bc(P->boundary) && valid_ip(P->interpolation), "invalid interpolation/boundary type");
  SVS_REQUIRE(!P->use_blur || (P->blur_K >= 1 && (P->blur_K & 1) && P->blur_weights),
              "blur enabled but kernel size/weights invalid");
  SVS_REQUIRE(P->huber_eps >= 0.0 && std::isfinite(P->huber_eps), "huber_eps must be finite and >= 0");
  SVS_REQUIRE(std::isfinite(P->relax_theta), "relax_theta must be finite");
  return 0;
}

static int validate_dims(const svs_tvl1_params* P, int ninst, const int* dims) {
  SVS_REQUIRE(ninst >= 1 && dims, "need at least one instance");
  // The per-voxel kernels of a plan launch one work-item per voxel of every
  // instance (whole work-groups per instance) at the finest level.
  int64_t items = 0;
  for (int k = 0; k < ninst; ++k)
    if (dims[3 * k] > 0 && dims[3 * k + 1] > 0 && dims[3 * k + 2] > 0)
      items += ((int64_t)dims[3 * k] * dims[3 * k + 1] * dims[3 * k + 2] + kBlock - 1) / kBlock * kBlock;
  SVS_REQUIRE_WORKITEMS(items, "plan (voxels of all instances at the finest level)");
  for (int k = 0; k < ninst; ++k) {
    int nz = dims[3 * k], ny = dims[3 * k + 1], nx = dims[3 * k + 2];
    SVS_REQUIRE_DIMS(nz, ny, nx);
    SVS_REQUIRE(nz < 32768 && ny < 32768 && nx < 32768, "every dimension must be < 32768");
    // The fused sweep addresses the 9 dual fields of an instance with 32-bit
    // byte offsets.
    SVS_REQUIRE((int64_t)nz * ny * nx <= (int64_t)0xFFFFFFFF / 36,
                "instance %d: %dx%dx%d exceeds the %lld voxels a plan instance may hold", k, nz, ny, nx,
                (long long)(0xFFFFFFFF / 36));
    for (int s = 0; s < P->num_scales; ++s) {
      if (s > 0) { nz = (nz + 1) / 2; ny = (ny + 1) / 2; nx = (nx + 1) / 2; }
      SVS_REQUIRE(nz >= 4 && ny >= 4 && nx >= 4,
                  "instance %d: pyramid level %d is %dx%dx%d; every level needs >= 4 voxels per axis", k, s, nz,
                  ny, nx);
    }
  }
  return 0;
}

static void fill_schedule(svs_plan* pl) {
  const svs_tvl1_params& P = pl->P;
  const int K = P.max_outer_iterations;
  pl->sig.resize(K); pl->tau.resize(K); pl->factor.resize(K); pl->hub.resize(K);
  // tvl13d.py:208-246: Python floats, reset per warp, rounded to float at use.
  double sg = P.sigma, ta = P.tau;
  for (int n = 0; n < K; ++n) {
    pl->sig[n] = (float)sg;
    pl->tau[n] = (float)ta;
    pl->factor[n] = (float)ta * P.weight_matching;  // opticalFlowOps.cu:442, FP32 product
    // Huber-TV (eps > 0; no reference counterpart): 1 + sigma_n eps / weight_tv
    // in double from the Python-float sigma_n, used as a float.
    pl->hub[n] = (P.huber_eps == 0.0 || P.weight_tv == 0.0f) ? 1.0f
                                                              : (float)(1.0 + sg * P.huber_eps / (double)P.weight_tv);
    if (P.pd_type == 2) {
      const double th = 1. / std::sqrt(1. + 2. * P.gamma * ta);
      ta *= th;
      sg /= th;
    }
  }
  for (int s = 0; s + 1 < pl->S; ++s)
    for (int k = 0; k < pl->ninst; ++k) {
      // tvl13d.py:151-153: ratio of one instance's FP32 lengths in Python
      // floats, applied as a float; (N_fine - 1)/(N_coarse - 1) for numDofs.
      LevelDesc& f = pl->levels[s].host[k];
      const LevelDesc& c = pl->levels[s + 1].host[k];
      f.upX = (float)((double)f.LX / (double)c.LX);
      f.upY = (float)((double)f.LY / (double)c.LY);
      f.upZ = (float)((double)f.LZ / (double)c.LZ);
    }
}

// Copies the host descriptor tables to the device in stream order without
// blocking the host: they are packed into the next pinned staging slot (after
// waiting, if needed, for the copy that last read that slot) and copied from
// there on `st`.
static int upload_tables(svs_plan* pl, hipStream_t st) {
  const int slot = pl->next_slot;
  pl->next_slot = (slot + 1) % svs_plan::kStagingSlots;
  SVS_HIP(hipEventSynchronize(pl->staged[slot]));
  char* stage = pl->staging + (size_t)slot * pl->table_bytes;
  size_t off = 0;
  for (int s = 0; s < pl->S; ++s) {
    LevelPlan& L = pl->levels[s];
    const size_t n = sizeof(LevelDesc) * pl->ninst;
    std::memcpy(stage + off, L.host.data(), n);
    SVS_HIP(hipMemcpyAsync(L.dev, stage + off, n, hipMemcpyHostToDevice, st));
    off += n;
  }
  SVS_HIP(hipEventRecord(pl->staged[slot], st));
  pl->tables_dirty = false;
  return 0;
}

// Orders work enqueued on `st` after the plan's previous work, whichever
// stream that ran on.
static int after_last_work(svs_plan* pl, hipStream_t st) {
  SVS_HIP(hipStreamWaitEvent(st, pl->last_work, 0));
  return 0;
}

static int mark_last_work(svs_plan* pl, hipStream_t st) {
  SVS_HIP(hipEventRecord(pl->last_work, st));
  return 0;
}

// Forgets the captured schedule (launch shapes or buffers changed); the next
// graph run captures it again.
static void drop_graph(svs_plan* pl) {
  if (pl->exec) { (void)hipGraphExecDestroy(pl->exec); pl->exec = nullptr; }
  if (pl->graph) { (void)hipGraphDestroy(pl->graph); pl->graph = nullptr; }
}

// Releases the plan's resources once its last work has completed.
static void free_plan(svs_plan* pl) {
  if (pl->last_work) (void)hipEventSynchronize(pl->last_work);
  for (int i = 0; i < svs_plan::kStagingSlots; ++i)
    if (pl->staged[i]) {
      (void)hipEventSynchronize(pl->staged[i]);
      (void)hipEventDestroy(pl->staged[i]);
    }
  if (pl->staging) (void)hipHostFree(pl->staging);
  if (pl->stats_part) (void)hipFree(pl->stats_part);
  if (pl->last_work) (void)hipEventDestroy(pl->last_work);
  if (pl->exec) (void)hipGraphExecDestroy(pl->exec);
  if (pl->graph) (void)hipGraphDestroy(pl->graph);
  if (pl->capture_stream) (void)hipStreamDestroy(pl->capture_stream);
  delete pl;
}

static int create_sync_objects(svs_plan* pl) {
  pl->table_bytes = sizeof(LevelDesc) * (size_t)pl->ninst * pl->S;
  SVS_HIP(hipHostMalloc((void**)&pl->staging, pl->table_bytes * svs_plan::kStagingSlots, hipHostMallocDefault));
  SVS_HIP(hipEventCreateWithFlags(&pl->last_work, hipEventDisableTiming));
  for (int i = 0; i < svs_plan::kStagingSlots; ++i)
    SVS_HIP(hipEventCreateWithFlags(&pl->staged[i], hipEventDisableTiming));
  return 0;
}

static int upload_tiles(svs_plan* pl, int s, hipStream_t st);

static int upload_static(svs_plan* pl, hipStream_t st) {
  for (int s = 0; s < pl->S; ++s) {
    LevelPlan& L = pl->levels[s];
    std::vector<int> bs(pl->ninst + 1);
    int acc = 0;
    for (int k = 0; k < pl->ninst; ++k) { bs[k] = acc; acc += grid_for(L.host[k].N, kBlock); }
    bs[pl->ninst] = acc;
    SVS_HIP(hipMemcpyAsync(L.block_start, bs.data(), sizeof(int) * bs.size(), hipMemcpyHostToDevice, st));

    SVS_HIP(hipStreamSynchronize(st));
    if (s + 1 < pl->S) {
      std::vector<int> bx(pl->ninst + 1), by(pl->ninst + 1);
      int ax = 0, ay = 0;
      for (int k = 0; k < pl->ninst; ++k) {
        const LevelDesc& c = pl->levels[s + 1].host[k];
        bx[k] = ax; ax += grid_for(prolong_x_domain(c, L.host[k]), kBlock);
        by[k] = ay; ay += grid_for(prolong_y_domain(c, L.host[k]), kBlock);
      }
      bx[pl->ninst] = ax;
      by[pl->ninst] = ay;
      SVS_HIP(hipMemcpyAsync(L.px_start, bx.data(), sizeof(int) * bx.size(), hipMemcpyHostToDevice, st));
      SVS_HIP(hipMemcpyAsync(L.py_start, by.data(), sizeof(int) * by.size(), hipMemcpyHostToDevice, st));
      SVS_HIP(hipStreamSynchronize(st));
    }
    if (int rc = upload_tiles(pl, s, st)) return rc;
  }
  if (pl->P.use_blur) {
    SVS_HIP(hipMemcpyAsync(pl->blur_dev, pl->blur_w.data(), sizeof(float) * pl->blur_w.size(),
                           hipMemcpyHostToDevice, st));
    SVS_HIP(hipStreamSynchronize(st));
  }
  if (pl->P.max_outer_iterations > 0) {
    std::vector<float> sched(4 * (size_t)pl->P.max_outer_iterations);
    for (int n = 0; n < pl->P.max_outer_iterations; ++n) {
      sched[4 * n] = pl->sig[n];
      sched[4 * n + 1] = pl->tau[n];
      sched[4 * n + 2] = pl->factor[n];
      sched[4 * n + 3] = pl->hub[n];
    }
    SVS_HIP(hipMemcpyAsync(pl->sched_dev, sched.data(), sizeof(float) * sched.size(), hipMemcpyHostToDevice, st));
    SVS_HIP(hipStreamSynchronize(st));
  }
  return 0;
}

// Builds and uploads level s's tile list for its current (bx, by, cz).
static int upload_tiles(svs_plan* pl, int s, hipStream_t st) {
  {
    LevelPlan& L = pl->levels[s];
    L.ntiles = count_tiles(L, L.bx, L.by, L.cz);
    SVS_REQUIRE(L.ntiles <= L.tile_capacity, "tile table overflow (%d > %d)", L.ntiles, L.tile_capacity);
    std::vector<SweepTile> natural;
    natural.reserve(L.ntiles);
    L.threads = 0;
    L.lds = 0;
    for (int k = 0; k < pl->ninst; ++k) {
      const LevelDesc& d = L.host[k];
      int tbx, tby;
      inst_tile(d, L.bx, L.by, tbx, tby);
      L.threads = std::max(L.threads, tbx * tby);
      L.lds = std::max(L.lds, sweep_lds_bytes(tbx, tby));
      const int tx = tbx - 2, ty = tby - 2;
      for (int z0 = 0; z0 < d.NZ; z0 += L.cz)
        for (int y0 = 0; y0 < d.NY; y0 += ty)
          for (int x0 = 0; x0 < d.NX; x0 += tx) {
            const int z1 = std::min(d.NZ, z0 + L.cz);
            natural.push_back(SweepTile{k, x0 | (y0 << 16), z0 | (z1 << 16), tbx | (tby << 16)});
          }
    }
    // Work-group b runs on XCD b % 8.  Deal contiguous runs of the natural
    // order (neighbouring tiles of one volume) to one XCD each, so the halo
    // rows a tile re-reads from its neighbours sit in the same L2.
    std::vector<SweepTile> tiles;
    tiles.reserve(natural.size());
    {
      const int nx = kXcds, n = (int)natural.size();
      std::vector<int> start(nx + 1);
      for (int g = 0; g <= nx; ++g) start[g] = (int)((int64_t)n * g / nx);
      const int maxlen = (n + nx - 1) / nx;
      for (int slot = 0; slot < maxlen; ++slot)
        for (int g = 0; g < nx; ++g)
          if (start[g] + slot < start[g + 1]) tiles.push_back(natural[start[g] + slot]);
    }
    SVS_HIP(hipMemcpyAsync(L.tiles, tiles.data(), sizeof(SweepTile) * tiles.size(), hipMemcpyHostToDevice, st));
    // Each instance's tiles in launch order, for the statistics reduction.
    std::vector<int> csr((size_t)pl->ninst + 1 + tiles.size(), 0);
    for (const SweepTile& t : tiles) csr[(size_t)t.inst + 1]++;
    for (int k = 0; k < pl->ninst; ++k) csr[(size_t)k + 1] += csr[k];
    {
      std::vector<int> fill(csr.begin(), csr.begin() + pl->ninst);
      for (int b = 0; b < (int)tiles.size(); ++b) csr[(size_t)pl->ninst + 1 + fill[tiles[b].inst]++] = b;
    }
    SVS_HIP(hipMemcpyAsync(L.tile_csr, csr.data(), sizeof(int) * csr.size(), hipMemcpyHostToDevice, st));
    SVS_HIP(hipStreamSynchronize(st));
  }
  return 0;
}

// The scalars of one sweep: schedule position n's steps, and for the extended
// model its Huber factor and eps, theta and whether z is still u (n == 0).
struct SweepScalars {
  float sg, ta, fac, wtv, hub, eps, theta;
  int zfirst;
  bool ext;
};

template <int BC, bool ANISO, bool STATS, bool EXT>
static void launch_sweep_s(const LevelPlan& L, int parity, const SweepScalars& a, float4* part, hipStream_t st) {
  if (L.unit_h)
    hipLaunchKernelGGL((k_sweep<BC, ANISO, true, STATS, EXT>), dim3(L.ntiles), dim3(L.threads), L.lds, st, L.dev,
                       L.tiles, parity, a.sg, a.ta, a.fac, a.wtv, a.hub, a.eps, a.theta, a.zfirst, part);
  else
    hipLaunchKernelGGL((k_sweep<BC, ANISO, false, STATS, EXT>), dim3(L.ntiles), dim3(L.threads), L.lds, st,
                       L.dev, L.tiles, parity, a.sg, a.ta, a.fac, a.wtv, a.hub, a.eps, a.theta, a.zfirst, part);
}

template <int BC, bool ANISO>
static void launch_sweep_t(const LevelPlan& L, int parity, const SweepScalars& a, float4* part, hipStream_t st) {
  if (part) {
    if (a.ext) launch_sweep_s<BC, ANISO, true, true>(L, parity, a, part, st);
    else launch_sweep_s<BC, ANISO, true, false>(L, parity, a, part, st);
  } else {
    if (a.ext) launch_sweep_s<BC, ANISO, false, true>(L, parity, a, nullptr, st);
    else launch_sweep_s<BC, ANISO, false, false>(L, parity, a, nullptr, st);
  }
}

// One sweep at schedule position n; part (null: no statistics) receives one
// float4 of statistics sums per work-group.
static void launch_sweep(const svs_plan* pl, const LevelPlan& L, int parity, int n, hipStream_t st,
                         float4* part = nullptr) {
  const svs_tvl1_params& P = pl->P;
  const SweepScalars a{pl->sig[n], pl->tau[n], pl->factor[n], P.weight_tv, pl->hub[n], (float)P.huber_eps,
                       (float)P.relax_theta, n == 0 ? 1 : 0, extended_model(P)};
  const bool an = P.use_aniso != 0;
  switch (P.boundary * 2 + (an ? 1 : 0)) {
    case 0: launch_sweep_t<BD_NEAREST, false>(L, parity, a, part, st); break;
    case 1: launch_sweep_t<BD_NEAREST, true>(L, parity, a, part, st); break;
    case 2: launch_sweep_t<BD_MIRROR, false>(L, parity, a, part, st); break;
    case 3: launch_sweep_t<BD_MIRROR, true>(L, parity, a, part, st); break;
    case 4: launch_sweep_t<BD_REFLECT, false>(L, parity, a, part, st); break;
    default: launch_sweep_t<BD_REFLECT, true>(L, parity, a, part, st); break;
  }
}

template <int BC, bool ANISO, int V>
static void launch_resident_v(const svs_plan* pl, const LevelPlan& L, int parity, int first, int K,
                              float4* part, hipStream_t st) {
  const svs_tvl1_params& P = pl->P;
  const float wtv = P.weight_tv;
  const int wz = P.weight_tv == 0.0f;
  const int len = P.max_outer_iterations;
  const float eps = (float)P.huber_eps, theta = (float)P.relax_theta;
  if (L.unit_h)
    hipLaunchKernelGGL((k_sweep_resident<BC, ANISO, true, V>), dim3(pl->ninst), dim3(L.res_threads), L.res_lds, st,
                       L.dev, pl->sched_dev, first, K, len, parity, wtv, wz, eps, theta, part);
  else
    hipLaunchKernelGGL((k_sweep_resident<BC, ANISO, false, V>), dim3(pl->ninst), dim3(L.res_threads), L.res_lds,
                       st, L.dev, pl->sched_dev, first, K, len, parity, wtv, wz, eps, theta, part);
}

template <int BC, bool ANISO>
static void launch_resident_t(const svs_plan* pl, const LevelPlan& L, int parity, int first, int K,
                              float4* part, hipStream_t st) {
  if (L.resident == 1) launch_resident_v<BC, ANISO, 1>(pl, L, parity, first, K, part, st);
  else launch_resident_v<BC, ANISO, 2>(pl, L, parity, first, K, part, st);
}

// K iterations of schedule positions first, first + 1, ... (cyclic) in one
// launch; the state moves from buffer `parity` to parity ^ (K & 1).  part
// (null: no statistics) receives K * ninst * (res_threads / 64) float4s.
static void launch_resident(const svs_plan* pl, const LevelPlan& L, int parity, int first, int K,
                            hipStream_t st, float4* part = nullptr) {
  const bool an = pl->P.use_aniso != 0;
  switch (pl->P.boundary * 2 + (an ? 1 : 0)) {
    case 0: launch_resident_t<BD_NEAREST, false>(pl, L, parity, first, K, part, st); break;
    case 1: launch_resident_t<BD_NEAREST, true>(pl, L, parity, first, K, part, st); break;
    case 2: launch_resident_t<BD_MIRROR, false>(pl, L, parity, first, K, part, st); break;
    case 3: launch_resident_t<BD_MIRROR, true>(pl, L, parity, first, K, part, st); break;
    case 4: launch_resident_t<BD_REFLECT, false>(pl, L, parity, first, K, part, st); break;
    default: launch_resident_t<BD_REFLECT, true>(pl, L, parity, first, K, part, st); break;
  }
}

template <int BC, bool ANISO>
static int allow_resident_lds() {
  const int v = 96 * 1024;
  const void* kernels[] = {
      (const void*)k_sweep_resident<BC, ANISO, false, 1>, (const void*)k_sweep_resident<BC, ANISO, true, 1>,
      (const void*)k_sweep_resident<BC, ANISO, false, 2>, (const void*)k_sweep_resident<BC, ANISO, true, 2>};
  for (const void* k : kernels) SVS_HIP(hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, v));
  return 0;
}

template <bool STATS, bool EXT>
static int allow_sweep_lds(int v) {
  const void* kernels[] = {
      (const void*)k_sweep<BD_NEAREST, false, false, STATS, EXT>,
      (const void*)k_sweep<BD_NEAREST, true, false, STATS, EXT>,
      (const void*)k_sweep<BD_MIRROR, false, false, STATS, EXT>,
      (const void*)k_sweep<BD_MIRROR, true, false, STATS, EXT>,
      (const void*)k_sweep<BD_REFLECT, false, false, STATS, EXT>,
      (const void*)k_sweep<BD_REFLECT, true, false, STATS, EXT>,
      (const void*)k_sweep<BD_NEAREST, false, true, STATS, EXT>,
      (const void*)k_sweep<BD_NEAREST, true, true, STATS, EXT>,
      (const void*)k_sweep<BD_MIRROR, false, true, STATS, EXT>,
      (const void*)k_sweep<BD_MIRROR, true, true, STATS, EXT>,
      (const void*)k_sweep<BD_REFLECT, false, true, STATS, EXT>,
      (const void*)k_sweep<BD_REFLECT, true, true, STATS, EXT>};
  for (const void* k : kernels) SVS_HIP(hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, v));
  return 0;
}

// Dynamic LDS above 64 KiB must be enabled per kernel.
static int allow_large_lds(const svs_plan* pl) {
  // shape_ok keeps every tile within 96 KiB; enable that much once.
  (void)pl;
  const int v = 96 * 1024;
  if (int rc = allow_sweep_lds<false, false>(v)) return rc;
  if (int rc = allow_sweep_lds<true, false>(v)) return rc;
  if (int rc = allow_sweep_lds<false, true>(v)) return rc;
  if (int rc = allow_sweep_lds<true, true>(v)) return rc;
  // The resident sweep holds 48 bytes per voxel of up to 2048 voxels.
  if (int rc = allow_resident_lds<BD_NEAREST, false>()) return rc;
  if (int rc = allow_resident_lds<BD_NEAREST, true>()) return rc;
  if (int rc = allow_resident_lds<BD_MIRROR, false>()) return rc;
  if (int rc = allow_resident_lds<BD_MIRROR, true>()) return rc;
  if (int rc = allow_resident_lds<BD_REFLECT, false>()) return rc;
  if (int rc = allow_resident_lds<BD_REFLECT, true>()) return rc;
  return 0;
}

// Bytes of partial sums one warp's statistics need at the current launch
// shapes: K iterations x (one float4 per marching work-group, or per wave of
// a resident work-group), the largest over the levels.
static size_t stats_part_need(const svs_plan* pl) {
  const size_t K = (size_t)std::max(1, pl->P.max_outer_iterations);
  size_t need = 0;
  for (const LevelPlan& L : pl->levels) {
    need = std::max(need, K * (size_t)L.ntiles * sizeof(float4));
    if (L.res_capable) need = std::max(need, K * (size_t)pl->ninst * (L.res_threads / 64) * sizeof(float4));
  }
  return need;
}

// With statistics on, (re)allocates the partial-sum buffer if the current
// launch shapes need more than it holds (after the plan's previous work, which
// may still read it; the captured schedule then holds the old address).
static int ensure_stats_part(svs_plan* pl) {
  if (!pl->stats_out) return 0;
  const size_t need = stats_part_need(pl);
  if (need <= pl->stats_part_bytes) return 0;
  SVS_HIP(hipEventSynchronize(pl->last_work));
  if (pl->stats_part) {
    (void)hipFree(pl->stats_part);
    pl->stats_part = nullptr;
    pl->stats_part_bytes = 0;
  }
  SVS_HIP(hipMalloc((void**)&pl->stats_part, need));
  pl->stats_part_bytes = need;
  drop_graph(pl);
  return 0;
}

// Enqueues the complete solve of every instance on one stream.
static int enqueue_solve(svs_plan* pl, hipStream_t st) {
  const svs_tvl1_params& P = pl->P;
  const int S = pl->S, n = pl->ninst;
  const int ip = P.interpolation, bc = P.boundary;
  if (P.use_blur) {
    LevelPlan& L0 = pl->levels[0];
    k_blur_pair<<<L0.nblocks, kBlock, 0, st>>>(L0.dev, L0.block_start, n, pl->blur_dev, P.blur_K);
    SVS_LAUNCHED();
  }
  for (int s = 1; s < S; ++s) {
    LevelPlan& L = pl->levels[s];
    k_restrict<<<L.nblocks, kBlock, 0, st>>>(pl->levels[s - 1].dev, L.dev, L.block_start, n, ip, bc);
    SVS_LAUNCHED();
  }
  int parity = 0;
  for (int s = S - 1; s >= 0; --s) {
    LevelPlan& L = pl->levels[s];
    k_level_setup<<<L.nblocks, kBlock, 0, st>>>(L.dev, L.block_start, n, P.use_aniso, P.alpha, P.beta, bc);
    SVS_LAUNCHED();
    if (s == S - 1) {
      const int from_input = (S == 1 && P.use_initial_state) ? 1 : 0;
      k_init_state<<<L.nblocks, kBlock, 0, st>>>(L.dev, L.block_start, n, from_input);
      SVS_LAUNCHED();
    }
    parity = 0;
    const int K = P.max_outer_iterations;
    const double elam = P.relax_theta != 0.0 ? std::nan("") : (double)P.weight_matching;
    for (int w = 0; w < P.max_warps; ++w) {
      k_linearise<<<L.nblocks, kBlock, 0, st>>>(L.dev, L.block_start, n, parity, ip, bc);
      SVS_LAUNCHED();
      // Statistics of this warp's K iterations: out[s][w][n][k][3].
      float* sout = pl->stats_out ? pl->stats_out + ((size_t)s * P.max_warps + w) * K * n * 3 : nullptr;
      float4* part = sout ? pl->stats_part : nullptr;
      if (L.resident && K > 0) {
        launch_resident(pl, L, parity, 0, K, st, part);
        SVS_LAUNCHED();
        if (sout) {
          k_stats_reduce<<<K * n, 64, 0, st>>>(part, nullptr, nullptr, n, L.res_threads / 64, n,
                                               elam, (double)P.weight_tv, sout);
          SVS_LAUNCHED();
        }
        parity ^= K & 1;
        continue;
      }
      for (int it = 0; it < K; ++it) {
        launch_sweep(pl, L, parity, it, st, part ? part + (size_t)it * L.ntiles : nullptr);
        SVS_LAUNCHED();
        parity ^= 1;
      }
      if (sout && K > 0) {
        k_stats_reduce<<<K * n, 64, 0, st>>>(part, L.tile_csr, L.tile_csr + n + 1, L.ntiles, 1, n,
                                             elam, (double)P.weight_tv, sout);
        SVS_LAUNCHED();
      }
    }
    if (s > 0) {
      LevelPlan& F = pl->levels[s - 1];
      if (ip == IP_CUBIC) {
        k_prolong_x<<<F.px_blocks, kBlock, 0, st>>>(L.dev, F.dev, F.px_start, n, parity, bc);
        SVS_LAUNCHED();
        k_prolong_y<<<F.py_blocks, kBlock, 0, st>>>(L.dev, F.dev, F.py_start, n, bc);
        SVS_LAUNCHED();
        k_prolong_z<<<F.nblocks, kBlock, 0, st>>>(L.dev, F.dev, F.block_start, n, bc);
      } else {
        k_prolong_state<<<F.nblocks, kBlock, 0, st>>>(L.dev, F.dev, F.block_start, n, parity, ip, bc);
      }
      SVS_LAUNCHED();
    }
  }
  LevelPlan& L0 = pl->levels[0];
  k_store_output<<<L0.nblocks, kBlock, 0, st>>>(L0.dev, L0.block_start, n, parity);
  SVS_LAUNCHED();
  return 0;
}

extern "C" {

size_t svs_tvl1_workspace_bytes(const svs_tvl1_params* P, int ninst, const int* dims) {
  if (validate_params(P) || validate_dims(P, ninst, dims)) return 0;
  svs_plan tmp;
  tmp.P = *P;
  tmp.ninst = ninst;
  tmp.S = P->num_scales;
  tmp.dims.assign(dims, dims + 3 * ninst);
  if (P->use_blur) tmp.blur_w.assign(P->blur_weights, P->blur_weights + (size_t)P->blur_K * P->blur_K * P->blur_K);
  size_t total = 0;
  plan_layout(&tmp, nullptr, &total);
  return total;
}

int svs_tvl1_plan_create(const svs_tvl1_params* P, int ninst, const int* dims, void* workspace, size_t ws_bytes,
                         void* stream, svs_plan** out) {
  if (int rc = validate_params(P)) return rc;
  if (int rc = validate_dims(P, ninst, dims)) return rc;
  SVS_REQUIRE(out, "null plan output");
  SVS_REQUIRE(workspace, "null workspace");
  svs_plan* pl = new (std::nothrow) svs_plan();
  SVS_REQUIRE(pl, "out of host memory");
  pl->P = *P;
  pl->ninst = ninst;
  pl->S = P->num_scales;
  pl->dims.assign(dims, dims + 3 * ninst);
  if (P->use_blur) pl->blur_w.assign(P->blur_weights, P->blur_weights + (size_t)P->blur_K * P->blur_K * P->blur_K);
  size_t need = 0;
  plan_layout(pl, nullptr, &need);
  if (ws_bytes < need) {
    delete pl;
    return fail(SVS_STATUS_INVALID, "workspace too small: %zu bytes given, %zu needed", ws_bytes, need);
  }
  pl->ws = (char*)workspace;
  pl->ws_bytes = ws_bytes;
  size_t total = 0;
  plan_layout(pl, pl->ws, &total);
  pl->P.blur_weights = nullptr;  // host copy kept in blur_w
  for (int s = 0; s < pl->S; ++s) {
    LevelPlan& L = pl->levels[s];
    if (!shape_ok(L, L.bx, L.by, L.cz)) {
      delete pl;
      return fail(SVS_STATUS_INVALID, "sweep shape %dx%d, z-chunk %d not launchable for level %d", L.bx, L.by,
                  L.cz, s);
    }
  }
  fill_schedule(pl);
  if (int rc = allow_large_lds(pl)) { delete pl; return rc; }
  if (int rc = create_sync_objects(pl)) { free_plan(pl); return rc; }
  hipStream_t st = (hipStream_t)stream;
  if (int rc = upload_static(pl, st)) { free_plan(pl); return rc; }
  if (int rc = mark_last_work(pl, st)) { free_plan(pl); return rc; }
  *out = pl;
  return 0;
}

int svs_tvl1_plan_bind(svs_plan* pl, int k, const float* I0, const float* I1, const float* u_in,
                       const float* p_in, float* u_out, float* p_out) {
  SVS_REQUIRE(pl && k >= 0 && k < pl->ninst, "plan/instance index out of range");
  SVS_REQUIRE(I0 && I1, "instance %d: null input frame", k);
  LevelDesc& d = pl->levels[0].host[k];
  d.src0 = I0;
  d.src1 = I1;
  if (!pl->P.use_blur) {
    d.I0 = const_cast<float*>(I0);
    d.I1 = const_cast<float*>(I1);
  }
  d.u_in = u_in;
  d.p_in = p_in;
  d.u_out = u_out;
  d.p_out = p_out;
  if (pl->S == 1 && pl->P.use_initial_state) SVS_REQUIRE(u_in && p_in, "instance %d: initial state required", k);
  pl->tables_dirty = true;
  return 0;
}

int svs_tvl1_plan_run(svs_plan* pl, int use_graph, void* stream) {
  SVS_REQUIRE(pl, "null plan");
  hipStream_t st = (hipStream_t)stream;
  for (int k = 0; k < pl->ninst; ++k)
    SVS_REQUIRE(pl->levels[0].host[k].src0, "instance %d has no bound input", k);
  if (int rc = ensure_stats_part(pl)) return rc;
  if (int rc = after_last_work(pl, st)) return rc;
  if (pl->tables_dirty)
    if (int rc = upload_tables(pl, st)) return rc;
  if (!use_graph) {
    if (int rc = enqueue_solve(pl, st)) return rc;
    return mark_last_work(pl, st);
  }
  if (!pl->exec) {
    if (!pl->capture_stream) SVS_HIP(hipStreamCreateWithFlags(&pl->capture_stream, hipStreamNonBlocking));
    SVS_HIP(hipStreamBeginCapture(pl->capture_stream, hipStreamCaptureModeThreadLocal));
    const int rc = enqueue_solve(pl, pl->capture_stream);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(pl->capture_stream, &g);
    if (rc) { if (g) (void)hipGraphDestroy(g); return rc; }
    if (e != hipSuccess) return fail(SVS_STATUS_DEVICE, "hipStreamEndCapture: %s", hipGetErrorString(e));
    pl->graph = g;
    SVS_HIP(hipGraphInstantiate(&pl->exec, pl->graph, nullptr, nullptr, 0));
  }
  SVS_HIP(hipGraphLaunch(pl->exec, st));
  return mark_last_work(pl, st);
}

int svs_tvl1_plan_destroy(svs_plan* pl) {
  if (!pl) return 0;
  free_plan(pl);
  return 0;
}

int svs_tvl1_plan_info(const svs_plan* pl, int level, int* out /* 8 ints */) {
  SVS_REQUIRE(pl && level >= 0 && level < pl->S && out, "bad plan/level");
  const LevelPlan& L = pl->levels[level];
  out[0] = L.bx; out[1] = L.by; out[2] = L.cz; out[3] = L.ntiles; out[4] = L.nblocks;
  out[5] = pl->ninst; out[6] = pl->S; out[7] = L.resident;
  return 0;
}

// Enqueues `count` fused sweeps of one level on the state a previous run
// left behind (ping-pong, iteration scalars cycling through the schedule).
// Used to time the sweep kernel on its own.
int svs_tvl1_plan_sweeps(svs_plan* pl, int level, int count, void* stream) {
  SVS_REQUIRE(pl && level >= 0 && level < pl->S && count >= 0, "bad plan/level/count");
  SVS_REQUIRE(!pl->tables_dirty, "run the plan once before timing its sweeps");
  SVS_REQUIRE(pl->P.max_outer_iterations > 0, "no iterations configured");
  hipStream_t st = (hipStream_t)stream;
  if (int rc = after_last_work(pl, st)) return rc;
  const LevelPlan& L = pl->levels[level];
  if (L.resident) {  // the level's iterations run in one resident launch
    if (count > 0) {
      launch_resident(pl, L, 0, 0, count, st);
      SVS_LAUNCHED();
    }
    return mark_last_work(pl, st);
  }
  int parity = 0;
  for (int i = 0; i < count; ++i) {
    launch_sweep(pl, L, parity, i % pl->P.max_outer_iterations, st);
    SVS_LAUNCHED();
    parity ^= 1;
  }
  return mark_last_work(pl, st);
}

// Times candidate sweep shapes per level on the plan's own state and keeps
// the fastest (stage 1: tile shape with a default z-chunk; stage 2: z-chunk).
// Results are remembered for later plans of the same geometry.  Drops the
// captured graph, which the next graph run re-captures.
int svs_tvl1_plan_autotune(svs_plan* pl, int reps, void* stream, int* report /* 3 per level, may be NULL */) {
  SVS_REQUIRE(pl && reps >= 1, "bad plan/reps");
  SVS_REQUIRE(!pl->tables_dirty, "run the plan once before autotuning");
  SVS_REQUIRE(pl->P.max_outer_iterations > 0, "no iterations configured");
  hipStream_t st = (hipStream_t)stream;
  if (int r = after_last_work(pl, st)) return r;
  hipEvent_t e0, e1;
  SVS_HIP(hipEventCreate(&e0));
  SVS_HIP(hipEventCreate(&e1));
  int rc = 0;
  for (int s = 0; s < pl->S && !rc; ++s) {
    LevelPlan& L = pl->levels[s];
    auto time_shape = [&](int bx, int by, int cz, float* ms) -> int {
      L.bx = bx; L.by = by; L.cz = cz;
      if (int r = upload_tiles(pl, s, st)) return r;
      launch_sweep(pl, L, 0, 0, st);
      SVS_LAUNCHED();
      SVS_HIP(hipEventRecord(e0, st));
      for (int i = 0; i < reps; ++i) launch_sweep(pl, L, i & 1, i % pl->P.max_outer_iterations, st);
      SVS_LAUNCHED();
      SVS_HIP(hipEventRecord(e1, st));
      SVS_HIP(hipEventSynchronize(e1));
      SVS_HIP(hipEventElapsedTime(ms, e0, e1));
      return 0;
    };
    // Stage 1: every tile with its default z-chunk.  Stage 2: every z-chunk
    // for the five fastest tiles of stage 1 (shapes a few per cent apart
    // change rank between repetitions).  Stage 3 below re-times the best.
    SweepShape best{L.bx, L.by, L.cz};
    float best_ms = 0.0f;
    if ((rc = time_shape(best.bx, best.by, best.cz, &best_ms))) break;
    std::vector<std::pair<float, SweepShape>> ranked = {{best_ms, best}};
    for (const SweepShape& c : tile_candidates(L)) {
      float ms = 0.0f;
      if ((rc = time_shape(c.bx, c.by, c.cz, &ms))) break;
      ranked.push_back({ms, c});
      if (ms < best_ms) { best_ms = ms; best = c; }
    }
    if (rc) break;
    std::sort(ranked.begin(), ranked.end(),
              [](const std::pair<float, SweepShape>& a, const std::pair<float, SweepShape>& b) { return a.first < b.first; });
    std::vector<std::pair<float, SweepShape>> timed = ranked;
    for (size_t r = 0; r < ranked.size() && r < 5 && !rc; ++r) {
      const SweepShape t = ranked[r].second;
      for (int cz : cz_candidates(L)) {
        float ms = 0.0f;
        if ((rc = time_shape(t.bx, t.by, cz, &ms))) break;
        timed.push_back({ms, SweepShape{t.bx, t.by, cz}});
        if (ms < best_ms) { best_ms = ms; best = SweepShape{t.bx, t.by, cz}; }
      }
    }
    if (rc) break;
    // Stage 3: the four fastest distinct shapes of stages 1 and 2 timed again,
    // twice each in alternating order, so that one lucky repetition does not
    // decide between shapes a few per cent apart.
    std::sort(timed.begin(), timed.end(),
              [](const std::pair<float, SweepShape>& a, const std::pair<float, SweepShape>& b) { return a.first < b.first; });
    std::vector<SweepShape> finalists;
    for (auto& e : timed) {
      bool dup = false;
      for (auto& f : finalists) dup |= f.bx == e.second.bx && f.by == e.second.by && f.cz == e.second.cz;
      if (!dup) finalists.push_back(e.second);
      if (finalists.size() == 4) break;
    }
    std::vector<float> total(finalists.size(), 0.0f);
    for (int round = 0; round < 2 && !rc; ++round)
      for (size_t f = 0; f < finalists.size() && !rc; ++f) {
        const size_t i = round ? finalists.size() - 1 - f : f;
        float ms = 0.0f;
        if ((rc = time_shape(finalists[i].bx, finalists[i].by, finalists[i].cz, &ms))) break;
        total[i] += ms;
      }
    if (rc) break;
    size_t win = 0;
    for (size_t f = 1; f < finalists.size(); ++f)
      if (total[f] < total[win]) win = f;
    best = finalists[win];
    best_ms = total[win] / 2;
    L.bx = best.bx; L.by = best.by; L.cz = best.cz;
    if ((rc = upload_tiles(pl, s, st))) break;
    if (report) { report[3 * s] = best.bx; report[3 * s + 1] = best.by; report[3 * s + 2] = best.cz; }
    // A level that qualifies for the resident sweep: one launch of a warp's
    // K iterations against K marching sweeps of the best shape, per iteration.
    bool resident_wins = false;
    if (L.res_capable) {
      const int K = pl->P.max_outer_iterations;
      auto time_resident = [&](float* ms) -> int {
        launch_resident(pl, L, 0, 0, K, st);
        SVS_LAUNCHED();
        SVS_HIP(hipEventRecord(e0, st));
        launch_resident(pl, L, 0, 0, K, st);
        SVS_LAUNCHED();
        SVS_HIP(hipEventRecord(e1, st));
        SVS_HIP(hipEventSynchronize(e1));
        SVS_HIP(hipEventElapsedTime(ms, e0, e1));
        return 0;
      };
      float ms_res = 0.0f;
      L.resident = L.res_capable;
      if ((rc = time_resident(&ms_res))) break;
      resident_wins = ms_res / K < best_ms / reps;
      L.resident = resident_wins ? L.res_capable : 0;
    }
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tuned[tune_key(pl->P, L)] = best;
    if (L.res_capable) g_tuned_resident[tune_key(pl->P, L)] = resident_wins;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  drop_graph(pl);
  if (!rc) rc = mark_last_work(pl, st);
  return rc;
}

int svs_tvl1_plan_set_shape(svs_plan* pl, int level, int bx, int by, int cz, void* stream) {
  SVS_REQUIRE(pl && level >= 0 && level < pl->S, "bad plan/level");
  LevelPlan& L = pl->levels[level];
  SVS_REQUIRE(shape_ok(L, bx, by, cz), "sweep shape %dx%d, z-chunk %d is outside the kernel's limits", bx, by, cz);
  SVS_REQUIRE(count_tiles(L, bx, by, cz) <= L.tile_capacity,
              "sweep shape %dx%d, z-chunk %d needs %d tiles; the plan reserved %d", bx, by, cz,
              count_tiles(L, bx, by, cz), L.tile_capacity);
  hipStream_t st = (hipStream_t)stream;
  if (int rc = after_last_work(pl, st)) return rc;
  L.bx = bx; L.by = by; L.cz = cz;
  if (int rc = upload_tiles(pl, level, st)) return rc;
  drop_graph(pl);
  return mark_last_work(pl, st);
}

int svs_tvl1_plan_set_stats(svs_plan* pl, float* out, void* stream) {
  SVS_REQUIRE(pl, "null plan");
  const int64_t groups = (int64_t)pl->P.max_outer_iterations * pl->ninst;
  SVS_REQUIRE(groups < ((int64_t)1 << 31), "max_outer_iterations x instances too large for statistics");
  hipStream_t st = (hipStream_t)stream;
  if (int rc = after_last_work(pl, st)) return rc;
  if (out != pl->stats_out) {
    // The previous output may still be written by enqueued runs (on any
    // stream): wait for them, so that the caller may free it on return.
    if (pl->stats_out) SVS_HIP(hipEventSynchronize(pl->last_work));
    drop_graph(pl);
  }
  pl->stats_out = out;
  if (!out && pl->stats_part) {  // off: release the partial sums once no work reads them
    SVS_HIP(hipEventSynchronize(pl->last_work));
    (void)hipFree(pl->stats_part);
    pl->stats_part = nullptr;
    pl->stats_part_bytes = 0;
  }
  if (int rc = ensure_stats_part(pl)) return rc;
  return mark_last_work(pl, st);
}

int svs_tvl1_plan_set_resident(svs_plan* pl, int level, int mode, void* stream) {
  SVS_REQUIRE(pl && level >= 0 && level < pl->S, "bad plan/level");
  SVS_REQUIRE(mode >= -1 && mode <= 1, "mode must be -1 (automatic), 0 (marching sweep) or 1 (resident)");
  LevelPlan& L = pl->levels[level];
  SVS_REQUIRE(mode != 1 || L.res_capable, "level %d: an instance exceeds the %d voxels the resident sweep holds",
              level, 1024 * kResidentPerThread);
  hipStream_t st = (hipStream_t)stream;
  if (int rc = after_last_work(pl, st)) return rc;
  L.resident = mode == 0 ? 0 : L.res_capable;
  drop_graph(pl);
  return mark_last_work(pl, st);
}

// One frame pair, enqueued directly (no graph): the plain drop-in.
int svs_tvl1_solve(const svs_tvl1_params* P, const float* I0, const float* I1, int NZ, int NY, int NX,
                   const float* u_in, const float* p_in, float* u_out, float* p_out, void* workspace,
                   size_t ws_bytes, void* stream) {
  const int dims[3] = {NZ, NY, NX};
  svs_plan* pl = nullptr;
  if (int rc = svs_tvl1_plan_create(P, 1, dims, workspace, ws_bytes, stream, &pl)) return rc;
  int rc = svs_tvl1_plan_bind(pl, 0, I0, I1, u_in, p_in, u_out, p_out);
  if (!rc) rc = svs_tvl1_plan_run(pl, 0, stream);
  svs_tvl1_plan_destroy(pl);
  return rc;
}

}  // extern "C"
