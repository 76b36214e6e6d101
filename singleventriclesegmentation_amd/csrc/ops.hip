// Drop-in grid operators behind the opticalFlow operator API: central
// differences and their adjoints, the anisotropic tensor, the two proxes,
// warping (scalar, vector field, batched CNN form and the adjoint),
// resampling between meshes and the Gaussian pre-blur.
//
// One thread per output voxel; x is the fastest index, so consecutive lanes
// of a 64-wide wavefront touch consecutive addresses.  Every kernel writes
// every element of its output, so callers may hand in uninitialised memory.
#include "svs_common.hpp"
#include "svs_device.hpp"
#include "svs_math.h"

#include <string.h>

namespace svs {

thread_local char g_last_error[1024] = "";

int fail(int code, const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_last_error, sizeof(g_last_error), fmt, ap);
  va_end(ap);
  return code;
}

constexpr int kBlock = 256;

// ------------------------------------------------------------------ fill
__global__ __launch_bounds__(kBlock) void k_fill(float* __restrict__ p, int64_t n, float v) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i < n) p[i] = v;
}

// ------------------------------------------------- central differences
// Nabla3D_CD.forward / forwardVectorField (differentialOps.cu:540-656,
// :782-907): b [Z,Y,X,C] -> out [Z,Y,X,C,3].
__global__ __launch_bounds__(kBlock) void k_nabla_fwd(const float* __restrict__ b, int C, int NZ, int NY, int NX, float hZ, float hY,
                            float hX, int bc, float* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= (int64_t)NZ * NY * NX) return;
  int z, y, x;
  unflatten(i, NY, NX, z, y, x);
  const int64_t sx = C, sy = (int64_t)NX * C, sz = (int64_t)NY * NX * C;
  for (int c = 0; c < C; ++c) {
    const float* p = b + i * C + c;
    const float b0 = p[0];
    const float xm = x > 0 ? p[-sx] : 0.0f, xp = x < NX - 1 ? p[sx] : 0.0f;
    const float ym = y > 0 ? p[-sy] : 0.0f, yp = y < NY - 1 ? p[sy] : 0.0f;
    const float zm = z > 0 ? p[-sz] : 0.0f, zp = z < NZ - 1 ? p[sz] : 0.0f;
    float* o = out + (i * C + c) * 3;
    o[0] = cd_fwd(xm, b0, xp, x, NX, hX, bc);
    o[1] = cd_fwd(ym, b0, yp, y, NY, hY, bc);
    o[2] = cd_fwd(zm, b0, zp, z, NZ, hZ, bc);
  }
}

// Nabla3D_CD.backward / backwardVectorField (differentialOps.cu:686-772,
// :909-1044): p [Z,Y,X,C,3] -> out [Z,Y,X,C] = ax/hX + ay/hY + az/hZ.
__global__ __launch_bounds__(kBlock) void k_nabla_bwd(const float* __restrict__ p, int C, int NZ, int NY, int NX, float hZ, float hY,
                            float hX, int bc, float* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= (int64_t)NZ * NY * NX) return;
  int z, y, x;
  unflatten(i, NY, NX, z, y, x);
  const int64_t sx = 3 * C, sy = (int64_t)NX * 3 * C, sz = (int64_t)NY * NX * 3 * C;
  for (int c = 0; c < C; ++c) {
    const float* q = p + (i * C + c) * 3;
    const float ax = cd_adj(x > 0 ? q[-sx] : 0.0f, q[0], x < NX - 1 ? q[sx] : 0.0f, x, NX, bc);
    const float ay = cd_adj(y > 0 ? q[1 - sy] : 0.0f, q[1], y < NY - 1 ? q[1 + sy] : 0.0f, y, NY, bc);
    const float az = cd_adj(z > 0 ? q[2 - sz] : 0.0f, q[2], z < NZ - 1 ? q[2 + sz] : 0.0f, z, NZ, bc);
    out[i * C + c] = ax / hX + ay / hY + az / hZ;
  }
}

// ------------------------------------------------ anisotropic tensor
// AnisotropicNabla3D.computeTangentVecs (anisotropicDifferentialOps.cu:90-141).
__global__ __launch_bounds__(kBlock) void k_aniso_tangents(const float* __restrict__ G, int64_t N, float alpha, float beta,
                                 float* __restrict__ s, float* __restrict__ n, float* __restrict__ t1,
                                 float* __restrict__ t2) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= N) return;
  const float g[3] = {G[3 * i], G[3 * i + 1], G[3 * i + 2]};
  float ns = 0.0f;
  for (int d = 0; d < 3; ++d) ns += g[d] * g[d];
  const float nrm = sqrtf(ns);
  if (nrm > 0.0f) {
    s[i] = svs_expf(-alpha * svs_powf(nrm, beta));
    float nn[3], tt[3];
    for (int d = 0; d < 3; ++d) nn[d] = g[d] / nrm;
    tangent1(nn, tt);
    for (int d = 0; d < 3; ++d) { n[3 * i + d] = nn[d]; t1[3 * i + d] = tt[d]; }
    t2[3 * i + 0] = nn[1] * tt[2] - nn[2] * tt[1];
    t2[3 * i + 1] = nn[2] * tt[0] - nn[0] * tt[2];
    t2[3 * i + 2] = nn[0] * tt[1] - nn[1] * tt[0];
  } else {
    s[i] = 0.0f;
    for (int d = 0; d < 3; ++d) { n[3 * i + d] = 0.0f; t1[3 * i + d] = 0.0f; t2[3 * i + d] = 0.0f; }
  }
}

// AnisotropicNabla3D.forwardVectorField / backwardVectorField
// (anisotropicDifferentialOps.cu:144-175); the t2 factor is a literal 0.
__global__ __launch_bounds__(kBlock) void k_aniso_apply(const float* __restrict__ in, const float* __restrict__ s,
                              const float* __restrict__ n, const float* __restrict__ t1,
                              const float* __restrict__ t2, int64_t N, float* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= N) return;
  const float nn[3] = {n[3 * i], n[3 * i + 1], n[3 * i + 2]};
  const float tt[3] = {t1[3 * i], t1[3 * i + 1], t1[3 * i + 2]};
  const float t2v[3] = {t2[3 * i], t2[3 * i + 1], t2[3 * i + 2]};
  const float si = s[i];
  const float t2f = 0.0f;
  for (int c = 0; c < 3; ++c) {
    const float q[3] = {in[9 * i + 3 * c], in[9 * i + 3 * c + 1], in[9 * i + 3 * c + 2]};
    float o[3];
    aniso_row(si, nn, tt, q, o);
    for (int d = 0; d < 3; ++d) out[9 * i + 3 * c + d] = o[d] + t2f * t2v[d];
  }
}

// ------------------------------------------------------------ proxes
// TVL1OF3D_proxPrimal (opticalFlowOps.cu:123-175).
__global__ __launch_bounds__(kBlock) void k_prox_primal(const float* __restrict__ primal, float factor, const float* __restrict__ rho,
                              const float* __restrict__ g, int64_t N, float* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= N) return;
  float v[3] = {primal[3 * i], primal[3 * i + 1], primal[3 * i + 2]};
  const float gg[3] = {g[3 * i], g[3 * i + 1], g[3 * i + 2]};
  prox_primal_point(v, rho[i], gg, factor);
  for (int c = 0; c < 3; ++c) out[3 * i + c] = v[c];
}

// TVL1OF3D_proxDual (opticalFlowOps.cu:179-208).
__global__ __launch_bounds__(kBlock) void k_prox_dual(const float* __restrict__ dual, float wtv, int64_t N, float* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= N) return;
  for (int c = 0; c < 3; ++c) {
    float v[3] = {dual[9 * i + 3 * c], dual[9 * i + 3 * c + 1], dual[9 * i + 3 * c + 2]};
    prox_dual_row(v, wtv);
    for (int d = 0; d < 3; ++d) out[9 * i + 3 * c + d] = v[d];
  }
}

// ------------------------------------------------------------- warping
// One strided scalar component of a [Z,Y,X,C] (or [B,C,Z,Y,X]) field.
struct StridedFetch {
  const float* base;
  int64_t stride;
  int NY, NX;
  __device__ void operator()(int z, int y, int x, float* v) const {
    v[0] = base[(((int64_t)z * NY + y) * NX + x) * stride];
  }
};

// Warping3D.forward / forwardVectorField (warpingOps.cu:572-596, :632-660):
// out(i) = u(i*h + phi(i)), u [Z,Y,X,C].
__global__ __launch_bounds__(kBlock) void k_warp(const float* __restrict__ u, int C, const float* __restrict__ phi, int NZ, int NY,
                       int NX, float hZ, float hY, float hX, int ip, int bc, float* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= (int64_t)NZ * NY * NX) return;
  int z, y, x;
  unflatten(i, NY, NX, z, y, x);
  const float cx = (float)x * hX + phi[3 * i + 0];
  const float cy = (float)y * hY + phi[3 * i + 1];
  const float cz = (float)z * hZ + phi[3 * i + 2];
  for (int c = 0; c < C; ++c) {
    StridedFetch f{u + c, C, NY, NX};
    float v[1];
    sample<1>(ip, f, NZ, NY, NX, hZ, hY, hX, bc, cz, cy, cx, v);
    out[i * C + c] = v[0];
  }
}

// WarpingCNN3D.forward (warpingOpsCNN.cu:97-128): u [B,C,Z,Y,X], phi [B,Z,Y,X,3].
__global__ __launch_bounds__(kBlock) void k_warp_cnn(const float* __restrict__ u, const float* __restrict__ phi, int B, int C, int NZ,
                           int NY, int NX, float hZ, float hY, float hX, int ip, int bc, float* __restrict__ out) {
  const int64_t N = (int64_t)NZ * NY * NX;
  const int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (t >= (int64_t)B * N) return;
  const int b = (int)(t / N);
  const int64_t i = t - (int64_t)b * N;
  int z, y, x;
  unflatten(i, NY, NX, z, y, x);
  const float* ph = phi + ((int64_t)b * N + i) * 3;
  const float cx = (float)x * hX + ph[0];
  const float cy = (float)y * hY + ph[1];
  const float cz = (float)z * hZ + ph[2];
  for (int c = 0; c < C; ++c) {
    const int64_t off = ((int64_t)b * C + c) * N;
    StridedFetch f{u + off, 1, NY, NX};
    float v[1];
    sample<1>(ip, f, NZ, NY, NX, hZ, hY, hX, bc, cz, cy, cx, v);
    out[off + i] = v[0];
  }
}

// Adjoint of one tricubic sample (interpolation_cubicHermiteSpline.cuh:527-636):
// scatters fwd * d out/d u into grad (atomics, 64 taps) and adds the three
// flow derivatives to gphi in the reference's order (z, y per dz, x per dy).
template <class Acc>
__device__ void warp_backward_cubic(const float* __restrict__ u, const Acc& acc, int NZ, int NY, int NX,
                                    float hZ, float hY, float hX, int bc, float cz, float cy, float cx, float fwd,
                                    float gphi[3]) {
  const Axis ax = make_axis(cx, hX), ay = make_axis(cy, hY), az = make_axis(cz, hZ);
  const float lx = ax.w + 1, ly = ay.w + 1, lz = az.w + 1;
  int xs[4], ys[4], zs[4];
  for (int k = 0; k < 4; ++k) {
    xs[k] = bidx(ax.f - 1 + k, NX, bc);
    ys[k] = bidx(ay.f - 1 + k, NY, bc);
    zs[k] = bidx(az.f - 1 + k, NZ, bc);
  }
  auto U = [&](int z, int y, int x) { return u[((int64_t)z * NY + y) * NX + x]; };
  float bz[4] = {0.0f, 0.0f, 0.0f, 0.0f}, by[4];
  float byz[4][4] = {};  // x-interpolated values per (dz, dy)
  // Plane by plane (a loop that is not unrolled, results placed by selects):
  // one plane's 16 taps in flight.
#pragma unroll 1
  for (int dz = 0; dz < 4; ++dz) {
    int z = zs[0];
#pragma unroll
    for (int k = 1; k < 4; ++k) z = dz == k ? zs[k] : z;
    float r[4];
#pragma unroll
    for (int dy = 0; dy < 4; ++dy) {
      const int y = ys[dy];
      r[dy] = cr_local(U(z, y, xs[0]), U(z, y, xs[1]), U(z, y, xs[2]), U(z, y, xs[3]), lx);
    }
    const float b = cr_local(r[0], r[1], r[2], r[3], ly);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      bz[k] = dz == k ? b : bz[k];
#pragma unroll
      for (int dy = 0; dy < 4; ++dy) byz[k][dy] = dz == k ? r[dy] : byz[k][dy];
    }
  }
  float gz[4];
  const float gwz = cr_local_backward(bz[0], bz[1], bz[2], bz[3], lz, fwd, gz);
  gphi[2] += gwz / hZ;
  // A loop that is not unrolled: the taps are read again per plane instead
  // of all 64 staying live from the forward pass (which spilled at the
  // work-group's 128-VGPR budget).  The plane's values are picked by selects;
  // the accumulation order is unchanged.
#pragma unroll 1
  for (int dz = 0; dz < 4; ++dz) {
    float gzd = gz[0];
    int z = zs[0];
#pragma unroll
    for (int k = 1; k < 4; ++k) {
      gzd = dz == k ? gz[k] : gzd;
      z = dz == k ? zs[k] : z;
    }
#pragma unroll
    for (int dy = 0; dy < 4; ++dy) {
      by[dy] = byz[0][dy];
#pragma unroll
      for (int k = 1; k < 4; ++k) by[dy] = dz == k ? byz[k][dy] : by[dy];
    }
    float gy[4];
    const float gwy = cr_local_backward(by[0], by[1], by[2], by[3], ly, gzd, gy);
    gphi[1] += gwy / hY;
#pragma unroll 1
    for (int dy = 0; dy < 4; ++dy) {
      int y = ys[0];
      float gyd = gy[0];
#pragma unroll
      for (int k = 1; k < 4; ++k) {
        y = dy == k ? ys[k] : y;
        gyd = dy == k ? gy[k] : gyd;
      }
      float gx[4];
      const float gwx = cr_local_backward(U(z, y, xs[0]), U(z, y, xs[1]), U(z, y, xs[2]), U(z, y, xs[3]), lx,
                                          gyd, gx);
      gphi[0] += gwx / hX;
      for (int dx = 0; dx < 4; ++dx) acc(z, y, xs[dx], gx[dx]);
    }
  }
}

// Adjoint of one trilinear sample (interpolation_linear.cuh:297-363).
template <class Acc>
__device__ void warp_backward_linear(const float* __restrict__ u, const Acc& acc, int NZ, int NY, int NX,
                                     float hZ, float hY, float hX, int bc, float cz, float cy, float cx, float fv,
                                     float gphi[3]) {
  const Axis ax = make_axis(cx, hX), ay = make_axis(cy, hY), az = make_axis(cz, hZ);
  const float wx = ax.w, wy = ay.w, wz = az.w;
  const int xf = bidx(ax.f, NX, bc), xc = bidx(ax.f + 1, NX, bc);
  const int yf = bidx(ay.f, NY, bc), yc = bidx(ay.f + 1, NY, bc);
  const int zf = bidx(az.f, NZ, bc), zc = bidx(az.f + 1, NZ, bc);
  auto I = [&](int z, int y, int x) { return ((int64_t)z * NY + y) * NX + x; };
  const float u_fff = u[I(zf, yf, xf)], u_ffc = u[I(zc, yf, xf)], u_fcf = u[I(zf, yc, xf)], u_fcc = u[I(zc, yc, xf)];
  const float u_cff = u[I(zf, yf, xc)], u_cfc = u[I(zc, yf, xc)], u_ccf = u[I(zf, yc, xc)], u_ccc = u[I(zc, yc, xc)];
  acc(zf, yf, xf, (1 - wz) * (1 - wy) * (1 - wx) * fv);
  acc(zf, yc, xf, (1 - wz) * wy * (1 - wx) * fv);
  acc(zf, yf, xc, (1 - wz) * (1 - wy) * wx * fv);
  acc(zf, yc, xc, (1 - wz) * wy * wx * fv);
  acc(zc, yf, xf, wz * (1 - wy) * (1 - wx) * fv);
  acc(zc, yc, xf, wz * wy * (1 - wx) * fv);
  acc(zc, yf, xc, wz * (1 - wy) * wx * fv);
  acc(zc, yc, xc, wz * wy * wx * fv);
  gphi[0] += ((1 - wz) * (1 - wy) * (u_cff - u_fff) + (1 - wz) * wy * (u_ccf - u_fcf) +
              wz * (1 - wy) * (u_cfc - u_ffc) + wz * wy * (u_ccc - u_fcc)) / hX * fv;
  gphi[1] += ((1 - wz) * (1 - wx) * (u_fcf - u_fff) + (1 - wz) * wx * (u_ccf - u_cff) +
              wz * (1 - wx) * (u_fcc - u_ffc) + wz * wx * (u_ccc - u_cfc)) / hY * fv;
  gphi[2] += ((1 - wy) * (1 - wx) * (u_ffc - u_fff) + (1 - wy) * wx * (u_cfc - u_cff) +
              wy * (1 - wx) * (u_fcc - u_fcf) + wy * wx * (u_ccc - u_ccf)) / hZ * fv;
}

// Warping3D.backward (warpingOps.cu:598-630, :1173-1262) generalised to the
// batched CNN layout: u/grad_u [B,C,Z,Y,X], phi/grad_phi [B,Z,Y,X,3].  The
// flow gradient sums the per-channel adjoints in channel order.  grad_u must
// be zero on entry (the launcher clears it).
//
// A work-group owns a kAdjTX x kAdjTY x kAdjTZ tile of voxels.  Per channel it
// sums the scattered grad_u contributions in an LDS window that covers the
// tile, a displacement margin of kAdjM voxels and the 2-voxel cubic footprint,
// then adds each non-zero window cell to grad_u with one global atomic.  Taps
// outside the window (larger displacements, far boundary reflections) go to
// grad_u directly.  The reference also accumulates with float atomics, so
// grad_u's low bits depend on the order in either case; grad_phi does not.
// A 16 x 16 x 1 tile (256 threads, 163 VGPRs, no spills, 29 KiB window) with a
// 2-voxel margin measured fastest on 4 x 4 x 80^3 (2.9 ms; 32 x 8 x 4 with a
// 3-voxel margin, 1024 threads capped at 128 VGPRs, spilled and took 4.0 ms).
constexpr int kAdjTX = 16, kAdjTY = 16, kAdjTZ = 1, kAdjM = 2;
constexpr int kAdjThreads = kAdjTX * kAdjTY * kAdjTZ;
constexpr int kAdjWX = kAdjTX + 2 * (kAdjM + 2), kAdjWY = kAdjTY + 2 * (kAdjM + 2), kAdjWZ = kAdjTZ + 2 * (kAdjM + 2);
constexpr int kAdjWin = kAdjWX * kAdjWY * kAdjWZ;

struct GlobalAcc {
  float* grad;
  int NY, NX;
  __device__ void operator()(int z, int y, int x, float v) const {
    unsafeAtomicAdd(grad + ((int64_t)z * NY + y) * NX + x, v);
  }
};

struct WindowAcc {
  float* win;       // LDS, kAdjWZ x kAdjWY x kAdjWX
  int z0, y0, x0;   // volume index of window cell (0, 0, 0)
  GlobalAcc outside;
  __device__ void operator()(int z, int y, int x, float v) const {
    const int lz = z - z0, ly = y - y0, lx = x - x0;
    if ((unsigned)lz < (unsigned)kAdjWZ && (unsigned)ly < (unsigned)kAdjWY && (unsigned)lx < (unsigned)kAdjWX)
      atomicAdd(win + (lz * kAdjWY + ly) * kAdjWX + lx, v);
    else
      outside(z, y, x, v);
  }
};

__global__ __launch_bounds__(kAdjThreads) void k_warp_backward(const float* __restrict__ u,
                                                               const float* __restrict__ phi,
                                                               const float* __restrict__ gout, int B, int C, int NZ,
                                                               int NY, int NX, float hZ, float hY, float hX, int ip,
                                                               int bc, float* __restrict__ grad_u,
                                                               float* __restrict__ grad_phi) {
  __shared__ float win[kAdjWin];
  const int tiles_x = (NX + kAdjTX - 1) / kAdjTX, tiles_y = (NY + kAdjTY - 1) / kAdjTY;
  const int tiles_z = (NZ + kAdjTZ - 1) / kAdjTZ;
  int tile = blockIdx.x;
  const int tx = tile % tiles_x;
  tile /= tiles_x;
  const int ty = tile % tiles_y;
  tile /= tiles_y;
  const int tz = tile % tiles_z;
  const int b = tile / tiles_z;
  const int t = threadIdx.x;
  const int x = tx * kAdjTX + t % kAdjTX, y = ty * kAdjTY + (t / kAdjTX) % kAdjTY, z = tz * kAdjTZ + t / (kAdjTX * kAdjTY);
  const bool valid = x < NX && y < NY && z < NZ;
  const int64_t N = (int64_t)NZ * NY * NX;
  const int64_t i = ((int64_t)z * NY + y) * NX + x;
  float cx = 0.0f, cy = 0.0f, cz = 0.0f;
  if (valid) {
    const float* ph = phi + ((int64_t)b * N + i) * 3;
    cx = (float)x * hX + ph[0];
    cy = (float)y * hY + ph[1];
    cz = (float)z * hZ + ph[2];
  }
  const int wz0 = tz * kAdjTZ - (kAdjM + 2), wy0 = ty * kAdjTY - (kAdjM + 2), wx0 = tx * kAdjTX - (kAdjM + 2);
  float gphi[3] = {0.0f, 0.0f, 0.0f};
  for (int c = 0; c < C; ++c) {
    const int64_t off = ((int64_t)b * C + c) * N;
    for (int k = t; k < kAdjWin; k += kAdjThreads) win[k] = 0.0f;
    __syncthreads();
    if (valid) {
      const float fv = gout[off + i];
      const WindowAcc acc{win, wz0, wy0, wx0, GlobalAcc{grad_u + off, NY, NX}};
      if (ip == IP_CUBIC) {
        warp_backward_cubic(u + off, acc, NZ, NY, NX, hZ, hY, hX, bc, cz, cy, cx, fv, gphi);
      } else if (ip == IP_LINEAR) {
        warp_backward_linear(u + off, acc, NZ, NY, NX, hZ, hY, hX, bc, cz, cy, cx, fv, gphi);
      } else {
        acc(nearest_axis(cz, hZ, NZ, bc), nearest_axis(cy, hY, NY, bc), nearest_axis(cx, hX, NX, bc), fv);
      }
    }
    __syncthreads();
    // Only in-volume cells ever receive a contribution (taps are mapped into
    // the volume first), so a non-zero cell is always a valid voxel.
    for (int k = t; k < kAdjWin; k += kAdjThreads) {
      const float v = win[k];
      if (v != 0.0f) {
        const int lx = k % kAdjWX, ly = (k / kAdjWX) % kAdjWY, lz = k / (kAdjWX * kAdjWY);
        unsafeAtomicAdd(grad_u + off + ((int64_t)(wz0 + lz) * NY + (wy0 + ly)) * NX + (wx0 + lx), v);
      }
    }
    __syncthreads();
  }
  if (valid) {
    float* gp = grad_phi + ((int64_t)b * N + i) * 3;
    gp[0] = gphi[0];
    gp[1] = gphi[1];
    gp[2] = gphi[2];
  }
}

// --------------------------------------------------------- prolongation
// Prolongation3D.forward / forwardVectorField / forwardMatrixField
// (prolongationOps.cu:533-625): source coordinate ((i * hP) * L) / LP.
__global__ __launch_bounds__(kBlock) void k_prolongate(const float* __restrict__ u, int C, int NZ, int NY, int NX, float LZ, float LY,
                             float LX, float hZ, float hY, float hX, int NZp, int NYp, int NXp, float LZp,
                             float LYp, float LXp, float hZp, float hYp, float hXp, int ip, int bc,
                             float* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= (int64_t)NZp * NYp * NXp) return;
  int z, y, x;
  unflatten(i, NYp, NXp, z, y, x);
  const float cxp = (float)x * hXp;
  const float cx = cxp * LX / LXp;
  const float cyp = (float)y * hYp;
  const float cy = cyp * LY / LYp;
  const float czp = (float)z * hZp;
  const float cz = czp * LZ / LZp;
  for (int c = 0; c < C; ++c) {
    StridedFetch f{u + c, C, NY, NX};
    float v[1];
    sample<1>(ip, f, NZ, NY, NX, hZ, hY, hX, bc, cz, cy, cx, v);
    out[i * C + c] = v[0];
  }
}

// ----------------------------------------------------------------- blur
// apply_gaussian_blur3d (flow_utils.py:94-112): zero-padded K^3 correlation,
// taps summed in (z, y, x) order.
__global__ __launch_bounds__(kBlock) void k_blur(const float* __restrict__ in, int NZ, int NY, int NX, const float* __restrict__ w, int K,
                       float* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= (int64_t)NZ * NY * NX) return;
  int z, y, x;
  unflatten(i, NY, NX, z, y, x);
  const int P = K / 2;
  float acc = 0.0f;
  for (int a = 0; a < K; ++a) {
    const int zz = z + a - P;
    if (zz < 0 || zz >= NZ) continue;
    for (int b = 0; b < K; ++b) {
      const int yy = y + b - P;
      if (yy < 0 || yy >= NY) continue;
      const float* row = in + ((int64_t)zz * NY + yy) * NX;
      const float* wr = w + (a * K + b) * K;
      for (int c = 0; c < K; ++c) {
        const int xx = x + c - P;
        if (xx < 0 || xx >= NX) continue;
        acc = acc + wr[c] * row[xx];
      }
    }
  }
  out[i] = acc;
}

// --------------------------------------------------------- median filter
// apply_median_filter (flow_utils.py:115-132) is scipy.ndimage.median_filter
// with size (ks, ks, ks, 1) on a [Z,Y,X,C] flow: for every voxel and
// component, the element of rank n/2 (0-based, ascending, n = ks^3) of the
// window at offsets -(ks/2) .. ks-1-(ks/2) per axis.  Outside the volume the
// window reads SciPy's default 'reflect' extension (d c b a | a b c d | d c b a),
// which repeats with period 2N.
constexpr int kMedianMaxK = 7;

__device__ __forceinline__ int reflect_index(int i, int N) {
  const int P = 2 * N;
  i %= P;
  if (i < 0) i += P;
  return i < N ? i : P - 1 - i;
}

// Batcher's odd-even merge sort for n inputs (any n), as a comparator list;
// every comparator moves the smaller value to the lower index.
template <int N>
struct OddEvenMergeNet {
  unsigned char a[N * N], b[N * N];
  int n;
  constexpr OddEvenMergeNet() : a{}, b{}, n(0) {
    for (int p = 1; p < N; p <<= 1)
      for (int k = p; k >= 1; k >>= 1)
        for (int j = k % p; j + k < N; j += 2 * k)
          for (int i = 0; i < k && i + j + k < N; ++i)
            if ((i + j) / (2 * p) == (i + j + k) / (2 * p)) {
              a[n] = (unsigned char)(i + j);
              b[n] = (unsigned char)(i + j + k);
              ++n;
            }
  }
};

// 3x3x3: the 27 window values in registers through the merge network; only
// the comparators that feed rank 13 survive dead-code elimination.  A window
// offset of at most one voxel needs a single reflection (scipy's 'reflect':
// -1 -> 0, N -> N - 1), two selects instead of reflect_index's modulo, and the
// element index is split in 32-bit arithmetic when the launch allows it.
__device__ __forceinline__ int reflect1(int i, int N) { return i < 0 ? -i - 1 : (i >= N ? 2 * N - 1 - i : i); }

template <typename I>
__global__ __launch_bounds__(kBlock) void k_median3(const float* __restrict__ in, int NZ, int NY, int NX, int C,
                                                    int64_t total, float* __restrict__ out) {
  const I t = (I)((int64_t)blockIdx.x * kBlock + threadIdx.x);
  if ((int64_t)t >= total) return;
  const I c = t % (I)C;
  const I i = t / (I)C;
  const I N = (I)NZ * NY * NX;
  const I b = i / N;
  const I r = i - b * N;
  const int x = (int)(r % NX);
  const I zy = r / NX;
  const int y = (int)(zy % NY);
  const int z = (int)(zy / NY);
  const float* src = in + (int64_t)b * N * C + c;
  int zs[3], ys[3], xs[3];
#pragma unroll
  for (int o = 0; o < 3; ++o) {
    zs[o] = reflect1(z + o - 1, NZ);
    ys[o] = reflect1(y + o - 1, NY);
    xs[o] = reflect1(x + o - 1, NX);
  }
  float v[27];
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int e = 0; e < 3; ++e)
#pragma unroll
      for (int f = 0; f < 3; ++f)
        v[(a * 3 + e) * 3 + f] = src[(((I)zs[a] * NY + ys[e]) * NX + xs[f]) * C];
  constexpr OddEvenMergeNet<27> net;
#pragma unroll
  for (int q = 0; q < net.n; ++q) {
    const float lo = fminf(v[net.a[q]], v[net.b[q]]);
    const float hi = fmaxf(v[net.a[q]], v[net.b[q]]);
    v[net.a[q]] = lo;
    v[net.b[q]] = hi;
  }
  out[t] = v[13];
}

// Any ks up to kMedianMaxK: the window in a private array, the answer is the
// value whose count of smaller elements is <= n/2 < its count of smaller or
// equal elements.
__global__ __launch_bounds__(kBlock) void k_median_any(const float* __restrict__ in, int NZ, int NY, int NX, int C,
                                                       int ks, int64_t total, float* __restrict__ out) {
  const int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (t >= total) return;
  const int c = (int)(t % C);
  const int64_t i = t / C;
  const int64_t N = (int64_t)NZ * NY * NX;
  const int64_t b = i / N;
  int z, y, x;
  unflatten(i - b * N, NY, NX, z, y, x);
  const float* src = in + b * N * C + c;
  float win[kMedianMaxK * kMedianMaxK * kMedianMaxK];
  const int h = ks / 2, n = ks * ks * ks;
  int m = 0;
  for (int a = 0; a < ks; ++a) {
    const int zz = reflect_index(z + a - h, NZ);
    for (int e = 0; e < ks; ++e) {
      const int yy = reflect_index(y + e - h, NY);
      for (int f = 0; f < ks; ++f) {
        const int xx = reflect_index(x + f - h, NX);
        win[m++] = src[(((int64_t)zz * NY + yy) * NX + xx) * C];
      }
    }
  }
  const int rank = n / 2;
  float res = win[0];
  for (int q = 0; q < n; ++q) {
    int less = 0, leq = 0;
    for (int r = 0; r < n; ++r) {
      less += win[r] < win[q];
      leq += win[r] <= win[q];
    }
    if (less <= rank && rank < leq) {
      res = win[q];
      break;
    }
  }
  out[t] = res;
}

int launch_fill(float* p, int64_t n, float v, hipStream_t s) {
  if (n <= 0) return SVS_STATUS_OK;
  k_fill<<<grid_for(n, kBlock), kBlock, 0, s>>>(p, n, v);
  SVS_LAUNCHED();
  return SVS_STATUS_OK;
}

}  // namespace svs

// ================================================================ C ABI
using namespace svs;

extern "C" {

const char* svs_last_error(void) { return g_last_error; }

int svs_fill(float* out, int64_t n, float value, void* stream) {
  SVS_REQUIRE(n >= 0 && (n == 0 || out), "svs_fill: bad arguments");
  SVS_REQUIRE_WORKITEMS(n, "svs_fill");
  return launch_fill(out, n, value, (hipStream_t)stream);
}

int svs_nabla3d_forward(const float* b, int C, int NZ, int NY, int NX, float hZ, float hY, float hX, int bdry,
                        float* out, void* stream) {
  SVS_REQUIRE(b && out, "Nabla3D_CD.forward: null pointer");
  SVS_REQUIRE(C >= 1, "Nabla3D_CD.forward: C must be >= 1");
  SVS_REQUIRE_DIMS(NZ, NY, NX);
  SVS_REQUIRE(valid_bc(bdry), "Nabla3D_CD.forward: invalid boundary type %d", bdry);
  const int64_t n = (int64_t)NZ * NY * NX;
  k_nabla_fwd<<<grid_for(n, kBlock), kBlock, 0, (hipStream_t)stream>>>(b, C, NZ, NY, NX, hZ, hY, hX, bdry, out);
  SVS_LAUNCHED();
  return SVS_STATUS_OK;
}

int svs_nabla3d_backward(const float* p, int C, int NZ, int NY, int NX, float hZ, float hY, float hX, int bdry,
                         float* out, void* stream) {
  SVS_REQUIRE(p && out, "Nabla3D_CD.backward: null pointer");
  SVS_REQUIRE(C >= 1, "Nabla3D_CD.backward: C must be >= 1");
  SVS_REQUIRE_DIMS(NZ, NY, NX);
  SVS_REQUIRE(valid_bc(bdry), "Nabla3D_CD.backward: invalid boundary type %d", bdry);
  const int64_t n = (int64_t)NZ * NY * NX;
  k_nabla_bwd<<<grid_for(n, kBlock), kBlock, 0, (hipStream_t)stream>>>(p, C, NZ, NY, NX, hZ, hY, hX, bdry, out);
  SVS_LAUNCHED();
  return SVS_STATUS_OK;
}

int svs_aniso_tangents(const float* G, int64_t N, float alpha, float beta, float* s, float* n, float* t1, float* t2,
                       void* stream) {
  SVS_REQUIRE(G && s && n && t1 && t2, "AnisotropicNabla3D.computeTangentVecs: null pointer");
  SVS_REQUIRE(N >= 0, "AnisotropicNabla3D.computeTangentVecs: negative size");
  SVS_REQUIRE_WORKITEMS(N, "AnisotropicNabla3D.computeTangentVecs");
  if (N == 0) return SVS_STATUS_OK;
  k_aniso_tangents<<<grid_for(N, kBlock), kBlock, 0, (hipStream_t)stream>>>(G, N, alpha, beta, s, n, t1, t2);
  SVS_LAUNCHED();
  return SVS_STATUS_OK;
}

int svs_aniso_apply(const float* in, const float* s, const float* n, const float* t1, const float* t2, int64_t N,
                    float* out, void* stream) {
  SVS_REQUIRE(in && s && n && t1 && t2 && out, "AnisotropicNabla3D.forwardVectorField: null pointer");
  SVS_REQUIRE(N >= 0, "AnisotropicNabla3D.forwardVectorField: negative size");
  SVS_REQUIRE_WORKITEMS(N, "AnisotropicNabla3D.forwardVectorField");
  if (N == 0) return SVS_STATUS_OK;
  k_aniso_apply<<<grid_for(N, kBlock), kBlock, 0, (hipStream_t)stream>>>(in, s, n, t1, t2, N, out);
  SVS_LAUNCHED();
  return SVS_STATUS_OK;
}

int svs_prox_primal(const float* primal, float tau, float lambda, const float* rho, const float* g, int64_t N,
                    float* out, void* stream) {
  SVS_REQUIRE(primal && rho && g && out, "TVL1OF3D_proxPrimal: null pointer");
  SVS_REQUIRE(N >= 0, "TVL1OF3D_proxPrimal: negative size");
  SVS_REQUIRE_WORKITEMS(N, "TVL1OF3D_proxPrimal");
  if (N == 0) return SVS_STATUS_OK;
  const float factor = tau * lambda;  // opticalFlowOps.cu:442
  k_prox_primal<<<grid_for(N, kBlock), kBlock, 0, (hipStream_t)stream>>>(primal, factor, rho, g, N, out);
  SVS_LAUNCHED();
  return SVS_STATUS_OK;
}

int svs_prox_dual(const float* dual, float sigma, float wtv, int64_t N, float* out, void* stream) {
  (void)sigma;  // the reference passes sigma but its kernel never reads it (opticalFlowOps.cu:475-521)
  SVS_REQUIRE(dual && out, "TVL1OF3D_proxDual: null pointer");
  SVS_REQUIRE(N >= 0, "TVL1OF3D_proxDual: negative size");
  SVS_REQUIRE_WORKITEMS(N, "TVL1OF3D_proxDual");
  if (N == 0) return SVS_STATUS_OK;
  k_prox_dual<<<grid_for(N, kBlock), kBlock, 0, (hipStream_t)stream>>>(dual, wtv, N, out);
  SVS_LAUNCHED();
  return SVS_STATUS_OK;
}

int svs_warp3d(const float* u, int C, const float* phi, int NZ, int NY, int NX, float hZ, float hY, float hX,
               int interp, int bdry, float* out, void* stream) {
  SVS_REQUIRE(u && phi && out, "Warping3D.forward: null pointer");
  SVS_REQUIRE(C >= 1, "Warping3D.forward: C must be >= 1");
  SVS_REQUIRE_DIMS(NZ, NY, NX);
  SVS_REQUIRE(valid_bc(bdry) && valid_ip(interp), "Warping3D.forward: invalid interpolation/boundary type");
  const int64_t n = (int64_t)NZ * NY * NX;
  k_warp<<<grid_for(n, kBlock), kBlock, 0, (hipStream_t)stream>>>(u, C, phi, NZ, NY, NX, hZ, hY, hX, interp, bdry,
                                                                  out);
  SVS_LAUNCHED();
  return SVS_STATUS_OK;
}

int svs_warp3d_cnn(const float* u, const float* phi, int B, int C, int NZ, int NY, int NX, float hZ, float hY,
                   float hX, int interp, int bdry, float* out, void* stream) {
  SVS_REQUIRE(u && phi && out, "WarpingCNN3D.forward: null pointer");
  SVS_REQUIRE(B >= 1 && C >= 1, "WarpingCNN3D.forward: batch and channels must be >= 1");
  SVS_REQUIRE_DIMS(NZ, NY, NX);
  SVS_REQUIRE(valid_bc(bdry) && valid_ip(interp), "WarpingCNN3D.forward: invalid interpolation/boundary type");
  const int64_t n = (int64_t)B * NZ * NY * NX;
  SVS_REQUIRE_WORKITEMS(n, "WarpingCNN3D.forward (batch x voxels)");
  k_warp_cnn<<<grid_for(n, kBlock), kBlock, 0, (hipStream_t)stream>>>(u, phi, B, C, NZ, NY, NX, hZ, hY, hX, interp,
                                                                      bdry, out);
  SVS_LAUNCHED();
  return SVS_STATUS_OK;
}

int svs_warp3d_backward(const float* u, const float* phi, const float* grad_out, int B, int C, int NZ, int NY,
                        int NX, float hZ, float hY, float hX, int interp, int bdry, float* grad_u, float* grad_phi,
                        void* stream) {
  SVS_REQUIRE(u && phi && grad_out && grad_u && grad_phi, "Warping3D.backward: null pointer");
  SVS_REQUIRE(B >= 1 && C >= 1, "Warping3D.backward: batch and channels must be >= 1");
  SVS_REQUIRE_DIMS(NZ, NY, NX);
  SVS_REQUIRE(valid_bc(bdry) && valid_ip(interp), "Warping3D.backward: invalid interpolation/boundary type");
  const int64_t n = (int64_t)NZ * NY * NX;
  int rc = launch_fill(grad_u, (int64_t)B * C * n, 0.0f, (hipStream_t)stream);
  if (rc) return rc;
  const int64_t tiles = (int64_t)B * ((NX + kAdjTX - 1) / kAdjTX) * ((NY + kAdjTY - 1) / kAdjTY) *
                        ((NZ + kAdjTZ - 1) / kAdjTZ);
  SVS_REQUIRE(tiles < ((int64_t)1 << 31), "Warping3D.backward: too many tiles");
  k_warp_backward<<<(int)tiles, kAdjThreads, 0, (hipStream_t)stream>>>(u, phi, grad_out, B, C, NZ, NY, NX, hZ, hY,
                                                                       hX, interp, bdry, grad_u, grad_phi);
  SVS_LAUNCHED();
  return SVS_STATUS_OK;
}

int svs_prolongate3d(const float* u, int C, int NZ, int NY, int NX, float LZ, float LY, float LX, int NZp, int NYp,
                     int NXp, float LZp, float LYp, float LXp, int interp, int bdry, float* out, void* stream) {
  SVS_REQUIRE(u && out, "Prolongation3D.forward: null pointer");
  SVS_REQUIRE(C >= 1, "Prolongation3D.forward: C must be >= 1");
  SVS_REQUIRE_DIMS(NZ, NY, NX);
  SVS_REQUIRE_DIMS(NZp, NYp, NXp);
  SVS_REQUIRE(valid_bc(bdry) && valid_ip(interp), "Prolongation3D.forward: invalid interpolation/boundary type");
  const float hZ = LZ / (float)(NZ - 1), hY = LY / (float)(NY - 1), hX = LX / (float)(NX - 1);
  const float hZp = LZp / (float)(NZp - 1), hYp = LYp / (float)(NYp - 1), hXp = LXp / (float)(NXp - 1);
  const int64_t n = (int64_t)NZp * NYp * NXp;
  k_prolongate<<<grid_for(n, kBlock), kBlock, 0, (hipStream_t)stream>>>(u, C, NZ, NY, NX, LZ, LY, LX, hZ, hY, hX,
                                                                        NZp, NYp, NXp, LZp, LYp, LXp, hZp, hYp,
                                                                        hXp, interp, bdry, out);
  SVS_LAUNCHED();
  return SVS_STATUS_OK;
}

int svs_blur3d(const float* in, int NZ, int NY, int NX, const float* weights, int K, float* out, void* stream) {
  SVS_REQUIRE(in && weights && out, "apply_gaussian_blur3d: null pointer");
  SVS_REQUIRE(K >= 1 && (K % 2) == 1, "apply_gaussian_blur3d: kernel size must be odd and positive (got %d)", K);
  SVS_REQUIRE_DIMS(NZ, NY, NX);
  const int64_t n = (int64_t)NZ * NY * NX;
  k_blur<<<grid_for(n, kBlock), kBlock, 0, (hipStream_t)stream>>>(in, NZ, NY, NX, weights, K, out);
  SVS_LAUNCHED();
  return SVS_STATUS_OK;
}

int svs_median3d(const float* in, int B, int NZ, int NY, int NX, int C, int ks, float* out, void* stream) {
  SVS_REQUIRE(in && out, "apply_median_filter: null pointer");
  SVS_REQUIRE(in != out, "apply_median_filter: the output must not alias the input");
  SVS_REQUIRE(B >= 1 && C >= 1, "apply_median_filter: batch and component counts must be positive (got %d, %d)", B, C);
  SVS_REQUIRE(ks >= 1 && ks <= kMedianMaxK, "apply_median_filter: kernel size must be in [1, %d] (got %d)",
              kMedianMaxK, ks);
  SVS_REQUIRE_DIMS(NZ, NY, NX);
  const int64_t total = (int64_t)B * NZ * NY * NX * C;
  SVS_REQUIRE_WORKITEMS(total, "apply_median_filter (batch x voxels x components)");
  if (ks == 3 && total < ((int64_t)1 << 31))
    k_median3<int32_t><<<grid_for(total, kBlock), kBlock, 0, (hipStream_t)stream>>>(in, NZ, NY, NX, C, total, out);
  else if (ks == 3)
    k_median3<int64_t><<<grid_for(total, kBlock), kBlock, 0, (hipStream_t)stream>>>(in, NZ, NY, NX, C, total, out);
  else
    k_median_any<<<grid_for(total, kBlock), kBlock, 0, (hipStream_t)stream>>>(in, NZ, NY, NX, C, ks, total, out);
  SVS_LAUNCHED();
  return SVS_STATUS_OK;
}

}  // extern "C"
