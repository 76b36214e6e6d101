// Device-side building blocks shared by the op kernels and the fused solver.
//
// Every routine keeps the reference's order of floating-point operations
// (the library is compiled with -ffp-contract=off, so nothing is fused into a
// multiply-add behind our back).  File:line citations point at the reference
// CUDA sources under svs/src/.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace svs {

enum { BD_NEAREST = 0, BD_MIRROR = 1, BD_REFLECT = 2 };
enum { IP_NEAREST = 0, IP_LINEAR = 1, IP_CUBIC = 2 };

// boundary.cuh:25-58 — out-of-range index map, one reflection only.  An
// index more than one period outside (a flow larger than the volume) would
// make the reference read out of bounds; here the result is clamped into the
// volume, which leaves every in-bounds case unchanged.
__device__ __forceinline__ int bidx(int i, int N, int bc) {
  int r;
  if (bc == BD_MIRROR) r = i < 0 ? -i - 1 : (i > N - 1 ? 2 * N - 1 - i : i);
  else if (bc == BD_NEAREST) r = i < 0 ? 0 : (i > N - 1 ? N - 1 : i);
  else r = i < 0 ? -i : (i > N - 1 ? 2 * N - 2 - i : i);
  return r < 0 ? 0 : (r > N - 1 ? N - 1 : r);
}

// interpolation_cubicHermiteSpline.cuh:20-56.  The local coordinate is
// lc = w + 1 with w in [0,1], so floor(lc) is 1, or 2 when w + 1 rounds up to
// 2.0f; in the second case the taps shift by one and the fourth reads 0.
__device__ __forceinline__ float cr_local(float b0, float b1, float b2, float b3, float lc) {
  const bool hi = lc >= 2.0f;
  const float u_f1 = hi ? b1 : b0;
  const float u_f = hi ? b2 : b1;
  const float u_c = hi ? b3 : b2;
  const float u_c1 = hi ? 0.0f : b3;
  const float p_f = u_f;
  const float pp_f = (u_c - u_f1) / 2;
  const float p_c = u_c;
  const float pp_c = (u_c1 - u_f) / 2;
  const float a = 2 * p_f - 2 * p_c + pp_f + pp_c;
  const float b = -3 * p_f + 3 * p_c - 2 * pp_f - pp_c;
  const float c = pp_f;
  const float d = p_f;
  const float w = lc - (hi ? 2.0f : 1.0f);
  return w * (w * (w * a + b) + c) + d;
}

// cr_local for lc < 2 (the usual case): the same operations without the
// selects of the rounded-up segment, hence the same bits whenever lc < 2.
__device__ __forceinline__ float cr_local_lo(float b0, float b1, float b2, float b3, float lc) {
  const float p_f = b1;
  const float pp_f = (b2 - b0) / 2;
  const float p_c = b2;
  const float pp_c = (b3 - b1) / 2;
  const float a = 2 * p_f - 2 * p_c + pp_f + pp_c;
  const float b = -3 * p_f + 3 * p_c - 2 * pp_f - pp_c;
  const float c = pp_f;
  const float d = p_f;
  const float w = lc - 1.0f;
  return w * (w * (w * a + b) + c) + d;
}

template <bool MAYBE_HI>
__device__ __forceinline__ float cr_eval(float b0, float b1, float b2, float b3, float lc) {
  return MAYBE_HI ? cr_local(b0, b1, b2, b3, lc) : cr_local_lo(b0, b1, b2, b3, lc);
}

// Adjoint of cr_local (interpolation_cubicHermiteSpline.cuh:59-117):
// g[k] = d out/d b[k] * fwd, returns d out/d w * fwd.  In the rounded-up case
// the tap that falls outside the 4-entry buffer contributes nothing and b[0]
// receives 0.
__device__ __forceinline__ float cr_local_backward(float b0, float b1, float b2, float b3, float lc, float fwd,
                                                   float g[4]) {
  const bool hi = lc >= 2.0f;
  const float w = lc - (hi ? 2.0f : 1.0f);
  const float w2 = w * w;
  const float w3 = w2 * w;
  const float d_f1 = -w3 / 2 + w2 - w / 2;
  const float d_f = (3 * w3) / 2 - (5 * w2) / 2 + 1;
  const float d_c = -(3 * w3) / 2 + 2 * w2 + w / 2;
  const float d_c1 = w3 / 2 - w2 / 2;
  if (hi) {
    g[0] = 0.0f; g[1] = d_f1 * fwd; g[2] = d_f * fwd; g[3] = d_c * fwd;
  } else {
    g[0] = d_f1 * fwd; g[1] = d_f * fwd; g[2] = d_c * fwd; g[3] = d_c1 * fwd;
  }
  const float u_f1 = hi ? b1 : b0;
  const float u_f = hi ? b2 : b1;
  const float u_c = hi ? b3 : b2;
  const float u_c1 = hi ? 0.0f : b3;
  const float p_f = u_f;
  const float pp_f = (u_c - u_f1) / 2;
  const float p_c = u_c;
  const float pp_c = (u_c1 - u_f) / 2;
  const float a = 2 * p_f - 2 * p_c + pp_f + pp_c;
  const float b = -3 * p_f + 3 * p_c - 2 * pp_f - pp_c;
  const float c = pp_f;
  return (3 * w2 * a + 2 * b * w + c) * fwd;
}

// Per-axis set-up of a tricubic sample: base index and fractional weight
// (interpolation_cubicHermiteSpline.cuh:490-498).
struct Axis {
  int f;    // floor(coord / h)
  float w;  // coord / h - f
};
// floor(v) as an int for every float: NaN gives 0, as the reference's CUDA
// float-to-int conversion (and gfx950's) does, and magnitudes are clamped to
// 2^30 so that the tap offsets added later cannot overflow (a coordinate that
// far outside a volume lands on its boundary map either way).
__device__ __forceinline__ int floor_int(float v) {
  const float f = floorf(v);
  if (f != f) return 0;
  return (int)fminf(fmaxf(f, -1073741824.0f), 1073741824.0f);
}

__device__ __forceinline__ Axis make_axis(float coord, float h) {
  Axis a;
  const float v = coord / h;
  a.f = floor_int(v);
  a.w = v - (float)a.f;
  return a;
}

// Tricubic Catmull-Rom sample of NC fields that share one geometry.  fetch(z,
// y, x, v) loads the NC values at an in-range voxel.  Loop order follows
// interpolation_cubicHermiteSpline.cuh:482-525: x-pass per (dz,dy), y-pass per
// dz, then the z-pass; each field is evaluated exactly as the scalar routine.
//
// LOOP sets how many taps are in flight: 0 unrolls everything (64 taps may be
// hoisted ahead of the first use), 1 runs the z-planes as a loop that is not
// unrolled (16 taps), 2 also the rows of a plane (4 taps).  Fewer taps in
// flight take fewer registers, so more waves share a SIMD; the partial
// results reach their slots through selects, no arithmetic moves.
template <int NC, bool MAYBE_HI, int LOOP, class Fetch>
__device__ __forceinline__ void cubic_row(const Fetch& fetch, int z, int y, const int (&xs)[4], float lx,
                                          float out[NC]) {
  float v[4][NC];
#pragma unroll
  for (int dx = 0; dx < 4; ++dx) fetch(z, y, xs[dx], v[dx]);
#pragma unroll
  for (int c = 0; c < NC; ++c) out[c] = cr_eval<MAYBE_HI>(v[0][c], v[1][c], v[2][c], v[3][c], lx);
}

template <int NC, bool MAYBE_HI, int LOOP, class Fetch>
__device__ __forceinline__ void cubic_plane(const Fetch& fetch, int z, const int (&xs)[4], const int (&ys)[4],
                                            float lx, float ly, float out[NC]) {
  float by[NC][4];
  if (LOOP >= 2) {
#pragma unroll
    for (int c = 0; c < NC; ++c) by[c][0] = by[c][1] = by[c][2] = by[c][3] = 0.0f;
#pragma unroll 1
    for (int dy = 0; dy < 4; ++dy) {
      float r[NC];
      cubic_row<NC, MAYBE_HI, LOOP>(fetch, z, ys[dy], xs, lx, r);
#pragma unroll
      for (int c = 0; c < NC; ++c)
#pragma unroll
        for (int k = 0; k < 4; ++k) by[c][k] = k == dy ? r[c] : by[c][k];
    }
  } else {
#pragma unroll
    for (int dy = 0; dy < 4; ++dy) {
      float r[NC];
      cubic_row<NC, MAYBE_HI, LOOP>(fetch, z, ys[dy], xs, lx, r);
#pragma unroll
      for (int c = 0; c < NC; ++c) by[c][dy] = r[c];
    }
  }
#pragma unroll
  for (int c = 0; c < NC; ++c) out[c] = cr_eval<MAYBE_HI>(by[c][0], by[c][1], by[c][2], by[c][3], ly);
}

template <int NC, bool MAYBE_HI, int LOOP, class Fetch>
__device__ __forceinline__ void sample_cubic_impl(const Fetch& fetch, int NZ, int NY, int NX, Axis az, Axis ay,
                                                  Axis ax, int bc, float out[NC]) {
  int xs[4], ys[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    xs[k] = bidx(ax.f - 1 + k, NX, bc);
    ys[k] = bidx(ay.f - 1 + k, NY, bc);
  }
  const float lx = ax.w + 1, ly = ay.w + 1, lz = az.w + 1;
  float bz[NC][4];
  if (LOOP >= 1) {
#pragma unroll
    for (int c = 0; c < NC; ++c) bz[c][0] = bz[c][1] = bz[c][2] = bz[c][3] = 0.0f;
#pragma unroll 1
    for (int dz = 0; dz < 4; ++dz) {
      float r[NC];
      cubic_plane<NC, MAYBE_HI, LOOP>(fetch, bidx(az.f - 1 + dz, NZ, bc), xs, ys, lx, ly, r);
#pragma unroll
      for (int c = 0; c < NC; ++c)
#pragma unroll
        for (int k = 0; k < 4; ++k) bz[c][k] = k == dz ? r[c] : bz[c][k];
    }
  } else {
#pragma unroll
    for (int dz = 0; dz < 4; ++dz) {
      float r[NC];
      cubic_plane<NC, MAYBE_HI, LOOP>(fetch, bidx(az.f - 1 + dz, NZ, bc), xs, ys, lx, ly, r);
#pragma unroll
      for (int c = 0; c < NC; ++c) bz[c][dz] = r[c];
    }
  }
#pragma unroll
  for (int c = 0; c < NC; ++c) out[c] = cr_eval<MAYBE_HI>(bz[c][0], bz[c][1], bz[c][2], bz[c][3], lz);
}

// The rounded-up segment (w + 1 == 2.0f, i.e. w = 1 - 2^-24) is rare: a
// wavefront in which no lane meets it takes the select-free evaluation.
template <int NC, class Fetch, int LOOP = 0>
__device__ __forceinline__ void sample_cubic(const Fetch& fetch, int NZ, int NY, int NX, Axis az, Axis ay, Axis ax,
                                             int bc, float out[NC]) {
  const bool hi = ax.w + 1 >= 2.0f || ay.w + 1 >= 2.0f || az.w + 1 >= 2.0f;
  if (__any(hi))
    sample_cubic_impl<NC, true, LOOP>(fetch, NZ, NY, NX, az, ay, ax, bc, out);
  else
    sample_cubic_impl<NC, false, LOOP>(fetch, NZ, NY, NX, az, ay, ax, bc, out);
}

// interpolation_linear.cuh:249-295.
template <int NC, class Fetch>
__device__ __forceinline__ void sample_linear(const Fetch& fetch, int NZ, int NY, int NX, Axis az, Axis ay, Axis ax,
                                              int bc, float out[NC]) {
  const float wx = ax.w, wy = ay.w, wz = az.w;
  const int xf = bidx(ax.f, NX, bc), xc = bidx(ax.f + 1, NX, bc);
  const int yf = bidx(ay.f, NY, bc), yc = bidx(ay.f + 1, NY, bc);
  const int zf = bidx(az.f, NZ, bc), zc = bidx(az.f + 1, NZ, bc);
  float fff[NC], ffc[NC], fcf[NC], fcc[NC], cff[NC], cfc[NC], ccf[NC], ccc[NC];
  fetch(zf, yf, xf, fff); fetch(zc, yf, xf, ffc); fetch(zf, yc, xf, fcf); fetch(zc, yc, xf, fcc);
  fetch(zf, yf, xc, cff); fetch(zc, yf, xc, cfc); fetch(zf, yc, xc, ccf); fetch(zc, yc, xc, ccc);
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    float o = (1 - wz) * (1 - wy) * (1 - wx) * fff[c];
    o += wz * (1 - wy) * (1 - wx) * ffc[c];
    o += (1 - wz) * (1 - wy) * wx * cff[c];
    o += wz * (1 - wy) * wx * cfc[c];
    o += (1 - wz) * wy * (1 - wx) * fcf[c];
    o += wz * wy * (1 - wx) * fcc[c];
    o += (1 - wz) * wy * wx * ccf[c];
    o += wz * wy * wx * ccc[c];
    out[c] = o;
  }
}

// interpolation_nearest.cuh:250-294 — ties go to the upper neighbour.
__device__ __forceinline__ int nearest_axis(float coord, float h, int N, int bc) {
  const int f = floor_int(coord / h);
  const float dist_f = coord / h - (float)f;
  const float dist_c = (float)(f + 1) - coord / h;
  return dist_f < dist_c ? bidx(f, N, bc) : bidx(f + 1, N, bc);
}

template <int NC, class Fetch, int LOOP = 0>
__device__ __forceinline__ void sample(int ip, const Fetch& fetch, int NZ, int NY, int NX, float hZ, float hY,
                                       float hX, int bc, float cz, float cy, float cx, float out[NC]) {
  if (ip == IP_CUBIC) {
    sample_cubic<NC, Fetch, LOOP>(fetch, NZ, NY, NX, make_axis(cz, hZ), make_axis(cy, hY), make_axis(cx, hX), bc, out);
  } else if (ip == IP_LINEAR) {
    sample_linear<NC>(fetch, NZ, NY, NX, make_axis(cz, hZ), make_axis(cy, hY), make_axis(cx, hX), bc, out);
  } else {
    fetch(nearest_axis(cz, hZ, NZ, bc), nearest_axis(cy, hY, NY, bc), nearest_axis(cx, hX, NX, bc), out);
  }
}

// differentialOps.cu:540-656 / :782-907 — central difference along one axis
// at index i of N from the values at i-1, i, i+1 (only in-range ones are used).
__device__ __forceinline__ float cd_fwd(float bm, float b0, float bp, int i, int N, float h, int bc) {
  if (bc == BD_MIRROR) return (i > 0 && i < N - 1) ? (0.5f * (bp - bm)) / h : 0.0f;
  if (i > 0) return (i < N - 1) ? (0.5f * (bp - bm)) / h : (0.5f * (b0 - bm)) / h;
  return (0.5f * (bp - b0)) / h;
}
// differentialOps.cu:686-772 / :909-1044 — the exact adjoint along one axis.
__device__ __forceinline__ float cd_adj(float pm, float p0, float pp, int i, int N, int bc) {
  if (bc == BD_MIRROR) {
    if (i > 1) return (i < N - 2) ? 0.5f * (pm - pp) : 0.5f * pm;
    return -0.5f * pp;
  }
  if (i > 0) return (i < N - 1) ? 0.5f * (pm - pp) : 0.5f * (pm + p0);
  return 0.5f * (-p0 - pp);
}

// anisotropicDifferentialOps.cu:144-175 applied to one row q (3 derivatives):
// out_d = s (n.q) n_d + (t1.q) t1_d.  The reference's t2 term is multiplied by
// a factor that stays 0 and is therefore omitted, and its dot products start
// from the first product instead of 0 + (the compiler must keep an add of +0,
// which turns -0 into +0).  Both change at most the sign of a zero, and no
// step of the solve lets the sign of a zero reach a nonzero value (every
// divisor is positive), so the results are the same values.
__device__ __forceinline__ void aniso_row(float s, const float n[3], const float t1[3], const float q[3],
                                          float out[3]) {
  float nf = n[0] * q[0], tf = t1[0] * q[0];
#pragma unroll
  for (int d = 1; d < 3; ++d) {
    nf += n[d] * q[d];
    tf += t1[d] * q[d];
  }
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    float o = s * nf * n[d];
    o += tf * t1[d];
    out[d] = o;
  }
}

// t1 from the unit normal, anisotropicDifferentialOps.cu:117-124.
__device__ __forceinline__ void tangent1(const float n[3], float t1[3]) {
  if (n[2] > 0.5f) {
    t1[0] = n[2]; t1[1] = 0.0f; t1[2] = -n[0];
  } else {
    t1[0] = n[1]; t1[1] = -n[0]; t1[2] = 0.0f;
  }
}

// opticalFlowOps.cu:187-205 — projection of one row onto the l2 ball.  hub
// (not in the reference, which is hub = 1): the Huber-TV step, v / max(hub,
// ||v|| / wtv) with hub = 1 + sigma eps / wtv, i.e. v / hub projected onto the
// ball, in one division; hub = 1 gives the reference's bits.
__device__ __forceinline__ void prox_dual_row(float v[3], float wtv, float hub = 1.0f) {
  float ns = 0.0f;
#pragma unroll
  for (int d = 0; d < 3; ++d) ns += v[d] * v[d];
  const float den = fmaxf(hub, sqrtf(ns) / wtv);
  // Inside the ball (hub = 1) den == 1 and v / 1 == v exactly: a wave whose
  // rows all lie inside skips the three divisions.
  if (den > 1.0f) {
#pragma unroll
    for (int d = 0; d < 3; ++d) v[d] = v[d] / den;
  }
}

// opticalFlowOps.cu:131-173 — the L1 data-term shift added to the primal
// point v; factor = tau * lambda.
__device__ __forceinline__ void prox_primal_point(float v[3], float r, const float g[3], float factor) {
  float g2 = 0.0f;
#pragma unroll
  for (int c = 0; c < 3; ++c) g2 += g[c] * g[c];
  float delta[3];
  if (r < -factor * g2) {
#pragma unroll
    for (int c = 0; c < 3; ++c) delta[c] = factor * g[c];
  } else if (r > factor * g2) {
#pragma unroll
    for (int c = 0; c < 3; ++c) delta[c] = -factor * g[c];
  } else if ((double)g2 > 1e-10) {
#pragma unroll
    for (int c = 0; c < 3; ++c) delta[c] = -r * g[c] / g2;
  } else {
#pragma unroll
    for (int c = 0; c < 3; ++c) delta[c] = 0.0f;
  }
#pragma unroll
  for (int c = 0; c < 3; ++c) v[c] = v[c] + delta[c];
}

typedef float svs_f4 __attribute__((ext_vector_type(4)));

// Raw buffer access for the fused sweep.  A 128-bit resource (base address and
// size in bytes) sits in scalar registers; each access adds a per-lane byte
// offset (VGPR) and a wave-uniform byte offset (SGPR), so the plane and field
// offsets cost no vector instructions.  Callers keep every access inside
// [base, base + bytes) or at kOutside.  0x00020000 is the gfx9 word 3 (32-bit
// data format).
typedef __amdgpu_buffer_rsrc_t svs_rsrc;
__device__ __forceinline__ svs_rsrc make_rsrc(const void* base, uint32_t bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0, (int)bytes, 0x00020000);
}
// Buffer offsets at or past 2^31 are outside every buffer these kernels
// address (each is smaller than 2^31 bytes, and offset plus wave offset stays
// below 2^32): a load there returns 0 and a store there is dropped (buffer
// range checking), with no memory access.
constexpr uint32_t kOutside = 0x80000000u;

// The wave offset goes through readfirstlane: the value is uniform by the
// callers' contract, and the (convergent) builtin keeps the compiler from
// merging equal offsets computed under different branches into a VGPR, which
// it would then read back lane by lane around every access.
__device__ __forceinline__ int wave_uniform(uint32_t v) { return __builtin_amdgcn_readfirstlane((int)v); }
__device__ __forceinline__ float bld(svs_rsrc r, uint32_t voff, uint32_t soff) {
  return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, (int)voff, wave_uniform(soff), 0));
}
__device__ __forceinline__ svs_f4 bld4(svs_rsrc r, uint32_t voff, uint32_t soff) {
  return __builtin_bit_cast(svs_f4, __builtin_amdgcn_raw_buffer_load_b128(r, (int)voff, wave_uniform(soff), 0));
}
__device__ __forceinline__ void bst(svs_rsrc r, uint32_t voff, uint32_t soff, float v) {
  __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned int, v), r, (int)voff, wave_uniform(soff), 0);
}

// The Huber function of a row norm t for the energy statistics: t^2 / (2 eps)
// up to eps, t - eps / 2 beyond; the identity for eps = 0.
__device__ __forceinline__ float huber_of_norm(float t, float eps) {
  if (eps == 0.0f) return t;
  return t <= eps ? t * t / (2.0f * eps) : t - 0.5f * eps;
}

// Sum over the 64 lanes of a wave, returned to every lane.  The butterfly
// pairs lanes symmetrically, so every lane gets the same bits, and the order is
// fixed, so every run does too.
__device__ __forceinline__ float wave_sum(float v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  return v;
}
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  return v;
}

// Linear voxel index -> (z, y, x).
__device__ __forceinline__ void unflatten(int64_t i, int NY, int NX, int& z, int& y, int& x) {
  x = (int)(i % NX);
  const int64_t t = i / NX;
  y = (int)(t % NY);
  z = (int)(t / NY);
}

}  // namespace svs
