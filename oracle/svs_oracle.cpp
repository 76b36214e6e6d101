// CPU oracle: a literal restatement of the reference's 3D TV-L1 optical-flow
// arithmetic, instantiated for float (the "reference FP32 result") and double
// (the FP64 oracle of the same iteration).
//
// TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
// cpu_baseline leg of bench.py may load this library, and only as the checker
// (or the timed CPU baseline).  The product path never calls it.
//
// Conventions (same as the reference):
//   volumes  [Z,Y,X]        x fastest
//   vectors  [Z,Y,X,3]      component order (x, y, z)
//   matrices [Z,Y,X,3,3]    m[..,c,d] = d/d(axis d) of component c
// Every function cites the reference file:line it restates.  Build with
// -ffp-contract=off so that no multiply-add is fused: the order of rounding is
// the order written here.
//
// Sensitivity variants (oracle/sensitivity.py builds them into
// oracle/_variants/; the default build defines none of these macros).  Each
// changes one FP32 rounding choice the reference's own build leaves open, so
// the spread of their solves measures how far the FP32 result depends on it:
//   SVS_ORACLE_DOT_ORDER=1 / =2   3-term sums of the driver as (a + b) + c or
//                                 a + (b + c) instead of torch's (a + c) + b
//   SVS_ORACLE_FMA                with -ffp-contract=fast: multiply-adds inside
//                                 the operators fused, as nvcc's default
//                                 --fmad=true may do (setup.py:43); the steps
//                                 the reference runs as separate torch kernels
//                                 stay rounded (round_kernel_output)
//   SVS_ORACLE_LIBM               glibc expf/powf for the anisotropic weight
//   SVS_ORACLE_WEIGHT_ULP         the weight moved by one ulp at every other
//                                 voxel (hash-chosen, either direction), an
//                                 expf/powf pair with a different 1-ulp error
//   SVS_ORACLE_BLUR_F64           the pre-blur accumulated in double and
//                                 rounded once (F.conv3d's summation order is
//                                 unspecified)
//
// Parity status per function (checks live in tests/test_oracle_*.py):
//   boundary maps ............ brute force over i in [-N, 2N)
//   Catmull-Rom interpolation  constant/linear/quadratic reproduction, identity
//                              for zero flow, integer shifts, FP64 brute force
//   nabla / adjoint .......... linear ramps, dot-product identity <Du,p>=<u,D^T p>
//   anisotropic tensor ....... per-voxel closed form, D symmetric rank <= 2
//   prox primal / dual ....... hand-computed branch cases, idempotence
//   warp adjoint ............. dot-product identity (image), FP64 central FD (flow)
//   blur ..................... torch.nn.functional.conv3d on CPU
//   median filter ............ scipy.ndimage.median_filter on CPU
//   mask metrics ............. scipy.ndimage binary_erosion + distance_transform_edt
//                              (what monai's Dice/Hausdorff call), closed forms
//   spatial augmentation ..... torch.nn.functional.grid_sample / torch.matmul /
//                              torch.einsum on the CPU (bit-exact), closed forms
//                              (180-degree turns = two flips, identity, rigid
//                              motion of an affine field)
//   solve .................... FP64 twin, EPE against a known displacement
#include <cmath>
#include <cstdint>
#include <cstring>
#include <cstdlib>
#include <vector>
#include <algorithm>

#include "svs_math.h"

namespace orc {

enum { BD_NEAREST = 0, BD_MIRROR = 1, BD_REFLECT = 2 };
enum { IP_NEAREST = 0, IP_LINEAR = 1, IP_CUBIC = 2 };

static inline float flr(float v) { return floorf(v); }
static inline double flr(double v) { return floor(v); }
// (int)floor(v) as the reference's CUDA build evaluates it (cvt.rzi.s32.f32:
// NaN gives 0, no undefined behaviour), with magnitudes clamped to 2^30 so the
// tap offsets added later cannot overflow; plain C++ leaves NaN and
// out-of-range conversions undefined.
template <class T> static inline int floor_int(T v) {
  const T f = flr(v);
  if (f != f) return 0;
  return (int)std::min(std::max(f, (T)-1073741824.0), (T)1073741824.0);
}
static inline float sqrt_t(float v) { return sqrtf(v); }
static inline double sqrt_t(double v) { return sqrt(v); }
// The FP32 anisotropic weight uses the faithful expf/powf of svs_math.h, the
// same source the GPU kernels compile, so both round it identically (the
// functions themselves are checked against glibc and double-precision libm);
// the FP64 twin uses libm.
#ifdef SVS_ORACLE_LIBM
static inline float exp_t(float v) { return expf(v); }
static inline float pow_t(float a, float b) { return powf(a, b); }
#else
static inline float exp_t(float v) { return svs_expf(v); }
static inline float pow_t(float a, float b) { return svs_powf(a, b); }
#endif
static inline double exp_t(double v) { return exp(v); }
static inline double pow_t(double a, double b) { return pow(a, b); }
// The weight as stored (SVS_ORACLE_WEIGHT_ULP: one ulp off at half the voxels).
static inline float weight_t(float s, size_t i) {
#ifdef SVS_ORACLE_WEIGHT_ULP
  uint32_t h = (uint32_t)i * 2654435761u;
  h ^= h >> 15;
  if (h & 1u) return nextafterf(s, (h & 2u) ? 2.0f : 0.0f);
#else
  (void)i;
#endif
  return s;
}
static inline double weight_t(double s, size_t) { return s; }
// A value the reference computes in a torch kernel of its own reaches the
// next step rounded to FP32.  In the SVS_ORACLE_FMA variant an empty asm
// with the value as an in/out register operand keeps the compiler from fusing
// a multiply across it; otherwise this is the identity.
template <class T> static inline T round_kernel_output(T v) {
#ifdef SVS_ORACLE_FMA
  asm("" : "+x"(v));
#endif
  return v;
}
// torch.sum over a last dimension of 3 on the GPU (see solve_level).
template <class T> static inline T torch_sum3(T a, T b, T c) {
#if SVS_ORACLE_DOT_ORDER == 1
  return (a + b) + c;
#elif SVS_ORACLE_DOT_ORDER == 2
  return a + (b + c);
#else
  return (a + c) + b;
#endif
}
static inline float fmax_t(float a, float b) { return fmaxf(a, b); }
static inline double fmax_t(double a, double b) { return fmax(a, b); }

// boundary.cuh:25-58 — one reflection only.  Beyond one period the
// reference reads out of bounds; the result is clamped into [0, N-1] here.
static inline int bidx(int i, int N, int bc) {
  int r;
  switch (bc) {
    case BD_NEAREST: r = i < 0 ? 0 : (i > N - 1 ? N - 1 : i); break;
    case BD_MIRROR:  r = i < 0 ? -i - 1 : (i > N - 1 ? 2 * N - 1 - i : i); break;
    default:         r = i < 0 ? -i : (i > N - 1 ? 2 * N - 2 - i : i); break;
  }
  return r < 0 ? 0 : (r > N - 1 ? N - 1 : r);
}

// Mesh spacing h = L/(N-1): coreDefines.h:72-75 (float there; T here).
template <class T> struct Mesh {
  int NZ, NY, NX;
  T LZ, LY, LX, hZ, hY, hX;
};
template <class T> static Mesh<T> make_mesh(int NZ, int NY, int NX, T LZ, T LY, T LX) {
  Mesh<T> m;
  m.NZ = NZ; m.NY = NY; m.NX = NX;
  m.LZ = LZ; m.LY = LY; m.LX = LX;
  m.hZ = LZ / (T)(NZ - 1); m.hY = LY / (T)(NY - 1); m.hX = LX / (T)(NX - 1);
  return m;
}

// interpolation_cubicHermiteSpline.cuh:20-56 — the local Catmull-Rom segment.
template <class T> static inline T cr_local(const T* buf, T lc) {
  const int f = floor_int(lc);
  const int c = f + 1, f1 = f - 1, c1 = c + 1;
  T u_f = 0, u_f1 = 0, u_c = 0, u_c1 = 0;
  if (f >= 0 && f < 4) u_f = buf[f];
  if (f1 >= 0 && f1 < 4) u_f1 = buf[f1];
  if (c >= 0 && c < 4) u_c = buf[c];
  if (c1 >= 0 && c1 < 4) u_c1 = buf[c1];
  const T p_f = u_f;
  const T pp_f = (u_c - u_f1) / 2;
  const T p_c = u_c;
  const T pp_c = (u_c1 - u_f) / 2;
  const T a = 2 * p_f - 2 * p_c + pp_f + pp_c;
  const T b = -3 * p_f + 3 * p_c - 2 * pp_f - pp_c;
  const T cc = pp_f;
  const T d = p_f;
  const T w = lc - (T)f;
  return w * (w * (w * a + b) + cc) + d;
}

// The 1D adjoint of cr_local (interpolation_cubicHermiteSpline.cuh:59-117):
// grad_buf[k] = d out / d buf[k] * fwd, grad_w = d out / d w * fwd.  Taps the
// reference would write outside [0,4) are dropped and untouched taps are 0.
template <class T> static inline void cr_local_backward(const T* buf, T lc, T fwd, T* grad_buf, T* grad_w) {
  const int f = floor_int(lc);
  const int c = f + 1, f1 = f - 1, c1 = c + 1;
  const T w = lc - (T)f;
  const T w2 = w * w;
  const T w3 = w2 * w;
  const T d_f1 = -w3 / 2 + w2 - w / 2;
  const T d_f = (3 * w3) / 2 - (5 * w2) / 2 + 1;
  const T d_c = -(3 * w3) / 2 + 2 * w2 + w / 2;
  const T d_c1 = w3 / 2 - w2 / 2;
  for (int k = 0; k < 4; ++k) grad_buf[k] = 0;
  T u_f = 0, u_f1 = 0, u_c = 0, u_c1 = 0;
  if (f >= 0 && f < 4) { u_f = buf[f]; grad_buf[f] = d_f * fwd; }
  if (f1 >= 0 && f1 < 4) { u_f1 = buf[f1]; grad_buf[f1] = d_f1 * fwd; }
  if (c >= 0 && c < 4) { u_c = buf[c]; grad_buf[c] = d_c * fwd; }
  if (c1 >= 0 && c1 < 4) { u_c1 = buf[c1]; grad_buf[c1] = d_c1 * fwd; }
  const T p_f = u_f;
  const T pp_f = (u_c - u_f1) / 2;
  const T p_c = u_c;
  const T pp_c = (u_c1 - u_f) / 2;
  const T a = 2 * p_f - 2 * p_c + pp_f + pp_c;
  const T b = -3 * p_f + 3 * p_c - 2 * pp_f - pp_c;
  const T cc = pp_f;
  *grad_w = (3 * w2 * a + 2 * b * w + cc) * fwd;
}

// A strided view of one scalar component of a [Z,Y,X,stride] field.
template <class T> struct Field {
  const T* v; int NZ, NY, NX, stride;
  T at(int z, int y, int x) const { return v[(((size_t)z * NY + y) * NX + x) * stride]; }
};

// interpolation_cubicHermiteSpline.cuh:482-525 (scalar); :645-689 (vector
// field), :695-739 (matrix field), :747-791 (CNN) are the same loop on a
// different accessor.
template <class T> static T sample_cubic(const Field<T>& u, T hZ, T hY, T hX, int bc, T cz, T cy, T cx) {
  const int ix_f = floor_int(cx / hX);
  const T wx = cx / hX - (T)ix_f;
  const int iy_f = floor_int(cy / hY);
  const T wy = cy / hY - (T)iy_f;
  const int iz_f = floor_int(cz / hZ);
  const T wz = cz / hZ - (T)iz_f;
  T bz[4], by[4], bx[4];
  for (int dz = -1; dz < 3; ++dz) {
    const int z = bidx(iz_f + dz, u.NZ, bc);
    for (int dy = -1; dy < 3; ++dy) {
      const int y = bidx(iy_f + dy, u.NY, bc);
      for (int dx = -1; dx < 3; ++dx) bx[dx + 1] = u.at(z, y, bidx(ix_f + dx, u.NX, bc));
      by[dy + 1] = cr_local(bx, wx + 1);
    }
    bz[dz + 1] = cr_local(by, wy + 1);
  }
  return cr_local(bz, wz + 1);
}

// interpolation_linear.cuh:249-295 (and the CNN copy :469-514).
template <class T> static T sample_linear(const Field<T>& u, T hZ, T hY, T hX, int bc, T cz, T cy, T cx) {
  const int ix_f = floor_int(cx / hX); const int ix_c = ix_f + 1; const T wx = cx / hX - (T)ix_f;
  const int xf = bidx(ix_f, u.NX, bc), xc = bidx(ix_c, u.NX, bc);
  const int iy_f = floor_int(cy / hY); const int iy_c = iy_f + 1; const T wy = cy / hY - (T)iy_f;
  const int yf = bidx(iy_f, u.NY, bc), yc = bidx(iy_c, u.NY, bc);
  const int iz_f = floor_int(cz / hZ); const int iz_c = iz_f + 1; const T wz = cz / hZ - (T)iz_f;
  const int zf = bidx(iz_f, u.NZ, bc), zc = bidx(iz_c, u.NZ, bc);
  const T u_fff = u.at(zf, yf, xf), u_ffc = u.at(zc, yf, xf), u_fcf = u.at(zf, yc, xf), u_fcc = u.at(zc, yc, xf);
  const T u_cff = u.at(zf, yf, xc), u_cfc = u.at(zc, yf, xc), u_ccf = u.at(zf, yc, xc), u_ccc = u.at(zc, yc, xc);
  T out = (1 - wz) * (1 - wy) * (1 - wx) * u_fff;
  out += wz * (1 - wy) * (1 - wx) * u_ffc;
  out += (1 - wz) * (1 - wy) * wx * u_cff;
  out += wz * (1 - wy) * wx * u_cfc;
  out += (1 - wz) * wy * (1 - wx) * u_fcf;
  out += wz * wy * (1 - wx) * u_fcc;
  out += (1 - wz) * wy * wx * u_ccf;
  out += wz * wy * wx * u_ccc;
  return out;
}

// interpolation_nearest.cuh:250-294: ties go to the upper neighbour.
template <class T> static inline int nearest_axis(T c, T h, int N, int bc) {
  const int f = floor_int(c / h);
  const T dist_f = c / h - (T)f;
  const int cidx = f + 1;
  const T dist_c = (T)cidx - c / h;
  return dist_f < dist_c ? bidx(f, N, bc) : bidx(cidx, N, bc);
}
template <class T> static T sample_nearest(const Field<T>& u, T hZ, T hY, T hX, int bc, T cz, T cy, T cx) {
  const int x = nearest_axis(cx, hX, u.NX, bc);
  const int y = nearest_axis(cy, hY, u.NY, bc);
  const int z = nearest_axis(cz, hZ, u.NZ, bc);
  return u.at(z, y, x);
}

template <class T> static T sample(int ip, const Field<T>& u, T hZ, T hY, T hX, int bc, T cz, T cy, T cx) {
  if (ip == IP_CUBIC) return sample_cubic(u, hZ, hY, hX, bc, cz, cy, cx);
  if (ip == IP_LINEAR) return sample_linear(u, hZ, hY, hX, bc, cz, cy, cx);
  return sample_nearest(u, hZ, hY, hX, bc, cz, cy, cx);
}

// ---------------------------------------------------------------- warping
// warpingOps.cu:572-596 (scalar) and :632-660 (vector field, C components):
// out(i) = u(i*h + phi(i)); N from the tensor, h from the mesh.
template <class T> static void warp(const T* u, int C, const T* phi, int NZ, int NY, int NX, T hZ, T hY, T hX,
                                    int ip, int bc, T* out) {
#pragma omp parallel for schedule(static)
  for (int z = 0; z < NZ; ++z)
    for (int y = 0; y < NY; ++y)
      for (int x = 0; x < NX; ++x) {
        const size_t i = ((size_t)z * NY + y) * NX + x;
        const T cx = (T)x * hX + phi[3 * i + 0];
        const T cy = (T)y * hY + phi[3 * i + 1];
        const T cz = (T)z * hZ + phi[3 * i + 2];
        for (int c = 0; c < C; ++c) {
          Field<T> f{u + c, NZ, NY, NX, C};
          out[i * C + c] = sample(ip, f, hZ, hY, hX, bc, cz, cy, cx);
        }
      }
}

// warpingOpsCNN.cu:97-128: u [B,C,Z,Y,X], phi [B,Z,Y,X,3].
template <class T> static void warp_cnn(const T* u, const T* phi, int B, int C, int NZ, int NY, int NX,
                                        T hZ, T hY, T hX, int ip, int bc, T* out) {
  const size_t N = (size_t)NZ * NY * NX;
  for (int b = 0; b < B; ++b) {
    const T* ph = phi + (size_t)b * N * 3;
#pragma omp parallel for schedule(static)
    for (int z = 0; z < NZ; ++z)
      for (int y = 0; y < NY; ++y)
        for (int x = 0; x < NX; ++x) {
          const size_t i = ((size_t)z * NY + y) * NX + x;
          const T cx = (T)x * hX + ph[3 * i + 0];
          const T cy = (T)y * hY + ph[3 * i + 1];
          const T cz = (T)z * hZ + ph[3 * i + 2];
          for (int c = 0; c < C; ++c) {
            Field<T> f{u + ((size_t)b * C + c) * N, NZ, NY, NX, 1};
            out[((size_t)b * C + c) * N + i] = sample(ip, f, hZ, hY, hX, bc, cz, cy, cx);
          }
        }
  }
}

// Adjoint of the tricubic warp, interpolation_cubicHermiteSpline.cuh:527-636
// and warpingOps.cu:598-630: grad_phi accumulates the z term, then one y term
// per dz, then one x term per (dz,dy); grad_u receives the 64 taps (the
// reference uses atomics; here voxels are visited in z,y,x order).
template <class T> static void warp_backward_cubic_one(const Field<T>& u, T hZ, T hY, T hX, int bc,
                                                       T cz, T cy, T cx, T fwd, T* grad_u, size_t gstride, T* gphi) {
  const int ix_f = floor_int(cx / hX); const T wx = cx / hX - (T)ix_f;
  const int iy_f = floor_int(cy / hY); const T wy = cy / hY - (T)iy_f;
  const int iz_f = floor_int(cz / hZ); const T wz = cz / hZ - (T)iz_f;
  T bz[4], by[4], bx[4], gz[4], gy[4], gx[4], gw;
  for (int dz = -1; dz < 3; ++dz) {
    const int z = bidx(iz_f + dz, u.NZ, bc);
    for (int dy = -1; dy < 3; ++dy) {
      const int y = bidx(iy_f + dy, u.NY, bc);
      for (int dx = -1; dx < 3; ++dx) bx[dx + 1] = u.at(z, y, bidx(ix_f + dx, u.NX, bc));
      by[dy + 1] = cr_local(bx, wx + 1);
    }
    bz[dz + 1] = cr_local(by, wy + 1);
  }
  cr_local_backward(bz, wz + 1, fwd, gz, &gw);
  gphi[2] += gw / hZ;
  for (int dz = -1; dz < 3; ++dz) {
    const int z = bidx(iz_f + dz, u.NZ, bc);
    for (int dy = -1; dy < 3; ++dy) {
      const int y = bidx(iy_f + dy, u.NY, bc);
      for (int dx = -1; dx < 3; ++dx) bx[dx + 1] = u.at(z, y, bidx(ix_f + dx, u.NX, bc));
      by[dy + 1] = cr_local(bx, wx + 1);
    }
    cr_local_backward(by, wy + 1, gz[dz + 1], gy, &gw);
    gphi[1] += gw / hY;
    for (int dy = -1; dy < 3; ++dy) {
      const int y = bidx(iy_f + dy, u.NY, bc);
      for (int dx = -1; dx < 3; ++dx) bx[dx + 1] = u.at(z, y, bidx(ix_f + dx, u.NX, bc));
      cr_local_backward(bx, wx + 1, gy[dy + 1], gx, &gw);
      gphi[0] += gw / hX;
      for (int dx = -1; dx < 3; ++dx) {
        const int x = bidx(ix_f + dx, u.NX, bc);
        grad_u[(((size_t)z * u.NY + y) * u.NX + x) * gstride] += gx[dx + 1];
      }
    }
  }
}

// interpolation_linear.cuh:297-363 and warpingOps.cu:390-415 (trilinear adjoint).
template <class T> static void warp_backward_linear_one(const Field<T>& u, T hZ, T hY, T hX, int bc,
                                                        T cz, T cy, T cx, T fv, T* grad_u, size_t gs, T* gphi) {
  const int ix_f = floor_int(cx / hX); const int ix_c = ix_f + 1; const T wx = cx / hX - (T)ix_f;
  const int xf = bidx(ix_f, u.NX, bc), xc = bidx(ix_c, u.NX, bc);
  const int iy_f = floor_int(cy / hY); const int iy_c = iy_f + 1; const T wy = cy / hY - (T)iy_f;
  const int yf = bidx(iy_f, u.NY, bc), yc = bidx(iy_c, u.NY, bc);
  const int iz_f = floor_int(cz / hZ); const int iz_c = iz_f + 1; const T wz = cz / hZ - (T)iz_f;
  const int zf = bidx(iz_f, u.NZ, bc), zc = bidx(iz_c, u.NZ, bc);
  const T u_fff = u.at(zf, yf, xf), u_ffc = u.at(zc, yf, xf), u_fcf = u.at(zf, yc, xf), u_fcc = u.at(zc, yc, xf);
  const T u_cff = u.at(zf, yf, xc), u_cfc = u.at(zc, yf, xc), u_ccf = u.at(zf, yc, xc), u_ccc = u.at(zc, yc, xc);
  auto G = [&](int z, int y, int x) -> T& { return grad_u[(((size_t)z * u.NY + y) * u.NX + x) * gs]; };
  G(zf, yf, xf) += (1 - wz) * (1 - wy) * (1 - wx) * fv;
  G(zf, yc, xf) += (1 - wz) * wy * (1 - wx) * fv;
  G(zf, yf, xc) += (1 - wz) * (1 - wy) * wx * fv;
  G(zf, yc, xc) += (1 - wz) * wy * wx * fv;
  G(zc, yf, xf) += wz * (1 - wy) * (1 - wx) * fv;
  G(zc, yc, xf) += wz * wy * (1 - wx) * fv;
  G(zc, yf, xc) += wz * (1 - wy) * wx * fv;
  G(zc, yc, xc) += wz * wy * wx * fv;
  gphi[0] += ((1 - wz) * (1 - wy) * (u_cff - u_fff) + (1 - wz) * wy * (u_ccf - u_fcf) + wz * (1 - wy) * (u_cfc - u_ffc) + wz * wy * (u_ccc - u_fcc)) / hX * fv;
  gphi[1] += ((1 - wz) * (1 - wx) * (u_fcf - u_fff) + (1 - wz) * wx * (u_ccf - u_cff) + wz * (1 - wx) * (u_fcc - u_ffc) + wz * wx * (u_ccc - u_cfc)) / hY * fv;
  gphi[2] += ((1 - wy) * (1 - wx) * (u_ffc - u_fff) + (1 - wy) * wx * (u_cfc - u_cff) + wy * (1 - wx) * (u_fcc - u_fcf) + wy * wx * (u_ccc - u_ccf)) / hZ * fv;
}

// Warping3D.backward (warpingOps.cu:1173-1262) for CUBIC and LINEAR; the
// nearest adjoint (interpolation_nearest.cuh:296-347) scatters g to the chosen
// tap and reports a zero flow gradient.
template <class T> static void warp_backward(const T* u, const T* phi, const T* gout, int NZ, int NY, int NX,
                                             T hZ, T hY, T hX, int ip, int bc, T* grad_u, T* grad_phi) {
  const size_t N = (size_t)NZ * NY * NX;
  std::fill(grad_u, grad_u + N, (T)0);
  std::fill(grad_phi, grad_phi + 3 * N, (T)0);
  Field<T> f{u, NZ, NY, NX, 1};
  for (int z = 0; z < NZ; ++z)
    for (int y = 0; y < NY; ++y)
      for (int x = 0; x < NX; ++x) {
        const size_t i = ((size_t)z * NY + y) * NX + x;
        const T cx = (T)x * hX + phi[3 * i + 0];
        const T cy = (T)y * hY + phi[3 * i + 1];
        const T cz = (T)z * hZ + phi[3 * i + 2];
        if (ip == IP_CUBIC) {
          warp_backward_cubic_one(f, hZ, hY, hX, bc, cz, cy, cx, gout[i], grad_u, 1, grad_phi + 3 * i);
        } else if (ip == IP_LINEAR) {
          warp_backward_linear_one(f, hZ, hY, hX, bc, cz, cy, cx, gout[i], grad_u, 1, grad_phi + 3 * i);
        } else {
          const int xx = nearest_axis(cx, hX, NX, bc), yy = nearest_axis(cy, hY, NY, bc), zz = nearest_axis(cz, hZ, NZ, bc);
          grad_u[((size_t)zz * NY + yy) * NX + xx] += gout[i];
        }
      }
}

// Batched adjoint of the CNN warp (no reference counterpart: the reference
// comments this out at warpingOpsCNN.h:36-41).  Per (b, c) it is the
// Warping3D.backward of that channel; the flow gradient sums over channels
// in channel order.
template <class T> static void warp_cnn_backward(const T* u, const T* phi, const T* gout, int B, int C, int NZ, int NY,
                                                 int NX, T hZ, T hY, T hX, int ip, int bc, T* grad_u, T* grad_phi) {
  const size_t N = (size_t)NZ * NY * NX;
  std::fill(grad_u, grad_u + (size_t)B * C * N, (T)0);
  std::fill(grad_phi, grad_phi + (size_t)B * N * 3, (T)0);
  for (int b = 0; b < B; ++b)
    for (int z = 0; z < NZ; ++z)
      for (int y = 0; y < NY; ++y)
        for (int x = 0; x < NX; ++x) {
          const size_t i = ((size_t)z * NY + y) * NX + x;
          const T* ph = phi + ((size_t)b * N + i) * 3;
          const T cx = (T)x * hX + ph[0], cy = (T)y * hY + ph[1], cz = (T)z * hZ + ph[2];
          T* gp = grad_phi + ((size_t)b * N + i) * 3;
          for (int c = 0; c < C; ++c) {
            const size_t off = ((size_t)b * C + c) * N;
            Field<T> f{u + off, NZ, NY, NX, 1};
            if (ip == IP_CUBIC)
              warp_backward_cubic_one(f, hZ, hY, hX, bc, cz, cy, cx, gout[off + i], grad_u + off, 1, gp);
            else if (ip == IP_LINEAR)
              warp_backward_linear_one(f, hZ, hY, hX, bc, cz, cy, cx, gout[off + i], grad_u + off, 1, gp);
            else {
              const int xx = nearest_axis(cx, hX, NX, bc), yy = nearest_axis(cy, hY, NY, bc), zz = nearest_axis(cz, hZ, NZ, bc);
              grad_u[off + ((size_t)zz * NY + yy) * NX + xx] += gout[off + i];
            }
          }
        }
}

// ------------------------------------------------------------ prolongation
// prolongationOps.cu:533-625: x_src = ((i * hP) * L) / LP on the source mesh.
template <class T> static void prolongate(const T* u, int C, const Mesh<T>& src, const Mesh<T>& dst, int ip, int bc, T* out) {
#pragma omp parallel for schedule(static)
  for (int z = 0; z < dst.NZ; ++z)
    for (int y = 0; y < dst.NY; ++y)
      for (int x = 0; x < dst.NX; ++x) {
        const T cxp = (T)x * dst.hX; const T cx = cxp * src.LX / dst.LX;
        const T cyp = (T)y * dst.hY; const T cy = cyp * src.LY / dst.LY;
        const T czp = (T)z * dst.hZ; const T cz = czp * src.LZ / dst.LZ;
        const size_t i = ((size_t)z * dst.NY + y) * dst.NX + x;
        for (int c = 0; c < C; ++c) {
          Field<T> f{u + c, src.NZ, src.NY, src.NX, C};
          out[i * C + c] = sample(ip, f, src.hZ, src.hY, src.hX, bc, cz, cy, cx);
        }
      }
}

// ----------------------------------------------------- central differences
// differentialOps.cu:540-656 (scalar) and :782-907 (vector field): one
// derivative along an axis of length N at index i from values at i-1, i, i+1.
template <class T> static inline T cd_fwd(T bm, T b0, T bp, int i, int N, T h, int bc) {
  if (bc == BD_MIRROR) return (i > 0 && i < N - 1) ? (T)((T)0.5 * (bp - bm)) / h : (T)0;
  if (i > 0) return (i < N - 1) ? (T)((T)0.5 * (bp - bm)) / h : (T)((T)0.5 * (b0 - bm)) / h;
  return (T)((T)0.5 * (bp - b0)) / h;
}
// Adjoint along one axis, differentialOps.cu:686-772 and :909-1044.  pm2..pp2
// are the values at i-1, i, i+1 (out-of-range entries are never read).
template <class T> static inline T cd_adj(T pm, T p0, T pp, int i, int N, int bc) {
  if (bc == BD_MIRROR) {
    if (i > 1) return (i < N - 2) ? (T)0.5 * (pm - pp) : (T)0.5 * pm;
    return (T)-0.5 * pp;
  }
  if (i > 0) return (i < N - 1) ? (T)0.5 * (pm - pp) : (T)0.5 * (pm + p0);
  return (T)0.5 * (-p0 - pp);
}

// Nabla3D_CD.forward / forwardVectorField: in [Z,Y,X,C] -> out [Z,Y,X,C,3].
template <class T> static void nabla_fwd(const T* b, int C, int NZ, int NY, int NX, T hZ, T hY, T hX, int bc, T* out) {
#pragma omp parallel for schedule(static)
  for (int z = 0; z < NZ; ++z)
    for (int y = 0; y < NY; ++y)
      for (int x = 0; x < NX; ++x) {
        auto B = [&](int zz, int yy, int xx, int c) -> T {
          if (zz < 0 || zz >= NZ || yy < 0 || yy >= NY || xx < 0 || xx >= NX) return (T)0;
          return b[(((size_t)zz * NY + yy) * NX + xx) * C + c];
        };
        const size_t i = ((size_t)z * NY + y) * NX + x;
        for (int c = 0; c < C; ++c) {
          const T b0 = B(z, y, x, c);
          out[(i * C + c) * 3 + 0] = cd_fwd(B(z, y, x - 1, c), b0, B(z, y, x + 1, c), x, NX, hX, bc);
          out[(i * C + c) * 3 + 1] = cd_fwd(B(z, y - 1, x, c), b0, B(z, y + 1, x, c), y, NY, hY, bc);
          out[(i * C + c) * 3 + 2] = cd_fwd(B(z - 1, y, x, c), b0, B(z + 1, y, x, c), z, NZ, hZ, bc);
        }
      }
}
// Nabla3D_CD.backward / backwardVectorField: p [Z,Y,X,C,3] -> out [Z,Y,X,C];
// out = ax/hX + ay/hY + az/hZ in that order.
template <class T> static void nabla_bwd(const T* p, int C, int NZ, int NY, int NX, T hZ, T hY, T hX, int bc, T* out) {
#pragma omp parallel for schedule(static)
  for (int z = 0; z < NZ; ++z)
    for (int y = 0; y < NY; ++y)
      for (int x = 0; x < NX; ++x) {
        auto P = [&](int zz, int yy, int xx, int c, int d) -> T {
          if (zz < 0 || zz >= NZ || yy < 0 || yy >= NY || xx < 0 || xx >= NX) return (T)0;
          return p[((((size_t)zz * NY + yy) * NX + xx) * C + c) * 3 + d];
        };
        const size_t i = ((size_t)z * NY + y) * NX + x;
        for (int c = 0; c < C; ++c) {
          const T ax = cd_adj(P(z, y, x - 1, c, 0), P(z, y, x, c, 0), P(z, y, x + 1, c, 0), x, NX, bc);
          const T ay = cd_adj(P(z, y - 1, x, c, 1), P(z, y, x, c, 1), P(z, y + 1, x, c, 1), y, NY, bc);
          const T az = cd_adj(P(z - 1, y, x, c, 2), P(z, y, x, c, 2), P(z + 1, y, x, c, 2), z, NZ, bc);
          out[i * C + c] = ax / hX + ay / hY + az / hZ;
        }
      }
}

// ----------------------------------------------------- anisotropic tensor
// anisotropicDifferentialOps.cu:90-141.
template <class T> static void aniso_tangents(const T* G, size_t N, T alpha, T beta, T* s, T* n, T* t1, T* t2) {
#pragma omp parallel for schedule(static)
  for (long long ii = 0; ii < (long long)N; ++ii) {
    const size_t i = (size_t)ii;
    T ns = 0;
    for (int d = 0; d < 3; ++d) ns += G[3 * i + d] * G[3 * i + d];
    const T nrm = sqrt_t(ns);
    if (nrm > 0) {
      s[i] = weight_t(exp_t(-alpha * pow_t(nrm, beta)), i);
      for (int d = 0; d < 3; ++d) n[3 * i + d] = G[3 * i + d] / nrm;
      if (n[3 * i + 2] > (T)0.5) {
        t1[3 * i + 0] = n[3 * i + 2]; t1[3 * i + 1] = 0; t1[3 * i + 2] = -n[3 * i + 0];
      } else {
        t1[3 * i + 0] = n[3 * i + 1]; t1[3 * i + 1] = -n[3 * i + 0]; t1[3 * i + 2] = 0;
      }
      const T* nn = n + 3 * i; const T* tt = t1 + 3 * i;
      t2[3 * i + 0] = nn[1] * tt[2] - nn[2] * tt[1];
      t2[3 * i + 1] = nn[2] * tt[0] - nn[0] * tt[2];
      t2[3 * i + 2] = nn[0] * tt[1] - nn[1] * tt[0];
    } else {
      s[i] = 0;
      for (int d = 0; d < 3; ++d) { n[3 * i + d] = 0; t1[3 * i + d] = 0; t2[3 * i + d] = 0; }
    }
  }
}
// anisotropicDifferentialOps.cu:144-175: out[c,d] = s (n.in_c) n_d + (t1.in_c) t1_d
// + 0 * t2_d (the t2 factor is never accumulated in the reference).
template <class T> static void aniso_apply(const T* in, const T* s, const T* n, const T* t1, const T* t2, size_t N, T* out) {
#pragma omp parallel for schedule(static)
  for (long long ii = 0; ii < (long long)N; ++ii) {
    const size_t i = (size_t)ii;
    for (int c = 0; c < 3; ++c) {
      T nf = 0, tf = 0;
      const T t2f = 0;
      for (int d = 0; d < 3; ++d) {
        nf += n[3 * i + d] * in[9 * i + 3 * c + d];
        tf += t1[3 * i + d] * in[9 * i + 3 * c + d];
      }
      for (int d = 0; d < 3; ++d) {
        T o = s[i] * nf * n[3 * i + d];
        o += tf * t1[3 * i + d];
        o += t2f * t2[3 * i + d];
        out[9 * i + 3 * c + d] = o;
      }
    }
  }
}

// ------------------------------------------------------------------ proxes
// opticalFlowOps.cu:123-175 (factor = tau * lambda, :442).
template <class T> static void prox_primal(const T* primal, T tau, T lam, const T* rho, const T* g, size_t N, T* out) {
  const T factor = tau * lam;
#pragma omp parallel for schedule(static)
  for (long long ii = 0; ii < (long long)N; ++ii) {
    const size_t i = (size_t)ii;
    const T r = rho[i];
    T g2 = 0;
    for (int c = 0; c < 3; ++c) g2 += g[3 * i + c] * g[3 * i + c];
    T delta[3];
    if (r < -factor * g2) {
      for (int c = 0; c < 3; ++c) delta[c] = factor * g[3 * i + c];
    } else if (r > factor * g2) {
      for (int c = 0; c < 3; ++c) delta[c] = -factor * g[3 * i + c];
    } else if ((double)g2 > 1e-10) {
      for (int c = 0; c < 3; ++c) delta[c] = -r * g[3 * i + c] / g2;
    } else {
      for (int c = 0; c < 3; ++c) delta[c] = 0;
    }
    for (int c = 0; c < 3; ++c) out[3 * i + c] = primal[3 * i + c] + delta[c];
  }
}
// opticalFlowOps.cu:179-208: row-wise projection onto the l2 ball of radius w.
// hub (not in the reference, which is hub = 1): the Huber-TV form of the same
// step, p_c = v_c / max(hub, ||v_c|| / lambda_tv) with hub = 1 + sigma eps /
// lambda_tv, i.e. v / hub projected onto the ball (the prox of the conjugate of
// lambda_tv H_eps(||.||)) written with one division; hub = 1 is the
// reference's projection, bit for bit.
template <class T> static void prox_dual(const T* dual, T wtv, size_t N, T* out, T hub = (T)1) {
#pragma omp parallel for schedule(static)
  for (long long ii = 0; ii < (long long)N; ++ii) {
    const size_t i = (size_t)ii;
    for (int c = 0; c < 3; ++c) {
      T ns = 0;
      for (int d = 0; d < 3; ++d) ns += dual[9 * i + 3 * c + d] * dual[9 * i + 3 * c + d];
      const T den = fmax_t(hub, sqrt_t(ns) / wtv);
      for (int d = 0; d < 3; ++d) out[9 * i + 3 * c + d] = dual[9 * i + 3 * c + d] / den;
    }
  }
}

// ---------------------------------------------------------------- blur
// flow_utils.py:94-112: zero-padded 3D correlation with a K^3 kernel (the
// weights are supplied by the caller; flow_utils.py:74-91 builds them).
template <class T> static void blur(const T* in, int NZ, int NY, int NX, const T* w, int K, T* out) {
  const int P = K / 2;
#pragma omp parallel for schedule(static)
  for (int z = 0; z < NZ; ++z)
    for (int y = 0; y < NY; ++y)
      for (int x = 0; x < NX; ++x) {
#ifdef SVS_ORACLE_BLUR_F64
        double acc = 0;
#else
        T acc = 0;
#endif
        for (int a = 0; a < K; ++a) {
          const int zz = z + a - P;
          if (zz < 0 || zz >= NZ) continue;
          for (int b = 0; b < K; ++b) {
            const int yy = y + b - P;
            if (yy < 0 || yy >= NY) continue;
            for (int c = 0; c < K; ++c) {
              const int xx = x + c - P;
              if (xx < 0 || xx >= NX) continue;
#ifdef SVS_ORACLE_BLUR_F64
              acc = acc + (double)w[((size_t)a * K + b) * K + c] * (double)in[((size_t)zz * NY + yy) * NX + xx];
#else
              acc = acc + w[((size_t)a * K + b) * K + c] * in[((size_t)zz * NY + yy) * NX + xx];
#endif
            }
          }
        }
        out[((size_t)z * NY + y) * NX + x] = (T)acc;
      }
}

// ---------------------------------------------------------------- solver
struct SolveParams {
  int num_scales, max_warps, max_iters;
  double weight_matching, weight_tv;
  int pd_type;
  double sigma, tau, theta, gamma;
  int use_aniso;
  double alpha, beta;
  int ip, bc;
  int length_fixed;          // 0: numDofs (L = N-1), 1: fixed lengths below
  double LZ, LY, LX;
  double huber_eps;          // 0: the reference's TV (see prox_dual)
  double relax_theta;        // 0: the reference's z = u (see solve_level)
};

// Per-iteration Huber factor hub_n = 1 + sigma_n eps / lambda_tv, formed in
// double from the Python-float sigma_n and the FP32 lambda_tv, used in T.
template <class T> static T huber_factor(const SolveParams& P, double sigma_n) {
  const T wtv = (T)P.weight_tv;
  if (P.huber_eps == 0.0 || wtv == (T)0) return (T)1;
  return (T)(1.0 + sigma_n * P.huber_eps / (double)wtv);
}

// The Huber function of a row norm t: t^2 / (2 eps) for t <= eps, t - eps / 2
// beyond (eps = 0: t), evaluated in double for the energy.
static double huber(double t, double eps) {
  if (eps == 0.0) return t;
  return t <= eps ? t * t / (2.0 * eps) : t - 0.5 * eps;
}

// The step-size schedule of tvl13d.py:208-246 in Python floats (double).
static void step_schedule(const SolveParams& P, std::vector<double>& sig, std::vector<double>& tau) {
  sig.resize(P.max_iters); tau.resize(P.max_iters);
  double s = P.sigma, t = P.tau;
  for (int n = 0; n < P.max_iters; ++n) {
    sig[n] = s; tau[n] = t;
    if (P.pd_type == 2) {
      const double th = 1. / std::sqrt(1. + 2. * P.gamma * t);
      t *= th;
      s /= th;
    }
  }
}

template <class T> static Mesh<T> level_mesh(const SolveParams& P, int NZ, int NY, int NX) {
  // flow_utils.py:48-71 (numDofs: L = N - 1; fixed: L as given).
  if (P.length_fixed) return make_mesh<T>(NZ, NY, NX, (T)P.LZ, (T)P.LY, (T)P.LX);
  return make_mesh<T>(NZ, NY, NX, (T)(NZ - 1), (T)(NY - 1), (T)(NX - 1));
}

// Per-iteration statistics of one iteration n (three doubles; stats may be
// null): the reference's residual norms tvl13d.py:222 (||p^{n+1} - p^n||),
// :237 (||u^{n+1} - u^n||; :253's ||z - z_old|| is the same number since z = u)
// and the energy of the iterate u^n entering the iteration (H_eps = huber(),
// the identity for the reference's eps = 0),
//   E(u^n) = lambda * sum_i |rho(u^n)_i| + lambda_tv * sum_i sum_c H_eps(||(D grad u^n)_{i,c}||_2),
// with rho and D grad u^n exactly as the iteration evaluates them, and lambda,
// lambda_tv the FP32 weights.  Sums and norms are taken in double.  With
// over-relaxation (theta != 0) the dual step differentiates z, not u^n, so the
// energy is not available and is NaN.
struct IterStats {
  double du2 = 0, dp2 = 0, rho_abs = 0, tv = 0;
};
static void store_stats(const IterStats& a, double lam, double wtv, bool relaxed, double* out) {
  out[0] = std::sqrt(a.du2);
  out[1] = std::sqrt(a.dp2);
  out[2] = relaxed ? std::nan("") : lam * a.rho_abs + wtv * a.tv;
}

// One pyramid level, tvl13d.py:164-255.  stats (optional) receives
// [max_warps][max_iters][3] per-iteration statistics (see IterStats).
// relax_theta (not in the reference, whose "relaxation" is z = u, tvl13d.py:
// 251-252): the dual step differentiates z = u^n + theta (u^n - u^{n-1}),
// formed after each primal step as d = u' - u, e = theta * d, z = u' + e in the
// solve's precision with theta rounded to it, and reset to u at every warp's
// first iteration (like sigma and tau).
template <class T> static void solve_level(const SolveParams& P, const Mesh<T>& m, const T* I0, const T* I1,
                                           std::vector<T>& u, std::vector<T>& p, double* stats = nullptr) {
  const size_t N = (size_t)m.NZ * m.NY * m.NX;
  std::vector<T> G(3 * N), s(N), n(3 * N), t1(3 * N), t2(3 * N);
  nabla_fwd(I1, 1, m.NZ, m.NY, m.NX, m.hZ, m.hY, m.hX, P.bc, G.data());
  if (P.use_aniso) aniso_tangents(G.data(), N, (T)P.alpha, (T)P.beta, s.data(), n.data(), t1.data(), t2.data());
  std::vector<double> sig, tau;
  step_schedule(P, sig, tau);
  std::vector<T> I1w(N), Gw(3 * N), rho_c(N), rho(N), zg(9 * N), Dz(9 * N), dual(9 * N), pn(9 * N),
      Dp(9 * N), div(3 * N), primal(3 * N), un(3 * N);
  const T lam = (T)P.weight_matching, wtv = (T)P.weight_tv;
  const bool relax = P.relax_theta != 0.0;
  const T theta = (T)P.relax_theta;
  std::vector<T> z(relax ? 3 * N : 0);
  for (int w = 0; w < P.max_warps; ++w) {
    warp(I1, 1, u.data(), m.NZ, m.NY, m.NX, m.hZ, m.hY, m.hX, P.ip, P.bc, I1w.data());
    warp(G.data(), 3, u.data(), m.NZ, m.NY, m.NX, m.hZ, m.hY, m.hX, P.ip, P.bc, Gw.data());
    // rho_c = I1w - sum(u * Gw, dim=3) - I0.  torch's GPU sum over a last
    // dimension of 3 rounds as (a + c) + b: its reduction gives that dimension
    // a block width of 2 (the largest power of two <= 3), so one lane adds
    // elements 0 and 2 and then combines with the lane holding element 1
    // (aten/src/ATen/native/cuda/Reduce.cuh, set_block_dimension and
    // block_x_reduce; the same source in the CUDA and ROCm builds).
    for (size_t i = 0; i < N; ++i) {
      const T a = round_kernel_output(u[3 * i] * Gw[3 * i]), b = round_kernel_output(u[3 * i + 1] * Gw[3 * i + 1]),
              c = round_kernel_output(u[3 * i + 2] * Gw[3 * i + 2]);
      const T sm = torch_sum3(a, b, c);
      rho_c[i] = (I1w[i] - sm) - I0[i];
    }
    if (relax) z = u;
    for (int it = 0; it < P.max_iters; ++it) {
      const T sg = (T)sig[it], ta = (T)tau[it];
      // dual step: p = prox_dual(p + sigma * D(grad z)), with z == u unless relaxed.
      nabla_fwd(relax ? z.data() : u.data(), 3, m.NZ, m.NY, m.NX, m.hZ, m.hY, m.hX, P.bc, zg.data());
      const T* dz = zg.data();
      if (P.use_aniso) { aniso_apply(zg.data(), s.data(), n.data(), t1.data(), t2.data(), N, Dz.data()); dz = Dz.data(); }
      for (size_t k = 0; k < 9 * N; ++k) { const T t = round_kernel_output(sg * dz[k]); dual[k] = p[k] + t; }
      prox_dual(dual.data(), wtv, N, pn.data(), huber_factor<T>(P, sig[it]));
      if (P.weight_tv == 0.) std::fill(pn.begin(), pn.end(), (T)0);
      IterStats st;
      if (stats) {
        for (size_t k = 0; k < 9 * N; ++k) {
          const double d = (double)pn[k] - (double)p[k];
          st.dp2 += d * d;
        }
        for (size_t r = 0; r < 3 * N; ++r) {
          const double a = dz[3 * r], b = dz[3 * r + 1], c = dz[3 * r + 2];
          st.tv += huber(std::sqrt(a * a + b * b + c * c), P.huber_eps);
        }
      }
      p.swap(pn);
      // primal step with rho evaluated at the current u.
      for (size_t i = 0; i < N; ++i) {
        const T a = round_kernel_output(u[3 * i] * Gw[3 * i]),
                b = round_kernel_output(u[3 * i + 1] * Gw[3 * i + 1]),
                c = round_kernel_output(u[3 * i + 2] * Gw[3 * i + 2]);
        rho[i] = rho_c[i] + torch_sum3(a, b, c);
      }
      const T* dp = p.data();
      if (P.use_aniso) { aniso_apply(p.data(), s.data(), n.data(), t1.data(), t2.data(), N, Dp.data()); dp = Dp.data(); }
      nabla_bwd(dp, 3, m.NZ, m.NY, m.NX, m.hZ, m.hY, m.hX, P.bc, div.data());
      for (size_t k = 0; k < 3 * N; ++k) { const T t = round_kernel_output(ta * div[k]); primal[k] = u[k] - t; }
      prox_primal(primal.data(), ta, lam, rho.data(), Gw.data(), N, un.data());
      if (stats) {
        for (size_t k = 0; k < 3 * N; ++k) {
          const double d = (double)un[k] - (double)u[k];
          st.du2 += d * d;
        }
        for (size_t i = 0; i < N; ++i) st.rho_abs += std::fabs((double)rho[i]);
        store_stats(st, (double)lam, (double)wtv, relax, stats + 3 * ((size_t)w * P.max_iters + it));
      }
      if (relax)
        for (size_t k = 0; k < 3 * N; ++k) {
          const T d = un[k] - u[k];
          const T e = round_kernel_output(theta * d);
          z[k] = un[k] + e;
        }
      u.swap(un);
    }
  }
}

// ------------------------------------------------------- median filter
// flow_utils.py:115-132: scipy.ndimage.median_filter(u, size=(ks,ks,ks,1)) with
// its default mode 'reflect' on B flows [B][Z,Y,X,C].  Per component, the
// window holds the ks^3 values at offsets -(ks/2) .. ks-1-(ks/2) per axis;
// the result is its element of rank ks^3/2 in ascending order.  Indices
// outside [0, N) reflect about the face (-1 -> 0, N -> N-1), with period 2N.
static int reflect_half_sample(int i, int N) {
  const int P = 2 * N;
  i %= P;
  if (i < 0) i += P;
  return i < N ? i : P - 1 - i;
}

template <class T>
static void median3d(const T* in, int B, int NZ, int NY, int NX, int C, int ks, T* out) {
  const int h = ks / 2, n = ks * ks * ks;
  const size_t N = (size_t)NZ * NY * NX;
#pragma omp parallel for schedule(static) collapse(2)
  for (int b = 0; b < B; ++b)
    for (int z = 0; z < NZ; ++z) {
      std::vector<T> win((size_t)n);
      const T* src = in + (size_t)b * N * C;
      for (int y = 0; y < NY; ++y)
        for (int x = 0; x < NX; ++x)
          for (int c = 0; c < C; ++c) {
            int m = 0;
            for (int a = 0; a < ks; ++a)
              for (int e = 0; e < ks; ++e)
                for (int f = 0; f < ks; ++f) {
                  const size_t zz = (size_t)reflect_half_sample(z + a - h, NZ);
                  const size_t yy = (size_t)reflect_half_sample(y + e - h, NY);
                  const size_t xx = (size_t)reflect_half_sample(x + f - h, NX);
                  win[(size_t)m++] = src[((zz * NY + yy) * NX + xx) * C + c];
                }
            std::nth_element(win.begin(), win.begin() + n / 2, win.end());
            out[(size_t)b * N * C + (((size_t)z * NY + y) * NX + x) * C + c] = win[(size_t)(n / 2)];
          }
    }
}

// ------------------------------------------------------- mask metrics
// warping.py:245-270 scores propagated masks with monai's compute_dice and
// compute_hausdorff_distance (include_background=False on one channel, which
// keeps the channel).  Masks are binary; a voxel is set when its value > 0.5.
//   counts: |A|, |B|, |A and B|  (Dice = 2|A and B| / (|A| + |B|)).
//   edges:  set voxels with a 6-neighbour that is unset or outside the volume
//           (mask XOR binary_erosion(mask) with scipy's default cross).
//   directed Hausdorff A -> B: the largest, over edge voxels of A, of the
//   squared Euclidean distance to the nearest edge voxel of B (voxel units);
//   -1 when A has no edge voxel, -2 when only B has none.
static void mask_counts(const float* a, const float* b, size_t n, long long* out) {
  long long na = 0, nb = 0, nab = 0;
  for (size_t i = 0; i < n; ++i) {
    const bool x = a[i] > 0.5f, y = b[i] > 0.5f;
    na += x; nb += y; nab += (x && y);
  }
  out[0] = na; out[1] = nb; out[2] = nab;
}

static std::vector<int> mask_edges(const float* m, int NZ, int NY, int NX) {
  std::vector<int> out;
  auto set = [&](int z, int y, int x) {
    return z >= 0 && z < NZ && y >= 0 && y < NY && x >= 0 && x < NX && m[((size_t)z * NY + y) * NX + x] > 0.5f;
  };
  for (int z = 0; z < NZ; ++z)
    for (int y = 0; y < NY; ++y)
      for (int x = 0; x < NX; ++x)
        if (set(z, y, x) && !(set(z - 1, y, x) && set(z + 1, y, x) && set(z, y - 1, x) && set(z, y + 1, x) &&
                              set(z, y, x - 1) && set(z, y, x + 1))) {
          out.push_back(z); out.push_back(y); out.push_back(x);
        }
  return out;
}

static long long directed_hausdorff2(const std::vector<int>& ea, const std::vector<int>& eb) {
  if (ea.empty()) return -1;
  if (eb.empty()) return -2;
  long long worst = 0;
  for (size_t i = 0; i < ea.size(); i += 3) {
    long long best = -1;
    for (size_t j = 0; j < eb.size(); j += 3) {
      const long long dz = ea[i] - eb[j], dy = ea[i + 1] - eb[j + 1], dx = ea[i + 2] - eb[j + 2];
      const long long d = dz * dz + dy * dy + dx * dx;
      if (best < 0 || d < best) best = d;
    }
    worst = std::max(worst, best);
  }
  return worst;
}

// ------------------------------------------------ spatial augmentation
// transforms/spatial.py.  RandomRotate (:13-128) samples every key with
// F.grid_sample on the grid G = identity_grid @ R (:56), identity_grid(z,y,x) =
// (lx[x], ly[y], lz[z]) with l = torch.linspace(-1, 1, n) (:93-105), and
// rotates flow vectors by R (:115-128).  RandomFlip (:131-168) is torch.flip
// plus a sign change of one flow component; it is exact and lives in oracle.py.
//
// torch.matmul and torch.einsum run on the CPU through the host's BLAS, whose
// 3-term dot products round in one of three ways (rounding):
//   0  fma(a2, b2, fma(a1, b1, fma(a0, b0, +0)))   MKL on Intel hosts
//   1  (a0 * b0 + a1 * b1) + a2 * b2                MKL on AMD hosts
//   2  ((+0 + a0 * b0) + a1 * b1) + a2 * b2
// (std::fma is exact; the +0 starts decide the sign of zero results).
// grid_sample is torch's CPU grid_sampler_3d
// (aten/src/ATen/native/GridSampler.cpp): scalar float code, restated
// expression by expression.
static float dot3(int rounding, float a0, float b0, float a1, float b1, float a2, float b2) {
  if (rounding == 0) return std::fma(a2, b2, std::fma(a1, b1, std::fma(a0, b0, 0.f)));
  if (rounding == 1) return (a0 * b0 + a1 * b1) + a2 * b2;
  return ((0.f + a0 * b0) + a1 * b1) + a2 * b2;
}

static void rotation_grid(const float* lz, const float* ly, const float* lx, int OZ, int OY, int OX,
                          const float* R, int rounding, float* grid) {
#pragma omp parallel for schedule(static)
  for (int z = 0; z < OZ; ++z)
    for (int y = 0; y < OY; ++y)
      for (int x = 0; x < OX; ++x) {
        float* g = grid + (((size_t)z * OY + y) * OX + x) * 3;
        for (int i = 0; i < 3; ++i) g[i] = dot3(rounding, lx[x], R[i], ly[y], R[3 + i], lz[z], R[6 + i]);
      }
}

// v[b, i] <- sum_j R[i][j] v[b, j] on [B, 3, N] (einsum 'ij,tjxyz->tixyz').
static void rotate_vectors(float* v, int B, size_t N, const float* R, int rounding) {
#pragma omp parallel for schedule(static)
  for (int b = 0; b < B; ++b)
    for (size_t n = 0; n < N; ++n) {
      float* p = v + (size_t)b * 3 * N + n;
      const float a0 = p[0], a1 = p[N], a2 = p[2 * N];
      for (int i = 0; i < 3; ++i) p[(size_t)i * N] = dot3(rounding, R[3 * i], a0, R[3 * i + 1], a1, R[3 * i + 2], a2);
    }
}

// grid_sampler_unnormalize / clip_coordinates / reflect_coordinates.
static float gs_unnormalize(float c, int size, bool align) {
  return align ? ((c + 1.f) / 2.f) * (float)(size - 1) : ((c + 1.f) * (float)size - 1.f) / 2.f;
}
static float gs_clip(float c, int size) { return std::min((float)(size - 1), std::max(c, 0.f)); }
// safe_downgrade_to_int_range: non-finite or out-of-int-range -> -100 (outside).
static float gs_int_range(float c) {
  return (c > (float)(INT32_MAX - 1) || c < (float)INT32_MIN || !std::isfinite(c)) ? -100.f : c;
}
static float gs_reflect(float in, int twice_low, int twice_high) {
  if (twice_low == twice_high) return 0.f;
  const float lo = (float)twice_low / 2.f, span = (float)(twice_high - twice_low) / 2.f;
  in = std::fabs(in - lo);
  const float extra = std::fmod(in, span);
  const int flips = (int)std::floor(in / span);
  return (flips % 2 == 0) ? extra + lo : span - extra + lo;
}
// padding: 0 zeros, 1 border, 2 reflection.
static float gs_source(float c, int size, int padding, bool align) {
  c = gs_unnormalize(c, size, align);
  if (padding == 1) {
    c = gs_clip(c, size);
  } else if (padding == 2) {
    c = align ? gs_reflect(c, 0, 2 * (size - 1)) : gs_reflect(c, -1, 2 * size - 1);
    c = gs_clip(c, size);
  }
  return gs_int_range(c);
}

// x [B, C, IZ, IY, IX] sampled at grid [OZ, OY, OX, 3] (shared by the batch,
// as the repeated grid of spatial.py:79) -> out [B, C, OZ, OY, OX].
// mode: 0 bilinear (trilinear in 3D), 1 nearest.
static void grid_sample3d(const float* x, int B, int C, int IZ, int IY, int IX, const float* grid, int OZ,
                          int OY, int OX, int mode, int padding, bool align, float* out) {
  const size_t Nin = (size_t)IZ * IY * IX, Nout = (size_t)OZ * OY * OX;
  auto inside = [&](long z, long y, long x) { return z >= 0 && z < IZ && y >= 0 && y < IY && x >= 0 && x < IX; };
#pragma omp parallel for schedule(static) collapse(2)
  for (int bc = 0; bc < B * C; ++bc)
    for (int z = 0; z < OZ; ++z) {
      const float* in = x + (size_t)bc * Nin;
      for (int y = 0; y < OY; ++y)
        for (int xo = 0; xo < OX; ++xo) {
          const size_t o = ((size_t)z * OY + y) * OX + xo;
          const float ix = gs_source(grid[3 * o], IX, padding, align);
          const float iy = gs_source(grid[3 * o + 1], IY, padding, align);
          const float iz = gs_source(grid[3 * o + 2], IZ, padding, align);
          auto at = [&](long zz, long yy, long xx) { return in[((size_t)zz * IY + yy) * IX + xx]; };
          float r = 0.f;
          if (mode == 1) {
            const long xn = (long)std::nearbyint(ix), yn = (long)std::nearbyint(iy), zn = (long)std::nearbyint(iz);
            r = inside(zn, yn, xn) ? at(zn, yn, xn) : 0.f;
          } else {
            const long x0 = (long)std::floor(ix), y0 = (long)std::floor(iy), z0 = (long)std::floor(iz);
            const long x1 = x0 + 1, y1 = y0 + 1, z1 = z0 + 1;
            const float wx0 = (float)x1 - ix, wx1 = ix - (float)x0;
            const float wy0 = (float)y1 - iy, wy1 = iy - (float)y0;
            const float wz0 = (float)z1 - iz, wz1 = iz - (float)z0;
            // (t|b)(n|s)(w|e): z0|z1, y0|y1, x0|x1; weights multiplied x, y, z.
            if (inside(z0, y0, x0)) r += at(z0, y0, x0) * (wx0 * wy0 * wz0);
            if (inside(z0, y0, x1)) r += at(z0, y0, x1) * (wx1 * wy0 * wz0);
            if (inside(z0, y1, x0)) r += at(z0, y1, x0) * (wx0 * wy1 * wz0);
            if (inside(z0, y1, x1)) r += at(z0, y1, x1) * (wx1 * wy1 * wz0);
            if (inside(z1, y0, x0)) r += at(z1, y0, x0) * (wx0 * wy0 * wz1);
            if (inside(z1, y0, x1)) r += at(z1, y0, x1) * (wx1 * wy0 * wz1);
            if (inside(z1, y1, x0)) r += at(z1, y1, x0) * (wx0 * wy1 * wz1);
            if (inside(z1, y1, x1)) r += at(z1, y1, x1) * (wx1 * wy1 * wz1);
          }
          out[(size_t)bc * Nout + o] = r;
        }
    }
}

// TVL13DOpticalFlow.compute, tvl13d.py:69-162.  I0/I1 are already blurred
// when the caller wants the blur (the blur weights come from flow_utils.py).
// stats (optional): [num_scales][max_warps][max_iters][3] per-iteration
// statistics, level 0 the finest (see IterStats).
template <class T> static void solve(const SolveParams& P, const T* I0, const T* I1, int NZ, int NY, int NX,
                                     const T* u_in, const T* p_in, T* u_out, T* p_out, double* stats = nullptr) {
  const int S = P.num_scales;
  std::vector<Mesh<T>> meshes;
  meshes.push_back(level_mesh<T>(P, NZ, NY, NX));
  int nz = NZ, ny = NY, nx = NX;
  for (int s = 1; s < S; ++s) {
    nz = (int)std::ceil(0.5 * nz); ny = (int)std::ceil(0.5 * ny); nx = (int)std::ceil(0.5 * nx);
    meshes.push_back(level_mesh<T>(P, nz, ny, nx));
  }
  std::vector<std::vector<T>> I0s(S), I1s(S), us(S), ps(S);
  const size_t N0 = (size_t)NZ * NY * NX;
  I0s[0].assign(I0, I0 + N0); I1s[0].assign(I1, I1 + N0);
  us[0].assign(u_in, u_in + 3 * N0); ps[0].assign(p_in, p_in + 9 * N0);
  for (int s = 1; s < S; ++s) {
    const size_t Ns = (size_t)meshes[s].NZ * meshes[s].NY * meshes[s].NX;
    I0s[s].resize(Ns); I1s[s].resize(Ns);
    prolongate(I0s[s - 1].data(), 1, meshes[s - 1], meshes[s], P.ip, P.bc, I0s[s].data());
    prolongate(I1s[s - 1].data(), 1, meshes[s - 1], meshes[s], P.ip, P.bc, I1s[s].data());
    us[s].assign(3 * Ns, (T)0); ps[s].assign(9 * Ns, (T)0);
  }
  for (int s = S - 1; s >= 0; --s) {
    solve_level(P, meshes[s], I0s[s].data(), I1s[s].data(), us[s], ps[s],
                stats ? stats + (size_t)s * P.max_warps * P.max_iters * 3 : nullptr);
    if (s == 0) break;
    const Mesh<T>& fine = meshes[s - 1];
    const Mesh<T>& coarse = meshes[s];
    const size_t Nf = (size_t)fine.NZ * fine.NY * fine.NX;
    us[s - 1].resize(3 * Nf); ps[s - 1].resize(9 * Nf);
    prolongate(us[s].data(), 3, coarse, fine, P.ip, P.bc, us[s - 1].data());
    // tvl13d.py:151-153: the ratio is a Python float of two FP32 lengths.
    const T fx = (T)((double)fine.LX / (double)coarse.LX);
    const T fy = (T)((double)fine.LY / (double)coarse.LY);
    const T fz = (T)((double)fine.LZ / (double)coarse.LZ);
    for (size_t i = 0; i < Nf; ++i) { us[s - 1][3 * i] *= fx; us[s - 1][3 * i + 1] *= fy; us[s - 1][3 * i + 2] *= fz; }
    prolongate(ps[s].data(), 9, coarse, fine, P.ip, P.bc, ps[s - 1].data());
  }
  std::copy(us[0].begin(), us[0].end(), u_out);
  std::copy(ps[0].begin(), ps[0].end(), p_out);
}

}  // namespace orc

// ============================================================ C interface
using namespace orc;

#define ORC_MESH(T) T hZ, T hY, T hX
#define ORC_INSTANTIATE(T, SFX)                                                                                  \
  extern "C" int orc_bdry_index_##SFX(int i, int N, int bc) { return bidx(i, N, bc); }                          \
  extern "C" T orc_cr_local_##SFX(const T* buf, T lc) { return cr_local<T>(buf, lc); }                          \
  extern "C" void orc_warp_##SFX(const T* u, int C, const T* phi, int NZ, int NY, int NX, ORC_MESH(T), int ip,  \
                                 int bc, T* out) { warp<T>(u, C, phi, NZ, NY, NX, hZ, hY, hX, ip, bc, out); }   \
  extern "C" void orc_warp_cnn_##SFX(const T* u, const T* phi, int B, int C, int NZ, int NY, int NX,            \
                                     ORC_MESH(T), int ip, int bc, T* out) {                                    \
    warp_cnn<T>(u, phi, B, C, NZ, NY, NX, hZ, hY, hX, ip, bc, out);                                            \
  }                                                                                                            \
  extern "C" void orc_warp_backward_##SFX(const T* u, const T* phi, const T* g, int NZ, int NY, int NX,         \
                                          ORC_MESH(T), int ip, int bc, T* gu, T* gphi) {                        \
    warp_backward<T>(u, phi, g, NZ, NY, NX, hZ, hY, hX, ip, bc, gu, gphi);                                     \
  }                                                                                                            \
  extern "C" void orc_warp_cnn_backward_##SFX(const T* u, const T* phi, const T* g, int B, int C, int NZ,       \
                                              int NY, int NX, ORC_MESH(T), int ip, int bc, T* gu, T* gphi) {   \
    warp_cnn_backward<T>(u, phi, g, B, C, NZ, NY, NX, hZ, hY, hX, ip, bc, gu, gphi);                           \
  }                                                                                                            \
  extern "C" void orc_prolongate_##SFX(const T* u, int C, int NZ, int NY, int NX, T LZ, T LY, T LX, int NZp,    \
                                       int NYp, int NXp, T LZp, T LYp, T LXp, int ip, int bc, T* out) {         \
    Mesh<T> a = make_mesh<T>(NZ, NY, NX, LZ, LY, LX), b = make_mesh<T>(NZp, NYp, NXp, LZp, LYp, LXp);          \
    prolongate<T>(u, C, a, b, ip, bc, out);                                                                    \
  }                                                                                                            \
  extern "C" void orc_nabla_fwd_##SFX(const T* b, int C, int NZ, int NY, int NX, ORC_MESH(T), int bc, T* out) { \
    nabla_fwd<T>(b, C, NZ, NY, NX, hZ, hY, hX, bc, out);                                                       \
  }                                                                                                            \
  extern "C" void orc_nabla_bwd_##SFX(const T* p, int C, int NZ, int NY, int NX, ORC_MESH(T), int bc, T* out) { \
    nabla_bwd<T>(p, C, NZ, NY, NX, hZ, hY, hX, bc, out);                                                       \
  }                                                                                                            \
  extern "C" void orc_aniso_tangents_##SFX(const T* G, long long N, T alpha, T beta, T* s, T* n, T* t1, T* t2) { \
    aniso_tangents<T>(G, (size_t)N, alpha, beta, s, n, t1, t2);                                                \
  }                                                                                                            \
  extern "C" void orc_aniso_apply_##SFX(const T* in, const T* s, const T* n, const T* t1, const T* t2,          \
                                        long long N, T* out) {                                                 \
    aniso_apply<T>(in, s, n, t1, t2, (size_t)N, out);                                                          \
  }                                                                                                            \
  extern "C" void orc_prox_primal_##SFX(const T* primal, T tau, T lam, const T* rho, const T* g, long long N,   \
                                        T* out) {                                                              \
    prox_primal<T>(primal, tau, lam, rho, g, (size_t)N, out);                                                  \
  }                                                                                                            \
  extern "C" void orc_prox_dual_##SFX(const T* dual, T wtv, long long N, T* out) {                              \
    prox_dual<T>(dual, wtv, (size_t)N, out);                                                                   \
  }                                                                                                            \
  extern "C" void orc_blur_##SFX(const T* in, int NZ, int NY, int NX, const T* w, int K, T* out) {              \
    blur<T>(in, NZ, NY, NX, w, K, out);                                                                        \
  }                                                                                                            \
  extern "C" void orc_median3d_##SFX(const T* in, int B, int NZ, int NY, int NX, int C, int ks, T* out) {      \
    median3d<T>(in, B, NZ, NY, NX, C, ks, out);                                                                \
  }                                                                                                            \
  extern "C" void orc_solve_stats_##SFX(const double* params, const T* I0, const T* I1, int NZ, int NY, int NX, \
                                        const T* u_in, const T* p_in, T* u_out, T* p_out, double* stats) {     \
    SolveParams P;                                                                                             \
    P.num_scales = (int)params[0]; P.max_warps = (int)params[1]; P.max_iters = (int)params[2];                 \
    P.weight_matching = params[3]; P.weight_tv = params[4]; P.pd_type = (int)params[5];                        \
    P.sigma = params[6]; P.tau = params[7]; P.theta = params[8]; P.gamma = params[9];                          \
    P.use_aniso = (int)params[10]; P.alpha = params[11]; P.beta = params[12];                                  \
    P.ip = (int)params[13]; P.bc = (int)params[14]; P.length_fixed = (int)params[15];                          \
    P.LZ = params[16]; P.LY = params[17]; P.LX = params[18]; P.huber_eps = params[19];                         \
    P.relax_theta = params[20];                                                                                \
    solve<T>(P, I0, I1, NZ, NY, NX, u_in, p_in, u_out, p_out, stats);                                          \
  }                                                                                                            \
  extern "C" void orc_solve_##SFX(const double* params, const T* I0, const T* I1, int NZ, int NY, int NX,       \
                                  const T* u_in, const T* p_in, T* u_out, T* p_out) {                          \
    orc_solve_stats_##SFX(params, I0, I1, NZ, NY, NX, u_in, p_in, u_out, p_out, nullptr);                      \
  }


ORC_INSTANTIATE(float, f32)
ORC_INSTANTIATE(double, f64)

// The step-size schedule (sigma_n, tau_n) as Python floats.
extern "C" float orc_svs_expf(float x) { return svs_expf(x); }
extern "C" float orc_svs_powf(float x, float y) { return svs_powf(x, y); }
extern "C" double orc_svs_exp_d(double x) { return svs_exp_d(x); }
extern "C" double orc_svs_log_d(double x) { return svs_log_d(x); }

extern "C" void orc_mask_counts(const float* a, const float* b, long long n, long long* out) {
  mask_counts(a, b, (size_t)n, out);
}
extern "C" void orc_mask_hausdorff2(const float* a, const float* b, int NZ, int NY, int NX, long long* out) {
  const std::vector<int> ea = mask_edges(a, NZ, NY, NX), eb = mask_edges(b, NZ, NY, NX);
  out[0] = directed_hausdorff2(ea, eb);
  out[1] = directed_hausdorff2(eb, ea);
}
extern "C" long long orc_mask_edge_count(const float* m, int NZ, int NY, int NX) {
  return (long long)(mask_edges(m, NZ, NY, NX).size() / 3);
}

extern "C" void orc_rotation_grid(const float* lz, const float* ly, const float* lx, int OZ, int OY, int OX,
                                  const float* R, int rounding, float* grid) {
  rotation_grid(lz, ly, lx, OZ, OY, OX, R, rounding, grid);
}
extern "C" void orc_rotate_vectors(float* v, int B, long long N, const float* R, int rounding) {
  rotate_vectors(v, B, (size_t)N, R, rounding);
}
extern "C" void orc_grid_sample3d(const float* x, int B, int C, int IZ, int IY, int IX, const float* grid, int OZ,
                                  int OY, int OX, int mode, int padding, int align, float* out) {
  grid_sample3d(x, B, C, IZ, IY, IX, grid, OZ, OY, OX, mode, padding, align != 0, out);
}

extern "C" void orc_step_schedule(const double* params, double* sig, double* tau) {
  SolveParams P;
  P.max_iters = (int)params[2]; P.pd_type = (int)params[5];
  P.sigma = params[6]; P.tau = params[7]; P.gamma = params[9];
  std::vector<double> s, t;
  step_schedule(P, s, t);
  std::copy(s.begin(), s.end(), sig);
  std::copy(t.begin(), t.end(), tau);
}
