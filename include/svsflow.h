/*
 * svsflow.h — C ABI of libsvsflow.so, the MI355X (gfx950) implementation of
 * the 3D TV-L1 optical flow and tricubic Catmull-Rom warping path of
 * singleVentricleSegmentation.
 *
 * Conventions
 *   - All pointers are device pointers (HIP global memory) unless a comment
 *     says "host".  Tensors are FP32, C-contiguous, x fastest:
 *       volume [Z,Y,X], vector field [Z,Y,X,C], matrix field [Z,Y,X,3,3],
 *       CNN batch [B,C,Z,Y,X] with flow [B,Z,Y,X,3] (components x, y, z).
 *   - `stream` is a hipStream_t (NULL = the legacy default stream).  Every call
 *     only enqueues work; nothing synchronises the host unless stated.
 *   - Outputs are fully overwritten; inputs are never modified.
 *   - Return value: 0 = success, 1 = invalid argument, 2 = HIP error.  The
 *     message of the last failure on the calling thread is svs_last_error().
 *   - Enum values match the reference (svs/src/bindings.cpp:51-55, :99-104):
 *       boundary       0 NEAREST, 1 MIRROR, 2 REFLECT
 *       interpolation  0 NEAREST, 1 LINEAR, 2 CUBIC_HERMITESPLINE
 *   - Mesh spacing h = L / (N - 1) computed in FP32 (coreDefines.h:72-75).
 */
#ifndef SVSFLOW_H
#define SVSFLOW_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Last error message of the calling thread ("" if none). */
const char* svs_last_error(void);

/* out[0:n] = value.  Helper used instead of hipMemset so graphs stay replayable. */
int svs_fill(float* out, int64_t n, float value, void* stream);

/* ------------------------------------------------ differential operators */

/* Nabla3D_CD.forward (C = 1: [Z,Y,X] -> [Z,Y,X,3]) and .forwardVectorField
 * (C = 3: [Z,Y,X,3] -> [Z,Y,X,3,3]); central differences, per-boundary face
 * rules.  Replaces svs/src/differentialOps.h:83-98 (cuda_nabla3d_cd_forward,
 * differentialOps.cu:1447; cuda_nabla3d_cd_forwardVectorField, :1588). */
int svs_nabla3d_forward(const float* b, int C, int NZ, int NY, int NX, float hZ, float hY, float hX, int bdry,
                        float* out, void* stream);

/* Nabla3D_CD.backward (C = 1: [Z,Y,X,3] -> [Z,Y,X]) and .backwardVectorField
 * (C = 3: [Z,Y,X,3,3] -> [Z,Y,X,3]); the exact adjoint of the forward
 * operator.  Replaces differentialOps.h:87-98 (differentialOps.cu:1517, :1659). */
int svs_nabla3d_backward(const float* p, int C, int NZ, int NY, int NX, float hZ, float hY, float hX, int bdry,
                         float* out, void* stream);

/* AnisotropicNabla3D.computeTangentVecs: G [N,3] -> s [N], n, t1, t2 [N,3].
 * Replaces anisotropicDifferentialOps.h:69-72 (anisotropicDifferentialOps.cu:275). */
int svs_aniso_tangents(const float* G, int64_t N, float alpha, float beta, float* s, float* n, float* t1, float* t2,
                       void* stream);

/* AnisotropicNabla3D.forwardVectorField / .backwardVectorField (the same
 * symmetric operator): in [N,3,3] -> out [N,3,3].  Replaces
 * anisotropicDifferentialOps.h:74-84 (anisotropicDifferentialOps.cu:322). */
int svs_aniso_apply(const float* in, const float* s, const float* n, const float* t1, const float* t2, int64_t N,
                    float* out, void* stream);

/* ------------------------------------------------------------- proxes */

/* TVL1OF3D_proxPrimal(primal [N,3], tau, lambda, rho [N], g [N,3]) -> [N,3].
 * Replaces opticalFlowOps.h:103-113 (opticalFlowOps.cu:421). */
int svs_prox_primal(const float* primal, float tau, float lambda, const float* rho, const float* g, int64_t N,
                    float* out, void* stream);

/* TVL1OF3D_proxDual(dual [N,3,3], sigma, weight_tv) -> [N,3,3].
 * Replaces opticalFlowOps.h:116-123 (opticalFlowOps.cu:475). */
int svs_prox_dual(const float* dual, float sigma, float weight_tv, int64_t N, float* out, void* stream);

/* ------------------------------------------------------------ warping */

/* Warping3D.forward (C = 1, u [Z,Y,X]) and .forwardVectorField (u [Z,Y,X,C]):
 * out(i) = u(i*h + phi(i)), phi [Z,Y,X,3].  Replaces warpingOps.h:92-110
 * (cuda_warp3d, warpingOps.cu:1086; cuda_warpVectorField3d, :1265). */
int svs_warp3d(const float* u, int C, const float* phi, int NZ, int NY, int NX, float hZ, float hY, float hX,
               int interpolation, int bdry, float* out, void* stream);

/* WarpingCNN3D.forward: u [B,C,Z,Y,X], phi [B,Z,Y,X,3] -> [B,C,Z,Y,X].
 * Replaces warpingOpsCNN.h:21-34 (warpingOpsCNN.cu:133). */
int svs_warp3d_cnn(const float* u, const float* phi, int B, int C, int NZ, int NY, int NX, float hZ, float hY,
                   float hX, int interpolation, int bdry, float* out, void* stream);

/* Adjoint of the warp with respect to the warped volume and the flow.
 * B = C = 1 is Warping3D.backward(u, phi, grad_out) -> [grad_u, grad_phi]
 * (warpingOps.h:98-103, warpingOps.cu:1173); B, C > 1 is the batched CNN form
 * (grad_phi summed over channels), which the reference leaves commented out
 * (warpingOpsCNN.h:36-41).  grad_u accumulates with float atomics, so its
 * low bits depend on the order of the 64 scatters per voxel. */
int svs_warp3d_backward(const float* u, const float* phi, const float* grad_out, int B, int C, int NZ, int NY,
                        int NX, float hZ, float hY, float hX, int interpolation, int bdry, float* grad_u,
                        float* grad_phi, void* stream);

/* -------------------------------------------------------- resampling */

/* Prolongation3D.forward / .forwardVectorField / .forwardMatrixField with C =
 * 1 / 3 / 9 values per voxel, from mesh (N, L) to mesh (Np, Lp).  Replaces
 * prolongationOps.h:86-103 (prolongationOps.cu:994, :1094, :1198). */
int svs_prolongate3d(const float* u, int C, int NZ, int NY, int NX, float LZ, float LY, float LX, int NZp, int NYp,
                     int NXp, float LZp, float LYp, float LXp, int interpolation, int bdry, float* out,
                     void* stream);

/* apply_gaussian_blur3d: zero-padded K^3 correlation with device weights
 * [K,K,K].  Replaces the F.conv3d call of svs/utils/flow_utils.py:94-112. */
int svs_blur3d(const float* in, int NZ, int NY, int NX, const float* weights, int K, float* out, void* stream);

/* apply_median_filter: for B flows [B][Z,Y,X,C], per component the median
 * (element of rank ks^3/2) of the ks^3 window, SciPy 'reflect' extension at
 * the faces, 1 <= ks <= 7.  Replaces the host round trip through
 * scipy.ndimage.median_filter(size=(ks,ks,ks,1)) of
 * svs/utils/flow_utils.py:115-132, applied per pair at
 * svs/modules/flow/processor.py:116-117.  `out` must not alias `in`. */
int svs_median3d(const float* in, int B, int NZ, int NY, int NX, int C, int ks, float* out, void* stream);

/* ------------------------------------------------ mask metrics (warp harness) */

/* |A|, |B| and |A and B| of two binary masks [Z,Y,X] (value > 0.5 is set) into
 * counts[0..2] (device, 64-bit).  Dice = 2|A and B| / (|A| + |B|): monai's
 * compute_dice as svs/modules/flow/warping.py:264,267 calls it. */
int svs_mask_counts(const float* a, const float* b, int NZ, int NY, int NX, unsigned long long* counts,
                    void* stream);

/* Workspace bytes of svs_mask_hausdorff2 for one volume size. */
size_t svs_mask_hausdorff_workspace_bytes(int NZ, int NY, int NX);

/* Directed squared Hausdorff distances between the edge voxels (6-connected
 * boundary: mask XOR its binary erosion) of two masks: out[0] = A -> B,
 * out[1] = B -> A (device, exact integers in voxel units).  The workspace's
 * first two 64-bit words receive the edge counts of A and B.  monai's
 * compute_hausdorff_distance (svs/modules/flow/warping.py:265,268) is the
 * square root of the larger one. */
int svs_mask_hausdorff2(const float* a, const float* b, int NZ, int NY, int NX, void* workspace, size_t ws_bytes,
                        unsigned long long* out, void* stream);

/* svs_mask_hausdorff2 without a workspace argument and without allocating:
 * no counterpart in the reference.  For callers that cannot allocate device
 * memory (code under hipGraph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  The edge lists live in a fixed 16 MiB device array
 * that every GPU holds until the process ends, so volumes are limited to
 * 2^21 voxels (128^3); larger ones are rejected with an error.  out[0..1] as
 * svs_mask_hausdorff2; edge_counts (device, 2 words, may be NULL) receives
 * the edge counts of A and B.  Every call on one GPU shares that array:
 * calls that overlap on one GPU (different streams, or threads) race. */
int svs_mask_hausdorff2_noalloc(const float* a, const float* b, int NZ, int NY, int NX, unsigned long long* out,
                                unsigned long long* edge_counts, void* stream);

/* ------------------------------------- flow-consistent spatial augmentation */

/* One key of a training sample: src [batch, channels, IZ, IY, IX] ->
 * dst [batch, channels, OZ, OY, OX], both device, float32, not overlapping.
 * Images are (T, C, Z, Y, X), masks (C, Z, Y, X) = batch 1, flows
 * (T, 3, Z, Y, X) with vector = 1 (channels hold the (x, y, z) components). */
typedef struct svs_aug_volume {
  const float* src;
  float* dst;
  int batch;
  int channels;
  int mode;    /* grid_sample mode when rotating: 0 bilinear, 1 nearest */
  int vector;  /* 1: rotate the 3-vectors by R and negate flipped components */
} svs_aug_volume;

/* RandomRotate (svs/modules/transforms/spatial.py:53-128) followed by
 * RandomFlips (spatial.py:142-168), applied to up to 8 keys in one launch.
 *   R      host 3x3 row-major float32 rotation, or NULL when no rotation was
 *          drawn (then O* must equal I*);
 *   grid   device [OZ, OY, OX, 3] = identity_grid @ R as the caller computed
 *          it (spatial.py:56), or NULL to compute it on the device from
 *   axes   device lx[OX] ++ ly[OY] ++ lz[OZ], torch.linspace(-1, 1, n) per
 *          axis (spatial.py:100), multiplied by R here with grid_rounding;
 *   padding 0 zeros, 1 border, 2 reflection (grid_sample padding_mode);
 *   flips  bit 0 X (axis 4), bit 1 Y (axis 3), bit 2 Z (axis 2): the XOR of
 *          the RandomFlips drawn after the rotation;
 *   grid_rounding, vector_rounding  how the host BLAS behind the reference's
 *          torch.matmul (grid) and torch.einsum (flow vectors) rounds its
 *          3-term dot products: 0 fma(a2,b2,fma(a1,b1,fma(a0,b0,+0))) (MKL
 *          on Intel hosts; the grid from ~110K points up), 1 (a0*b0 + a1*b1)
 *          + a2*b2 (MKL on AMD hosts), 2 as 1 but summed from +0.
 * Sampling is torch's CPU grid_sampler_3d, rounding for rounding. */
int svs_spatial_augment(const svs_aug_volume* volumes, int nvol, int IZ, int IY, int IX, int OZ, int OY, int OX,
                        const float* R, const float* grid, const float* axes, int padding, int align_corners,
                        int flips, int grid_rounding, int vector_rounding, void* stream);

/* ---------------------------------------------------- the whole solve */

/* Constructor arguments of TVL13DOpticalFlow (svs/modules/flow/tvl13d.py:14-67)
 * plus launch-shape overrides.  Replaces TVL13DOpticalFlow.compute
 * (tvl13d.py:122-255) and the ~27 launches + 3 host syncs per iteration it
 * makes through the operator API. */
typedef struct svs_tvl1_params {
  int num_scales;            /* pyramid levels, sizes ceil(N/2) per level */
  int max_warps;             /* warps per level */
  int max_outer_iterations;  /* primal-dual iterations per warp */
  float weight_matching;     /* lambda */
  float weight_tv;
  int pd_type;               /* 2: theta = 1/sqrt(1+2 gamma tau), tau *= theta, sigma /= theta */
  double sigma, tau, theta, gamma;
  int use_aniso;
  float alpha, beta;
  int use_blur;
  int blur_K;                 /* odd kernel size */
  const float* blur_weights;  /* host pointer, K^3 FP32 weights (flow_utils.py:74-91, 108-109) */
  int interpolation, boundary;
  int length_fixed;           /* 0: numDofs (L = N - 1); 1: the fixed lengths below */
  float LZ, LY, LX;
  int use_initial_state;      /* num_scales == 1 only: start from the bound u_in, p_in */
  int sweep_bx, sweep_by, sweep_cz; /* fused-sweep shape policy (see svs_tvl1_plan_set_shape); 0 = automatic */
  double huber_eps;           /* >= 0; 0 is the reference's TV.  eps > 0: Huber-TV, each dual row becomes
                                 v_c / max(1 + sigma_n eps / weight_tv, ||v_c|| / weight_tv) (no reference
                                 counterpart: tvl13d.py has no eps) */
  double relax_theta;         /* 0 is the reference's z = u (tvl13d.py:251-252).  theta != 0: the dual step
                                 differentiates z = u^n + theta (u^n - u^{n-1}), reset to u at each warp's first
                                 iteration; the energy statistic is then NaN (no reference counterpart) */
} svs_tvl1_params;

/* Opaque solve plan: pyramid geometry, workspace layout, device tables and an
 * optional captured hipGraph for a fixed list of instance sizes. */
typedef struct svs_plan svs_plan;

/* Bytes of device workspace a plan for `ninst` instances needs; dims holds
 * (NZ, NY, NX) per instance.  0 on invalid arguments. */
size_t svs_tvl1_workspace_bytes(const svs_tvl1_params* params, int ninst, const int* dims);

/* Builds a plan inside `workspace` (>= svs_tvl1_workspace_bytes).  Uploads the
 * static tables on `stream` and synchronises it (plan creation blocks; runs
 * do not). */
int svs_tvl1_plan_create(const svs_tvl1_params* params, int ninst, const int* dims, void* workspace,
                         size_t ws_bytes, void* stream, svs_plan** plan);

/* Binds instance k to its frames I0, I1 [Z,Y,X], optional initial state
 * u_in [Z,Y,X,3] / p_in [Z,Y,X,3,3] (read only when num_scales == 1 and
 * use_initial_state) and outputs u_out [Z,Y,X,3] / p_out [Z,Y,X,3,3]
 * (either may be NULL).  Takes effect at the next run. */
int svs_tvl1_plan_bind(svs_plan* plan, int k, const float* I0, const float* I1, const float* u_in,
                       const float* p_in, float* u_out, float* p_out);

/* Enqueues every instance's solve on `stream` and returns without waiting
 * for it.  use_graph = 1 captures the fixed level x warp x iteration schedule
 * as a hipGraph on first use and replays it; re-binding instances keeps the
 * graph.  If bindings changed since the last run, the descriptor tables are
 * copied in stream order from a ring of 4 pinned staging slots (the host
 * waits only when a slot's previous copy has not executed yet, i.e. when it
 * is more than 4 re-bound runs ahead of the device).  The run waits for the
 * plan's previous work on whatever stream that was enqueued, since it reuses
 * the workspace.  The bound frames and outputs must stay allocated until the
 * run completes. */
int svs_tvl1_plan_run(svs_plan* plan, int use_graph, void* stream);

/* Waits for the plan's last enqueued work, then frees its host side (the
 * workspace belongs to the caller and is free to reuse on return). */
int svs_tvl1_plan_destroy(svs_plan* plan);

/* Launch shape of one level: out[0..7] = bx, by, cz, tiles, blocks, ninst,
 * levels, and the resident sweep's voxels per thread (0: the level runs the
 * marching sweep; see svs_tvl1_plan_set_resident). */
int svs_tvl1_plan_info(const svs_plan* plan, int level, int* out);

/* Enqueues `count` fused primal-dual sweeps of pyramid level `level` on the
 * state left by a previous run (measurement aid: times the sweep kernel on
 * its own with events around it).  No reference counterpart. */
int svs_tvl1_plan_sweeps(svs_plan* plan, int level, int count, void* stream);

/* Autotuner: times candidate fused-sweep shapes (work-group tile and
 * z-chunk) per level on the plan's state (needs one prior run), keeps the
 * fastest; at levels that qualify for the resident sweep it also times one
 * resident launch of a warp's iterations against that many marching sweeps
 * and keeps the faster per iteration (iterations per launch: K or 1).  It
 * remembers both for later plans of the same geometry in this process and
 * drops the captured graph.  report (optional) receives
 * bx, by, cz per level.  Results do not depend on the shape.  No reference
 * counterpart. */
int svs_tvl1_plan_autotune(svs_plan* plan, int reps, void* stream, int* report);

/* Sets the fused-sweep shape of one level and drops the captured graph: for
 * reproducing a measured configuration.  The shape is a policy: tiles (plus a
 * one-voxel ring) at most bx wide within a budget of bx * by threads, marching
 * cz planes; each instance splits its rows evenly and takes as many rows as
 * the budget allows, so an instance of width bx - 2 gets exactly bx x by.
 * Fails if a resulting tile exceeds the kernel's limits or the plan's tile
 * table.  Results do not depend on the shape.  No reference counterpart. */
int svs_tvl1_plan_set_shape(svs_plan* plan, int level, int bx, int by, int cz, void* stream);

/* Chooses how a level runs its primal-dual iterations and drops the captured
 * graph: mode 1 runs each warp's iterations in one resident launch (one
 * work-group per instance holding the state on chip; only for levels whose
 * instances have at most 2048 voxels), 0 launches the marching sweep once per
 * iteration, -1 restores the automatic choice (resident wherever the level
 * qualifies).  Results do not depend on the mode.  No reference counterpart. */
int svs_tvl1_plan_set_resident(svs_plan* plan, int level, int mode, void* stream);

/* Per-iteration statistics, the reference's residual norms made usable:
 * tvl13d.py:222, 237 and 253 compute ||p^{n+1} - p^n||, ||u^{n+1} - u^n|| and
 * ||z^{n+1} - z^n|| (the same number as the second, since z = u) every
 * iteration with a host sync each and then discard them.  With out != NULL,
 * every later run also writes, for pyramid level s (0 = finest), warp w,
 * iteration n and instance k, three floats at
 * out[(((s * max_warps + w) * max_outer_iterations + n) * ninst + k) * 3 + j]:
 *   j = 0  ||u^{n+1} - u^n||_2 over the instance's voxels,
 *   j = 1  ||p^{n+1} - p^n||_2,
 *   j = 2  the energy of the iterate entering the iteration,
 *          E(u^n) = lambda sum_i |rho(u^n)_i| + lambda_tv sum_i sum_c ||(D grad u^n)_{i,c}||_2
 *          with rho and D grad u^n as the iteration evaluates them.
 * The sweeps add the terms per work-group and wave while they update u and p
 * (the same u and p bits as without), a reduction after each warp sums them in
 * a fixed order in double, so the values are the same on every run.  out holds
 * num_scales * max_warps * max_outer_iterations * ninst * 3 floats and must
 * stay allocated while statistics are on; out = NULL turns them off.  The plan
 * allocates its own buffer for the partial sums (freed when statistics are
 * turned off or the plan is destroyed) and re-captures its graph at the next
 * run.  Replacing or removing an output waits for the plan's enqueued work,
 * so the previous output may be freed on return.  No reference counterpart
 * beyond the discarded norms. */
int svs_tvl1_plan_set_stats(svs_plan* plan, float* out, void* stream);

/* One frame pair without a graph: TVL13DOpticalFlow.compute(I0, I1, u, p).
 * Builds a plan in `workspace`, runs it and frees it, so it returns after the
 * solve has completed; the cached-plan calls above are the asynchronous form. */
int svs_tvl1_solve(const svs_tvl1_params* params, const float* I0, const float* I1, int NZ, int NY, int NX,
                   const float* u_in, const float* p_in, float* u_out, float* p_out, void* workspace,
                   size_t ws_bytes, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* SVSFLOW_H */
