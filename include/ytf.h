/* ytf.h -- C-ABI of the MI355X (gfx950) layered-media Green's-function hot
 * path.  Plain pointers and sizes only; no C++ or torch types.
 *
 * Every entry point replaces a piece of the reference Julia code
 * (HortenseBeaudoin/YellowThinFilms); the file:line it stands in for is
 * given with each declaration.  The reference's own Julia API
 * (mod_GreenOperator.jl:56-57, mod_CavityTransfer.jl:48-49) stays in the
 * caller; a ccall shim binds the functions below.
 *
 * Conventions (the reference's):
 *   - complex numbers are interleaved (re, im) doubles = Julia ComplexF64;
 *   - arrays are column-major (first index fastest), as Julia's;
 *   - layer indices are 1-based, as Julia's;
 *   - time convention exp(-i w t), loss Im(eps) > 0, eps0 = mu0 = c = 1.
 *
 * Return codes: YTF_OK; YTF_EINVAL (contract violation, message in
 * ytf_last_error; the reference raises ErrorException there); YTF_EHIP (HIP
 * runtime failure).  Device pointers (suffix _dev) must be allocated on the
 * context's device; work is enqueued on the context's stream and is
 * asynchronous unless the function says otherwise.
 */
#ifndef YTF_H
#define YTF_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define YTF_OK 0
#define YTF_EINVAL 1
#define YTF_EHIP 2

/* Material kinds: which branch of _eigen_QV a layer takes
 * (mod_ModeSolver.jl:203-265): isotropic/uniaxial eps with isotropic mu use
 * the closed form, everything else the general Berreman eigen-solve. */
#define YTF_EPS_ISOTROPIC 0
#define YTF_EPS_UNIAXIAL 1
#define YTF_EPS_BIAXIAL 2
#define YTF_EPS_GENERAL 3
#define YTF_MU_ISOTROPIC 0
#define YTF_MU_GENERAL 1

/* One layer: FlatOpticalLayer / FlatLayer (mod_ModeSolver.jl:71-80).
 * eps, mu: 3x3 complex, column-major, interleaved; d: thickness (also the
 * reference plane of the semi-infinite first and last layers). */
typedef struct ytf_layer {
  double eps[18];
  double mu[18];
  double d;
  int32_t eps_kind;
  int32_t mu_kind;
} ytf_layer;

/* Per-node status bits written to flags_dev. */
#define YTF_NODE_VALID 1u    /* CavityTransferCache.valid  (mod_CavityTransfer.jl:167) */
#define YTF_NODE_TVALID 2u   /* CavityTransferCache.Tvalid[o]                          */
#define YTF_NODE_SINGULAR 4u /* an exactly singular source projection (W'V) solve; a    */
                             /* singular interface solve takes the pivoted-QR fallback */
                             /* (mod_ScatteringMatrix.jl:35-44) and is not flagged      */
#define YTF_NODE_NOMODES 8u  /* mode solve failed (no QR convergence / >2 modes a side) */

typedef struct ytf_ctx ytf_ctx;

/* Context = one device + one stream (the reference's per-task thread pool). */
int ytf_ctx_create(int device, ytf_ctx **ctx);
int ytf_ctx_destroy(ytf_ctx *ctx);
/* Message of the last failed call on ctx (ctx may be NULL for ctx_create). */
const char *ytf_last_error(const ytf_ctx *ctx);
/* Enqueue all later work on this hipStream_t; NULL selects the HIP null
 * (legacy default) stream.  A new context uses a stream of its own. */
int ytf_ctx_set_stream(ytf_ctx *ctx, void *hip_stream);
/* Block until all work enqueued on the context's stream is done. */
int ytf_ctx_synchronize(ytf_ctx *ctx);
/* Kernel timing (instrumentation; the reference times whole calls with
 * @elapsed, Bench_GreenOperator.jl:188-197).  While on, each kernel launch
 * of the context is bracketed by HIP events on the context's stream.
 * set_timing clears the totals.  kernel_time synchronizes the stream and
 * returns the summed time (ms) and launch count of one kernel name:
 * "node_solve", "spectral_apply", "synth_coef", "synth_fp64", "synth_reduce",
 * "synth_rotate", "radial_terms", "bessel_table", "split_columns",
 * "bessel_gemm", "mfma_combine", "batch_solve", "mlp_gemm", "mlp_adam",
 * "mlp_fused", "ktable", "nufft_sort", "nufft_spread", "nufft_fft",
 * "lateral_direct", "lateral_segments", "rot_fill", "op_fft_fwd", "op_mix",
 * "op_fft_inv", "cavity_response", "pole_candidates", "pole_refine", and for
 * whole calls "nufft_apply" (ytf_nufft2d1) and "op_apply"
 * (ytf_operator_apply).  The events are created for timing only
 * (hipEventDisableSystemFence). */
int ytf_ctx_set_timing(ytf_ctx *ctx, int on);
/* As ytf_ctx_set_timing(ctx, 1), but only the launches whose names appear in
 * the comma-separated list `names` are bracketed, so the events (a few
 * microseconds each) stay out of the other launches of a timed loop;
 * ytf_ctx_set_timing ends it.  Besides the per-kernel names above,
 * "nufft_apply" brackets a whole ytf_nufft2d1 call and "op_apply" a whole
 * ytf_operator_apply call. */
int ytf_ctx_set_timing_only(ytf_ctx *ctx, const char *names);
int ytf_ctx_kernel_time(ytf_ctx *ctx, const char *name, double *total_ms, int64_t *launches);

/* Device cache outputs of the node solve, structure-of-arrays with the node
 * index fastest (X[e * n_nodes + node], complex).  Any pointer may be NULL.
 *   Qsrc_dev  [4]  Qf1 Qf2 Qb1 Qb2 of the source layer       (Qf_src, Qb_src)
 *   sf_dev    [6]  forward seed, 2x3 column-major (mode, moment x/y/z)
 *   sb_dev    [6]  backward seed, Modal_amp convention sb = -sib_raw
 *   T_dev     [16] T^o, 4x4 column-major
 *   basis_dev [16] obs-layer basis Exf Eyf Ezf Exb Eyb Ezb KZf KZb (2 modes each)
 *   flags_dev [1]  uint8 YTF_NODE_* bits
 * (CavityTransferCache1D/2D fields, mod_CavityTransfer.jl:160-178, 273-294.) */
/* Term classes of the spectral dyadic (_term_amplitudes). */
#define YTF_TERMS_ALL 0       /* direct (same layer) + scattered            */
#define YTF_TERMS_DIRECT 1    /* the same-layer direct term only            */
#define YTF_TERMS_SCATTERED 2 /* everything but the direct term             */
#define YTF_TERMS_TOEPLITZ 3  /* direct + f->f and b->b reflected terms     */
#define YTF_TERMS_HANKEL 4    /* the cross-branch reflected terms b->f, f->b */

typedef struct ytf_node_cache {
  double *Qsrc_dev;
  double *sf_dev;
  double *sb_dev;
  double *T_dev;
  double *basis_dev;
  uint8_t *flags_dev;
} ytf_node_cache;

/* Batched Green k-node solve: for every node (kx[k], ky[k]) of one stack,
 * modes -> S-matrix cascade -> T^o -> point seed -> obs-layer basis, and
 * optionally the spectral dyadic g~(k; zsrc[p], zobs[p]) of each depth pair.
 *
 * Replaces the per-node body of build_cavity_transfer_1d/_2d
 * (mod_CavityTransfer.jl:224-251, 383-388: Modes_Stack, S_all_ingredients,
 * S_all_Matrix, _cavity_transfer_columns, Source_projection_Berreman_xyz,
 * _fill_obs_basis!) for one observation layer, fused with the per-node
 * strength of _strengths_row! (mod_ClosedFormField.jl:616-659; amplitude_at
 * :539-566 + direct_amplitude :577-602 + obs phases) divided by its
 * quadrature weight.
 *
 *   layers: host array of n_layers (>= 2); omega: angular frequency;
 *   kx_dev, ky_dev: n_nodes doubles each;
 *   source_layer in 1..n_layers-1, obs_layer in 1..n_layers (1-based);
 *   zsrc, zobs: host arrays of n_pairs local depths (zsrc as Dipole.z, zobs
 *     as closed_form_G's z_obs);
 *   ghat_dev: [n_pairs][9][n_nodes] complex, entry c + 3q = field component c
 *     of unit moment q, or NULL;
 *   cache: device outputs (or NULL);
 *   terms: YTF_TERMS_* class of g~ (closed_form_G's `terms`,
 *     mod_ClosedFormField.jl:143-160); DIRECT, TOEPLITZ and HANKEL need
 *     obs_layer == source_layer. */
int ytf_green_nodes(ytf_ctx *ctx, const ytf_layer *layers, int n_layers, double omega, const double *kx_dev,
                    const double *ky_dev, int64_t n_nodes, int source_layer, int obs_layer, int n_pairs,
                    const double *zsrc, const double *zobs, double *ghat_dev, const ytf_node_cache *cache,
                    int terms);
/* The same for box sources of thickness dz_box centred at each zsrc
 * (closed_form_G's Δz_box): the seed factor D in face-difference form
 * (_boxphase, mod_ClosedFormField.jl:86-91), the exterior direct term
 * (|z_obs - zpos| >= dz_box/2), inside the box's own slab the face ramps
 * (interior_slab_at, :815-843) and, if ez_local, the local Ez
 * -p_z / (eps_zz dz_box) (ez_local_density, :861-862).  dz_box = 0 is
 * ytf_green_nodes.  The box must fit in its layer (_check_box_in_layer). */
int ytf_green_nodes_box(ytf_ctx *ctx, const ytf_layer *layers, int n_layers, double omega, const double *kx_dev,
                        const double *ky_dev, int64_t n_nodes, int source_layer, int obs_layer, int n_pairs,
                        const double *zsrc, const double *zobs, double *ghat_dev, const ytf_node_cache *cache,
                        int terms, double dz_box, int ez_local);

/* Cavity response of a source layer at in-plane wavevectors (the scan
 * function of real_axis_poles, mod_KSpaceFeatures.jl:109-127 / 151-214):
 *   out[i] = max(sigma(RL), sigma(RR) [, sigma(inv(I - PhiF RL PhiB RR)) for s > 1])
 * with RL = SLeft[s][1:2,3:4] (the stack above, seen from the layer's top
 * edge), RR = (Sint[s] * SRight[s+1])[3:4,1:2] (the stack below, at its bottom
 * edge), sigma the largest singular value of a 2x2, PhiF = e^{i Qf d_s},
 * PhiB = e^{-i Qb d_s}.  It peaks at the real-axis poles (guided modes,
 * plasmons, polaritons) of the field of a source in layer s.  0 where the
 * mode solve fails; an exactly singular interface solve takes the pivoted-QR
 * fallback of mod_ScatteringMatrix.jl:35-44.
 *   source_layer: 1 .. n_layers - 1; kx_dev, ky_dev, out_dev: n_nodes doubles. */
int ytf_cavity_response(ytf_ctx *ctx, const ytf_layer *layers, int n_layers, double omega, int source_layer,
                        const double *kx_dev, const double *ky_dev, int64_t n_nodes, double *out_dev);

/* Prominent maxima of a real-axis scan (real_axis_poles, mod_KSpaceFeatures.jl:
 * 151-214, the candidate step).  R_dev[m][i] (n_dirs x n doubles) is the
 * response along direction m at k = ks_dev[i]; y = log(R + 1e-300).  A
 * candidate is a local maximum 1 <= i <= n-2 (y[i] > y[i-1], y[i] >= y[i+1],
 * y[i] >= max(max_i y[m][i] + log_floor_rel, log_floor_abs)) whose prominence
 * y[i] - base is >= log_prom, base the higher of its two valleys: left, the
 * lowest sample after the last one above y[i] within W steps (the last lowest),
 * right, the lowest before the first one above y[i] within W steps (the first
 * lowest).  Writes, in no particular order, key_dev[c] = m n + i and
 * val_dev[5c .. 5c+4] = (base, prominence, k of the left valley, k of the right
 * valley, y[i]) for c < min(*n_out, max_out); *n_out = number of candidates
 * (may exceed max_out: call again with room for them).  Synchronizes. */
int ytf_pole_candidates(ytf_ctx *ctx, int n_dirs, int64_t n, const double *R_dev, const double *ks_dev, int W,
                        double log_prom, double log_floor_rel, double log_floor_abs, int max_out, int64_t *key_dev,
                        double *val_dev, int64_t *n_out);

/* Location and half-width of pole candidates (real_axis_poles, mod_KSpaceFeatures.jl:
 * 151-214, step 3), every response evaluated on the device.  Candidate c is
 * cand_dev[7c .. 7c+6] = (phi, a, b, k_left_valley, k_right_valley, base, y):
 *   q: golden-section search for the maximum of f(k) = cavity response at
 *      (k cos phi, k sin phi) on [a, b] (at most 60 steps, until b - a <= 1e-12 b),
 *      q = (a + b)/2;
 *   level = e^base + (max(f(q), e^y) - e^base)/sqrt(2);
 *   the crossings of the level by bisection between q and each valley (at most
 *      50 steps, until |hi - lo| <= 1e-12 q).
 * Each search runs on 8 lanes, 3 steps per round (the decision tree of the
 * next 3 points evaluated at once), with the points and results of the
 * one-step-at-a-time search.
 * out_dev[3c .. 3c+2] = (q, left crossing, right crossing).  The stack arguments
 * are those of ytf_cavity_response.  n_evals may be NULL; otherwise the call
 * waits for the stream and sets *n_evals = the number of responses the
 * searches took. */
int ytf_pole_refine(ytf_ctx *ctx, const ytf_layer *layers, int n_layers, double omega, int source_layer,
                    int64_t n_cand, const double *cand_dev, double *out_dev, int64_t *n_evals);

/* Spectral convolution on the uniform DFT grid (the "G(kx,ky) F{J}" route of
 * mod_UniformSource.jl:229-248):
 *     E = scale * bfft2( g~ .* fft2(J) )        per field component c:
 *     E_c = scale * bfft2( sum_q g~_{c,q} fft2(J_q) ).
 * With scale = dx*dy * dkx*dky*nx*ny/(2pi)^2 / (nx*ny) this is
 * _inverse_transform(g~ .* _forward_transform(J)) exactly (|ramp| = 1).
 *   nx, ny: powers of two in 8..2048, or any length 1..1024 (mixed-radix
 *     transforms; uniform_kxky_grid's Nx = ceil(2L/dx) is any integer);
 *   ghat_dev: [9][nx*ny] complex with node index y + ny*x (y fastest);
 *   J_dev, E_dev: (nx, ny, 3) complex column-major (E may not alias J);
 *   work_dev: scratch of 3*nx*ny complex. */
int ytf_spectral_apply(ytf_ctx *ctx, int nx, int ny, const double *ghat_dev, const double *J_dev, double *E_dev,
                       double *work_dev, double scale);
/* No counterpart in the reference: ytf_spectral_apply for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  It allocates nothing and takes no work
 * buffer: the work array and the root tables live in one file-scope device
 * array of 12.06 MiB that each GPU the library runs on holds until the process
 * ends, set up by ytf_ctx_create.  nx, ny: powers of two in 8..2048 with
 * nx*ny <= 512*512 (larger grids return YTF_EINVAL; use ytf_spectral_apply).
 * Same arguments and result, to the bit, as ytf_spectral_apply.  Every call on
 * one GPU shares that array: calls that overlap on one GPU (other contexts,
 * other streams) race. */
int ytf_spectral_apply_noalloc(ytf_ctx *ctx, int nx, int ny, const double *ghat_dev, const double *J_dev,
                               double *E_dev, double scale);

/* Precision of the radial synthesis. */
#define YTF_SYNTH_FP64 0  /* complex FP64 throughout (the reference's arithmetic) */
/* Bessel values (FP64-evaluated) and node coefficients (power-of-two column
 * scaled) each split into two fp16 numbers; three matrix-core products per
 * term (hi*hi + hi*lo + lo*hi), FP32 accumulation, FP64 combination.
 * Relative error of order 1e-6 of the FP64 result (reported by the tests). */
#define YTF_SYNTH_F16X3 1

/* Radial Bessel synthesis: the 3x3 response of every depth pair at distinct
 * radii, for a pixel on the local x axis.  Replaces the J0/J1/J2 contraction
 * of _integrate_pixel_1d (mod_ClosedFormField.jl:344-390) over the Bessel row
 * cache (_build_bessel_row_cache_vol :304-319) with the node weights of
 * _weighted_mcache (:287-294):
 *   R_p(rho) = sum_k (i wt_k / 2pi) kx_k sum_n J_n(kx_k rho) M_n[g~_p(kx_k, 0)]
 * where g~_p is ytf_green_nodes' spectral dyadic of the 1D nodes (ky = 0).
 *   kx_dev, wt_dev: n_nodes doubles (the CavityTransferCache1D's kx_nodes and
 *     the quadrature weights);
 *   ghat_dev: [n_pairs][9][n_nodes] complex (ytf_green_nodes output);
 *   rho_dev: n_rho radii; R_dev: [n_pairs][9][n_rho] complex, entry c + 3q;
 *   mode: YTF_SYNTH_*. */
int ytf_radial_synthesis(ytf_ctx *ctx, int64_t n_nodes, const double *kx_dev, const double *wt_dev, int n_pairs,
                         const double *ghat_dev, int n_rho, const double *rho_dev, double *R_dev, int mode);

/* Per-node terms of ytf_radial_synthesis: T_dev[k][p][c + 3q][r] = node k's
 * term of R_p(rho_r), which is R_p(rho_r) of a one-node cache (kx_k, wt_k).
 * Replaces the per-k_x one-node caches of point_kx_grid's probes
 * (cache_at / closed_form_pixel, mod_ClosedFormGrid.jl:407-411): with
 * wt_k = 1 a pixel at (rho_r, 0) sees T[k] · moment.
 *   kx_dev, wt_dev: n_nodes doubles; ghat_dev: [n_pairs][9][n_nodes] complex
 *   (ytf_green_nodes output); rho_dev: n_rho radii;
 *   T_dev: [n_nodes][n_pairs][9][n_rho] complex. */
int ytf_radial_terms(ytf_ctx *ctx, int64_t n_nodes, const double *kx_dev, const double *wt_dev, int n_pairs,
                     const double *ghat_dev, int n_rho, const double *rho_dev, double *T_dev);

/* closed_form_G for a CavityTransferCache1D (mod_ClosedFormField.jl:435-487):
 * the radial synthesis at the distinct radii of the pixel grid, rotated to
 * every pixel's azimuth.  Depth pair p is the z row p.
 *   xs, ys: host arrays of pixel offsets (nx, ny);
 *   G_dev: (nx, ny, n_pairs, 3, 3) complex column-major (x, y, z, component,
 *     moment). */
int ytf_closed_form_G_1d(ytf_ctx *ctx, int64_t n_nodes, const double *kx_dev, const double *wt_dev, int n_pairs,
                         const double *ghat_dev, const double *xs, int nx, const double *ys, int ny, double *G_dev,
                         int mode);

/* Uneven batches: many independent stacks, each with its own node list,
 * solved in one persistent launch (captured as a HIP graph).  Per node: the
 * Berreman modes of every layer with continuity matching (Modes_Stack,
 * mod_ModeSolver.jl:358-413) and the stack's total S-matrix
 * SLeft[N] = SRight[1] of S_all_Matrix (mod_ScatteringMatrix.jl:112-154), with
 * the first and last layers' modes.  Replaces the per-stack, per-node loop of
 * Modes_Stack + S_all_Matrix over a batch of stacks.
 *   layers: all stacks' layers concatenated (host); layer_counts[s] >= 2;
 *   node_counts[s]: nodes of stack s (its nodes are contiguous in kx/ky, stack
 *     after stack). */
typedef struct ytf_batch ytf_batch;
int ytf_batch_create(ytf_ctx *ctx, const ytf_layer *layers, const int32_t *layer_counts, const int64_t *node_counts,
                     int n_stacks, double omega, ytf_batch **out);
int ytf_batch_destroy(ytf_ctx *ctx, ytf_batch *batch);
/* Outputs (device, complex, node index fastest, any but S may be NULL):
 *   S_dev [16][n]: total S, 4x4 column-major (zero where flags lack VALID);
 *   V1_dev, VN_dev [16][n]: [Vf Vb] of the first / last layer, 4x4 column-major;
 *   Q1_dev, QN_dev [4][n]: Qf1 Qf2 Qb1 Qb2; flags_dev [n]: YTF_NODE_* bits. */
int ytf_batch_solve(ytf_ctx *ctx, ytf_batch *batch, const double *kx_dev, const double *ky_dev, double *S_dev,
                    double *V1_dev, double *Q1_dev, double *VN_dev, double *QN_dev, uint8_t *flags_dev);
/* Launch shape: nodes per work item (multiple of 64), waves per block
 * (1, 2, 4), persistent grid size. */
int ytf_batch_set_config(ytf_ctx *ctx, ytf_batch *batch, int nodes_per_item, int waves_per_block, int blocks);
int ytf_batch_get_config(ytf_ctx *ctx, const ytf_batch *batch, int *nodes_per_item, int *waves_per_block,
                         int *blocks);
/* Autotune the launch shape on this batch's inputs (writes S and flags).
 * tune_db: optional text file; a stored choice for the same batch signature
 * (device, stack/node/work counts) is reused, a new one appended.  report:
 * optional buffer for a one-line summary. */
int ytf_batch_tune(ytf_ctx *ctx, ytf_batch *batch, const double *kx_dev, const double *ky_dev, double *S_dev,
                   uint8_t *flags_dev, const char *tune_db, char *report, int report_len);

/* Learned surrogate of the spectral dyadic (no counterpart in the reference,
 * which evaluates every node with the exact solver; that solve remains the
 * fallback and the arbiter of accuracy).  A fully connected network n_in -> width x depth
 * (SiLU) -> n_out, trained and evaluated on the matrix cores in BF16 with FP32
 * accumulation (FP32 master weights, Adam).  Inputs and targets are raw FP32
 * rows (x[n][n_in], y[n][n_out], device); the network sees
 * (x - in_shift) * in_scale and predicts (y - out_shift) / out_scale.
 * Parameters in flat order: for each layer W [n_out_l][n_in_l] row-major, then
 * b [n_out_l]. */
typedef struct ytf_mlp ytf_mlp;
/* flags: YTF_MLP_KZ_FEATURES -- the first layer also sees Re and Im of
 * sqrt(1 - x0^2 - x1^2) (kz / omega of the air when x0, x1 are kx / omega,
 * ky / omega), so its weights are [width][n_in + 2]. */
#define YTF_MLP_KZ_FEATURES 1
int ytf_mlp_create(ytf_ctx *ctx, int n_in, int width, int depth, int n_out, int flags, uint64_t seed, ytf_mlp **out);
int ytf_mlp_destroy(ytf_ctx *ctx, ytf_mlp *mlp);
int64_t ytf_mlp_num_params(const ytf_mlp *mlp);
int ytf_mlp_get_params(ytf_ctx *ctx, ytf_mlp *mlp, float *params_host);
/* Also resets the optimizer state. */
int ytf_mlp_set_params(ytf_ctx *ctx, ytf_mlp *mlp, const float *params_host);
int ytf_mlp_set_normalization(ytf_ctx *ctx, ytf_mlp *mlp, const float *in_shift, const float *in_scale,
                              const float *out_shift, const float *out_scale);
/* Inference: y_dev[n][n_out] = network(x_dev[n][n_in]). */
int ytf_mlp_forward(ytf_ctx *ctx, ytf_mlp *mlp, int64_t n, const float *x_dev, float *y_dev);
/* Mean-squared loss of the normalised outputs over the batch and its gradient
 * (flat order, host), without an update (synchronizes). */
int ytf_mlp_loss_grad(ytf_ctx *ctx, ytf_mlp *mlp, int64_t n, const float *x_dev, const float *y_dev, float *loss_host,
                      float *grad_host);
/* One Adam step (beta 0.9, 0.999, eps 1e-8) on the batch; the batch loss
 * (before the step) is copied to loss_dev when not NULL. */
int ytf_mlp_train_step(ytf_ctx *ctx, ytf_mlp *mlp, int64_t n, const float *x_dev, const float *y_dev, float lr,
                       float *loss_dev);
/* The pole flag of ytf_mlp_surrogate_ghat from the cavity response: nodes
 * with k_rho inside one of n_windows windows [lo, hi] (windows: lo0 hi0 lo1
 * hi1 ..., host) are flagged from their inputs and solved exactly.  The caller
 * places the windows around the poles of _cavity_response found by
 * real_axis_poles (mod_KSpaceFeatures.jl:109-127, 151-214) -- the fallback
 * indicator of sigma_max(RL), sigma_max(RR), sigma_max((I - PF RL PB RR)^-1);
 * n_windows = 0 clears them.  At most 4096 windows. */
int ytf_mlp_set_pole_windows(ytf_ctx *ctx, ytf_mlp *mlp, int n_windows, const double *windows);
/* Surrogate evaluation of the scattered spectral dyadic above the stack, with
 * the exact solve as fallback (no reference counterpart: the reference runs
 * the exact solve, build_cavity_transfer_2d + closed_form_G's
 * terms=:scattered, on every node).  One kernel evaluates the network on
 * inputs (kx/omega, ky/omega, xconst[0 .. n_in-3]) -- n_in >= 2, 18 outputs,
 * hidden width 128 or 256 -- and forms, per node, in FP64,
 *     g~_s = t omega (1 + k^2/omega^2) exp(i kz0 zsum) / kz0,
 *     t_e = y_e + i y_{9+e},  kz0 = sqrt(omega^2 - k^2) (Im >= 0),
 * into ghat_dev [9][n] complex (entry c + 3q, as ytf_green_nodes).  A node is
 * flagged (flags_dev[i] = 1) when |k - omega| < branch_tol omega (the air
 * branch point), k < origin_tol omega, k lies in a pole window
 * (ytf_mlp_set_pole_windows), or |y|_2 > t_big (a large predicted response).  With
 * fallback, the flagged nodes are solved exactly (ytf_green_nodes with source
 * and observation layer 1, depths zsrc / zobs, YTF_TERMS_SCATTERED) and their
 * g~ replaced (the nodes flagged from their inputs alone before the network,
 * the others after it; every count stays on the device).  n_flagged may be
 * NULL: the call then only queues work on the context's stream; otherwise it
 * waits for that stream and sets *n_flagged = number of flagged nodes. */
int ytf_mlp_surrogate_ghat(ytf_ctx *ctx, ytf_mlp *mlp, const ytf_layer *layers, int n_layers, double omega,
                           const double *kx_dev, const double *ky_dev, int64_t n, const float *xconst, double zsum,
                           double zsrc, double zobs, double branch_tol, double origin_tol, double t_big, int fallback,
                           double *ghat_dev, unsigned char *flags_dev, int64_t *n_flagged);

/* The tabulated k_rho surrogate of the same scattered spectral dyadic, for a
 * stack whose every layer is invariant under rotations about z (isotropic,
 * uniaxial about z, or gyrotropic about z: eps and mu of the form
 * [[a, b, 0], [-b, a, 0], [0, 0, c]]; the caller checks).  No counterpart in
 * the reference, which solves every node; the table is built from exact
 * solves at phi = 0 (the reference's 1D k_rho treatment of such stacks,
 * CavityTransferCache1D, mod_CavityTransfer.jl), and the exact solve stays the
 * fallback.  Per node, with k_rho = |(kx, ky)|, phi its angle and R(phi) the
 * rotation about z:
 *     g~_s = R(phi) [t(k_rho) omega (1 + k^2/omega^2) exp(i kz0 zsum) / kz0] R(phi)^T,
 * t from panel p (kb_dev[p] <= k_rho < kb_dev[p + 1], n_panels contiguous
 * panels, at most 2048): four-point (cubic) Lagrange interpolation of the
 * m_dev[p] >= 4 uniform samples of t over the panel variable s from
 * sab_dev[3p] to sab_dev[3p + 1]; kind_dev[p] & 3 selects s around the
 * panel's real branch point k_b = sab_dev[3p + 2] (0: k_rho, 1: sqrt(k_b^2 -
 * k_rho^2), 2: sqrt(k_rho^2 - k_b^2)).  Sample q of panel p is row
 * r = row0_dev[p] + q of the table, 18 doubles (Re t_e for e < 9, Im t_{e-9}
 * for e >= 9, entry c + 3q) stored entry-major: entry e of row r at
 * val_dev[e * n_rows + r], n_rows = row0_dev[n_panels - 1] + m_dev[n_panels - 1]
 * (the four rows a node reads are then adjacent per entry).  A node is flagged (flags_dev[i] = 1) outside
 * [kb_dev[0], kb_dev[n_panels]], in a panel with kind bit 4 (not converged),
 * or at kz0 = 0; with fallback the flagged nodes are solved exactly as in
 * ytf_mlp_surrogate_ghat (source and observation layer 1, depths zsrc / zobs,
 * YTF_TERMS_SCATTERED).  All table arrays are device arrays; n_flagged as in
 * ytf_mlp_surrogate_ghat. */
int ytf_ktable_ghat(ytf_ctx *ctx, const ytf_layer *layers, int n_layers, double omega, int n_panels,
                    const double *kb_dev, const int32_t *kind_dev, const double *sab_dev, const int64_t *row0_dev,
                    const int32_t *m_dev, const double *val_dev, const double *kx_dev, const double *ky_dev,
                    int64_t n, double zsum, double zsrc, double zobs, int fallback, double *ghat_dev,
                    unsigned char *flags_dev, int64_t *n_flagged);

/* 2D type-1 nonuniform FFT (FINUFFT's type 1 with iflag +1, the transform of
 * closed_form_G's 2D lateral sum with lateral=:nufft,
 * mod_ClosedFormField.jl:758-780):
 *   f[t][k2][k1] = sum_j c[t][j] exp(+i (k1 x_j + k2 y_j)),
 *   k1 = -floor(N1/2) .. , k2 = -floor(N2/2) ..   (k1 fastest: Julia's (N1, N2, ntrans)),
 * to relative accuracy about eps (1e-15 <= eps < 1).
 *   x_dev, y_dev: M doubles (any real; used mod 2 pi);
 *   c_dev: [ntrans][M] complex; f_dev: [ntrans][N2][N1] complex; N1, N2 <= 2048. */
int ytf_nufft2d1(ytf_ctx *ctx, int64_t n_points, const double *x_dev, const double *y_dev, int ntrans,
                 const double *c_dev, int N1, int N2, double eps, double *f_dev);

/* Lateral sum of the 2D path. */
#define YTF_LATERAL_DIRECT 0 /* exact separable sum (lateral=:direct)         */
#define YTF_LATERAL_NUFFT 1  /* type-1 NUFFT, accuracy nufft_eps (:nufft)     */

/* closed_form_G for a CavityTransferCache2D and a point source
 * (mod_ClosedFormField.jl:711-782):
 *   G(x_i, y_j, p, c, q) = sum_n W_n g~_p(n)[c + 3q] exp(i (kx_n x_i + ky_n y_j))
 * with W_n the node weights (k_rho w_rho w_phi / 4 pi^2, times
 * box_lateral_factor for a box footprint, :723-727); non-finite weighted
 * strengths are zeroed (_finite_or_zero).
 *   kx_dev, ky_dev: n_nodes doubles; w_dev: n_nodes complex;
 *   ghat_dev: [n_pairs][9][n_nodes] complex (ytf_green_nodes), pair p = z row p;
 *   xs, ys: host arrays of lateral offsets; YTF_LATERAL_NUFFT needs uniform
 *     xs, ys with nx, ny <= 2048, and 1e-15 <= nufft_eps < 1;
 *   G_dev: (nx, ny, n_pairs, 3, 3) complex column-major. */
int ytf_closed_form_G_2d(ytf_ctx *ctx, int64_t n_nodes, const double *kx_dev, const double *ky_dev,
                         const double *w_dev, int n_pairs, const double *ghat_dev, const double *xs, int nx,
                         const double *ys, int ny, int lateral, double nufft_eps, double *G_dev);

/* Nodes of one-k_rho-node caches (build_cavity_transfer_2d with one k_rho of
 * weight 1, mod_ClosedFormGrid.jl:209) for a batch of k_rho, on the device:
 * node k = i n_phi + j (k_rho i, phi j) at kx = krho_i cos phi_j,
 * ky = krho_i sin phi_j, with the complex weight
 *   w_k = krho_i wphi_j / (4 pi^2) x sinc(kx box_dx / 2) sinc(ky box_dy / 2)
 * (the node weights and box footprint of mod_ClosedFormField.jl:517-524,
 * 723-727; box_dx = box_dy = 0 for a point source).  All arrays on the device. */
int ytf_polar_nodes(ytf_ctx *ctx, int64_t n_krho, const double *krho_dev, int n_phi, const double *phi_dev,
                    const double *wphi_dev, double box_dx, double box_dy, double *kx_dev, double *ky_dev,
                    double *w_dev);

/* The probe values of the adaptive k_rho grids (mod_ClosedFormGrid.jl:204-218:
 * closed_form_G of a one-k_rho-node CavityTransferCache2D at one point per
 * probe, for every visited k_rho), many k_rho per call.  The nodes come in
 * n_seg segments of seg_len consecutive nodes (one k_rho, every phi); for
 * segment s and depth pair p:
 *   out[(s n_pairs + p) 9 + c + 3q] = sum_{j < seg_len} W_k g~_p(k)[c + 3q]
 *                                     exp(i (kx_k x_p + ky_k y_p)),  k = s seg_len + j,
 * with the weights, zeroing and layouts of ytf_closed_form_G_2d (ghat_dev:
 * [n_pairs][9][n_seg seg_len]); x_p, y_p (host arrays, n_pairs) is the probe
 * point of pair p.  out_dev: n_seg x n_pairs x 9 complex. */
int ytf_segment_lateral_sum(ytf_ctx *ctx, int64_t n_seg, int seg_len, const double *kx_dev, const double *ky_dev,
                            const double *w_dev, int n_pairs, const double *ghat_dev, const double *xp,
                            const double *yp, double *out_dev);

/* The pixel grid of uniform_kxky_grid (mod_UniformSource.jl:89-116) as the
 * operator uses it: Nx x Ny pixels of pitch dx, dy.  The operator samples G
 * on the 2Nx x 2Ny padded offsets (padded_offsets,
 * mod_ClosedFormAssembly.jl:75-79); Nx, Ny must be powers of two in 4..1024. */
typedef struct ytf_grid {
  int32_t Nx, Ny;
  double dx, dy;
} ytf_grid;

#define YTF_STRUCTURE_AUTO 0       /* structure=:auto        */
#define YTF_STRUCTURE_STRUCTURED 1 /* structure=:structured  */
#define YTF_STRUCTURE_DENSE 2      /* structure=:dense       */

/* Keyword arguments of build_layer_operator (mod_GreenOperator.jl:430-433)
 * and of the closed_form_G calls it makes. */
typedef struct ytf_operator_options {
  int32_t structure;   /* YTF_STRUCTURE_*                              (:auto)  */
  int32_t check;       /* numerical Toeplitz+Hankel check under :auto  (true)   */
  double check_tol;    /* residual above which the check fails         (1e-8)   */
  int32_t skeleton;    /* skeleton planes for blocks without z structure (true) */
  double skeleton_tol; /* interpolation error per node term            (1e-6)   */
  int32_t lateral;     /* YTF_LATERAL_*                                (:direct)*/
  double nufft_eps;    /*                                              (1e-9)   */
  /* 1: z grids must be uniform (build_layer_operator); 0: any depths, dense
   * blocks only (closed_form_field / precompute_ghats!, mod_ClosedFormAssembly.jl). */
  int32_t uniform_z;
  double dz_box;       /* box sources of this thickness (Δz_box; 0: points)  */
  int32_t ez_local;    /* local Ez inside a box's own slab          (true)   */
  /* skeletons=: precomputed z skeletons (ytf_z_skeletons) shared by the
   * blocks without z structure; NULL builds them from this build's nodes.
   * They must cover the source planes and every observation layer's planes. */
  const struct ytf_skeletons *skeletons;
  /* phi-rotation fill (build_cavity_transfer_2d(...; rotation), mod_CavityTransfer.jl:361-378,
   * with _rotate_modes, mod_ModeSolver.jl:471-482).  rot_n_phi > 0 declares the nodes the
   * product of rot_n_krho values k_rho (host array rot_krho) and rot_n_phi angles (host array
   * rot_phi), k_rho fastest: n_nodes = rot_n_krho rot_n_phi and node k + rot_n_krho q sits at
   * (k_rho_k cos phi_q, k_rho_k sin phi_q) -- the caller keeps kx, ky, w consistent with it.
   * Every layer must take the closed-form branch (isotropic or uniaxial eps, isotropic mu;
   * YTF_EINVAL otherwise).  Each k_rho is then solved once, at phi = 0, and every angle is its
   * rotation, g~(k_rho, phi) = R(phi) g~(k_rho, 0) R(phi)^T with R the rotation about z.
   * rot_n_phi = 0: every node is solved. */
  int32_t rot_n_phi;
  int64_t rot_n_krho;
  const double *rot_krho;
  const double *rot_phi;
} ytf_operator_options;

/* Block kinds (ZBlock.kind, mod_GreenOperator.jl:85-96). */
#define YTF_BLOCK_STRUCTURED 0
#define YTF_BLOCK_TOEPLITZ_DENSE 1
#define YTF_BLOCK_DENSE 2
#define YTF_BLOCK_SKELETON 3
#define YTF_BLOCK_TOEPLITZ_SKELETON 4

/* Bits of build_layer_operator's `warnings` output. */
#define YTF_WARN_STRUCTURE_CHECK 1u /* a Toeplitz+Hankel check failed: dense form used (the @warn of :462) */

/* Precomputed Green operator of one source layer onto a set of observation
 * layers on uniform z grids: LayerGreenOperator + build_layer_operator
 * (mod_GreenOperator.jl:105-113, 430-480).  The nodes are those of a
 * CavityTransferCache2D (every (k_rho, phi) node, its weight W_n as for
 * ytf_closed_form_G_2d); the node solve runs inside the build.  A grouped
 * build (build_layer_operator_grouped, :497-531) is one build on the
 * concatenated nodes of all groups, since every block is linear in the
 * quadrature sum.
 *   z_src: n_src local source depths (Dipole.z), in the caller's order;
 *   obs_layers: n_obs_layers layer indices (1-based, any order, no repeats);
 *     n_obs[l] depths of layer obs_layers[l], concatenated in z_obs;
 *   warnings: YTF_WARN_* bits (may be NULL).
 * Blocks are stored in ascending observation layer order; a layer with no
 * depths gets no block.  Grid sizes: Nx, Ny powers of two up to 1024, or
 * any size up to 512 (the padded 2Nx x 2Ny transforms are mixed-radix). */
typedef struct ytf_operator ytf_operator;
int ytf_operator_build(ytf_ctx *ctx, const ytf_layer *layers, int n_layers, double omega, int source_layer,
                       int64_t n_nodes, const double *kx_dev, const double *ky_dev, const double *w_dev,
                       const ytf_grid *grid, const double *z_src, int n_src, int n_obs_layers,
                       const int32_t *obs_layers, const int32_t *n_obs, const double *z_obs,
                       const ytf_operator_options *opt, uint32_t *warnings, ytf_operator **out);
int ytf_operator_destroy(ytf_ctx *ctx, ytf_operator *op);
/* apply_operator (mod_GreenOperator.jl:674-712): FFTs, per-node z coupling
 * and inverse FFTs, all on the device.
 *   P_dev: (Nx, Ny, n_src, 3) complex column-major dipole moments per pixel
 *     and source plane (caller's order of z_src);
 *   E_dev[b]: (Nx, Ny, n_obs(b), 3) complex for block b. */
int ytf_operator_apply(ytf_ctx *ctx, ytf_operator *op, const double *P_dev, double *const *E_dev);

typedef struct ytf_block_info {
  int32_t obs_layer;   /* 1-based                                          */
  int32_t kind;        /* YTF_BLOCK_*                                      */
  int32_t n_obs;       /* observation planes                               */
  int32_t n_src_skel;  /* source skeleton planes (n_src when kept whole)   */
  int32_t n_obs_skel;  /* observation skeleton planes                      */
  int32_t n_toeplitz;  /* stored Toeplitz planes (0 if none)               */
  int32_t n_hankel;    /* stored Hankel planes (0 if none)                 */
  int32_t n_dense;     /* stored dense plane pairs (0 if none)             */
  int64_t bytes;       /* device bytes of the block                        */
  double build_ms;     /* wall time of the block's build                   */
  double check_residual; /* Toeplitz+Hankel check residual, or -1 if not run */
} ytf_block_info;
/* operator_summary (mod_GreenOperator.jl:735-747). */
int ytf_operator_num_blocks(const ytf_operator *op);
int ytf_operator_block_info(ytf_ctx *ctx, const ytf_operator *op, int block, ytf_block_info *info);
/* The z skeleton of one side of a block (ZSkeleton, mod_GreenOperator.jl:
 * 202-208): side 0 the source planes (in increasing cavity-frame depth),
 * side 1 the observation planes.  n_planes receives k, the skeleton planes
 * (N when the side is kept whole); idx (k, 0-based, increasing), P (N x k
 * complex column-major, host; the identity when whole) and err (the largest
 * interpolation error over every node term) may be NULL. */
int ytf_operator_skeleton(ytf_ctx *ctx, const ytf_operator *op, int block, int side, int32_t *n_planes,
                          int32_t *idx, double *P, double *err);
/* One stored Ĝ plane of a block in the reference layout (Px, Py, 3, 3),
 * Px = 2Nx, Py = 2Ny: part 0 Toeplitz plane m + n_src - 1 (m = j - i),
 * part 1 Hankel plane i + j, part 2 dense pair i + n_src_skel * j (source
 * planes in increasing cavity-frame depth). */
int ytf_operator_ghat(ytf_ctx *ctx, const ytf_operator *op, int block, int part, int slot, double *ghat_dev);

/* z_skeletons (mod_GreenOperator.jl:335-359) and z_skeletons_grouped
 * (:372-382): the interpolative decomposition of each side's z-exponential
 * family (ZSkeleton, :186-208) from the mode wavenumbers of the source layer and
 * of every observation layer at the given nodes, to `tol` per node term (sides
 * of more than 64 planes are kept whole).  For the grouped form pass the nodes
 * of every group (phi = 0 only on an azimuthally symmetric stack).  Planes as
 * for ytf_operator_build: z_src in the caller's order, n_obs[l] depths of layer
 * obs_layers[l] concatenated in z_obs.  The handle goes to
 * ytf_operator_options.skeletons. */
typedef struct ytf_skeletons ytf_skeletons;
int ytf_z_skeletons(ytf_ctx *ctx, const ytf_layer *layers, int n_layers, double omega, int source_layer,
                    int64_t n_nodes, const double *kx_dev, const double *ky_dev, const double *z_src, int n_src,
                    int n_obs_layers, const int32_t *obs_layers, const int32_t *n_obs, const double *z_obs,
                    double tol, ytf_skeletons **out);
/* z_skeletons in the reference's exported form (mod_GreenOperator.jl:335-345):
 * the mode wavenumbers given on every quadrature node, [4][n_nodes] complex
 * (Qf1 Qf2 Qb1 Qb2 per node, device) for the source layer (Qsrc_dev) and for
 * each observation layer (Qobs_dev[l], the (KZf, KZb) of the cache's basis);
 * zs_int: the source planes in increasing cavity-frame depth (as the operator
 * stores them).  Only the layers' thicknesses are read from `layers`. */
int ytf_z_skeletons_q(ytf_ctx *ctx, const ytf_layer *layers, int n_layers, int source_layer, int64_t n_nodes,
                      const double *Qsrc_dev, const double *zs_int, int n_src, int n_obs_layers,
                      const int32_t *obs_layers, const double *const *Qobs_dev, const int32_t *n_obs,
                      const double *z_obs, double tol, ytf_skeletons **out);
/* One side of a skeleton set: layer 0 the source side (planes in increasing
 * cavity-frame depth), otherwise observation layer `layer` (1-based).
 * n_planes receives N, n_skel k (N when kept whole); idx (k, 0-based,
 * increasing), P (N x k complex column-major, host) and err may be NULL. */
int ytf_skeletons_get(ytf_ctx *ctx, const ytf_skeletons *sk, int layer, int32_t *n_planes, int32_t *n_skel,
                      int32_t *idx, double *P, double *err);
int ytf_skeletons_destroy(ytf_ctx *ctx, ytf_skeletons *sk);

/* The block-by-block sum of build_layer_operator_grouped (mod_GreenOperator.jl:
 * 512-527): adds every stored plane of `src` into `dst`.  Both must come from
 * builds with the same grid, planes, block kinds and z skeletons (the grouped
 * build passes one skeleton set to every chunk); otherwise the reference's
 * errors ("groups chose different block kinds", "groups have different z
 * skeletons").  Build times add up. */
int ytf_operator_accumulate(ytf_ctx *ctx, ytf_operator *dst, const ytf_operator *src);

/* apply_operator (mod_GreenOperator.jl:674-712) with the reference's host
 * arrays: P_host (Nx, Ny, n_src, 3) complex column-major, E_host[b] (Nx, Ny,
 * n_obs(b), 3) per block.  The operator stages them in device buffers it keeps;
 * returns after E has been copied back. */
int ytf_operator_apply_host(ytf_ctx *ctx, ytf_operator *op, const double *P_host, double *const *E_host);

/* Device buffers for callers without a GPU array type of their own (the
 * Julia binding's node arrays).  Copies run on the context's stream and
 * return when done. */
int ytf_buffer_alloc(ytf_ctx *ctx, size_t bytes, void **dev);
int ytf_buffer_free(ytf_ctx *ctx, void *dev);
int ytf_buffer_upload(ytf_ctx *ctx, void *dev, const void *host, size_t bytes);
int ytf_buffer_download(ytf_ctx *ctx, void *host, const void *dev, size_t bytes);

/* Library version string. */
const char *ytf_version(void);

#ifdef __cplusplus
}
#endif
#endif /* YTF_H */
