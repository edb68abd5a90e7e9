// ytf_kernels.h -- internal structures shared by the HIP kernels and the
// C-ABI implementation (not part of the public interface).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "ytf_device.hpp"

namespace ytf {

// One layer as the kernels read it: eps, mu column-major 3x3 complex.
struct LayerDev {
  cd eps[9];
  cd mu[9];
  double d;
  int ek;  // 0 isotropic, 1 uniaxial, 2 biaxial, 3 general 3x3
  int mk;  // 0 isotropic mu, 1 general
};

// Arguments of the node kernel.  All pointers are device pointers; outputs
// are structure-of-arrays, node index fastest: X[e * nn + node].
struct NodeParams {
  const LayerDev *L;
  int nl;
  double omega;
  const double *kx, *ky;
  int64_t nn;
  int src, obs;  // 0-based
  int npairs;
  const double *zsrc, *zobs;  // device, npairs each
  cd *ghat;                   // [npairs][9][nn] (c + 3q) or null
  cd *Qsrc;                   // [4][nn]  Qf1 Qf2 Qb1 Qb2
  cd *sf, *sb;                // [6][nn]  2x3 column-major
  cd *T;                      // [16][nn] 4x4 column-major
  cd *basis;                  // [16][nn] Exf Eyf Ezf Exb Eyb Ezb KZf KZb
  unsigned char *flags;       // [nn] bit0 valid, bit1 Tvalid, bit2 singular solve, bit3 mode solve failed
  bool needs_general;         // some layer needs the general 4x4 eigen-solver
  int terms;                  // TERMS_*: which terms of g~
  double dz;                  // box thickness (0: point source)
  int ez_local;               // add the local Ez inside the box's slab
  // Optional row list: node k < *n_rows (at most nn) solves row rows[k] of
  // the inputs and outputs (nn is then their row count); null: node k, k < nn.
  const int *rows = nullptr;
  const int *n_rows = nullptr;
  // Redo marks of the rows whose fast solve met an exactly singular interface
  // (word 0: some row marked, word 1: block counter, one byte per row from
  // byte 8; zero on entry and cleared again by the redo launch).  With it the
  // launch solves those rows again with the pivoted-QR fallback; without it
  // they keep the fast solve's answer: invalid, with the singular-solve bit.
  int *redo = nullptr;
};

// term classes of the spectral dyadic (mod_ClosedFormField.jl:143-160)
enum { TERMS_ALL = 0, TERMS_DIRECT = 1, TERMS_SCATTERED = 2, TERMS_TOEPLITZ = 3, TERMS_HANKEL = 4 };

// pair_groups > 1 splits the depth pairs into that many groups (grid.y), each
// solving the node again for its share of g~ (at most npairs groups)
hipError_t launch_node_top(const NodeParams &P, int blocks, hipStream_t stream, int pair_groups = 1);
hipError_t launch_node_general(const NodeParams &P, int blocks, hipStream_t stream, int pair_groups = 1);

// 2D spectral convolution E = scale * bfft2(g~ . fft2(J)) on a power-of-two
// nx x ny grid; see fft_kernels.hip for layouts.
hipError_t launch_spectral_apply(int nx, int ny, const cd *ghat_xy, const cd *J, cd *E, cd *work, const cd *tw_x,
                                 const cd *tw_y, double scale, hipStream_t stream);

// Radial Bessel synthesis (synth.hip): node coefficients from g~, the FP64
// contraction (Bessel table, node chunks, fixed-order sum), the nine entries,
// the pixel rotation.
hipError_t launch_synth_coef(int64_t n, int np, const double *kx, const double *wt, const cd *g, cd *coef,
                             hipStream_t s);
// FP64 contraction plan: nchunks node chunks of `chunk` nodes (about 2048
// waves in all), Bessel tables of chunks_per_batch chunks (batch_nodes nodes,
// at most 2^23 entries per order).
struct SynthPlan {
  int64_t chunk, batch_nodes;
  int nchunks, chunks_per_batch;
};
SynthPlan synth_fp64_plan(int64_t n, int np, int nr);
// Bessel table J (3 x batch_nodes x nr doubles), chunk partials part
// (nchunks x np x 9 x nr complex), their fixed-order sum S (np x 9 x nr).
hipError_t launch_synth_fp64(int64_t n, int np, int nr, const SynthPlan &P, const double *kx, const double *rho,
                             const cd *coef, double *J, cd *part, cd *S, hipStream_t s);
hipError_t launch_synth_terms(int64_t n, int np, int nr, const double *kx, const double *wt, const cd *g,
                              const double *rho, cd *T, hipStream_t s);
hipError_t launch_synth_reduce(int np, int nr, int nks, const cd *part, cd *R, hipStream_t s);
hipError_t launch_synth_rotate(int nx, int ny, int np, int nr, const double *xs, const double *ys, const int *rmap,
                               const cd *R, cd *G, hipStream_t s);

// The same contraction on the matrix cores with split fp16 products
// (synth_mfma.hip): workspace sizes, then Bessel table + column split + GEMM +
// FP64 combine into R.
struct MfmaPlan {
  int64_t Kpad;
  int Mpad;
  int64_t out_slot;  // floats per output slot (the GEMM writes two)
  size_t jtab_bytes, ctab_bytes, cexp_bytes, out_bytes, split_bytes;
};
MfmaPlan mfma_plan(int64_t n, int np, int nr);
hipError_t launch_bessel_table(int64_t n, int np, int nr, const double *kx, const double *rho, void *jtab,
                               hipStream_t s);
hipError_t launch_split_columns(int64_t n, int np, int nr, const cd *coef, void *ctab, int *cexp, hipStream_t s);
// n_cu: compute units of the device (the GEMM runs two blocks per CU over
// equal contiguous ranges of all tiles' K steps); split: MfmaPlan::split_bytes
hipError_t launch_bessel_gemm(int64_t n, int np, int nr, const void *jtab, const void *ctab, float *out, int *split,
                              int n_cu, hipStream_t s);
hipError_t launch_mfma_combine(int64_t n, int np, int nr, const float *out, const int *split, const int *cexp, cd *R,
                               hipStream_t s);

// Uneven batch of stacks (batch.hip): total S-matrix + end-layer modes per
// node, one persistent launch (counter reset + kernel).
struct BatchLaunch {
  const LayerDev *L;
  const int *stack_l0, *stack_nl;
  const int64_t *stack_n0, *stack_nn;
  const int *item_stack, *item_first;
  int n_items, nodes_per_item;
  double omega;
  const double *kx, *ky;
  cd *S, *V1, *Q1, *VN, *QN;
  unsigned char *flags;
  int64_t ntot;
  unsigned int *counter;
  int *redo;  // redo marks: word 0 some node marked, word 1 block counter, a byte per node from byte 8
  int n_stacks;
  bool general;
  int waves_per_block, blocks;
};
hipError_t launch_batch(const BatchLaunch &B, hipStream_t s);

// Cavity response of a source layer (response.hip): the scan function of the
// real-axis pole search; s 1-based.
// redo: redo marks of the n nodes (word 0, word 1, a byte per node from byte
// 8); zero on entry and left zero.
hipError_t launch_cavity_response(const LayerDev *L, int nl, int s, double omega, const double *kx, const double *ky,
                                  int64_t n, double *out, bool general, int *redo, hipStream_t st);
// Golden-section location and half-level crossings of nc candidates, every
// response evaluated in the kernel (response.hip).
// n_cu: compute units of the device (the lanes per search are chosen so
// that the waves fit one per SIMD when they can)
hipError_t launch_pole_refine(const LayerDev *L, int nl, int s, double omega, int64_t nc, const double *cand,
                              double *out, unsigned long long *evals, bool general, int n_cu, hipStream_t st);
// Prominent maxima of a real-axis scan R[nphi][n] (response.hip); scratch Y
// (nphi x n doubles) and ymax (nphi doubles), *count zeroed by the caller.
hipError_t launch_pole_candidates(int nphi, int64_t n, int W, double log_prom, double log_floor_rel,
                                  double log_floor_abs, const double *R, const double *ks, double *Y, double *ymax,
                                  int max_out, int *count, int64_t *key, double *val, hipStream_t s);

// Surrogate network (mlp.hip): one BF16 matrix-core GEMM with fused epilogues,
// plus the small elementwise kernels of training and inference.
enum { MLP_EPI_F32 = 0, MLP_EPI_BIAS_SILU = 1, MLP_EPI_BIAS_F32 = 2, MLP_EPI_DSILU = 3 };
struct MlpGemm {
  const __bf16 *A;
  int64_t lda;
  const __bf16 *B;
  int64_t ldb;
  int M, N, K;  // multiples of 64, 64, 32
  float *C32;
  int64_t ldc;
  __bf16 *C16;
  const float *bias;
  const float *Zin;
  int ksplit;            // > 1: MLP_EPI_F32 only, partial sums at C32 + z * split_stride
  int64_t split_stride;
};
hipError_t launch_mlp_gemm(const MlpGemm &p, int epi, bool at, bool bt, hipStream_t s);
hipError_t launch_mlp_inputs(int64_t n, int64_t rows, int nin, int ldx, int kzf, const float *x, const float *shift,
                             const float *scale, void *X, hipStream_t s);
hipError_t launch_mlp_mse(int64_t n, int64_t rows, int nout, int ldy, const float *Y, const float *y,
                          const float *oshift, const float *oscale, void *dZ, float *loss, hipStream_t s);
hipError_t launch_mlp_colsum(int64_t rows, int ld, const void *dZ, float *part, float *db, hipStream_t s);
hipError_t launch_mlp_outputs(int64_t n, int nout, int ldy, const float *Y, const float *oshift, const float *oscale,
                              float *y, hipStream_t s);
hipError_t launch_mlp_adam(int64_t n, float *w, float *m, float *v, const float *g, void *w16, float lr, float b1,
                           float b2, float eps, float c1, float c2, hipStream_t s);
hipError_t launch_mlp_to_bf16(int64_t n, const float *w, void *w16, hipStream_t s);
hipError_t launch_mlp_zero(int64_t n, float *x, hipStream_t s);
hipError_t launch_mlp_reduce(int64_t n, int splits, int64_t stride, const float *part, float *dst, hipStream_t s);

// Whole-network forward in one kernel (hidden width 128 or 256, n_out <= 64,
// at most 8 hidden layers): 128 rows per block, activations kept in LDS.
constexpr int MLP_FUSED_MAX_LAYERS = 9;  // hidden biases staged in LDS
struct MlpFused {
  int64_t n;                      // rows
  int nin, kzf, nl, nout;         // raw inputs, light-line features, layers, outputs
  int kp0;                        // padded K of layer 0 (row stride of its weights)
  const __bf16 *W16;              // bf16 weights in fragment order (launch_mlp_pack)
  const float *P32;               // FP32 parameters (biases)
  int64_t offW[MLP_FUSED_MAX_LAYERS], offb[MLP_FUSED_MAX_LAYERS];
  const float *norm;              // in_shift[nin] in_scale[nin] out_shift[nout] out_scale[nout]
  // generic forward: x[n][nin] -> y[n][nout]
  const float *x;
  float *y;
  // surrogate evaluation: inputs (kx/omega, ky/omega, xconst[0 .. nin-3]);
  // output t (18 reals) -> g~_s = t omega (1 + k^2/omega^2) exp(i kz0 zsum) / kz0
  // as [9][n] complex (not written for rows with flags[row] != 0 when
  // skip_flagged); rows with flags[row] == 0 and |t|_2 > tbig are flagged and
  // appended to the list of flagged rows
  const double *kx, *ky;
  float xconst[62];
  double omega, zsum, tbig;
  int skip_flagged;
  double *ghat;
  unsigned char *flags;
  int *n_flagged, *flagged;
};
bool mlp_fused_supported(int width, int nin_total, int nout, int nl);
// W [N][K] (row-major bf16) -> fragment order [N/32][K/16][64 lanes][8]: lane L
// of (tile, step) holds row 32 tile + L % 32, columns 16 step + 8 (L / 32) .. + 8.
hipError_t launch_mlp_pack(int N, int K, const __bf16 *W, __bf16 *Wpk, hipStream_t s);
hipError_t launch_mlp_fused(const MlpFused &p, int width, bool surrogate, hipStream_t s);
// Input flags of the surrogate rows (near the light line or the origin) into
// flags[n] (0/1), the flagged rows appended to list through *count.
hipError_t launch_sur_preflag(int64_t n, const double *kx, const double *ky, double w, double branch_tol,
                              double origin_tol, const double *win, int nwin, unsigned char *flags, int *count,
                              int *list, hipStream_t s);

// The tabulated k_rho surrogate (ktable.hip): np contiguous panels with
// boundaries kb[np + 1] (increasing k_rho), kind[np] (0: s = k_rho,
// 1: s = sqrt(k_b^2 - k_rho^2), 2: s = sqrt(k_rho^2 - k_b^2); bit 4: exact
// solve), sab[3 np] (s range, k_b), and per panel m[p] >= 4 uniform samples
// of Re, Im of t's 9 entries in s, rows row0[p] .. row0[p] + m[p] - 1 of
// val[rows][18].
struct KTable {
  int np;
  const double *kb;
  const int *kind;
  const double *sab;
  const int64_t *row0;
  const int *m;
  const double *val;
  double inv_w;  // omega^-1 (set by the launch)
};
// g~_s [9][n] of the nodes from the table; flagged nodes (outside the table,
// unconverged panels, kz0 = 0) get flags 1, no g~, and go to list via *count.
hipError_t launch_ktable_ghat(int64_t n, const double *kx, const double *ky, double w, double zsum, const KTable &T,
                              cd *ghat, unsigned char *flags, int *count, int *list, hipStream_t s);

// 2D type-1 nonuniform FFT (nufft.hip).
int nu_width(double eps);
// Upper bound on the spreading subproblems of M points on an nf x nf fine grid.
int nu_max_subproblems(int64_t M, int nf);
// Binning blocks of the sort for M points (rows of its histogram matrix).
int nu_bin_blocks(int64_t M);
// Sort (once per point set): fine-grid coordinates, tile bins, the sorted
// order and the subproblem table (count, start: ntiles; substart: ntiles + 1;
// sub_tile / sub_p0: nu_max_subproblems entries).  Scratch: rank (M) and H
// (nu_bin_blocks(M) x ntiles).
hipError_t launch_nufft_sort(int64_t M, const double *x, const double *y, int nf, double *uv, int *tile, int *rank,
                             int *H, int *count, int *start, int *substart, int *sub_tile, int *sub_p0, int *order,
                             hipStream_t s);
// nb transforms at once: strengths c + t * c_stride, fine grids grid + t * nf^2;
// hc: the kernel's piecewise-polynomial coefficients, (w + 3) x w, d-major.
hipError_t launch_nufft_spread(int64_t M, int nb, int nf, int w, const double *hc, const int *start,
                               const int *substart, const int *sub_tile, const int *sub_p0, const int *order,
                               const double *uv, const cd *c, int64_t c_stride, cd *grid, hipStream_t s);
// nb transforms at once: grids grid + t * nf^2, T1 + t * N1 nf, f + t * N1 N2.
hipError_t launch_nufft_fft(int nb, int nf, const cd *grid, cd *T1, cd *f, int N1, int N2, const double *d1,
                            const double *d2, const cd *tw, hipStream_t s);

// 2D lateral sum and the padded Green operator (operator.hip).
hipError_t launch_lat_points(int64_t n, const double *kx, const double *ky, const cd *W, double px, double py,
                             double x0, double y0, double *xf, double *yf, cd *wr, hipStream_t s);
hipError_t launch_lat_strengths(int64_t n, int np, const cd *g, const cd *w, cd *out, int64_t tp, int64_t te,
                                hipStream_t s);
// Nodes and box-footprint weights of one-k_rho-node caches (operator.hip).
hipError_t launch_polar_nodes(int64_t nk, int L, const double *kr, const double *phi, const double *wphi, double dx,
                              double dy, double *kx, double *ky, cd *W, hipStream_t s);
// Per-segment lateral sums at one point per depth pair (operator.hip).  The
// points (x, y) come from device arrays xp, yp, or, when those are null, from
// pts (at most LatPts::kMax points, passed by value with the launch).
struct LatPts {
  static constexpr int kMax = 64;
  double x[kMax], y[kMax];
};
hipError_t launch_lat_segments(int64_t n_seg, int L, const double *kx, const double *ky, const cd *W, int np,
                               const cd *g, const double *xp, const double *yp, const LatPts &pts, cd *out,
                               hipStream_t s);
hipError_t launch_lat_direct(int64_t n, const double *kx, const double *ky, const cd *W, int np, const cd *g,
                             const double *xs, int nx, const double *ys, int ny, cd *G, int64_t sp, int64_t se,
                             hipStream_t s);
hipError_t launch_opx(bool inv, int PX, const cd *in, int pdiv, int64_t pa, int64_t pb, int row, int nx, int ny_rows,
                      cd *out, int py, int64_t plane, int planes, const cd *tw, hipStream_t s);
hipError_t launch_opy(bool inv, int PY, cd *work, int px, int64_t plane, int planes, int ny_in, int ny_out,
                      double scale, int sign, const cd *tw, hipStream_t s);
enum { OP_MIX_TOEPLITZ = 0, OP_MIX_HANKEL = 1, OP_MIX_DENSE = 2 };
// H: Hankel planes added in a Toeplitz pass (kind OP_MIX_TOEPLITZ only), or NULL.
hipError_t launch_op_mix(int64_t nn, const cd *J, int ns, cd *E, int no, const cd *K, const cd *H, int kind,
                         int accumulate, hipStream_t s);
hipError_t launch_op_zcomb(int64_t nn, const cd *in, int nin, const cd *M, cd *out, int nout, int accumulate,
                           hipStream_t s);
hipError_t launch_op_add(int64_t n, const cd *src, cd *dst, hipStream_t s);
// g [np][9][nk*nphi] = R(phi_q) g0[p][.][k] R(phi_q)^T at node k + nk*q
// (g0 [np][9][nk]; cs: cos phi at [q], sin phi at [nphi + q]).
hipError_t launch_rot_fill(int64_t nk, int nphi, int np, const cd *g0, const double *cs, cd *g, hipStream_t s);
hipError_t launch_opy_mix1(int PY, const cd *in, cd *out, int px, const cd *K, int ny_in, int ny_out, const cd *tw,
                           hipStream_t s);
// The fused single-plane y pass exists for power-of-two PY 8..2048 with px a
// multiple of its columns per block.
bool opy_mix1_supported(int PY, int px);
// Lengths the LDS FFT passes take: powers of two 8..2048, or 1..1024.
bool fft_length_ok(int n);

// z skeletons (skeleton.hip): the exponential family of one side, its
// tall-skinny QR factor and the interpolation error of a column ID.
struct SkFamRows {
  const cd *Q;  // [4][n] mode wavenumbers
  int64_t n;
  int fam[4];   // family of each entry, -1 skipped
  double sig[2], c[2];
  const double *z;  // N planes
};
hipError_t launch_sk_tsqr(const SkFamRows &F, const cd *Rin, int mode, int N, int64_t nrows, cd *Rout, int blocks,
                          hipStream_t s);
hipError_t launch_sk_err(const SkFamRows &F, int64_t nrows, int N, int k, const int *idx, const cd *P,
                         unsigned long long *err, hipStream_t s);
hipError_t launch_op_export(int px, int py, const cd *slot, cd *out, hipStream_t s);

}  // namespace ytf
