/*
 * fca_nn.h -- C-ABI of the flood surrogate's gfx950 kernels
 * (floodca_amd/libfca_nn.so).
 *
 * The surrogate has no counterpart in the reference: it consumes the
 * reference's per-scenario product, the hmax field (src/common/grid.c:308-311),
 * and its scenario parameters (analysis/scenarios.py:18,
 * src/hybrid/ens_common.h:15-22), and the plain CA stays the judge of its
 * error. Every pointer is device memory unless stated otherwise; `stream` is a
 * hipStream_t (NULL = the null stream). Calls return 0, -1 on a launch error
 * and -2 on arguments the kernels do not support. BF16 buffers hold raw
 * bfloat16 bits (uint16, round to nearest even).
 */
#ifndef FCA_NN_H
#define FCA_NN_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* C[M][N] (+)= sum over k of A_can[m][k] * B_can[n][k] with BF16 operands and
 * FP32 accumulation on the matrix cores. Operand modes (a_mode / b_mode):
 *   0  plain, stored [rows][K], leading dimension lda / ldb (multiple of 8)
 *   1  plain, stored [K][rows]
 *   2  (A) 3x3 convolution, padding 1, stride s, read in place from an NHWC
 *      source [B][sh][sw][2^lc]: row m = output pixel (b, y, x) of a
 *      [B][2^lh][2^lw] pixel space, k = (3*ky + kx) * 2^lc + c
 *   3  (A) its transpose (input gradient): row m = pixel of the conv's input
 *      [B][2^lh][2^lw], source = the conv's output gradient [B][sh][sw][2^lc]
 *   4  (B) the transpose of mode 2: k = output pixel, n = (tap, c)
 * Epilogue per element: v = acc (+ bias[n], or bias[(m / bias_rows) * N + n]
 * when bias_rows > 0), ReLU if relu; written as FP32 to C (added to C when
 * accumulate) and/or as BF16 to Cb, both with row stride ldc. kchunk > 0
 * splits K into chunks (multiples of 32) whose FP32 partials go to
 * C + z * M * ldc (z = chunk index; no bias, ReLU or Cb), to be summed with
 * fca_nn_splitk_sum. K must be a multiple of 8 unless A is in mode 1 and B
 * in mode 1 or 4. */
typedef struct FcaNnGemm {
    int M, N, K;
    const void *A;
    int lda, a_mode;
    int a_lh, a_lw, a_lc, a_sh, a_sw, a_s;
    const void *B;
    int ldb, b_mode;
    int b_lh, b_lw, b_lc, b_sh, b_sw, b_s;
    float *C;
    void *Cb;
    int ldc, accumulate;
    const float *bias;
    int bias_rows, relu, kchunk;
} FcaNnGemm;

int fca_nn_gemm(const FcaNnGemm *args, void *stream);
/* C[i] (+)= sum over z < nz of part[z * mn + i], in z order. */
int fca_nn_splitk_sum(const float *part, int nz, long long mn, float *C, int accumulate,
                      void *stream);

/* dY[i] = 0 where relu and bf16 Y[i] <= 0; dYb[i] = bf16(dY[i]). */
int fca_nn_relu_back(float *dY, const void *Y, long long n, int relu, void *dYb, void *stream);
int fca_nn_to_bf16(const float *x, void *y, long long n, void *stream);
int fca_nn_add(float *y, const float *x, long long n, void *stream);
/* ps[b][c] = sum over the P pixels of sample b of dY[b][p][c]. */
int fca_nn_colsum_sample(const float *dY, int B, int P, int C, float *ps, void *stream);
/* Per-sample bias bs[i][n] = bias[n] + sum_j Wf[n][j] * s[i][j], j < 4 (Wf may
 * be NULL), and its gradient: db[n] = sum_i ps[i][n],
 * dWf[n][j] = sum_i ps[i][n] * s[i][j]. */
int fca_nn_film(const float *bias, const float *Wf, const float *s, int B, int N, float *bs,
                void *stream);
int fca_nn_film_back(const float *ps, const float *s, int B, int N, float *db, float *dWf,
                     void *stream);
/* Wt[ci][tap * Cout + co] = bf16(W[co][tap * Cin + ci]) for the 9 taps. */
int fca_nn_wt(const float *W, int Cout, int Cin, void *Wt, void *stream);
/* X = bf16(up2(Y) + S): Y [B][h][w][C], S and X [B][2h][2w][C] (S may be NULL);
 * and dY[b][y][x][c] = sum of dX over the 2x2 block of (y, x). */
int fca_nn_up2_add(const void *Y, const void *S, int B, int h, int w, int C, void *X,
                   void *stream);
int fca_nn_up2_back(const float *dX, int B, int h, int w, int C, float *dY, void *stream);

/* Head maps O[B][R/4][R/4][32]: output pixel (y, x) has its mean at channel
 * ((y%4)*4 + x%4)*2 of head pixel (y/4, x/4) and its log-variance s next to
 * it. Loss against T[B][R][R] with weights W[B][R][R]:
 *   l = W * (alpha * r^2 + 0.5 * (exp(-s') * r^2 + s')),  r = mean - T,
 *   s' = s clamped to [-12, 6] (no gradient outside).
 * The mean learns from the squared error alone and the log-variance from the
 * Gaussian likelihood at that mean: dO(mean) = 2*alpha*W*r*inv_wsum,
 * dO(s) = 0.5*W*(1 - exp(-s')*r^2)*inv_wsum (every channel written);
 * part[j] = sum of l over the j-th run of 256 pixels. */
int fca_nn_loss(const float *O, const float *T, const float *W, int B, int R, float alpha,
                float inv_wsum, float *dO, float *part, void *stream);
/* Adam step t (1-based) over n parameters. */
int fca_nn_adam(float *p, const float *g, float *m, float *v, long long n, float lr, float b1,
                float b2, float eps, int t, void *stream);
/* Per sample, with mean mu = O*ys[b] and sigma = exp(s'/2)*ys[b] (ys device):
 * st4[b] = {sum Lf*[mu >= thr], sum Lf*sigma, sum Lf*max(mu, 0), sum Lf}. */
int fca_nn_pred_stats(const float *O, const float *Lf, int B, int R, const float *ys, float thr,
                      float *st4, void *stream);
/* Per row r of n values: mo[r] = {sum W, sum W*T, sum W*T^2} (double). */
int fca_nn_row_moments(const float *T, const float *W, int rows, long long n, double *mo,
                       void *stream);
/* se[b] = {sum W*(max(mu, 0) - T)^2, sum W} (double). */
int fca_nn_sq_err(const float *O, const float *T, const float *W, int B, int R, float ys,
                  double *se, void *stream);

/* Terrain pyramid: level 0 has one record per grid cell, level l+1 one per
 * 2x2 block of level l (mean, min and max of z, LAND count, the Manning class
 * of the child with the most LAND). */
#define FCA_NN_NCLASS 8
#define FCA_NN_LEVELS 6
typedef struct FcaNnPyrRec {
    float zmean, zmin, zmax;
    uint16_t land; /* LAND cells of the 4^level the record covers */
    uint8_t cls, pad;
} FcaNnPyrRec;

typedef struct FcaNnPyr {
    const FcaNnPyrRec *lev[FCA_NN_LEVELS];
    int nr, nc; /* level 0 (multiples of 2^FCA_NN_LEVELS) */
} FcaNnPyr;

/* One network input: the square crop (r0, c0, side) of the grid sampled from
 * pyramid level `level`, and the scale of each Manning class. */
typedef struct FcaNnSample {
    int r0, c0, side, level;
    float nscale[FCA_NN_NCLASS];
} FcaNnSample;

int fca_nn_pyr0(const float *z, const uint8_t *mask, const float *n, long long N,
                const float *ncls, int nclass, FcaNnPyrRec *out, void *stream);
int fca_nn_pyr_down(const FcaNnPyrRec *in, int nr, int nc, FcaNnPyrRec *out, void *stream);
/* Inputs of B samples (smp is device memory) at R x R (R multiple of 4),
 * each pixel the overlap-area-weighted combination of the records of the
 * sample's level that its footprint covers:
 * X [B][R/4][R/4][64] BF16 space-to-depth with channel (i*4 + j)*4 + ch for
 * pixel (4Y+i, 4X+j), ch = {(mean z - 30)/20, (mean z - zref)/0.5, LAND
 * fraction, (LAND-mean n - 0.06)/0.04}, zref the bilinear mean z of level
 * min(level + 2, FCA_NN_LEVELS - 1) at the pixel centre (local relief, which
 * BF16 keeps to ~1 cm where it could not keep z itself); Lf [B][R][R] = LAND
 * fraction (FP32). */
int fca_nn_inputs(const FcaNnPyr *pyr, const FcaNnSample *smp, int B, const float *ncls, int R,
                  void *X, float *Lf, void *stream);
/* Target of one simulated side x side crop: T[y][x] = tscale x the LAND-mean
 * hmax over the footprint of pixel (y, x) of an R x R map (area weighted, 0
 * without LAND), W[y][x] = wscale x its LAND fraction (R <= side <= 8R). */
int fca_nn_target(const float *hmax, const uint8_t *mask, int side, int R, float tscale,
                  float wscale, float *T, float *W, void *stream);
/* dst[i] = src[idx[i]] for n rows of row_bytes (a multiple of 16); idx is
 * device memory. */
int fca_nn_gather(const void *src, const int *idx, int n, long long row_bytes, void *dst,
                  void *stream);

#ifdef __cplusplus
}
#endif

#endif
