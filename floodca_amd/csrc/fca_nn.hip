/*
 * fca_nn.hip -- gfx950 kernels of the flood surrogate network: one BF16 GEMM
 * on the matrix cores (v_mfma_f32_16x16x32_bf16, FP32 accumulation) whose
 * operand loaders also read 3x3 convolutions in place (implicit GEMM), and the
 * bias, activation, loss, optimiser and data kernels around it.
 *
 * Tensors are NHWC. Stored activations and GEMM operands are BF16 (raw
 * uint16 bits, round to nearest even); accumulators, gradients of activations,
 * master weights and optimiser state are FP32. Every reduction runs over a
 * fixed partition in a fixed order, so results repeat bit for bit.
 *
 * The C-ABI (include/fca_nn.h) takes device pointers and a hipStream_t; the
 * caller owns all memory.
 */
#include <hip/hip_runtime.h>

#include <math.h>
#include <stdint.h>

#include "fca_nn.h"

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

#define NN_BLOCK 256

/* float -> bf16 bits, round to nearest even (finite inputs) */
__device__ __forceinline__ uint16_t f2bf(float f)
{
    uint32_t u = __float_as_uint(f);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

__device__ __forceinline__ float bf2f(uint16_t b)
{
    return __uint_as_float((uint32_t)b << 16);
}

/* ---- GEMM ---------------------------------------------------------------------
 * C[M][N] = sum_k A_can[m][k] * B_can[n][k] with the operand modes of
 * fca_nn.h. Pixels index as p = (b*H + y)*W + x over the pixel space of the
 * mode; H, W and C are powers of two. Outside [0, M) x [0, N) x [0, K)
 * operands read as 0.
 *
 * Workgroup: 4 wavefronts on a 64x64 tile of C, each 32x32 as 2x2 MFMA blocks
 * of 16x16; K advances 32 per LDS stage. Tiles sit in LDS as [row][k] whatever
 * their layout in memory, so a lane reads its 8 consecutive k of one row as
 * one 16-byte LDS load; A and B fragments take the same k order, so the dot
 * products do not depend on how the matrix core orders k inside a lane. */
#define GT 64
#define GK 32
#define LDSK (GK + 8) /* row pitch in bf16 (80 bytes) */

struct Geo {
    int lh, lw, lc; /* log2 of the pixel space's H, W and of C */
    int sh, sw;     /* the source's H and W */
    int S;          /* stride */
};

/* src[b][ys][xs][c .. c+7] as 8 bf16, or zeros. */
__device__ __forceinline__ uint4 ld8(const uint16_t *src, const Geo &g, int b, int ys, int xs,
                                     int c, bool ok)
{
    if (ok && ys >= 0 && ys < g.sh && xs >= 0 && xs < g.sw) {
        return *(const uint4 *)(src + ((((size_t)b * g.sh + ys) * g.sw + xs) << g.lc) + c);
    }
    return make_uint4(0, 0, 0, 0);
}

/* Input pixel that tap (ky, kx) of output pixel (y, x) reads. */
__device__ __forceinline__ void fwd_src(const Geo &g, int y, int x, int ky, int kx, int &ys,
                                        int &xs)
{
    ys = y * g.S + ky - 1;
    xs = x * g.S + kx - 1;
}

/* Output pixel whose tap (ky, kx) reads input pixel (y, x), -1 if none. */
__device__ __forceinline__ void tr_src(const Geo &g, int y, int x, int ky, int kx, int &ys,
                                       int &xs)
{
    const int ty = y + 1 - ky, tx = x + 1 - kx;
    ys = (ty >= 0 && ty % g.S == 0) ? ty / g.S : -1;
    xs = (tx >= 0 && tx % g.S == 0) ? tx / g.S : -1;
}

template <int MODE>
struct Loader {
    int r, kq, row;   /* modes 0, 2, 3: one row, 8 consecutive k per thread */
    int b, y, x;      /* modes 2, 3: the row's pixel */
    int kk, rq;       /* modes 1, 4: one k, 8 consecutive rows per thread */
    int ky, kx, c;    /* mode 4: the tap and first channel of the 8 rows */

    __device__ void init(int tid, int r0, const Geo &g)
    {
        if (MODE == 0 || MODE == 2 || MODE == 3) {
            r = tid >> 2;
            kq = (tid & 3) * 8;
            row = r0 + r;
            if (MODE != 0) {
                x = row & ((1 << g.lw) - 1);
                y = (row >> g.lw) & ((1 << g.lh) - 1);
                b = row >> (g.lw + g.lh);
            }
        } else {
            kk = tid >> 3;
            rq = (tid & 7) * 8;
            row = r0 + rq;
            if (MODE == 4) {
                const int tap = row >> g.lc;
                c = row & ((1 << g.lc) - 1);
                ky = tap / 3;
                kx = tap - 3 * ky;
            }
        }
    }

    __device__ void stage(uint16_t (*S)[LDSK], const uint16_t *X, int ld, int rows, int kend,
                          int k0, const Geo &g)
    {
        if (MODE == 0) {
            const int gk = k0 + kq;
            uint4 q = make_uint4(0, 0, 0, 0);
            if (row < rows && gk < kend) { /* gk, kend and ld are multiples of 8 */
                q = *(const uint4 *)(X + (size_t)row * ld + gk);
            }
            *(uint4 *)&S[r][kq] = q;
        } else if (MODE == 2 || MODE == 3) {
            const int gk = k0 + kq;
            const int tap = gk >> g.lc, ch = gk & ((1 << g.lc) - 1);
            const int ty = tap / 3, tx = tap - 3 * ty;
            int ys, xs;
            if (MODE == 2) {
                fwd_src(g, y, x, ty, tx, ys, xs);
            } else {
                tr_src(g, y, x, ty, tx, ys, xs);
            }
            *(uint4 *)&S[r][kq] = ld8(X, g, b, ys, xs, ch, row < rows && gk < kend);
        } else {
            const int gk = k0 + kk;
            uint16_t v[8];
            if (MODE == 1) {
                if (gk < kend && row + 8 <= rows) {
                    *(uint4 *)v = *(const uint4 *)(X + (size_t)gk * ld + row);
                } else {
                    for (int j = 0; j < 8; j++) {
                        v[j] = (gk < kend && row + j < rows) ? X[(size_t)gk * ld + row + j]
                                                             : (uint16_t)0;
                    }
                }
            } else { /* k = output pixel; the 8 rows are 8 channels of one tap */
                const int xo = gk & ((1 << g.lw) - 1);
                const int yo = (gk >> g.lw) & ((1 << g.lh) - 1);
                const int bb = gk >> (g.lw + g.lh);
                int ys, xs;
                fwd_src(g, yo, xo, ky, kx, ys, xs);
                *(uint4 *)v = ld8(X, g, bb, ys, xs, c, gk < kend && row < rows);
            }
            for (int j = 0; j < 8; j++) {
                S[rq + j][kk] = v[j];
            }
        }
    }
};

template <int AM, int BM>
__global__ void __launch_bounds__(NN_BLOCK) k_gemm(FcaNnGemm a)
{
    __shared__ uint16_t As[GT][LDSK];
    __shared__ uint16_t Bs[GT][LDSK];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int m0 = blockIdx.y * GT, n0 = blockIdx.x * GT;
    const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
    const int kbeg = blockIdx.z * a.kchunk;
    const int kend = min(a.K, kbeg + a.kchunk);
    const Geo ga = {a.a_lh, a.a_lw, a.a_lc, a.a_sh, a.a_sw, a.a_s};
    const Geo gb = {a.b_lh, a.b_lw, a.b_lc, a.b_sh, a.b_sw, a.b_s};
    Loader<AM> la;
    Loader<BM> lb;
    la.init(tid, m0, ga);
    lb.init(tid, n0, gb);
    f32x4 acc[2][2];
    for (int i = 0; i < 2; i++) {
        for (int j = 0; j < 2; j++) {
            acc[i][j] = (f32x4){0.f, 0.f, 0.f, 0.f};
        }
    }
    const uint16_t *A = (const uint16_t *)a.A, *B = (const uint16_t *)a.B;
    for (int k0 = kbeg; k0 < kend; k0 += GK) {
        la.stage(As, A, a.lda, a.M, kend, k0, ga);
        lb.stage(Bs, B, a.ldb, a.N, kend, k0, gb);
        __syncthreads();
        const int kq = (lane >> 4) * 8;
        bf16x8 af[2], bfr[2];
        for (int i = 0; i < 2; i++) {
            uint4 q = *(const uint4 *)&As[wm + 16 * i + (lane & 15)][kq];
            __builtin_memcpy(&af[i], &q, 16);
            q = *(const uint4 *)&Bs[wn + 16 * i + (lane & 15)][kq];
            __builtin_memcpy(&bfr[i], &q, 16);
        }
        for (int i = 0; i < 2; i++) {
            for (int j = 0; j < 2; j++) {
                acc[i][j] =
                    __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[i], bfr[j], acc[i][j], 0, 0, 0);
            }
        }
        __syncthreads();
    }
    /* lane holds D[4*(lane/16) + r][lane%16] of each 16x16 block */
    float *C = a.C ? a.C + (size_t)blockIdx.z * a.M * a.ldc : nullptr;
    for (int i = 0; i < 2; i++) {
        for (int j = 0; j < 2; j++) {
            const int n = n0 + wn + 16 * j + (lane & 15);
            for (int r = 0; r < 4; r++) {
                const int m = m0 + wm + 16 * i + 4 * (lane >> 4) + r;
                if (m >= a.M || n >= a.N) {
                    continue;
                }
                float v = acc[i][j][r];
                if (a.bias) {
                    v += a.bias[a.bias_rows ? (size_t)(m / a.bias_rows) * a.N + n : (size_t)n];
                }
                if (a.relu && !(v > 0.f)) {
                    v = 0.f;
                }
                const size_t o = (size_t)m * a.ldc + n;
                if (C) {
                    C[o] = a.accumulate ? C[o] + v : v;
                }
                if (a.Cb) {
                    ((uint16_t *)a.Cb)[o] = f2bf(v);
                }
            }
        }
    }
}

/* ---- elementwise and reductions -------------------------------------------------- */

#define GRID_STRIDE(i, n)                                                                    \
    for (size_t i = (size_t)blockIdx.x * NN_BLOCK + threadIdx.x; i < (size_t)(n);             \
         i += (size_t)gridDim.x * NN_BLOCK)

/* C[i] (+)= sum over z of part[z][i], in z order. */
__global__ void __launch_bounds__(NN_BLOCK)
    k_splitk_sum(const float *part, int nz, size_t mn, float *C, int accumulate)
{
    GRID_STRIDE(i, mn)
    {
        float s = 0.f;
        for (int z = 0; z < nz; z++) {
            s += part[(size_t)z * mn + i];
        }
        C[i] = accumulate ? C[i] + s : s;
    }
}

__global__ void __launch_bounds__(NN_BLOCK)
    k_relu_back(float *dY, const uint16_t *Y, size_t n, int relu, uint16_t *dYb)
{
    GRID_STRIDE(i, n)
    {
        float d = dY[i];
        if (relu && !(bf2f(Y[i]) > 0.f)) {
            d = 0.f;
            dY[i] = 0.f;
        }
        dYb[i] = f2bf(d);
    }
}

__global__ void __launch_bounds__(NN_BLOCK) k_f2bf(const float *x, uint16_t *y, size_t n)
{
    GRID_STRIDE(i, n)
    {
        y[i] = f2bf(x[i]);
    }
}

__global__ void __launch_bounds__(NN_BLOCK) k_addf(float *y, const float *x, size_t n)
{
    GRID_STRIDE(i, n)
    {
        y[i] += x[i];
    }
}

/* ps[b][c] = sum over the P pixels of sample b of dY[b][p][c]: 4 pixel lanes
 * x 64 channels per workgroup, the lanes combined in a fixed order. */
__global__ void __launch_bounds__(NN_BLOCK)
    k_colsum_sample(const float *dY, int P, int C, float *ps)
{
    __shared__ float sh[4][64];
    const int b = blockIdx.x, l = threadIdx.x & 63, q = threadIdx.x >> 6;
    const int c = blockIdx.y * 64 + l;
    float s = 0.f;
    if (c < C) {
        const float *src = dY + (size_t)b * P * C + c;
        for (int p = q; p < P; p += 4) {
            s += src[(size_t)p * C];
        }
    }
    sh[q][l] = s;
    __syncthreads();
    if (q == 0 && c < C) {
        ps[(size_t)b * C + c] = (sh[0][l] + sh[1][l]) + (sh[2][l] + sh[3][l]);
    }
}

__global__ void __launch_bounds__(NN_BLOCK)
    k_film(const float *bias, const float *Wf, const float *s, int B, int N, float *bs)
{
    GRID_STRIDE(t, (size_t)B * N)
    {
        const int i = (int)(t / N), n = (int)(t % N);
        float v = bias[n];
        if (Wf) {
            for (int j = 0; j < 4; j++) {
                v += Wf[n * 4 + j] * s[i * 4 + j];
            }
        }
        bs[t] = v;
    }
}

__global__ void __launch_bounds__(NN_BLOCK)
    k_film_back(const float *ps, const float *s, int B, int N, float *db, float *dWf)
{
    const int n = blockIdx.x * NN_BLOCK + threadIdx.x;
    if (n >= N) {
        return;
    }
    float g = 0.f, gw[4] = {0.f, 0.f, 0.f, 0.f};
    for (int i = 0; i < B; i++) {
        const float p = ps[(size_t)i * N + n];
        g += p;
        for (int j = 0; j < 4; j++) {
            gw[j] += p * s[i * 4 + j];
        }
    }
    db[n] = g;
    if (dWf) {
        for (int j = 0; j < 4; j++) {
            dWf[n * 4 + j] = gw[j];
        }
    }
}

/* dst[i] = src[idx[i]] for rows of row_bytes (a multiple of 16). */
__global__ void __launch_bounds__(NN_BLOCK)
    k_gather(const uint4 *src, const int *idx, int n, size_t row16, uint4 *dst)
{
    GRID_STRIDE(t, (size_t)n * row16)
    {
        const size_t i = t / row16, j = t % row16;
        dst[t] = src[(size_t)idx[i] * row16 + j];
    }
}

__global__ void __launch_bounds__(NN_BLOCK)
    k_wt(const float *W, int Cout, int Cin, uint16_t *Wt)
{
    GRID_STRIDE(t, (size_t)Cout * 9 * Cin)
    {
        const int ci = (int)(t % Cin);
        const int tap = (int)((t / Cin) % 9);
        const int co = (int)(t / ((size_t)9 * Cin));
        Wt[((size_t)ci * 9 + tap) * Cout + co] = f2bf(W[t]);
    }
}

__global__ void __launch_bounds__(NN_BLOCK)
    k_up2_add(const uint16_t *Y, const uint16_t *S, int B, int h, int w, int C, uint16_t *X)
{
    GRID_STRIDE(t, (size_t)B * 4 * h * w * C)
    {
        const int c = (int)(t % C);
        const size_t p = t / C;
        const int x = (int)(p % (2 * w)), y = (int)((p / (2 * w)) % (2 * h));
        const int b = (int)(p / ((size_t)4 * h * w));
        float v = bf2f(Y[(((size_t)b * h + y / 2) * w + x / 2) * C + c]);
        if (S) {
            v += bf2f(S[t]);
        }
        X[t] = f2bf(v);
    }
}

__global__ void __launch_bounds__(NN_BLOCK)
    k_up2_back(const float *dX, int B, int h, int w, int C, float *dY)
{
    GRID_STRIDE(t, (size_t)B * h * w * C)
    {
        const int c = (int)(t % C);
        const size_t p = t / C;
        const int x = (int)(p % w), y = (int)((p / w) % h), b = (int)(p / ((size_t)w * h));
        const float *s = dX + (((size_t)b * 2 * h + 2 * y) * 2 * w + 2 * x) * C + c;
        const size_t row = (size_t)2 * w * C;
        dY[t] = (s[0] + s[C]) + (s[row] + s[row + C]);
    }
}

/* Offset of output pixel (y, x)'s mean in a head map [B][R/4][R/4][32]. */
__device__ __forceinline__ size_t head_at(int b, int y, int x, int R)
{
    const int q = R / 4;
    return ((((size_t)b * q + y / 4) * q + x / 4) * 32) + (size_t)(((y & 3) * 4 + (x & 3)) * 2);
}

#define LS_LO (-12.0f)
#define LS_HI (6.0f)

__global__ void __launch_bounds__(NN_BLOCK)
    k_nll(const float *O, const float *T, const float *Wp, int B, int R, float alpha,
          float inv_wsum, float *dO, float *part)
{
    __shared__ float sh[NN_BLOCK];
    const size_t t = (size_t)blockIdx.x * NN_BLOCK + threadIdx.x;
    float l = 0.f;
    if (t < (size_t)B * R * R) {
        const int x = (int)(t % R), y = (int)((t / R) % R), b = (int)(t / ((size_t)R * R));
        const size_t o = head_at(b, y, x, R);
        const float mu = O[o], s = O[o + 1], w = Wp[t];
        const float sc = fminf(fmaxf(s, LS_LO), LS_HI);
        const float r = mu - T[t], e = expf(-sc), r2 = r * r;
        l = w * (alpha * r2 + 0.5f * (e * r2 + sc));
        dO[o] = w * (2.f * alpha * r) * inv_wsum;
        dO[o + 1] = (s >= LS_LO && s <= LS_HI) ? w * 0.5f * (1.f - e * r2) * inv_wsum : 0.f;
    }
    sh[threadIdx.x] = l;
    __syncthreads();
    for (int o = NN_BLOCK / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) {
            sh[threadIdx.x] += sh[threadIdx.x + o];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[blockIdx.x] = sh[0];
    }
}

__global__ void __launch_bounds__(NN_BLOCK)
    k_adam(float *p, const float *g, float *m, float *v, size_t n, float lr, float b1, float b2,
           float eps, float c1, float c2)
{
    GRID_STRIDE(i, n)
    {
        const float gi = g[i];
        const float mi = b1 * m[i] + (1.f - b1) * gi;
        const float vi = b2 * v[i] + (1.f - b2) * (gi * gi);
        m[i] = mi;
        v[i] = vi;
        p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    }
}

__global__ void __launch_bounds__(NN_BLOCK)
    k_pred_stats(const float *O, const float *Lf, int R, const float *ysv, float thr, float *st)
{
    __shared__ float sh[4][NN_BLOCK];
    const int b = blockIdx.x;
    const float ys = ysv[b];
    float a[4] = {0.f, 0.f, 0.f, 0.f};
    for (int p = threadIdx.x; p < R * R; p += NN_BLOCK) {
        const int y = p / R, x = p % R;
        const size_t o = head_at(b, y, x, R);
        const float w = Lf[(size_t)b * R * R + p];
        const float mu = O[o] * ys;
        const float sc = fminf(fmaxf(O[o + 1], LS_LO), LS_HI);
        a[0] += (mu >= thr) ? w : 0.f;
        a[1] += w * expf(0.5f * sc) * ys;
        a[2] += w * fmaxf(mu, 0.f);
        a[3] += w;
    }
    for (int j = 0; j < 4; j++) {
        sh[j][threadIdx.x] = a[j];
    }
    __syncthreads();
    for (int o = NN_BLOCK / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) {
            for (int j = 0; j < 4; j++) {
                sh[j][threadIdx.x] += sh[j][threadIdx.x + o];
            }
        }
        __syncthreads();
    }
    if (threadIdx.x < 4) {
        st[(size_t)b * 4 + threadIdx.x] = sh[threadIdx.x][0];
    }
}

__global__ void __launch_bounds__(NN_BLOCK)
    k_sq_err(const float *O, const float *T, const float *Wp, int R, float ys, double *se)
{
    __shared__ double sh[2][NN_BLOCK];
    const int b = blockIdx.x;
    double e = 0.0, ws = 0.0;
    for (int p = threadIdx.x; p < R * R; p += NN_BLOCK) {
        const int y = p / R, x = p % R;
        const size_t q = (size_t)b * R * R + p;
        const double pred = fmax((double)O[head_at(b, y, x, R)] * ys, 0.0);
        const double d = pred - (double)T[q];
        e += (double)Wp[q] * d * d;
        ws += (double)Wp[q];
    }
    sh[0][threadIdx.x] = e;
    sh[1][threadIdx.x] = ws;
    __syncthreads();
    for (int o = NN_BLOCK / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) {
            sh[0][threadIdx.x] += sh[0][threadIdx.x + o];
            sh[1][threadIdx.x] += sh[1][threadIdx.x + o];
        }
        __syncthreads();
    }
    if (threadIdx.x < 2) {
        se[(size_t)b * 2 + threadIdx.x] = sh[threadIdx.x][0];
    }
}

/* Per row r of n values: mo[r] = {sum W, sum W*T, sum W*T^2} (double, one
 * workgroup per row, fixed order). */
__global__ void __launch_bounds__(NN_BLOCK)
    k_row_moments(const float *T, const float *Wp, size_t n, double *mo)
{
    __shared__ double sh[3][NN_BLOCK];
    const size_t r = blockIdx.x;
    double a = 0.0, b = 0.0, c = 0.0;
    for (size_t i = threadIdx.x; i < n; i += NN_BLOCK) {
        const double w = Wp[r * n + i], t = T[r * n + i];
        a += w;
        b += w * t;
        c += w * t * t;
    }
    sh[0][threadIdx.x] = a;
    sh[1][threadIdx.x] = b;
    sh[2][threadIdx.x] = c;
    __syncthreads();
    for (int o = NN_BLOCK / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) {
            for (int j = 0; j < 3; j++) {
                sh[j][threadIdx.x] += sh[j][threadIdx.x + o];
            }
        }
        __syncthreads();
    }
    if (threadIdx.x < 3) {
        mo[r * 3 + threadIdx.x] = sh[threadIdx.x][0];
    }
}

/* ---- terrain pyramid, inputs and targets -------------------------------------------- */

__global__ void __launch_bounds__(NN_BLOCK)
    k_pyr0(const float *z, const uint8_t *mask, const float *n, size_t N, const float *ncls,
           int nclass, FcaNnPyrRec *out)
{
    GRID_STRIDE(i, N)
    {
        FcaNnPyrRec r;
        r.zmean = r.zmin = r.zmax = z[i];
        r.land = mask[i] == 0;
        int best = 0;
        float bd = 3.0e38f;
        for (int c = 0; c < nclass; c++) {
            const float d = fabsf(n[i] - ncls[c]);
            if (d < bd) {
                bd = d;
                best = c;
            }
        }
        r.cls = r.land ? (uint8_t)best : 0;
        r.pad = 0;
        out[i] = r;
    }
}

__global__ void __launch_bounds__(NN_BLOCK)
    k_pyr_down(const FcaNnPyrRec *in, int nr, int nc, FcaNnPyrRec *out)
{
    const int onc = nc / 2;
    GRID_STRIDE(i, (size_t)(nr / 2) * onc)
    {
        const int r = (int)(i / onc), c = (int)(i % onc);
        const FcaNnPyrRec *a = in + (size_t)(2 * r) * nc + 2 * c;
        const FcaNnPyrRec q[4] = {a[0], a[1], a[nc], a[nc + 1]};
        FcaNnPyrRec o;
        o.zmean = ((q[0].zmean + q[1].zmean) + (q[2].zmean + q[3].zmean)) * 0.25f;
        o.zmin = fminf(fminf(q[0].zmin, q[1].zmin), fminf(q[2].zmin, q[3].zmin));
        o.zmax = fmaxf(fmaxf(q[0].zmax, q[1].zmax), fmaxf(q[2].zmax, q[3].zmax));
        int land = 0, best = -1, cls = 0;
        for (int j = 0; j < 4; j++) {
            land += q[j].land;
            if ((int)q[j].land > best) {
                best = q[j].land;
                cls = q[j].cls;
            }
        }
        o.land = (uint16_t)land;
        o.cls = (uint8_t)(land ? cls : 0);
        o.pad = 0;
        out[i] = o;
    }
}

/* Cells of [lo, hi) on an axis of lim cells: the first and, in w, the overlap
 * length of each (at most FCA_NN_MAXSPAN cells). */
#define FCA_NN_MAXSPAN 10
__device__ __forceinline__ int span(double lo, double hi, int lim, int &first, float *w)
{
    first = (int)floor(lo);
    int last = (int)ceil(hi) - 1;
    if (last >= lim) {
        last = lim - 1;
    }
    int n = last - first + 1;
    if (n > FCA_NN_MAXSPAN) {
        n = FCA_NN_MAXSPAN;
    }
    for (int j = 0; j < n; j++) {
        const double a = fmax(lo, (double)(first + j)), b = fmin(hi, (double)(first + j + 1));
        w[j] = (float)(b - a);
    }
    return n;
}

/* One workgroup row per sample; each pixel combines the level records its
 * footprint covers, weighted by overlap area. */
/* Bilinear value of level lv's mean z at (cy, cx) in that level's cell units
 * (cell centres at +0.5), clamped at the edges. */
__device__ __forceinline__ float zref_at(const FcaNnPyr &pyr, int lv, double cy, double cx)
{
    const int n_r = pyr.nr >> lv, n_c = pyr.nc >> lv;
    const FcaNnPyrRec *rec = pyr.lev[lv];
    double u = cy - 0.5, v = cx - 0.5;
    int i0 = (int)floor(u), j0 = (int)floor(v);
    double fu = u - i0, fv = v - j0;
    if (i0 < 0) {
        i0 = 0;
        fu = 0.0;
    }
    if (i0 > n_r - 2) {
        i0 = n_r - 2;
        fu = 1.0;
    }
    if (j0 < 0) {
        j0 = 0;
        fv = 0.0;
    }
    if (j0 > n_c - 2) {
        j0 = n_c - 2;
        fv = 1.0;
    }
    const FcaNnPyrRec *a = rec + (size_t)i0 * n_c + j0;
    const double top = a[0].zmean * (1.0 - fv) + a[1].zmean * fv;
    const double bot = a[n_c].zmean * (1.0 - fv) + a[n_c + 1].zmean * fv;
    return (float)(top * (1.0 - fu) + bot * fu);
}

/* One workgroup row per sample; each pixel combines the level records its
 * footprint covers, weighted by overlap area. */
__global__ void __launch_bounds__(NN_BLOCK)
    k_inputs(FcaNnPyr pyr, const FcaNnSample *smp, const float *ncls, int R, uint16_t *X,
             float *Lf)
{
    const int b = blockIdx.y;
    const FcaNnSample s = smp[b];
    const int lev = s.level;
    const int lref = min(lev + 2, FCA_NN_LEVELS - 1);
    const FcaNnPyrRec *rec = pyr.lev[lev];
    const int lnr = pyr.nr >> lev, lnc = pyr.nc >> lev;
    const double inv = 1.0 / (double)(1 << lev), f = (double)s.side / R;
    const double iref = 1.0 / (double)(1 << lref);
    const float cell = (float)(1 << (2 * lev));
    for (int p = blockIdx.x * NN_BLOCK + threadIdx.x; p < R * R; p += gridDim.x * NN_BLOCK) {
        const int y = p / R, x = p % R;
        float wy[FCA_NN_MAXSPAN], wx[FCA_NN_MAXSPAN];
        int fy, fx;
        const int ny = span((s.r0 + y * f) * inv, (s.r0 + (y + 1) * f) * inv, lnr, fy, wy);
        const int nx = span((s.c0 + x * f) * inv, (s.c0 + (x + 1) * f) * inv, lnc, fx, wx);
        float W = 0.f, zs = 0.f, lf = 0.f, nn = 0.f;
        for (int j = 0; j < ny; j++) {
            const FcaNnPyrRec *row = rec + (size_t)(fy + j) * lnc + fx;
            for (int i = 0; i < nx; i++) {
                const FcaNnPyrRec r = row[i];
                const float w = wy[j] * wx[i];
                const float land = (float)r.land / cell;
                W += w;
                zs += w * r.zmean;
                lf += w * land;
                nn += w * land * (ncls[r.cls] * s.nscale[r.cls]);
            }
        }
        const float zp = W > 0.f ? zs / W : 0.f;
        const float zr = zref_at(pyr, lref, (s.r0 + (y + 0.5) * f) * iref,
                                 (s.c0 + (x + 0.5) * f) * iref);
        float ch[4];
        ch[0] = (zp - 30.f) / 20.f;
        ch[1] = (zp - zr) / 0.5f;
        ch[2] = W > 0.f ? lf / W : 0.f;
        ch[3] = lf > 0.f ? (nn / lf - 0.06f) / 0.04f : 0.f;
        const int q = R / 4;
        uint16_t *o = X + ((((size_t)b * q + y / 4) * q + x / 4) * 64) +
                      (size_t)(((y & 3) * 4 + (x & 3)) * 4);
        uint2 v;
        v.x = (uint32_t)f2bf(ch[0]) | ((uint32_t)f2bf(ch[1]) << 16);
        v.y = (uint32_t)f2bf(ch[2]) | ((uint32_t)f2bf(ch[3]) << 16);
        *(uint2 *)o = v;
        Lf[(size_t)b * R * R + p] = ch[2];
    }
}

__global__ void __launch_bounds__(NN_BLOCK)
    k_target(const float *hmax, const uint8_t *mask, int side, int R, float tscale,
             float wscale, float *T, float *Wl)
{
    const double f = (double)side / R;
    for (int p = blockIdx.x * NN_BLOCK + threadIdx.x; p < R * R; p += gridDim.x * NN_BLOCK) {
        const int y = p / R, x = p % R;
        float wy[FCA_NN_MAXSPAN], wx[FCA_NN_MAXSPAN];
        int fy, fx;
        const int ny = span(y * f, (y + 1) * f, side, fy, wy);
        const int nx = span(x * f, (x + 1) * f, side, fx, wx);
        double W = 0.0, L = 0.0, H = 0.0;
        for (int j = 0; j < ny; j++) {
            const size_t row = (size_t)(fy + j) * side + fx;
            for (int i = 0; i < nx; i++) {
                const double w = (double)wy[j] * wx[i];
                W += w;
                if (mask[row + i] == 0) {
                    L += w;
                    H += w * (double)hmax[row + i];
                }
            }
        }
        T[p] = L > 0.0 ? (float)(H / L * tscale) : 0.f;
        Wl[p] = W > 0.0 ? (float)(L / W * wscale) : 0.f;
    }
}

/* ---- C ABI ----------------------------------------------------------------------- */

static inline unsigned grid_for(size_t n)
{
    size_t b = (n + NN_BLOCK - 1) / NN_BLOCK;
    if (b > (1u << 20)) {
        b = 1u << 20;
    }
    return (unsigned)(b ? b : 1);
}

static int rc(void)
{
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

static bool geo_ok(int mode, int lh, int lw, int lc, int sh, int sw, int S)
{
    if (mode < 2) {
        return true;
    }
    return lh >= 0 && lw >= 0 && lc >= 3 && lh < 16 && lw < 16 && lc < 12 && sh > 0 && sw > 0 &&
           (S == 1 || S == 2);
}

template <int AM>
static void launch_b(const FcaNnGemm &a, dim3 g, hipStream_t st)
{
    switch (a.b_mode) {
    case 0: hipLaunchKernelGGL((k_gemm<AM, 0>), g, dim3(NN_BLOCK), 0, st, a); break;
    case 1: hipLaunchKernelGGL((k_gemm<AM, 1>), g, dim3(NN_BLOCK), 0, st, a); break;
    default: hipLaunchKernelGGL((k_gemm<AM, 4>), g, dim3(NN_BLOCK), 0, st, a); break;
    }
}

extern "C" int fca_nn_gemm(const FcaNnGemm *p, void *stream)
{
    if (!p) {
        return -2;
    }
    FcaNnGemm a = *p;
    /* modes 0, 2 and 3 load 8 consecutive k at a time */
    const bool kvec = a.a_mode != 1 || a.b_mode == 0;
    if (a.M <= 0 || a.N <= 0 || a.K <= 0 || (kvec && a.K % 8) || (!a.C && !a.Cb) ||
        a.ldc < a.N ||
        a.a_mode < 0 || a.a_mode > 3 || (a.b_mode != 0 && a.b_mode != 1 && a.b_mode != 4) ||
        !geo_ok(a.a_mode, a.a_lh, a.a_lw, a.a_lc, a.a_sh, a.a_sw, a.a_s) ||
        !geo_ok(a.b_mode, a.b_lh, a.b_lw, a.b_lc, a.b_sh, a.b_sw, a.b_s) ||
        (a.a_mode < 2 && a.lda % 8) || (a.b_mode < 2 && a.ldb % 8)) {
        return -2;
    }
    /* the in-place conv modes must cover exactly the rows and k they index */
    if (a.a_mode >= 2 && (a.M != (a.M >> (a.a_lh + a.a_lw)) << (a.a_lh + a.a_lw) ||
                          a.K != 9 << a.a_lc)) {
        return -2;
    }
    if (a.b_mode == 4 && (a.N != 9 << a.b_lc || a.K != (a.K >> (a.b_lh + a.b_lw))
                                                            << (a.b_lh + a.b_lw))) {
        return -2;
    }
    if (a.kchunk <= 0) {
        a.kchunk = ((a.K + GK - 1) / GK) * GK;
    }
    if (a.kchunk % GK) {
        return -2;
    }
    const int nz = (a.K + a.kchunk - 1) / a.kchunk;
    if (nz > 1 && (a.Cb || a.bias || a.relu || !a.C || a.accumulate)) {
        return -2; /* split K writes plain FP32 partials */
    }
    const dim3 g((a.N + GT - 1) / GT, (a.M + GT - 1) / GT, nz);
    const hipStream_t st = (hipStream_t)stream;
    switch (a.a_mode) {
    case 0: launch_b<0>(a, g, st); break;
    case 1: launch_b<1>(a, g, st); break;
    case 2: launch_b<2>(a, g, st); break;
    default: launch_b<3>(a, g, st); break;
    }
    return rc();
}

extern "C" int fca_nn_splitk_sum(const float *part, int nz, long long mn, float *C,
                                 int accumulate, void *st)
{
    hipLaunchKernelGGL(k_splitk_sum, dim3(grid_for((size_t)mn)), dim3(NN_BLOCK), 0,
                       (hipStream_t)st, part, nz, (size_t)mn, C, accumulate);
    return rc();
}

extern "C" int fca_nn_relu_back(float *dY, const void *Y, long long n, int relu, void *dYb,
                                void *st)
{
    hipLaunchKernelGGL(k_relu_back, dim3(grid_for((size_t)n)), dim3(NN_BLOCK), 0,
                       (hipStream_t)st, dY, (const uint16_t *)Y, (size_t)n, relu,
                       (uint16_t *)dYb);
    return rc();
}

extern "C" int fca_nn_to_bf16(const float *x, void *y, long long n, void *st)
{
    hipLaunchKernelGGL(k_f2bf, dim3(grid_for((size_t)n)), dim3(NN_BLOCK), 0, (hipStream_t)st, x,
                       (uint16_t *)y, (size_t)n);
    return rc();
}

extern "C" int fca_nn_add(float *y, const float *x, long long n, void *st)
{
    hipLaunchKernelGGL(k_addf, dim3(grid_for((size_t)n)), dim3(NN_BLOCK), 0, (hipStream_t)st, y,
                       x, (size_t)n);
    return rc();
}

extern "C" int fca_nn_colsum_sample(const float *dY, int B, int P, int C, float *ps, void *st)
{
    if (B <= 0 || P <= 0 || C <= 0) {
        return -2;
    }
    hipLaunchKernelGGL(k_colsum_sample, dim3(B, (C + 63) / 64), dim3(NN_BLOCK), 0,
                       (hipStream_t)st, dY, P, C, ps);
    return rc();
}

extern "C" int fca_nn_film(const float *bias, const float *Wf, const float *s, int B, int N,
                           float *bs, void *st)
{
    hipLaunchKernelGGL(k_film, dim3(grid_for((size_t)B * N)), dim3(NN_BLOCK), 0,
                       (hipStream_t)st, bias, Wf, s, B, N, bs);
    return rc();
}

extern "C" int fca_nn_film_back(const float *ps, const float *s, int B, int N, float *db,
                                float *dWf, void *st)
{
    hipLaunchKernelGGL(k_film_back, dim3((N + NN_BLOCK - 1) / NN_BLOCK), dim3(NN_BLOCK), 0,
                       (hipStream_t)st, ps, s, B, N, db, dWf);
    return rc();
}

extern "C" int fca_nn_wt(const float *W, int Cout, int Cin, void *Wt, void *st)
{
    hipLaunchKernelGGL(k_wt, dim3(grid_for((size_t)Cout * 9 * Cin)), dim3(NN_BLOCK), 0,
                       (hipStream_t)st, W, Cout, Cin, (uint16_t *)Wt);
    return rc();
}

extern "C" int fca_nn_up2_add(const void *Y, const void *S, int B, int h, int w, int C, void *X,
                              void *st)
{
    hipLaunchKernelGGL(k_up2_add, dim3(grid_for((size_t)B * 4 * h * w * C)), dim3(NN_BLOCK), 0,
                       (hipStream_t)st, (const uint16_t *)Y, (const uint16_t *)S, B, h, w, C,
                       (uint16_t *)X);
    return rc();
}

extern "C" int fca_nn_up2_back(const float *dX, int B, int h, int w, int C, float *dY, void *st)
{
    hipLaunchKernelGGL(k_up2_back, dim3(grid_for((size_t)B * h * w * C)), dim3(NN_BLOCK), 0,
                       (hipStream_t)st, dX, B, h, w, C, dY);
    return rc();
}

extern "C" int fca_nn_loss(const float *O, const float *T, const float *W, int B, int R,
                           float alpha, float inv_wsum, float *dO, float *part, void *st)
{
    if (R % 4 || B <= 0) {
        return -2;
    }
    const size_t n = (size_t)B * R * R;
    hipLaunchKernelGGL(k_nll, dim3((unsigned)((n + NN_BLOCK - 1) / NN_BLOCK)), dim3(NN_BLOCK), 0,
                       (hipStream_t)st, O, T, W, B, R, alpha, inv_wsum, dO, part);
    return rc();
}

extern "C" int fca_nn_adam(float *p, const float *g, float *m, float *v, long long n, float lr,
                           float b1, float b2, float eps, int t, void *st)
{
    if (t < 1) {
        return -2;
    }
    const float c1 = 1.f - powf(b1, (float)t), c2 = 1.f - powf(b2, (float)t);
    hipLaunchKernelGGL(k_adam, dim3(grid_for((size_t)n)), dim3(NN_BLOCK), 0, (hipStream_t)st, p,
                       g, m, v, (size_t)n, lr, b1, b2, eps, c1, c2);
    return rc();
}

extern "C" int fca_nn_pred_stats(const float *O, const float *Lf, int B, int R, const float *ys,
                                 float thr, float *st4, void *st)
{
    if (R % 4 || B <= 0) {
        return -2;
    }
    hipLaunchKernelGGL(k_pred_stats, dim3(B), dim3(NN_BLOCK), 0, (hipStream_t)st, O, Lf, R, ys,
                       thr, st4);
    return rc();
}

extern "C" int fca_nn_sq_err(const float *O, const float *T, const float *W, int B, int R,
                             float ys, double *se, void *st)
{
    if (R % 4 || B <= 0) {
        return -2;
    }
    hipLaunchKernelGGL(k_sq_err, dim3(B), dim3(NN_BLOCK), 0, (hipStream_t)st, O, T, W, R, ys,
                       se);
    return rc();
}

extern "C" int fca_nn_pyr0(const float *z, const uint8_t *mask, const float *n, long long N,
                           const float *ncls, int nclass, FcaNnPyrRec *out, void *st)
{
    if (nclass < 1 || nclass > FCA_NN_NCLASS || N <= 0) {
        return -2;
    }
    hipLaunchKernelGGL(k_pyr0, dim3(grid_for((size_t)N)), dim3(NN_BLOCK), 0, (hipStream_t)st, z,
                       mask, n, (size_t)N, ncls, nclass, out);
    return rc();
}

extern "C" int fca_nn_pyr_down(const FcaNnPyrRec *in, int nr, int nc, FcaNnPyrRec *out, void *st)
{
    if (nr % 2 || nc % 2 || nr <= 0 || nc <= 0) {
        return -2;
    }
    hipLaunchKernelGGL(k_pyr_down, dim3(grid_for((size_t)(nr / 2) * (nc / 2))), dim3(NN_BLOCK),
                       0, (hipStream_t)st, in, nr, nc, out);
    return rc();
}

extern "C" int fca_nn_inputs(const FcaNnPyr *pyr, const FcaNnSample *smp, int B,
                             const float *ncls, int R, void *X, float *Lf, void *st)
{
    if (!pyr || B <= 0 || R <= 0 || R % 4) {
        return -2;
    }
    const dim3 g((unsigned)((R * R + NN_BLOCK - 1) / NN_BLOCK), (unsigned)B);
    hipLaunchKernelGGL(k_inputs, g, dim3(NN_BLOCK), 0, (hipStream_t)st, *pyr, smp, ncls, R,
                       (uint16_t *)X, Lf);
    return rc();
}

extern "C" int fca_nn_target(const float *hmax, const uint8_t *mask, int side, int R,
                             float tscale, float wscale, float *T, float *W, void *st)
{
    if (R <= 0 || side < R || side > 8 * R) {
        return -2;
    }
    hipLaunchKernelGGL(k_target, dim3((unsigned)((R * R + NN_BLOCK - 1) / NN_BLOCK)),
                       dim3(NN_BLOCK), 0, (hipStream_t)st, hmax, mask, side, R, tscale, wscale,
                       T, W);
    return rc();
}

extern "C" int fca_nn_gather(const void *src, const int *idx, int n, long long row_bytes,
                             void *dst, void *st)
{
    if (n <= 0 || row_bytes <= 0 || row_bytes % 16) {
        return -2;
    }
    const size_t row16 = (size_t)row_bytes / 16;
    hipLaunchKernelGGL(k_gather, dim3(grid_for((size_t)n * row16)), dim3(NN_BLOCK), 0,
                       (hipStream_t)st, (const uint4 *)src, idx, n, row16, (uint4 *)dst);
    return rc();
}

extern "C" int fca_nn_row_moments(const float *T, const float *W, int rows, long long n,
                                  double *mo, void *st)
{
    if (rows <= 0 || n <= 0) {
        return -2;
    }
    hipLaunchKernelGGL(k_row_moments, dim3(rows), dim3(NN_BLOCK), 0, (hipStream_t)st, T, W,
                       (size_t)n, mo);
    return rc();
}
