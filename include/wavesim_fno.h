/* wavesim_fno.h -- kernels of the learned warm start (no counterpart in the
 * reference).  All pointers are device pointers; `stream` is a hipStream_t
 * (NULL = default stream).  Functions return 0 or a negative code;
 * wfno_last_error() describes the last failure. */
#ifndef WAVESIM_FNO_H
#define WAVESIM_FNO_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* C[b] = alpha * A[b] (M x K) * B[b] (K x N) + beta * C[b] for the batch
 * index b = (b1, b2, b3) < (nb1, nb2, nb3): element (m,k) of A at
 * A + b1*sab1 + b2*sab2 + b3*sab3 + m*sam + k*sak, likewise B (k,n) and C (m,n).
 * FP32 storage, BF16 MFMA products, FP32 accumulation. */
typedef struct {
  long long M, N, K;
  int nb1, nb2, nb3;
  const float* A;
  long long sam, sak, sab1, sab2, sab3;
  const float* B;
  long long sbk, sbn, sbb1, sbb2, sbb3;
  float* C;
  long long scm, scn, scb1, scb2, scb3;
  float alpha, beta;
} wfno_gemm_args;

int wfno_gemm(const wfno_gemm_args* args, void* stream);
/* pre[b][c][v] += bias[c]; out = act ? GELU(pre) : pre (tanh form) */
int wfno_bias_act(float* pre, float* out, const float* bias, long long C, long long V, long long nb, int act,
                  void* stream);
/* dpre = dout * GELU'(pre) */
int wfno_gelu_bwd(const float* dout, const float* pre, float* dpre, long long n, void* stream);
/* dbias[c] (+)= sum_{b,v} d[b][c][v] */
int wfno_bias_grad(const float* d, float* dbias, long long C, long long V, long long nb, int accumulate,
                   void* stream);
/* dpred = scale * (pred - target); partial[i] = per-block sum of (pred - target)^2 */
int wfno_mse(const float* pred, const float* target, float* dpred, double* partial, int nblocks, long long n,
             float scale, void* stream);
/* one Adam step (t = 1, 2, ...) */
int wfno_adam(float* p, const float* g, float* m, float* v, long long n, float lr, float b1, float b2,
              float eps, int t, void* stream);
/* warm-start field on the solver grid from the envelope A ([2][Z][Y][X] at
 * 1/factor of the ROI) and the prior u(z) (complex64, grid[2] entries):
 * E0 = u(z) (1 + A) in the ROI, u(z) on the source footprint
 * foot = [y_begin, y_end, x_begin, x_end) elsewhere.  grid = [Ny Nx Nz],
 * roi_lo / roi_n = 0-based ROI origin and extent; E0 in the solver layout. */
int wfno_compose_warm(void* E0, const float* A, const void* u, const int grid[3], const int roi_lo[3],
                      const int roi_n[3], int factor, const int foot[4], void* stream);
/* Straight-ray prior on the ROI (refractive index n, [Lz][Ly][Lx], x fastest,
 * roi_n = [Lz Ly Lx]): phi = k0 h (sum_{z'<z}(n - n_bg) + (n(z) - n_bg)/2),
 * P = u(z) exp(i phi) (complex64, same layout), u = background field at the
 * ROI's z planes. */
int wfno_wkb_prior(const float* n, float* phase, void* P, const void* u, const int roi_n[3], float k0h,
                   float n_bg, void* stream);
/* Network input [6][L^3] on the grid pooled by `factor`: block means of
 * n^2 - n_bg^2, cos phi, sin phi, then the normalised coordinates z, y, x. */
int wfno_features(const float* n, const float* phase, float* X, const int roi_n[3], int factor, float n_bg,
                  void* stream);
/* E0 = P (1 + A) in the ROI (P from wfno_wkb_prior, A the envelope [2][Z][Y][X]
 * at 1/factor resolution), u(z) on the source footprint elsewhere; grid,
 * roi_lo and roi_n in [y x z] order, E0 in the solver layout. */
int wfno_compose_prior(void* E0, const float* A, const void* P, const void* u, const int grid[3],
                       const int roi_lo[3], const int roi_n[3], int factor, const int foot[4], void* stream);
/* Real form of the per-mode complex channel mixing: for each of npairs
 * weight pairs (Wr [modes][width][width] followed by Wi of the same shape,
 * pair g at W + (g / per_group) * group_stride + (g % per_group) * pair_stride
 * elements) and each mode k, Wc[g][k] = [[Wr_k, -Wi_k], [Wi_k, Wr_k]]
 * (2 width x 2 width, row-major), so that [Yr; Yi] = Wc_k [Fr; Fi] is
 * Y = (Wr + i Wi) F in one wfno_gemm. */
int wfno_complex_weights(const float* W, float* Wc, long long pair_stride, int per_group, long long group_stride,
                         int npairs, int modes, int width, void* stream);
/* dst[i] (+)= sum_{c < nchunks} src[c*n + i]  (fixed order) */
int wfno_reduce_chunks(const float* src, float* dst, int nchunks, long long n, int accumulate, void* stream);
const char* wfno_last_error(void);

#ifdef __cplusplus
}
#endif
#endif /* WAVESIM_FNO_H */
