// Kernels of the learned warm start (factorised 3-D Fourier neural operator).
//
// Everything the network needs, forward and backward, is one of:
//   * wfno_gemm: strided, two-level-batched real GEMM  C = alpha A B + beta C
//     on BF16 matrix cores (v_mfma_f32_32x32x16_bf16, FP32 accumulate); A and B
//     are FP32 in memory with arbitrary strides, read four consecutive
//     elements at a time along their unit-stride dimension and rounded to BF16
//     (round-to-nearest-even) when staged into LDS.  Transposes are stride
//     swaps, so the backward GEMMs (dA = dC B^T, dB = A^T dC) use the same kernel.
//     The truncated DFTs of the spectral layers are GEMMs with cos/sin matrices.
//   * elementwise kernels: bias + GELU forward, GELU backward, per-channel bias
//     gradient, MSE loss gradient, Adam update, and the composition of the warm
//     start field from the predicted envelope.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cmath>
#include <string>

#include "wavesim_fno.h"

namespace wfno {

typedef float f32x16 __attribute__((ext_vector_type(16)));

static thread_local std::string g_err;

// FP32 -> BF16 bits, round to nearest even (finite inputs)
__device__ __forceinline__ unsigned short to_bf16(float f) {
  uint32_t u = __float_as_uint(f);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (unsigned short)(u >> 16);
}

constexpr int TM = 64, TN = 64, TK = 32, KP = TK + 8;  // LDS row pitch in bf16 (80 B: 16-byte aligned rows)

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

// One operand tile of the K step starting at k0 in registers: 512 groups of
// four FP32 values (two per thread), each four consecutive elements along
// the operand's unit-stride dimension (the rows of the tile, R = 64, or its
// K = 32 columns), read as one 16-byte load where the group is aligned and
// inside the matrix.
struct TileRegs {
  float v[2][4];
};

// element (r, k) of an operand with strides (s_r, s_k) and extents (nr, nk);
// `kfast`: the K direction has unit stride (groups run along k)
__device__ __forceinline__ void load_tile(TileRegs& t, const float* P, long long s_r, long long s_k, long long r0,
                                          long long k0, long long nr, long long nk, bool kfast, int tid) {
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const int c = tid + h * 256;  // group 0..511
    // kfast: group (r = c / 8, k = 4 (c % 8)); else (k = c / 16, r = 4 (c % 16))
    const int r = kfast ? c >> 3 : (c & 15) * 4;
    const int k = kfast ? (c & 7) * 4 : c >> 4;
    const long long gr = r0 + r, gk = k0 + k;
    const float* q = P + gr * s_r + gk * s_k;
    const bool whole = kfast ? (gr < nr && gk + 3 < nk) : (gr + 3 < nr && gk < nk);
    const bool unit = kfast ? s_k == 1 : s_r == 1;
    if (unit && whole && ((uintptr_t)q & 15) == 0) {
      const float4 x = *reinterpret_cast<const float4*>(q);
      t.v[h][0] = x.x;
      t.v[h][1] = x.y;
      t.v[h][2] = x.z;
      t.v[h][3] = x.w;
    } else {
      const long long step = kfast ? s_k : s_r;
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const bool in = kfast ? (gr < nr && gk + e < nk) : (gr + e < nr && gk < nk);
        t.v[h][e] = in ? q[e * step] : 0.f;
      }
    }
  }
}

// the registers of load_tile into the LDS tile [R][KP] as BF16 (round to
// nearest even): a k-run is one 8-byte store, an r-run four 2-byte stores
__device__ __forceinline__ void store_tile(const TileRegs& t, unsigned short* S, bool kfast, int tid) {
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const int c = tid + h * 256;
    const int r = kfast ? c >> 3 : (c & 15) * 4;
    const int k = kfast ? (c & 7) * 4 : c >> 4;
    if (kfast) {
      const uint2 w = make_uint2((uint32_t)to_bf16(t.v[h][0]) | ((uint32_t)to_bf16(t.v[h][1]) << 16),
                                 (uint32_t)to_bf16(t.v[h][2]) | ((uint32_t)to_bf16(t.v[h][3]) << 16));
      *reinterpret_cast<uint2*>(S + r * KP + k) = w;
    } else {
#pragma unroll
      for (int e = 0; e < 4; ++e) S[(r + e) * KP + k] = to_bf16(t.v[h][e]);
    }
  }
}

// Workgroup: 4 waves, each a 32x32 output block of the 64x64 tile, on
// v_mfma_f32_32x32x16_bf16 (gfx950: 16 products of the K step per
// instruction, two instructions per wave and K step).  The next K step's
// operands are read from HBM into registers while the current step's
// products run; the LDS tiles are double-buffered, one barrier per step.
__global__ void __launch_bounds__(256) gemm_kernel(wfno_gemm_args g) {
  __shared__ __align__(16) unsigned short As[2][TM * KP];  // [m][k]
  __shared__ __align__(16) unsigned short Bs[2][TN * KP];  // [n][k]
  const int tid = threadIdx.x;
  const int wave = tid >> 6, lane = tid & 63;
  const int wm = wave >> 1, wn = wave & 1;
  const long long m0 = (long long)blockIdx.y * TM, n0 = (long long)blockIdx.x * TN;
  const long long b3 = blockIdx.z % g.nb3, b12 = blockIdx.z / g.nb3;
  const long long b1 = b12 / g.nb2, b2 = b12 % g.nb2;
  const float* A = g.A + b1 * g.sab1 + b2 * g.sab2 + b3 * g.sab3;
  const float* B = g.B + b1 * g.sbb1 + b2 * g.sbb2 + b3 * g.sbb3;
  float* Cm = g.C + b1 * g.scb1 + b2 * g.scb2 + b3 * g.scb3;
  // groups of four along k where K has unit stride, else along the rows
  const bool akf = g.sak == 1 || g.sam != 1;
  const bool bkf = g.sbk == 1 || g.sbn != 1;

  f32x16 acc;
#pragma unroll
  for (int r = 0; r < 16; ++r) acc[r] = 0.f;

  TileRegs ra, rb;
  load_tile(ra, A, g.sam, g.sak, m0, 0, g.M, g.K, akf, tid);
  load_tile(rb, B, g.sbn, g.sbk, n0, 0, g.N, g.K, bkf, tid);
  store_tile(ra, As[0], akf, tid);
  store_tile(rb, Bs[0], bkf, tid);
  __syncthreads();
  const int nk = (int)((g.K + TK - 1) / TK);
  for (int s = 0; s < nk; ++s) {
    const int cur = s & 1;
    const bool more = s + 1 < nk;
    if (more) {
      load_tile(ra, A, g.sam, g.sak, m0, (long long)(s + 1) * TK, g.M, g.K, akf, tid);
      load_tile(rb, B, g.sbn, g.sbk, n0, (long long)(s + 1) * TK, g.N, g.K, bkf, tid);
    }
    // lane holds A[i][kk + 8 (lane / 32) + j] and B[kk + 8 (lane / 32) + j][n], j < 8
    const int i = wm * 32 + (lane & 31);
    const int n = wn * 32 + (lane & 31);
#pragma unroll
    for (int kk = 0; kk < TK; kk += 16) {
      const int kb = kk + 8 * (lane >> 5);
      const bf16x8 a = *reinterpret_cast<const bf16x8*>(&As[cur][i * KP + kb]);
      const bf16x8 b = *reinterpret_cast<const bf16x8*>(&Bs[cur][n * KP + kb]);
      acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
    }
    if (more) {
      store_tile(ra, As[cur ^ 1], akf, tid);
      store_tile(rb, Bs[cur ^ 1], bkf, tid);
    }
    __syncthreads();
  }
  // acc[r] holds row 8*(r/4) + 4*(lane/32) + r%4, column lane%32 of the wave's block
  const long long n = n0 + wn * 32 + (lane & 31);
  if (n >= g.N) return;
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    const long long m = m0 + wm * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
    if (m < g.M) {
      float* c = Cm + m * g.scm + n * g.scn;
      *c = g.alpha * acc[r] + (g.beta != 0.f ? g.beta * *c : 0.f);
    }
  }
}

// Wc[g][k] = [[Wr_k, -Wi_k], [Wi_k, Wr_k]] (2w x 2w) for every mode k of
// pair g: the complex mixing Y_k = (Wr_k + i Wi_k) F_k as one real product
// [Yr; Yi] = Wc_k [Fr; Fi].  Pair g's Wr [m][w][w] starts at
// W + (g / per_group) group_stride + (g % per_group) pair_stride, its Wi
// right after it.
__global__ void complex_weights_kernel(const float* W, float* Wc, long long pair_stride, int per_group,
                                       long long group_stride, long long total, int m, int w) {
  const long long blk = 4LL * w * w, per = (long long)m * blk;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < total;
       i += (long long)gridDim.x * blockDim.x) {
    const long long g = i / per, k = (i % per) / blk;
    const int rc = (int)(i % blk), r = rc / (2 * w), c = rc % (2 * w);
    const float* Wr = W + (g / per_group) * group_stride + (g % per_group) * pair_stride;
    const bool rh = r >= w, ch = c >= w;
    const long long e = k * w * w + (long long)(r - (rh ? w : 0)) * w + (c - (ch ? w : 0));
    const float v = rh == ch ? Wr[e] : Wr[(long long)m * w * w + e];
    Wc[i] = (!rh && ch) ? -v : v;
  }
}

__global__ void reduce_chunks_kernel(const float* src, float* dst, int nchunks, long long n, int accumulate) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    float s = accumulate ? dst[i] : 0.f;
    for (int c = 0; c < nchunks; ++c) s += src[(long long)c * n + i];
    dst[i] = s;
  }
}

// GELU, tanh form: 0.5 x (1 + tanh(s (x + 0.044715 x^3))), s = sqrt(2/pi)
__device__ __forceinline__ float gelu(float x) {
  const float s = 0.7978845608028654f;
  return 0.5f * x * (1.f + tanhf(s * (x + 0.044715f * x * x * x)));
}
__device__ __forceinline__ float gelu_grad(float x) {
  const float s = 0.7978845608028654f;
  const float u = s * (x + 0.044715f * x * x * x);
  const float t = tanhf(u);
  return 0.5f * (1.f + t) + 0.5f * x * (1.f - t * t) * s * (1.f + 3.f * 0.044715f * x * x);
}

// pre[c][v] += bias[c]; out[c][v] = act ? gelu(pre) : pre   (batch: stride C*V)
__global__ void bias_act_kernel(float* pre, float* out, const float* bias, long long C, long long V,
                                long long total, int act) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < total;
       i += (long long)gridDim.x * blockDim.x) {
    const long long c = (i / V) % C;
    const float x = pre[i] + bias[c];
    pre[i] = x;
    out[i] = act ? gelu(x) : x;
  }
}

// dpre = dout * gelu'(pre)
__global__ void gelu_bwd_kernel(const float* dout, const float* pre, float* dpre, long long total) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < total;
       i += (long long)gridDim.x * blockDim.x)
    dpre[i] = dout[i] * gelu_grad(pre[i]);
}

// dbias[c] (+)= sum over batch and v of d[b][c][v]; one block per channel, fixed order
__global__ void __launch_bounds__(256) bias_grad_kernel(const float* d, float* dbias, long long C, long long V,
                                                        long long nb, int accumulate) {
  __shared__ double red[256];
  const long long c = blockIdx.x;
  double s = 0.0;
  for (long long b = 0; b < nb; ++b)
    for (long long v = threadIdx.x; v < V; v += 256) s += d[(b * C + c) * V + v];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) dbias[c] = (accumulate ? dbias[c] : 0.f) + (float)red[0];
}

// loss partials of 0.5 mean((pred - target)^2) and its gradient
__global__ void __launch_bounds__(256) mse_kernel(const float* pred, const float* target, float* dpred,
                                                  double* partial, long long total, float scale) {
  __shared__ double red[256];
  double s = 0.0;
  for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < total; i += (long long)gridDim.x * 256) {
    const float e = pred[i] - target[i];
    dpred[i] = e * scale;
    s += (double)e * e;
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

// Adam (Kingma & Ba 2015) with bias correction
__global__ void adam_kernel(float* p, const float* g, float* m, float* v, long long n, float lr, float b1,
                            float b2, float eps, float c1, float c2) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
  }
}

// E0[y,x,z] = u(z) * (1 + A(y/f, x/f, z/f)) inside the ROI box, u(z) on the
// source footprint outside it, 0 elsewhere.  A is [2][Z][Y][X] at 1/f resolution
// (channel 0 real, 1 imaginary; x fastest), u is the prior along z.
__global__ void compose_warm_kernel(float2* E0, const float* A, const float2* u, int Ny, int Nx, int Nz,
                                    int y0, int x0, int z0, int ny, int nx, int nz, int f, int fy0, int fy1,
                                    int fx0, int fx1) {
  const long long total = (long long)Ny * Nx * Nz;
  const int Xl = nx / f, Yl = ny / f, Zl = nz / f;
  const long long plane = (long long)Xl * Yl * Zl;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < total;
       i += (long long)gridDim.x * blockDim.x) {
    const int y = (int)(i % Ny);
    const int x = (int)((i / Ny) % Nx);
    const int z = (int)(i / ((long long)Ny * Nx));
    float2 v = make_float2(0.f, 0.f);
    if (y >= fy0 && y < fy1 && x >= fx0 && x < fx1) v = u[z];
    const int ly = y - y0, lx = x - x0, lz = z - z0;
    if (ly >= 0 && ly < ny && lx >= 0 && lx < nx && lz >= 0 && lz < nz) {
      const long long j = ((long long)(lz / f) * Yl + ly / f) * Xl + lx / f;
      const float ar = 1.f + A[j], ai = A[plane + j];
      v = make_float2(v.x * ar - v.y * ai, v.x * ai + v.y * ar);
    }
    E0[i] = v;
  }
}

// Straight-ray (WKB) prior on the ROI, one thread per (y, x) column:
//   phi(z) = k0 h (sum_{z'<z} (n - n_bg) + (n(z) - n_bg)/2),  P = u(z) exp(i phi)
// n, phase: [Lz][Ly][Lx] (x fastest); u: the background field at the ROI's z.
__global__ void wkb_kernel(const float* n, float* phase, float2* P, const float2* u, int Lz, int Ly, int Lx,
                           float k0h, float n_bg) {
  const long long col = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long plane = (long long)Ly * Lx;
  if (col >= plane) return;
  float acc = 0.f;
  for (int z = 0; z < Lz; ++z) {
    const long long i = z * plane + col;
    const float d = n[i] - n_bg;
    const float ph = k0h * (acc + 0.5f * d);
    acc += d;
    phase[i] = ph;
    float sn, cs;
    sincosf(ph, &sn, &cs);
    const float2 uz = u[z];
    P[i] = make_float2(uz.x * cs - uz.y * sn, uz.x * sn + uz.y * cs);
  }
}

// network input on the pooled grid (f^3 blocks): [6][L^3] = contrast
// n^2 - n_bg^2, cos phi, sin phi (block means) and the coordinates z, y, x
__global__ void features_kernel(const float* n, const float* phase, float* X, int Lz, int Ly, int Lx, int f,
                                float nbg2) {
  const int Z = Lz / f, Y = Ly / f, Xn = Lx / f;
  const long long V = (long long)Z * Y * Xn;
  for (long long j = (long long)blockIdx.x * blockDim.x + threadIdx.x; j < V;
       j += (long long)gridDim.x * blockDim.x) {
    const int x = (int)(j % Xn), y = (int)((j / Xn) % Y), z = (int)(j / ((long long)Xn * Y));
    float de = 0.f, c = 0.f, sn = 0.f;
    for (int dz = 0; dz < f; ++dz)
      for (int dy = 0; dy < f; ++dy)
        for (int dx = 0; dx < f; ++dx) {
          const long long i = ((long long)(z * f + dz) * Ly + (y * f + dy)) * Lx + (x * f + dx);
          const float nv = n[i];
          de += nv * nv - nbg2;
          float s1, c1;
          sincosf(phase[i], &s1, &c1);
          c += c1;
          sn += s1;
        }
    const float inv = 1.f / (float)(f * f * f);
    X[j] = de * inv;
    X[V + j] = c * inv;
    X[2 * V + j] = sn * inv;
    X[3 * V + j] = (z + 0.5f) / Z - 0.5f;
    X[4 * V + j] = (y + 0.5f) / Y - 0.5f;
    X[5 * V + j] = (x + 0.5f) / Xn - 0.5f;
  }
}

// E0 on the solver grid ([z][x][y], y fastest): P (1 + A) in the ROI, with P
// the ROI prior ([z][y][x]) and A the envelope at 1/f resolution ([2][Z][Y][X]);
// u(z) on the source footprint outside the ROI, 0 elsewhere.
__global__ void compose_prior_kernel(float2* E0, const float* A, const float2* P, const float2* u, int Ny, int Nx,
                                     int Nz, int y0, int x0, int z0, int ny, int nx, int nz, int f, int fy0,
                                     int fy1, int fx0, int fx1) {
  const long long total = (long long)Ny * Nx * Nz;
  const int Xl = nx / f, Yl = ny / f, Zl = nz / f;
  const long long plane = (long long)Xl * Yl * Zl;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < total;
       i += (long long)gridDim.x * blockDim.x) {
    const int y = (int)(i % Ny);
    const int x = (int)((i / Ny) % Nx);
    const int z = (int)(i / ((long long)Ny * Nx));
    float2 v = make_float2(0.f, 0.f);
    const int ly = y - y0, lx = x - x0, lz = z - z0;
    if (ly >= 0 && ly < ny && lx >= 0 && lx < nx && lz >= 0 && lz < nz) {
      const float2 p = P[((long long)lz * ny + ly) * nx + lx];
      const long long j = ((long long)(lz / f) * Yl + ly / f) * Xl + lx / f;
      const float ar = 1.f + A[j], ai = A[plane + j];
      v = make_float2(p.x * ar - p.y * ai, p.x * ai + p.y * ar);
    } else if (y >= fy0 && y < fy1 && x >= fx0 && x < fx1) {
      v = u[z];
    }
    E0[i] = v;
  }
}

static unsigned grid_for(long long n) {
  long long b = (n + 255) / 256;
  return (unsigned)(b < 1 ? 1 : (b > 65536 ? 65536 : b));
}

template <class F>
static int guard(F&& f) {
  const int rc = f();
  if (rc != 0) return rc;
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    g_err = hipGetErrorString(e);
    return -2;
  }
  return 0;
}

}  // namespace wfno

using namespace wfno;

extern "C" {

const char* wfno_last_error(void) { return g_err.c_str(); }

int wfno_gemm(const wfno_gemm_args* a, void* stream) {
  return guard([&]() -> int {
    if (!a || a->M <= 0 || a->N <= 0 || a->K <= 0 || a->nb1 <= 0 || a->nb2 <= 0 || a->nb3 <= 0) {
      g_err = "bad gemm shape";
      return -1;
    }
    const long long nb = (long long)a->nb1 * a->nb2 * a->nb3;
    if (nb > 65535 || (a->M + TM - 1) / TM > 65535) {
      g_err = "gemm grid too large";
      return -1;
    }
    dim3 grid((unsigned)((a->N + TN - 1) / TN), (unsigned)((a->M + TM - 1) / TM), (unsigned)nb);
    hipLaunchKernelGGL(gemm_kernel, grid, dim3(256), 0, (hipStream_t)stream, *a);
    return 0;
  });
}

int wfno_reduce_chunks(const float* src, float* dst, int nchunks, long long n, int accumulate, void* stream) {
  return guard([&]() -> int {
    hipLaunchKernelGGL(reduce_chunks_kernel, dim3(grid_for(n)), dim3(256), 0, (hipStream_t)stream, src, dst,
                       nchunks, n, accumulate);
    return 0;
  });
}

int wfno_complex_weights(const float* W, float* Wc, long long pair_stride, int per_group, long long group_stride,
                         int npairs, int modes, int width, void* stream) {
  return guard([&]() -> int {
    if (npairs <= 0 || per_group <= 0 || modes <= 0 || width <= 0) {
      g_err = "bad complex weight shape";
      return -1;
    }
    const long long n = (long long)npairs * modes * 4 * width * width;
    hipLaunchKernelGGL(complex_weights_kernel, dim3(grid_for(n)), dim3(256), 0, (hipStream_t)stream, W, Wc,
                       pair_stride, per_group, group_stride, n, modes, width);
    return 0;
  });
}

int wfno_bias_act(float* pre, float* out, const float* bias, long long C, long long V, long long nb, int act,
                  void* stream) {
  return guard([&]() -> int {
    const long long n = C * V * nb;
    hipLaunchKernelGGL(bias_act_kernel, dim3(grid_for(n)), dim3(256), 0, (hipStream_t)stream, pre, out, bias,
                       C, V, n, act);
    return 0;
  });
}

int wfno_gelu_bwd(const float* dout, const float* pre, float* dpre, long long n, void* stream) {
  return guard([&]() -> int {
    hipLaunchKernelGGL(gelu_bwd_kernel, dim3(grid_for(n)), dim3(256), 0, (hipStream_t)stream, dout, pre, dpre, n);
    return 0;
  });
}

int wfno_bias_grad(const float* d, float* dbias, long long C, long long V, long long nb, int accumulate,
                   void* stream) {
  return guard([&]() -> int {
    hipLaunchKernelGGL(bias_grad_kernel, dim3((unsigned)C), dim3(256), 0, (hipStream_t)stream, d, dbias, C, V,
                       nb, accumulate);
    return 0;
  });
}

int wfno_mse(const float* pred, const float* target, float* dpred, double* partial, int nblocks, long long n,
             float scale, void* stream) {
  return guard([&]() -> int {
    hipLaunchKernelGGL(mse_kernel, dim3((unsigned)nblocks), dim3(256), 0, (hipStream_t)stream, pred, target,
                       dpred, partial, n, scale);
    return 0;
  });
}

int wfno_adam(float* p, const float* g, float* m, float* v, long long n, float lr, float b1, float b2, float eps,
              int t, void* stream) {
  return guard([&]() -> int {
    const float c1 = 1.f - std::pow(b1, (float)t), c2 = 1.f - std::pow(b2, (float)t);
    hipLaunchKernelGGL(adam_kernel, dim3(grid_for(n)), dim3(256), 0, (hipStream_t)stream, p, g, m, v, n, lr, b1,
                       b2, eps, c1, c2);
    return 0;
  });
}

int wfno_compose_warm(void* E0, const float* A, const void* u, const int grid[3], const int roi_lo[3],
                      const int roi_n[3], int factor, const int foot[4], void* stream) {
  return guard([&]() -> int {
    if (factor < 1 || roi_n[0] % factor || roi_n[1] % factor || roi_n[2] % factor) {
      g_err = "ROI not divisible by the network's down-sampling factor";
      return -1;
    }
    const long long n = (long long)grid[0] * grid[1] * grid[2];
    hipLaunchKernelGGL(compose_warm_kernel, dim3(grid_for(n)), dim3(256), 0, (hipStream_t)stream, (float2*)E0, A,
                       (const float2*)u, grid[0], grid[1], grid[2], roi_lo[0], roi_lo[1], roi_lo[2], roi_n[0],
                       roi_n[1], roi_n[2], factor, foot[0], foot[1], foot[2], foot[3]);
    return 0;
  });
}

int wfno_wkb_prior(const float* n, float* phase, void* P, const void* u, const int roi_n[3], float k0h,
                   float n_bg, void* stream) {
  return guard([&]() -> int {
    const long long cols = (long long)roi_n[1] * roi_n[2];
    hipLaunchKernelGGL(wkb_kernel, dim3((unsigned)((cols + 255) / 256)), dim3(256), 0, (hipStream_t)stream, n, phase,
                       (float2*)P, (const float2*)u, roi_n[0], roi_n[1], roi_n[2], k0h, n_bg);
    return 0;
  });
}

int wfno_features(const float* n, const float* phase, float* X, const int roi_n[3], int factor, float n_bg,
                  void* stream) {
  return guard([&]() -> int {
    if (factor < 1 || roi_n[0] % factor || roi_n[1] % factor || roi_n[2] % factor) {
      g_err = "ROI not divisible by the pooling factor";
      return -1;
    }
    const long long V = (long long)(roi_n[0] / factor) * (roi_n[1] / factor) * (roi_n[2] / factor);
    hipLaunchKernelGGL(features_kernel, dim3(grid_for(V)), dim3(256), 0, (hipStream_t)stream, n, phase, X,
                       roi_n[0], roi_n[1], roi_n[2], factor, n_bg * n_bg);
    return 0;
  });
}

int wfno_compose_prior(void* E0, const float* A, const void* P, const void* u, const int grid[3],
                       const int roi_lo[3], const int roi_n[3], int factor, const int foot[4], void* stream) {
  return guard([&]() -> int {
    if (factor < 1 || roi_n[0] % factor || roi_n[1] % factor || roi_n[2] % factor) {
      g_err = "ROI not divisible by the network's down-sampling factor";
      return -1;
    }
    const long long n = (long long)grid[0] * grid[1] * grid[2];
    // roi_lo / roi_n are [y x z] like grid; P is [z][y][x] over the ROI
    hipLaunchKernelGGL(compose_prior_kernel, dim3(grid_for(n)), dim3(256), 0, (hipStream_t)stream, (float2*)E0, A,
                       (const float2*)P, (const float2*)u, grid[0], grid[1], grid[2], roi_lo[0], roi_lo[1], roi_lo[2],
                       roi_n[0], roi_n[1], roi_n[2], factor, foot[0], foot[1], foot[2], foot[3]);
    return 0;
  });
}

}  // extern "C"
