// Small-radix DFT butterflies for the LDS Stockham FFT (gfx950).
//
// DIR = -1 computes the forward transform  X[k] = sum_n x[n] exp(-2 pi i n k / R),
// DIR = +1 the unnormalised inverse.  Every butterfly works in place on R
// registers; radix 2/4/8/16 use the split-radix identities (no multiplies for
// 2 and 4), odd radices 3/5/7/11 use the pair-symmetric real/imaginary form.
#pragma once
#include <hip/hip_runtime.h>

namespace wsim {

__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cconj(float2 a) { return make_float2(a.x, -a.y); }
__device__ __forceinline__ float2 cscale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }
// a * (DIR * i):  DIR=-1 -> multiply by -i, DIR=+1 -> multiply by +i
template <int DIR>
__device__ __forceinline__ float2 mul_dir_i(float2 a) {
  return DIR < 0 ? make_float2(a.y, -a.x) : make_float2(-a.y, a.x);
}

template <int DIR>
__device__ __forceinline__ void dft2(float2* v) {
  float2 a = v[0], b = v[1];
  v[0] = cadd(a, b);
  v[1] = csub(a, b);
}

template <int DIR>
__device__ __forceinline__ void dft4(float2& a0, float2& a1, float2& a2, float2& a3) {
  float2 s02 = cadd(a0, a2), d02 = csub(a0, a2);
  float2 s13 = cadd(a1, a3), d13 = mul_dir_i<DIR>(csub(a1, a3));
  a0 = cadd(s02, s13);
  a2 = csub(s02, s13);
  a1 = cadd(d02, d13);
  a3 = csub(d02, d13);
}

template <int DIR>
__device__ __forceinline__ void dft4(float2* v) { dft4<DIR>(v[0], v[1], v[2], v[3]); }

template <int DIR>
__device__ __forceinline__ void dft8(float2* v) {
  constexpr float H = 0.70710678118654752f;
  float2 e0 = v[0], e1 = v[2], e2 = v[4], e3 = v[6];
  float2 o0 = v[1], o1 = v[3], o2 = v[5], o3 = v[7];
  dft4<DIR>(e0, e1, e2, e3);
  dft4<DIR>(o0, o1, o2, o3);
  // w8^k = exp(DIR * 2 pi i k / 8)
  float2 t1 = make_float2(H * (o1.x - DIR * o1.y), H * (o1.y + DIR * o1.x));   // o1 * H(1 + DIR i)
  float2 t2 = mul_dir_i<DIR>(o2);                                              // o2 * DIR i
  float2 t3 = make_float2(H * (-o3.x - DIR * o3.y), H * (-o3.y + DIR * o3.x)); // o3 * H(-1 + DIR i)
  v[0] = cadd(e0, o0);
  v[4] = csub(e0, o0);
  v[1] = cadd(e1, t1);
  v[5] = csub(e1, t1);
  v[2] = cadd(e2, t2);
  v[6] = csub(e2, t2);
  v[3] = cadd(e3, t3);
  v[7] = csub(e3, t3);
}

template <int DIR>
__device__ __forceinline__ void dft16(float2* v) {
  // 4 x 4 decomposition: n = 4 n1 + n2, k = k1 + 4 k2
  float2 a[4][4];
#pragma unroll
  for (int n2 = 0; n2 < 4; ++n2) {
    a[n2][0] = v[n2];
    a[n2][1] = v[n2 + 4];
    a[n2][2] = v[n2 + 8];
    a[n2][3] = v[n2 + 12];
    dft4<DIR>(a[n2][0], a[n2][1], a[n2][2], a[n2][3]);  // over n1 -> k1
  }
  // twiddle w16^(n2 k1)
  constexpr float C1 = 0.92387953251128676f, S1 = 0.38268343236508977f, H = 0.70710678118654752f;
  const float cs[4] = {1.f, C1, H, S1};
  const float sn[4] = {0.f, S1, H, C1};
#pragma unroll
  for (int n2 = 1; n2 < 4; ++n2) {
#pragma unroll
    for (int k1 = 1; k1 < 4; ++k1) {
      const int m = n2 * k1;  // 1..9
      // exp(DIR i 2 pi m / 16)
      float c, s;
      if (m < 4) { c = cs[m]; s = sn[m]; }
      else if (m == 4) { c = 0.f; s = 1.f; }
      else if (m < 8) { c = -cs[8 - m]; s = sn[8 - m]; }
      else if (m == 8) { c = -1.f; s = 0.f; }
      else { c = -cs[m - 8]; s = -sn[m - 8]; }
      a[n2][k1] = cmul(a[n2][k1], make_float2(c, DIR * s));
    }
  }
#pragma unroll
  for (int k1 = 0; k1 < 4; ++k1) {
    float2 b0 = a[0][k1], b1 = a[1][k1], b2 = a[2][k1], b3 = a[3][k1];
    dft4<DIR>(b0, b1, b2, b3);  // over n2 -> k2
    v[k1] = b0;
    v[k1 + 4] = b1;
    v[k1 + 8] = b2;
    v[k1 + 12] = b3;
  }
}

// cos/sin(2 pi m / R) for m = 1..(R-1)/2, odd R
template <int R> struct OddTrig;
template <> struct OddTrig<3> {
  static constexpr float c[1] = {-0.49999999999999978f};
  static constexpr float s[1] = {0.86602540378443871f};
};
template <> struct OddTrig<5> {
  static constexpr float c[2] = {0.30901699437494745f, -0.80901699437494734f};
  static constexpr float s[2] = {0.95105651629515353f, 0.58778525229247325f};
};
template <> struct OddTrig<7> {
  static constexpr float c[3] = {0.62348980185873359f, -0.22252093395631434f, -0.90096886790241903f};
  static constexpr float s[3] = {0.7818314824680298f, 0.97492791218182362f, 0.43388373911755823f};
};
template <> struct OddTrig<11> {
  static constexpr float c[5] = {0.84125353283118121f, 0.41541501300188644f, -0.142314838273285f,
                                 -0.65486073394528499f, -0.95949297361449737f};
  static constexpr float s[5] = {0.54064081745559756f, 0.90963199535451833f, 0.9898214418809328f,
                                 0.75574957435425827f, 0.28173255684142967f};
};

// cos(2 pi m / R) and sin(2 pi m / R) for any m (reduced mod R)
template <int R>
__device__ __forceinline__ constexpr float odd_cos(int m) {
  m %= R;
  return m == 0 ? 1.f : (m <= (R - 1) / 2 ? OddTrig<R>::c[m - 1] : OddTrig<R>::c[R - m - 1]);
}
template <int R>
__device__ __forceinline__ constexpr float odd_sin(int m) {
  m %= R;
  return m == 0 ? 0.f : (m <= (R - 1) / 2 ? OddTrig<R>::s[m - 1] : -OddTrig<R>::s[R - m - 1]);
}

template <int R, int DIR>
__device__ __forceinline__ void dft_odd(float2* v) {
  constexpr int H = (R - 1) / 2;
  float2 a[H], b[H];
#pragma unroll
  for (int n = 1; n <= H; ++n) {
    a[n - 1] = cadd(v[n], v[R - n]);
    b[n - 1] = csub(v[n], v[R - n]);
  }
  float2 x0 = v[0];
  float2 sum = x0;
#pragma unroll
  for (int n = 0; n < H; ++n) sum = cadd(sum, a[n]);
  v[0] = sum;
#pragma unroll
  for (int k = 1; k <= H; ++k) {
    float2 t1 = x0, t2 = make_float2(0.f, 0.f);
#pragma unroll
    for (int n = 1; n <= H; ++n) {
      const float c = odd_cos<R>(n * k), s = odd_sin<R>(n * k);
      t1.x += a[n - 1].x * c;
      t1.y += a[n - 1].y * c;
      t2.x += b[n - 1].x * s;
      t2.y += b[n - 1].y * s;
    }
    // forward: X[k] = t1 - i t2, X[R-k] = t1 + i t2 ; inverse swaps the sign
    float2 it2 = mul_dir_i<DIR>(t2);  // DIR*i*t2
    v[k] = cadd(t1, it2);
    v[R - k] = csub(t1, it2);
  }
}

template <int R, int DIR>
__device__ __forceinline__ void dft(float2* v) {
  if constexpr (R == 2) dft2<DIR>(v);
  else if constexpr (R == 4) dft4<DIR>(v);
  else if constexpr (R == 8) dft8<DIR>(v);
  else if constexpr (R == 16) dft16<DIR>(v);
  else dft_odd<R, DIR>(v);
}

}  // namespace wsim
