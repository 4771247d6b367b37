// Small-radix DFT butterflies for the LDS Stockham FFT (gfx950).
//
// DIR = -1 computes the forward transform  X[k] = sum_n x[n] exp(-2 pi i n k / R),
// DIR = +1 the unnormalised inverse.  Every butterfly works in place on R
// registers; radix 2/4/8/16 use the split-radix identities (no multiplies for
// 2 and 4), odd radices 3/5/7/11 use the pair-symmetric real/imaginary form.
#pragma once
#include <hip/hip_runtime.h>

namespace wsim {

// complex values are float2 (complex64) or double2 (complex128); every
// helper and butterfly below is written once for both
template <class V> struct RealOf;
template <> struct RealOf<float2> { using type = float; };
template <> struct RealOf<double2> { using type = double; };
template <class V> using real_of = typename RealOf<V>::type;
template <class S> struct VecOf;
template <> struct VecOf<float> { using type = float2; };
template <> struct VecOf<double> { using type = double2; };
template <class S> using vec_of = typename VecOf<S>::type;

template <class V>
__device__ __forceinline__ V mkv(real_of<V> x, real_of<V> y) {
  V r;
  r.x = x;
  r.y = y;
  return r;
}
template <class V>
__device__ __forceinline__ V cadd(V a, V b) { return mkv<V>(a.x + b.x, a.y + b.y); }
template <class V>
__device__ __forceinline__ V csub(V a, V b) { return mkv<V>(a.x - b.x, a.y - b.y); }
template <class V>
__device__ __forceinline__ V cmul(V a, V b) {
  return mkv<V>(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
template <class V>
__device__ __forceinline__ V cconj(V a) { return mkv<V>(a.x, -a.y); }
template <class V>
__device__ __forceinline__ V cscale(V a, real_of<V> s) { return mkv<V>(a.x * s, a.y * s); }
// a * (DIR * i):  DIR=-1 -> multiply by -i, DIR=+1 -> multiply by +i
template <int DIR, class V>
__device__ __forceinline__ V mul_dir_i(V a) {
  return DIR < 0 ? mkv<V>(a.y, -a.x) : mkv<V>(-a.y, a.x);
}

template <int DIR, class V>
__device__ __forceinline__ void dft2(V* v) {
  V a = v[0], b = v[1];
  v[0] = cadd(a, b);
  v[1] = csub(a, b);
}

template <int DIR, class V>
__device__ __forceinline__ void dft4(V& a0, V& a1, V& a2, V& a3) {
  V s02 = cadd(a0, a2), d02 = csub(a0, a2);
  V s13 = cadd(a1, a3), d13 = mul_dir_i<DIR>(csub(a1, a3));
  a0 = cadd(s02, s13);
  a2 = csub(s02, s13);
  a1 = cadd(d02, d13);
  a3 = csub(d02, d13);
}

template <int DIR, class V>
__device__ __forceinline__ void dft4(V* v) { dft4<DIR>(v[0], v[1], v[2], v[3]); }

template <int DIR, class V>
__device__ __forceinline__ void dft8(V* v) {
  using S = real_of<V>;
  const S H = S(0.70710678118654752440);
  V e0 = v[0], e1 = v[2], e2 = v[4], e3 = v[6];
  V o0 = v[1], o1 = v[3], o2 = v[5], o3 = v[7];
  dft4<DIR>(e0, e1, e2, e3);
  dft4<DIR>(o0, o1, o2, o3);
  // w8^k = exp(DIR * 2 pi i k / 8)
  V t1 = mkv<V>(H * (o1.x - DIR * o1.y), H * (o1.y + DIR * o1.x));   // o1 * H(1 + DIR i)
  V t2 = mul_dir_i<DIR>(o2);                                          // o2 * DIR i
  V t3 = mkv<V>(H * (-o3.x - DIR * o3.y), H * (-o3.y + DIR * o3.x)); // o3 * H(-1 + DIR i)
  v[0] = cadd(e0, o0);
  v[4] = csub(e0, o0);
  v[1] = cadd(e1, t1);
  v[5] = csub(e1, t1);
  v[2] = cadd(e2, t2);
  v[6] = csub(e2, t2);
  v[3] = cadd(e3, t3);
  v[7] = csub(e3, t3);
}

template <int DIR, class V>
__device__ __forceinline__ void dft16(V* v) {
  using S = real_of<V>;
  // 4 x 4 decomposition: n = 4 n1 + n2, k = k1 + 4 k2
  V a[4][4];
#pragma unroll
  for (int n2 = 0; n2 < 4; ++n2) {
    a[n2][0] = v[n2];
    a[n2][1] = v[n2 + 4];
    a[n2][2] = v[n2 + 8];
    a[n2][3] = v[n2 + 12];
    dft4<DIR>(a[n2][0], a[n2][1], a[n2][2], a[n2][3]);  // over n1 -> k1
  }
  // twiddle w16^(n2 k1)
  const S C1 = S(0.92387953251128676), S1 = S(0.38268343236508977), H = S(0.70710678118654752440);
  const S cs[4] = {S(1), C1, H, S1};
  const S sn[4] = {S(0), S1, H, C1};
#pragma unroll
  for (int n2 = 1; n2 < 4; ++n2) {
#pragma unroll
    for (int k1 = 1; k1 < 4; ++k1) {
      const int m = n2 * k1;  // 1..9
      // exp(DIR i 2 pi m / 16)
      S c, s;
      if (m < 4) { c = cs[m]; s = sn[m]; }
      else if (m == 4) { c = S(0); s = S(1); }
      else if (m < 8) { c = -cs[8 - m]; s = sn[8 - m]; }
      else if (m == 8) { c = S(-1); s = S(0); }
      else { c = -cs[m - 8]; s = -sn[m - 8]; }
      a[n2][k1] = cmul(a[n2][k1], mkv<V>(c, DIR * s));
    }
  }
#pragma unroll
  for (int k1 = 0; k1 < 4; ++k1) {
    V b0 = a[0][k1], b1 = a[1][k1], b2 = a[2][k1], b3 = a[3][k1];
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
  static constexpr double c[1] = {-0.49999999999999978};
  static constexpr double s[1] = {0.86602540378443871};
};
template <> struct OddTrig<5> {
  static constexpr double c[2] = {0.30901699437494745, -0.80901699437494734};
  static constexpr double s[2] = {0.95105651629515353, 0.58778525229247325};
};
template <> struct OddTrig<7> {
  static constexpr double c[3] = {0.62348980185873359, -0.22252093395631434, -0.90096886790241903};
  static constexpr double s[3] = {0.7818314824680298, 0.97492791218182362, 0.43388373911755823};
};
template <> struct OddTrig<11> {
  static constexpr double c[5] = {0.84125353283118121, 0.41541501300188644, -0.142314838273285,
                                 -0.65486073394528499, -0.95949297361449737};
  static constexpr double s[5] = {0.54064081745559756, 0.90963199535451833, 0.9898214418809328,
                                 0.75574957435425827, 0.28173255684142967};
};

// cos(2 pi m / R) and sin(2 pi m / R) for any m (reduced mod R)
template <int R>
__device__ __forceinline__ constexpr double odd_cos(int m) {
  m %= R;
  return m == 0 ? 1.0 : (m <= (R - 1) / 2 ? OddTrig<R>::c[m - 1] : OddTrig<R>::c[R - m - 1]);
}
template <int R>
__device__ __forceinline__ constexpr double odd_sin(int m) {
  m %= R;
  return m == 0 ? 0.0 : (m <= (R - 1) / 2 ? OddTrig<R>::s[m - 1] : -OddTrig<R>::s[R - m - 1]);
}

template <int R, int DIR, class V>
__device__ __forceinline__ void dft_odd(V* v) {
  using S = real_of<V>;
  constexpr int H = (R - 1) / 2;
  V a[H], b[H];
#pragma unroll
  for (int n = 1; n <= H; ++n) {
    a[n - 1] = cadd(v[n], v[R - n]);
    b[n - 1] = csub(v[n], v[R - n]);
  }
  V x0 = v[0];
  V sum = x0;
#pragma unroll
  for (int n = 0; n < H; ++n) sum = cadd(sum, a[n]);
  v[0] = sum;
#pragma unroll
  for (int k = 1; k <= H; ++k) {
    V t1 = x0, t2 = mkv<V>(S(0), S(0));
#pragma unroll
    for (int n = 1; n <= H; ++n) {
      const S c = S(odd_cos<R>(n * k)), s = S(odd_sin<R>(n * k));
      t1.x += a[n - 1].x * c;
      t1.y += a[n - 1].y * c;
      t2.x += b[n - 1].x * s;
      t2.y += b[n - 1].y * s;
    }
    // forward: X[k] = t1 - i t2, X[R-k] = t1 + i t2 ; inverse swaps the sign
    V it2 = mul_dir_i<DIR>(t2);  // DIR*i*t2
    v[k] = cadd(t1, it2);
    v[R - k] = csub(t1, it2);
  }
}

// radix 20 = 4 x 5 in registers (Cooley-Tukey, n = 5 n1 + n2, k = k1 + 4 k2):
// 4-point DFTs over n1, twiddles w20^(n2 k1), 5-point DFTs over n2.  Two
// stages of 16 and 20 make a 320-point transform with one LDS exchange.
template <int DIR, class V>
__device__ __forceinline__ void dft20(V* v) {
  using S = real_of<V>;
  V a[5][4];  // a[n2][k1]
#pragma unroll
  for (int n2 = 0; n2 < 5; ++n2) {
    a[n2][0] = v[n2];
    a[n2][1] = v[n2 + 5];
    a[n2][2] = v[n2 + 10];
    a[n2][3] = v[n2 + 15];
    dft4<DIR>(a[n2][0], a[n2][1], a[n2][2], a[n2][3]);
  }
  // cos/sin(2 pi m / 20), m = 0..5 (m = n2 k1 <= 12 folds onto them)
  constexpr double c20[6] = {1.0, 0.95105651629515357, 0.80901699437494742, 0.58778525229247314,
                             0.30901699437494745, 0.0};
#pragma unroll
  for (int n2 = 1; n2 < 5; ++n2) {
#pragma unroll
    for (int k1 = 1; k1 < 4; ++k1) {
      const int m = n2 * k1;  // 1..12
      // exp(DIR i 2 pi m / 20) from the first quadrant
      const int q = m / 5, r = m % 5;
      const S cr = S(c20[r]), sr = S(c20[5 - r]);
      S c, sn;
      if (q == 0) { c = cr; sn = sr; }
      else if (q == 1) { c = -sr; sn = cr; }
      else { c = -cr; sn = -sr; }
      a[n2][k1] = cmul(a[n2][k1], mkv<V>(c, DIR * sn));
    }
  }
#pragma unroll
  for (int k1 = 0; k1 < 4; ++k1) {
    V b[5] = {a[0][k1], a[1][k1], a[2][k1], a[3][k1], a[4][k1]};
    dft_odd<5, DIR>(b);
#pragma unroll
    for (int k2 = 0; k2 < 5; ++k2) v[k1 + 4 * k2] = b[k2];
  }
}

// cos and sin of 2 pi m / R as compile-time constants (Taylor series on the
// octant-reduced angle, double precision)
constexpr double kPi = 3.14159265358979323846;
constexpr double taylor_sin(double x) {
  double t = x, s = x;
  for (int k = 1; k < 14; ++k) {
    t *= -x * x / ((2.0 * k) * (2.0 * k + 1.0));
    s += t;
  }
  return s;
}
constexpr double taylor_cos(double x) {
  double t = 1.0, s = 1.0;
  for (int k = 1; k < 14; ++k) {
    t *= -x * x / ((2.0 * k - 1.0) * (2.0 * k));
    s += t;
  }
  return s;
}
// (cos, sin)(2 pi m / R) from the angle's octant: m/R in eighths of a turn
constexpr double unit_cos(int m, int R) {
  m = ((m % R) + R) % R;
  const int o = (8 * m) / R;          // octant 0..7
  const double r = kPi * (double)(8 * m - o * R) / (4.0 * R);  // 0 <= r < pi/4
  switch (o) {
    case 0: return taylor_cos(r);
    case 1: return taylor_sin(kPi / 4 - r);
    case 2: return -taylor_sin(r);
    case 3: return -taylor_cos(kPi / 4 - r);
    case 4: return -taylor_cos(r);
    case 5: return -taylor_sin(kPi / 4 - r);
    case 6: return taylor_sin(r);
    default: return taylor_cos(kPi / 4 - r);
  }
}
constexpr double unit_sin(int m, int R) {
  m = ((m % R) + R) % R;
  const int o = (8 * m) / R;
  const double r = kPi * (double)(8 * m - o * R) / (4.0 * R);
  switch (o) {
    case 0: return taylor_sin(r);
    case 1: return taylor_cos(kPi / 4 - r);
    case 2: return taylor_cos(r);
    case 3: return taylor_sin(kPi / 4 - r);
    case 4: return -taylor_sin(r);
    case 5: return -taylor_cos(kPi / 4 - r);
    case 6: return -taylor_cos(r);
    default: return -taylor_sin(kPi / 4 - r);
  }
}

template <int R, int DIR, class V>
__device__ __forceinline__ void dft(V* v);

// radix R = R1 x R2 in registers (Cooley-Tukey, n = R2 n1 + n2, k = k1 + R1 k2):
// R1-point DFTs over n1, twiddles w_R^(n2 k1), R2-point DFTs over n2.  Used
// for the composite radices of plan-specialised kernels (9, 12, 24, 32, 36),
// which shorten long axes to two or three stages.
template <int R1, int R2, int DIR, class V>
__device__ __forceinline__ void dft_ct(V* v) {
  using S = real_of<V>;
  constexpr int R = R1 * R2;
  V a[R2][R1];
#pragma unroll
  for (int n2 = 0; n2 < R2; ++n2) {
#pragma unroll
    for (int n1 = 0; n1 < R1; ++n1) a[n2][n1] = v[R2 * n1 + n2];
    dft<R1, DIR>(a[n2]);
  }
#pragma unroll
  for (int n2 = 1; n2 < R2; ++n2)
#pragma unroll
    for (int k1 = 1; k1 < R1; ++k1) {
      const S c = S(unit_cos(n2 * k1, R)), sn = S(unit_sin(n2 * k1, R));
      a[n2][k1] = cmul(a[n2][k1], mkv<V>(c, DIR * sn));
    }
#pragma unroll
  for (int k1 = 0; k1 < R1; ++k1) {
    V b[R2];
#pragma unroll
    for (int n2 = 0; n2 < R2; ++n2) b[n2] = a[n2][k1];
    dft<R2, DIR>(b);
#pragma unroll
    for (int k2 = 0; k2 < R2; ++k2) v[k1 + R1 * k2] = b[k2];
  }
}

template <int R, int DIR, class V>
__device__ __forceinline__ void dft(V* v) {
  if constexpr (R == 2) dft2<DIR>(v);
  else if constexpr (R == 4) dft4<DIR>(v);
  else if constexpr (R == 8) dft8<DIR>(v);
  else if constexpr (R == 16) dft16<DIR>(v);
  else if constexpr (R == 20) dft20<DIR>(v);
  else if constexpr (R == 9) dft_ct<3, 3, DIR>(v);
  else if constexpr (R == 12) dft_ct<4, 3, DIR>(v);
  else if constexpr (R == 24) dft_ct<8, 3, DIR>(v);
  else if constexpr (R == 32) dft_ct<4, 8, DIR>(v);
  else if constexpr (R == 36) dft_ct<4, 9, DIR>(v);
  else dft_odd<R, DIR>(v);
}

}  // namespace wsim
