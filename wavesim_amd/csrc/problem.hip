// Problem set-up on the host: the constructor of WaveSim / WaveSimVector
// (WaveSimBase.m:40-82) with Medium (Medium.m:15-72, 126-132),
// SimulationGrid (SimulationGrid.m:20-80), the ARL edge filters
// (WaveSimBase.m:217-258) and the wiggle descriptors (WaveSimBase.m:272-316,
// private/wiggle_perm.m:1-29).  Everything is computed in double precision in
// the reference's order of operations and cast to single at the end, as the
// reference's data_array does (Simulation.m:221-252): gamma, epsilon and
// k02e fix the iteration count, so they must come out bit for bit.
#include <algorithm>
#include <array>
#include <cmath>
#include <complex>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "wavesim_mbs.h"

// products and sums stay separately rounded unless a fused multiply-add is
// written out: the two places that use one (the real part of a complex
// product, and |z|) round as numpy does on x86-64, which the restatement of
// the reference these results are checked against computes with
#pragma clang fp contract(off)

namespace wsim {

void set_last_error(const std::string& msg);

namespace {

struct ArgError : std::runtime_error {
  using std::runtime_error::runtime_error;
};

void check(bool ok, const std::string& msg) {
  if (!ok) throw ArgError(msg);
}

// prime factors in ascending order (MATLAB factor; factor(1) = 1)
std::vector<int64_t> factor(int64_t n) {
  std::vector<int64_t> f;
  for (int64_t p = 2; p * p <= n; ++p)
    while (n % p == 0) {
      f.push_back(p);
      n /= p;
    }
  if (n > 1 || f.empty()) f.push_back(n);
  return f;
}

// SimulationGrid.efficient_size (SimulationGrid.m:64-80): the next size whose
// largest prime factor is at most 11 and, with more than two factors, whose
// second largest is at most 5
int64_t efficient_size(int64_t s) {
  for (;; ++s) {
    const auto f = factor(s);
    const size_t k = f.size();
    if (!(f[k - 1] > 11 || (k > 2 && f[k - 2] > 5))) return s;
  }
}

// 4-term symmetric Nuttall window of length L (MATLAB nuttallwin)
std::vector<double> nuttall(int64_t L) {
  if (L == 1) return {1.0};
  std::vector<double> w((size_t)L);
  for (int64_t i = 0; i < L; ++i) {
    const double t = 2 * M_PI * (double)i / (double)(L - 1);
    w[(size_t)i] = 0.3635819 - 0.4891775 * std::cos(t) + 0.1365995 * std::cos(2 * t) - 0.0106411 * std::cos(3 * t);
  }
  return w;
}

// one ARL edge filter (WaveSimBase.m:230-252): zeros, the rising window,
// ones over the medium, the falling window, zeros
std::vector<double> edge_filter(int64_t N, int64_t bl, int64_t br, int64_t L, bool three_d) {
  auto win = [&](int64_t B) {
    std::vector<double> w((size_t)std::max<int64_t>(B, 0));
    if (three_d) {
      for (int64_t i = 0; i < B; ++i) w[(size_t)i] = ((double)(i + 1) - 0.21) / ((double)B + 0.66);
    } else {
      const auto nw = nuttall(2 * L - 1);
      for (int64_t i = 0; i < B; ++i) w[(size_t)i] = nw[(size_t)i];
    }
    return w;
  };
  std::vector<double> smooth;
  if (br == bl) {
    smooth.push_back(0.0);
    const auto w = win(L - 1);
    smooth.insert(smooth.end(), w.begin(), w.end());
  } else {
    smooth = win(L);
  }
  check(bl >= L && br >= L, "boundary layer wider than the padding");
  std::vector<double> f;
  f.reserve((size_t)N);
  f.insert(f.end(), (size_t)(bl - L), 0.0);
  f.insert(f.end(), smooth.begin(), smooth.end());
  f.insert(f.end(), (size_t)(N - bl - br), 1.0);
  f.insert(f.end(), smooth.rbegin(), smooth.rend());
  f.insert(f.end(), (size_t)(br - L), 0.0);
  check((int64_t)f.size() == N, "edge filter length mismatch");
  return f;
}

// private/wiggle_perm.m: the +-1 sign table of the enabled axes (rows
// (-1)^(ceil(j/2^(r-1))+1)), columns reordered [1 2 4 3] / [1 2 3 5 4 6 7 8]
std::vector<std::array<double, 3>> wiggle_set(const bool acc[3]) {
  int n = 0;
  for (int d = 0; d < 3; ++d) n += acc[d] ? 1 : 0;
  const int nw = 1 << n;
  std::vector<std::array<double, 3>> t((size_t)nw, {0.0, 0.0, 0.0});
  int r = 0;
  for (int d = 0; d < 3; ++d) {
    if (!acc[d]) continue;
    for (int c = 1; c <= nw; ++c) {
      const double e = std::ceil((double)c / std::pow(2.0, r)) + 1;
      t[(size_t)c - 1][(size_t)d] = std::fmod(e, 2.0) == 0.0 ? 1.0 : -1.0;
    }
    ++r;
  }
  std::vector<int> order;
  if (n == 2)
    order = {1, 2, 4, 3};
  else if (n == 3)
    order = {1, 2, 3, 5, 4, 6, 7, 8};
  else
    for (int c = 1; c <= nw; ++c) order.push_back(c);
  std::vector<std::array<double, 3>> out;
  for (int c : order) out.push_back(t[(size_t)c - 1]);
  return out;
}

// |z| as max * sqrt(1 + (min/max)^2), the square with one rounding
double cabs(double re, double im) {
  const double a = std::fabs(re), b = std::fabs(im);
  const double m = std::max(a, b);
  if (m == 0.0 || std::isinf(m)) return m;
  const double r = std::min(a, b) / m;
  return m * std::sqrt(std::fma(r, r, 1.0));
}

// SimulationGrid.symrange (SimulationGrid.m:55-63): fftshift(0..n-1 - floor(n/2))
double symrange(int64_t n, int64_t i) {
  const int64_t j = ((i - n / 2) % n + n) % n;  // fftshift: roll by floor(n/2)
  return (double)(j - n / 2);
}

struct Layout {
  int64_t sz[3], N[3], Bl[3], Br[3], bw[3];
  bool periodic[3], acc[3];
  int nw;
  int64_t roi[2][4];
};

Layout layout(const wsim_problem_desc* d) {
  check(d != nullptr, "null desc");
  Layout L{};
  for (int a = 0; a < 3; ++a) {
    check(d->shape[a] >= 1, "map shape entries must be >= 1");
    check(d->boundary_widths[a] >= 0, "boundary widths must be >= 0");
    L.sz[a] = d->shape[a];
    L.bw[a] = d->boundary_widths[a];
    const int64_t min_size = L.sz[a] + 2 * L.bw[a];
    L.periodic[a] = L.bw[a] == 0;
    L.N[a] = L.periodic[a] ? min_size : efficient_size(min_size);
    L.Bl[a] = (L.N[a] - L.sz[a] + 1) / 2;
    L.Br[a] = (L.N[a] - L.sz[a]) / 2;
  }
  const bool dflt = d->acc[0] == -1 && d->acc[1] == -1 && d->acc[2] == -1;
  for (int a = 0; a < 3; ++a) {
    check(dflt || d->acc[a] == 0 || d->acc[a] == 1, "acc entries must be 0 or 1 (or all -1)");
    L.acc[a] = dflt ? !L.periodic[a] : d->acc[a] == 1;
  }
  L.nw = 1;
  for (int a = 0; a < 3; ++a) L.nw *= L.acc[a] ? 2 : 1;
  // Simulation.m:117-124: the ROI is the medium unless given (relative to it)
  for (int a = 0; a < 3; ++a) {
    L.roi[0][a] = d->has_roi ? d->roi[0][a] + L.Bl[a] : L.Bl[a] + 1;
    L.roi[1][a] = d->has_roi ? d->roi[1][a] + L.Bl[a] : L.N[a] - L.Br[a];
  }
  L.roi[0][3] = d->has_roi ? d->roi[0][3] : 1;
  L.roi[1][3] = d->has_roi ? d->roi[1][3] : (d->vector ? 3 : 1);
  const int64_t npol = d->vector ? 3 : 1;
  for (int a = 0; a < 4; ++a) {
    const int64_t hi = a < 3 ? L.N[a] : npol;
    check(L.roi[0][a] >= 1 && L.roi[0][a] <= L.roi[1][a] && L.roi[1][a] <= hi, "roi outside the grid");
  }
  return L;
}

void fill_info(const Layout& L, const wsim_problem_desc* d, wsim_problem_info* info) {
  for (int a = 0; a < 3; ++a) {
    info->N[a] = L.N[a];
    info->Bl[a] = L.Bl[a];
    info->Br[a] = L.Br[a];
  }
  info->N[3] = d->vector ? 3 : 1;
  for (int r = 0; r < 2; ++r)
    for (int a = 0; a < 4; ++a) info->roi[r][a] = L.roi[r][a];
  info->nwiggles = L.nw;
}

template <class F>
int guarded(F&& f) {
  try {
    f();
    return WSIM_OK;
  } catch (const ArgError& e) {
    set_last_error(e.what());
    return WSIM_ERR_ARG;
  } catch (const std::bad_alloc&) {
    set_last_error("host allocation failed");
    return WSIM_ERR_NOMEM;
  } catch (const std::exception& e) {
    set_last_error(e.what());
    return WSIM_ERR_ARG;
  }
}

}  // namespace
}  // namespace wsim

using namespace wsim;

extern "C" {

int wsim_problem_query(const wsim_problem_desc* desc, wsim_problem_info* info) {
  return guarded([&] {
    check(info != nullptr, "null info");
    const Layout L = layout(desc);
    *info = wsim_problem_info{};
    fill_info(L, desc, info);
  });
}

int wsim_problem_build(const wsim_problem_desc* desc, const wsim_cdouble* n, wsim_problem_info* info,
                       wsim_cfloat* gamma, wsim_wiggle_out* wiggles) {
  return guarded([&] {
    check(n != nullptr && info != nullptr && gamma != nullptr && wiggles != nullptr, "null argument");
    const Layout L = layout(desc);
    *info = wsim_problem_info{};
    fill_info(L, desc, info);
    for (int w = 0; w < L.nw; ++w)
      check(wiggles[w].pxe && wiggles[w].pye && wiggles[w].pze && wiggles[w].gx && wiggles[w].gy && wiggles[w].gz,
            "null wiggle table");
    const double lambda = desc->lambda > 0 ? desc->lambda : 1.0;
    const double dx = desc->pixel_size;
    check(dx > 0, "pixel_size must be > 0");
    const int64_t sy = L.sz[0], sx = L.sz[1], sz = L.sz[2];
    const int64_t Ny = L.N[0], Nx = L.N[1], Nz = L.N[2];
    // e_r = n.^2 (Medium.m:37), its real range over the map (Medium.m:45-47)
    const size_t nm = (size_t)(sy * sx * sz);
    std::vector<std::complex<double>> er(nm);
    double emin = INFINITY, emax = -INFINITY;
    for (size_t i = 0; i < nm; ++i) {
      const double a = n[i].re, b = n[i].im;
      check(b >= 0.0, "Medium cannot have gain, imaginary part of refractive index should be positive");
      er[i] = {std::fma(a, a, -(b * b)), std::fma(a, b, b * a)};  // n^2
      emin = std::min(emin, er[i].real());
      emax = std::max(emax, er[i].real());
    }
    const double ecen = (emin + emax) / 2;
    const double k0 = 2 * M_PI / lambda;
    const double k0sq = k0 * k0;
    const double k0c = std::sqrt(ecen) * k0;
    const double k0c2 = k0c * k0c;
    // the padded medium replicates the map's edges (Medium.m:130-131)
    auto src = [&](int64_t i, int64_t a) { return std::min(std::max(i - L.Bl[a], (int64_t)0), L.sz[a] - 1); };
    auto er_at = [&](int64_t y, int64_t x, int64_t z) {
      return er[(size_t)(src(y, 0) + sy * (src(x, 1) + sx * src(z, 2)))];
    };
    // calculate_epsilon (WaveSimBase.m:201-215)
    double eps = desc->epsilon;
    if (!(eps > 0)) {
      double m = 0;
      for (int64_t z = 0; z < Nz; ++z)
        for (int64_t x = 0; x < Nx; ++x)
          for (int64_t y = 0; y < Ny; ++y) {
            const auto e = er_at(y, x, z);
            m = std::max(m, cabs(e.real() * k0sq - k0c2, e.imag() * k0sq));
          }
      eps = std::max(m / 0.95, desc->epsilonmin > 0 ? desc->epsilonmin : 3.0);
    }
    // ARL edge filters (WaveSimBase.m:217-258)
    std::vector<double> filt[3];
    const bool any_pad = L.Bl[0] > 0 || L.Bl[1] > 0 || L.Bl[2] > 0;
    const bool three_d = Ny > 1 && Nx > 1 && Nz > 1;
    for (int a = 0; a < 3; ++a)
      if (any_pad && L.Bl[a] > 0) filt[a] = edge_filter(L.N[a], L.Bl[a], L.Br[a], L.bw[a], three_d);
    // V = e_r k0^2 - k0c^2 - i eps, times the filters; gamma = (i/eps) V (WaveSimBase.m:69-73)
    const double c = 1.0 / eps;
    for (int64_t z = 0; z < Nz; ++z)
      for (int64_t x = 0; x < Nx; ++x)
        for (int64_t y = 0; y < Ny; ++y) {
          const auto e = er_at(y, x, z);
          double vr = e.real() * k0sq - k0c2 - 0.0;
          double vi = e.imag() * k0sq - eps;
          if (!filt[0].empty()) { vr = vr * filt[0][(size_t)y]; vi = vi * filt[0][(size_t)y]; }
          if (!filt[1].empty()) { vr = vr * filt[1][(size_t)x]; vi = vi * filt[1][(size_t)x]; }
          if (!filt[2].empty()) { vr = vr * filt[2][(size_t)z]; vi = vi * filt[2][(size_t)z]; }
          const double gr = 0.0 * vr - c * vi;
          const double gi = 0.0 * vi + c * vr;
          gamma[(size_t)(y + Ny * (x + Nx * z))] = {(float)gr, (float)gi};
        }
    const float epsf = (float)eps;
    info->epsilon = epsf;
    info->k02e = {(float)(k0c2 / eps), 1.0f};
    info->k0c = k0c;
    info->iterations_per_cycle = lambda / (2 * k0c / eps);
    // wiggle descriptors (WaveSimBase.m:300-316): Fourier coordinates shifted
    // by s/4 of a pixel, divided by the single-precision sqrt(epsilon), and
    // the real-space ramps exp(i pi/2 s x / (dx N))
    const auto table = wiggle_set(L.acc);
    const float sq = std::sqrt(epsf);
    const double dp[3] = {2 * M_PI / (dx * (double)Ny), 2 * M_PI / (dx * (double)Nx), 2 * M_PI / (dx * (double)Nz)};
    for (int w = 0; w < L.nw; ++w) {
      const wsim_wiggle_out& o = wiggles[w];
      float* pe[3] = {o.pye, o.pxe, o.pze};
      wsim_cfloat* g[3] = {o.gy, o.gx, o.gz};
      for (int a = 0; a < 3; ++a) {
        const double s = table[(size_t)w][(size_t)a];
        const int64_t N = L.N[a];
        for (int64_t i = 0; i < N; ++i) {
          const double pr = dp[a] * symrange(N, i);
          const float t = (float)(pr - dp[a] * s / 4);
          pe[a][i] = t / sq;
          const double r = (double)i * dx / dx / (double)N;
          const double th = M_PI / 2 * s * r;
          g[a][i] = {(float)std::cos(th), (float)std::sin(th)};
        }
      }
    }
  });
}

}  // extern "C"
