// Synthetic benchmark systems (host side): sphere density fields, random RHS
// and the pressure Expr of the projection step.
//
// Random numbers: splitmix64 (Steele, Lea, Flood 2014) used as a counter-based
// generator: u01(key, i) = (splitmix64(key + i) >> 11) * 2^-53, with
// key(seed, stream) = splitmix64(splitmix64(seed) + stream).  Stream 1 draws
// the spheres (4 numbers each: cx, cy, cz, radius), stream 2 the RHS.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

#include "../../include/aphpois.h"

namespace {

inline uint64_t sm64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}

inline uint64_t key(uint64_t seed, uint64_t stream) { return sm64(sm64(seed) + stream); }

inline double u01(uint64_t k, uint64_t i) { return (double)(sm64(k + i) >> 11) * 0x1.0p-53; }

template <class F>
void parallel_planes(int nz, F f) {
  unsigned nt = std::thread::hardware_concurrency();
  if (nt == 0) nt = 1;
  if (nt > 32) nt = 32;
  if ((int)nt > nz) nt = (unsigned)nz;
  std::vector<std::thread> th;
  for (unsigned t = 0; t < nt; ++t) {
    th.emplace_back([=]() {
      for (int z = (int)t; z < nz; z += (int)nt) f(z);
    });
  }
  for (auto& x : th) x.join();
}

}  // namespace

extern "C" {

int aphpois_synth_density(int n, int nspheres, double rmin, double rmax, double ratio,
                          unsigned long long seed, double* rho) {
  if (n < 1 || nspheres < 0 || !rho || !(rmin >= 0) || !(rmax >= rmin)) return APHPOIS_EINVAL;
  const double h = 1.0 / n;
  const uint64_t ks = key(seed, 1);
  std::vector<double> cx(nspheres), cy(nspheres), cz(nspheres), rr(nspheres);
  for (int s = 0; s < nspheres; ++s) {
    cx[s] = u01(ks, 4ull * s);
    cy[s] = u01(ks, 4ull * s + 1);
    cz[s] = u01(ks, 4ull * s + 2);
    rr[s] = rmin + (rmax - rmin) * u01(ks, 4ull * s + 3);
  }
  const long long plane = (long long)n * n;
  parallel_planes(n, [&](int z) {
    double* pz = rho + (long long)z * plane;
    for (long long i = 0; i < plane; ++i) pz[i] = ratio;
    const double zc = (z + 0.5) * h;
    for (int s = 0; s < nspheres; ++s) {
      double dz = zc - cz[s];
      dz -= std::nearbyint(dz);
      const double r2 = rr[s] * rr[s];
      if (dz * dz >= r2) continue;
      // candidate box in y and x (periodic), then the exact test
      const int w = (int)std::ceil(rr[s] / h) + 2;
      const int yc = (int)std::floor(cy[s] / h), xc = (int)std::floor(cx[s] / h);
      const int wy = 2 * w + 1 > n ? n : 2 * w + 1;
      const int wx = wy;
      for (int jy = 0; jy < wy; ++jy) {
        const int y = ((yc - (wy == n ? 0 : w) + jy) % n + n) % n;
        double dy = (y + 0.5) * h - cy[s];
        dy -= std::nearbyint(dy);
        for (int jx = 0; jx < wx; ++jx) {
          const int x = ((xc - (wx == n ? 0 : w) + jx) % n + n) % n;
          double dx = (x + 0.5) * h - cx[s];
          dx -= std::nearbyint(dx);
          if (dx * dx + dy * dy + dz * dz < r2) pz[(long long)y * n + x] = 1.0;
        }
      }
    }
  });
  return APHPOIS_OK;
}

int aphpois_synth_rhs(int n, unsigned long long seed, double* f) {
  if (n < 1 || !f) return APHPOIS_EINVAL;
  const uint64_t kf = key(seed, 2);
  const long long plane = (long long)n * n;
  std::vector<double> psum(n);
  parallel_planes(n, [&](int z) {
    double s = 0;
    const long long o = (long long)z * plane;
    for (long long i = 0; i < plane; ++i) {
      const double v = 2.0 * u01(kf, (uint64_t)(o + i)) - 1.0;
      f[o + i] = v;
      s += v;
    }
    psum[z] = s;
  });
  double tot = 0;
  for (int z = 0; z < n; ++z) tot += psum[z];
  const double mean = tot / ((double)plane * n);
  parallel_planes(n, [&](int z) {
    const long long o = (long long)z * plane;
    for (long long i = 0; i < plane; ++i) f[o + i] -= mean;
  });
  return APHPOIS_OK;
}

// Face coefficient of the face between cells cm (lower) and cp (upper):
// k = (-1/h) * s, s = -area / dens * dt with dens = 1 / ((1/rho_cp + 1/rho_cm) * 0.5)
// (harmonic face density); e0 gathers +k over the six faces in neighbour
// order, e_q = -k, e7 = -f * vol.
int aphpois_synth_assemble(int n, const double* rho, const double* f, double* expr) {
  if (n < 1 || !rho || !f || !expr) return APHPOIS_EINVAL;
  const double h = 1.0 / n;
  const double vol = h * h * h;
  const double area = vol / h;
  const double a = 1.0 / h;
  const long long plane = (long long)n * n;
  auto face_k = [&](long long cm, long long cp) {
    const double inv = (1.0 / rho[cp] + 1.0 / rho[cm]) * 0.5;
    const double dens = 1.0 / inv;
    const double s = -area / dens * 1.0;
    return (-a) * s;
  };
  parallel_planes(n, [&](int z) {
    const int zm = z == 0 ? n - 1 : z - 1, zp = z == n - 1 ? 0 : z + 1;
    for (int y = 0; y < n; ++y) {
      const int ym = y == 0 ? n - 1 : y - 1, yp = y == n - 1 ? 0 : y + 1;
      for (int x = 0; x < n; ++x) {
        const int xm = x == 0 ? n - 1 : x - 1, xp = x == n - 1 ? 0 : x + 1;
        const long long c = (long long)z * plane + (long long)y * n + x;
        const long long nb[6] = {(long long)z * plane + (long long)y * n + xm,
                                 (long long)z * plane + (long long)y * n + xp,
                                 (long long)z * plane + (long long)ym * n + x,
                                 (long long)z * plane + (long long)yp * n + x,
                                 (long long)zm * plane + (long long)y * n + x,
                                 (long long)zp * plane + (long long)y * n + x};
        double* e = expr + 8 * c;
        double diag = 0;
        for (int q = 0; q < 6; ++q) {
          const bool plus = q & 1;
          const double k = plus ? face_k(c, nb[q]) : face_k(nb[q], c);
          diag += k;
          e[1 + q] = -k;
        }
        e[0] = diag;
        e[7] = 0.0 - f[c] * vol;
      }
    }
  });
  return APHPOIS_OK;
}

}  // extern "C"
