// host_util.cpp -- host-side helpers whose rounding must match GPUE's host
// code (built by g++ with -ffp-contract=off, as GPUE's gcc host pass is).
#include <cmath>
#include <vector>

namespace gpue {

// GPUE's grids (src/operators.cu:277-432) and box (src/init.cu:139-169).
void make_grids(int dimnum, int xDim, int yDim, int zDim, double xMax, double yMax, double zMax,
                std::vector<double>* g /* x y z px py pz */, double* d /* dx dy dz */) {
  const double PI = 3.141592653589793;   // include/constants.h:19
  int dims[3] = {xDim, yDim, zDim};
  double mx[3] = {xMax, yMax, zMax};
  for (int a = 0; a < 3; ++a) {
    int n = dims[a];
    g[a].assign(n, 0.0);
    g[a + 3].assign(n, 0.0);
    double pMax = (PI / mx[a]) * (n >> 1);
    double dd = mx[a] / (n >> 1);
    double dp = PI / (mx[a]);
    d[a] = dd;
    if (a < dimnum) {
      for (int i = 0; i < n / 2; ++i) {
        g[a][i] = -mx[a] + i * dd;
        g[a][i + n / 2] = i * dd;
        g[a + 3][i] = i * dp;
        g[a + 3][i + n / 2] = -pMax + i * dp;
      }
    }
  }
  if (dimnum < 3) d[2] = 1;
  if (dimnum < 2) d[1] = 1;
}

// In-mask plaquettes of Tracker::findVortex: sqrt(x[i]^2 + x[j]^2) < radius
// (src/tracker.cu:113), i = slow (row) index, j = fast index.
void vortex_mask(const double* x, int xDim, double radius, std::vector<unsigned char>& m) {
  const int P = xDim - 1;
  m.assign((size_t)P * P, 0);
  for (int i = 0; i < P; ++i)
    for (int j = 0; j < P; ++j) m[(size_t)i * P + j] = std::sqrt(x[i] * x[i] + x[j] * x[j]) < radius;
}

// exp(-2 pi i m / n) in long double, rounded once
void twiddles(int n, std::vector<double>& re, std::vector<double>& im) {
  re.resize(n);
  im.resize(n);
  const long double pi = 3.141592653589793238462643383279502884L;
  for (int m = 0; m < n; ++m) {
    long double a = 2.0L * pi * (long double)m / (long double)n;
    re[m] = (double)cosl(a);
    im[m] = (double)-sinl(a);
  }
}

}  // namespace gpue
