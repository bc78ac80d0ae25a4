// ORACLE — test infrastructure only.
//
// GLL quadrature shared by the 2D and 3D CPU restatements (include inside an
// anonymous namespace).  Nodes/weights are the closed forms of the N = 4
// Lobatto rule; h' and the Lagrange interpolants follow the reference's
// recurrences in the reference's order.
#pragma once

constexpr int NGLL = 5;

// ---------------------------------------------------------------- GLL ----
// Legendre recurrence, src/quadrature/gll/gll_library.cpp:11-33.
template <typename real> real pnleg(real z, int n) {
  real p1 = 1.0, p2 = z, p3 = p2;
  for (int k = 1; k < n; k++) {
    double dk = static_cast<double>(k);
    p3 = ((2.0 * dk + 1.0) * z * p2 - dk * p1) / (dk + 1.0);
    p1 = p2;
    p2 = p3;
  }
  return p3;
}

// Derivative of the Legendre polynomial, src/quadrature/gll/gll_library.cpp:53-93.
template <typename real> real pndleg(real z, int n) {
  real p1 = 1.0, p2 = z, p1d = 0.0, p2d = 1.0, p3 = p2, p3d = 1.0;
  for (int k = 1; k < n; k++) {
    double dk = static_cast<double>(k);
    p3 = ((2.0 * dk + 1.0) * z * p2 - dk * p1) / (dk + 1.0);
    p3d = ((2.0 * dk + 1.0) * p2 + (2.0 * dk + 1.0) * z * p2d - dk * p1d) /
          (dk + 1.0);
    p1 = p2;
    p2 = p3;
    p1d = p2d;
    p2d = p3d;
  }
  return p3d;
}

template <typename real> struct GLL {
  real xi[NGLL];
  real w[NGLL];
  real hprime[NGLL][NGLL]; // hprime[i][j] = l'_j(xi_i)
  GLL() {
    // Lobatto nodes and weights for N = 4 (closed forms of the zwgljd roots,
    // src/quadrature/gll/gll_library.cpp:102-147).
    const double s = std::sqrt(3.0 / 7.0);
    const double xd[NGLL] = {-1.0, -s, 0.0, s, 1.0};
    const double wd[NGLL] = {0.1, 49.0 / 90.0, 32.0 / 45.0, 49.0 / 90.0, 0.1};
    for (int i = 0; i < NGLL; ++i) {
      xi[i] = static_cast<real>(xd[i]);
      w[i] = static_cast<real>(wd[i]);
    }
    // src/quadrature/gll/lagrange_poly.cpp:161-190
    const int deg = NGLL - 1;
    for (int i = 0; i < NGLL; i++) {
      for (int j = 0; j < NGLL; j++) {
        if (i == 0 && j == 0) {
          hprime[i][j] = -1.0 * static_cast<real>(deg) *
                         (static_cast<real>(deg) + 1.0) * 0.25;
        } else if (i == deg && j == deg) {
          hprime[i][j] = 1.0 * static_cast<real>(deg) *
                         (static_cast<real>(deg) + 1.0) * 0.25;
        } else if (i == j) {
          hprime[i][j] = 0.0;
        } else {
          hprime[i][j] =
              pnleg<real>(xi[i], deg) /
                  (pnleg<real>(xi[j], deg) * (xi[i] - xi[j])) +
              (1.0 - xi[i] * xi[i]) * pndleg<real>(xi[i], deg) /
                  (static_cast<real>(deg) * (static_cast<real>(deg) + 1.0) *
                   pnleg<real>(xi[j], deg) * (xi[i] - xi[j]) *
                   (xi[i] - xi[j]));
        }
      }
    }
  }
};

// Lagrange interpolants at xi, src/quadrature/gll/lagrange_poly.cpp:11-43.
template <typename real>
void lagrange_interpolants(real xi, const real *xigll, real *h, real *hp) {
  for (int dgr = 0; dgr < NGLL; dgr++) {
    real prod1 = 1.0, prod2 = 1.0, sum = 0.0;
    const real x0 = xigll[dgr];
    for (int i = 0; i < NGLL; i++) {
      if (i != dgr) {
        const real x = xigll[i];
        prod1 = prod1 * (xi - x);
        prod2 = prod2 * (x0 - x);
        real prod3 = 1.0;
        for (int j = 0; j < NGLL; j++)
          if (j != dgr && j != i)
            prod3 = prod3 * (xi - xigll[j]);
        sum = sum + prod3;
      }
    }
    const real prod2_inv = 1.0 / prod2;
    h[dgr] = prod1 * prod2_inv;
    if (hp)
      hp[dgr] = sum * prod2_inv;
  }
}

