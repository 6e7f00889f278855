// ep2d_analytic.cpp -- see ep2d_analytic.h.
//
// Elastic zone: complex potentials of a plane with an elliptic plastic zone,
// mapped from the unit circle by w(zeta) = c0 (zeta - kappa / zeta)
// (getAnalytic, EffPlast2D.cu:999-1015; displacement :1016-1029; stresses
// :1039-1051).  Plastic zone: logarithmic radial stresses (:1052-1059).
// Hydrostatic load: closed forms (:1030-1032, :1193-1200).  Every expression
// keeps the reference's operation order; |x| of a double is its magnitude.
#include "ep2d_analytic.h"

#include <algorithm>
#include <cmath>
#include <complex>
#include <limits>

namespace ep2d {

namespace {

using cplx = std::complex<double>;

struct Potentials {
  cplx zeta, w, dw, wv, Phi, Psi;
};

// conformal map and potentials at z = x + iy
Potentials potentials(double x, double y, double xi, double kappa, double c0, double Y,
                      double rad) {
  const cplx z(x, y);
  const double sgn = (std::fabs(x) < std::numeric_limits<double>::epsilon() || x > 0.0) ? 1.0 : -1.0;
  Potentials p;
  p.zeta = (z + sgn * std::sqrt(z * z + 4.0 * c0 * c0 * kappa)) / 2.0 / c0;
  p.w = c0 * (p.zeta - kappa / p.zeta);
  p.dw = c0 * (1.0 + kappa / (p.zeta * p.zeta));
  p.wv = c0 * (1.0 / p.zeta - kappa * p.zeta);
  p.Phi = -Y * xi / 2.0 - Y * xi * std::log(p.w / p.zeta / rad);
  p.Psi = -Y * xi / p.zeta * p.wv / p.dw;
  return p;
}

double ur_hydro(double r, double deltaP, double Y, double rad, double G0) {
  return -0.5 * Y * rad * rad * std::exp((deltaP - Y) / Y) / (G0 * r);
}

cplx u_elastic(double x, double y, double tauInfty, double xi, double kappa, double c0,
               const AnGeometry& g) {
  const Potentials p = potentials(x, y, xi, kappa, c0, g.Y, g.rad);
  const cplx phi = -g.Y * xi * p.w * (std::log(p.w / p.zeta / g.rad) + 0.5) - 2.0 * c0 * tauInfty / p.zeta;
  const cplx psi = c0 * g.Y * xi * (1.0 / p.zeta + kappa * p.zeta);
  const cplx dphi = p.Phi * p.dw;
  return ((1.0 + 6.0 * g.G0 / (g.G0 + 3.0 * g.K0)) * phi - p.w / std::conj(p.dw) * std::conj(dphi) -
          std::conj(psi)) / 2.0 / g.G0;
}

struct Inv {
  double J1, J2;
};
Inv invariants(double s11, double s22, double s12) {
  return {0.5 * (s11 + s22), 0.5 * ((s11 - s22) * (s11 - s22) + 4.0 * s12 * s12)};
}

Inv j_elastic(double x, double y, double xi, double kappa, double c0, double Y, double rad) {
  const Potentials p = potentials(x, y, xi, kappa, c0, Y, rad);
  const cplx z(x, y);
  const cplx dPhi = -2.0 * xi * Y * kappa / p.zeta / (p.zeta * p.zeta - kappa);
  const cplx F = 2.0 * (std::conj(p.w) / p.dw * dPhi + p.Psi) / std::exp(-2.0 * std::arg(z) * cplx(0.0, 1.0));
  const double Srr = 2.0 * std::real(p.Phi) - std::real(F) / 2.0;
  const double Sff = 2.0 * std::real(p.Phi) + std::real(F) / 2.0;
  return invariants(Srr, Sff, std::imag(F) / 2.0);
}

Inv j_plastic(double r, double xi, double Y, double rad) {
  const double Srr = -2.0 * xi * Y * std::log(r / rad);
  const double Sff = -2.0 * xi * Y * (1.0 + std::log(r / rad));
  return invariants(Srr, Sff, 0.0);
}

}  // namespace

void an_static_fields(const AnGeometry& g, const double* P, const double* txx,
                      const double* tyy, const double* txy, const double* Ux, const double* Uy,
                      double deltaP, double tauInfty, const double lt[3], ep2d_an_static* o) {
  const long nX = g.nX, nY = g.nY;
  const double dX = g.dX, dY = g.dY, rad = g.rad, Y = g.Y;
  const double tiny = std::numeric_limits<double>::epsilon();
  const bool hydro = std::fabs(lt[0] - lt[1]) < tiny && std::fabs(lt[2]) < tiny;
  const double Rmin = rad, Rmax = 0.5 * dX * (nX - 1) - dX * 20.0, eps = 1.0e-18;
  const double xi = deltaP > 0.0 ? 1.0 : -1.0;
  const double kappa = tauInfty / Y * xi;
  const double c0 = rad * std::exp(std::fabs(deltaP) / 2.0 / Y - 0.5);
  const double rx = rad + ur_hydro(rad, deltaP, Y, rad, g.G0), ry = rx;  // deformed pore
  const double Rx = c0 * (1.0 - kappa), Ry = c0 * (1.0 + kappa);          // plastic zone
  const bool an = g.nPores == 1;
  auto cut = [](double& e) {
    if (std::fabs(e) > 0.5) e = -0.5;
  };
  auto put = [](double* a, long k, double v) {
    if (a) a[k] = v;
  };

  double eU = 0, eUs = 0, e1 = 0, e1s = 0, e2 = 0, e2s = 0;
  size_t nU = 0, nJ = 0;
  for (long i = 0; i < nX; i++) {
    for (long j = 0; j < nY; j++) {
      const double ux = 0.5 * (Ux[(nX + 1) * j + i] + Ux[(nX + 1) * j + (i + 1)]);
      const double uy = 0.5 * (Uy[nX * j + i] + Uy[nX * (j + 1) + i]);
      const double unu = std::sqrt(ux * ux + uy * uy);
      put(o->UnuAbs, j * nX + i, unu);
      const bool vert = i < nX - 1 && j < nY - 1;
      const long v = j * (nX - 1) + i;
      Inv nu{0, 0};
      if (vert) {
        // vertex stress: mean of the four cells' -P + tau (reference term order)
        auto cellsum = [&](const double* t) {
          return 0.25 * (-P[j * nX + i] + t[j * nX + i] + -P[j * nX + (i + 1)] + t[j * nX + (i + 1)] +
                         -P[(j + 1) * nX + i] + t[(j + 1) * nX + i] + -P[(j + 1) * nX + (i + 1)] +
                         t[(j + 1) * nX + (i + 1)]);
        };
        nu = invariants(cellsum(txx), cellsum(tyy), txy[v]);
        put(o->J1nu, v, nu.J1);
        put(o->J2nu, v, nu.J2);
      }
      if (!an) continue;

      // radial displacement at the cell centre
      {
        const double x = -0.5 * dX * (nX - 1) + dX * i;
        const double y = -0.5 * dY * (nY - 1) + dY * j;
        const double r = std::sqrt(x * x + y * y);
        double uan = 0.0;
        if (x * x / (Rx * Rx) + y * y / (Ry * Ry) > 1.0)
          uan = hydro ? std::fabs(ur_hydro(r, deltaP, Y, rad, g.G0))
                      : std::abs(u_elastic(x, y, tauInfty, xi, kappa, c0, g));
        else if (x * x / (rx * rx) + y * y / (ry * ry) > 1.0)
          uan = std::fabs(ur_hydro(r, deltaP, Y, rad, g.G0));
        double e = 0.0;
        if (x * x + y * y > Rmin * Rmin && x * x + y * y < Rmax * Rmax && std::fabs(unu) > eps) {
          e = std::fabs((uan - unu) / uan);
          eU = std::max(eU, e);
          eUs += e;
          nU++;
          cut(e);
        }
        put(o->UanAbs, j * nX + i, uan);
        put(o->errorUabs, j * nX + i, e);
      }
      // stress invariants at the vertex
      if (vert) {
        const double x = -0.5 * dX * (nX - 1) + dX * i + 0.5 * dX;
        const double y = -0.5 * dY * (nY - 1) + dY * j + 0.5 * dY;
        const double r = std::sqrt(x * x + y * y);
        Inv a{0.0, 0.0};
        if (x * x / (Rx * Rx) + y * y / (Ry * Ry) > 1.0) {
          if (hydro) {
            const double relR = rad / r;
            const double Srr = -deltaP + relR * relR * Y * std::exp(deltaP / Y - 1);
            const double Sff = -deltaP - relR * relR * Y * std::exp(deltaP / Y - 1);
            a = invariants(Srr, Sff, 0.0);
          } else {
            a = j_elastic(x, y, xi, kappa, c0, Y, rad);
          }
        } else if (x * x / (rx * rx) + y * y / (ry * ry) > 1.0) {
          a = j_plastic(r, xi, Y, rad);
        }
        double q1 = 0.0, q2 = 0.0;
        if (x * x + y * y > Rmin * Rmin && x * x + y * y < Rmax * Rmax) {
          if (std::fabs(nu.J1) > eps) {
            q1 = std::fabs((a.J1 - nu.J1) / a.J1);
            e1 = std::max(e1, q1);
            e1s += q1;
            cut(q1);
          }
          if (std::fabs(nu.J2) > eps) {
            q2 = std::fabs((a.J2 - nu.J2) / a.J2);
            e2 = std::max(e2, q2);
            e2s += q2;
            cut(q2);
          }
          nJ++;
        }
        put(o->J1an, v, a.J1);
        put(o->J2an, v, a.J2);
        put(o->errorJ1, v, q1);
        put(o->errorJ2, v, q2);
      }
    }
  }
  o->analytic = an ? 1 : 0;
  o->errorUabsMax = eU;
  o->errorUabsAvg = an ? eUs / (double)nU : 0.0;
  o->errorJ1Max = e1;
  o->errorJ1Avg = an ? e1s / (double)nJ : 0.0;
  o->errorJ2Max = e2;
  o->errorJ2Avg = an ? e2s / (double)nJ : 0.0;
}

}  // namespace ep2d
