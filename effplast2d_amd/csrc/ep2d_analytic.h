// ep2d_analytic.h -- analytic static solution of the single-pore problem and
// its comparison with the numerical fields (SaveAnStatic2D,
// EffPlast2D.cu:988-1292).  Host FP64, run once after a solve.
#pragma once

#include "ep2d.h"

namespace ep2d {

struct AnGeometry {
  long nX, nY;
  double dX, dY, rad, Y, G0, K0;  // Y = pa[8]/sqrt(2), EffPlast2D.cu:1572
  int nPores;
};

// Fields of one state: P, tauXX, tauYY (nX*nY), tauXY ((nX-1)*(nY-1)) and the
// displacements Ux ((nX+1)*nY), Uy (nX*(nY+1)) with the pore nodes already
// zeroed.  Fills the non-null arrays of out and its error statistics.
void an_static_fields(const AnGeometry& g, const double* P, const double* txx,
                      const double* tyy, const double* txy, const double* Ux, const double* Uy,
                      double deltaP, double tauInfty, const double loadType[3],
                      ep2d_an_static* out);

}  // namespace ep2d
