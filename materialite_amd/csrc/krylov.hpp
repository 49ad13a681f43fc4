// Device-resident MINRES state and the geometry shared by the FFT passes.
#pragma once

#include "common.hpp"

namespace mtl {

// Grid of the periodic cell: point p = (x*ny + y)*nz + z (z fastest).  The
// real-to-complex spectrum keeps kz in [0, nzc) with nzc = (nz+1)/2: for even
// nz the Nyquist plane kz = nz/2 is dropped because the projection zeroes it.
struct Grid {
  int nx, ny, nz, nzc;
  long long n;      // nx*ny*nz
  long long nspec;  // nx*ny*nzc complex values per component
};

// Scalars of the Paige-Saunders MINRES recurrence (SciPy's minres with x0 = 0,
// no preconditioner, no shift), kept in device memory so that a whole chunk of
// iterations is enqueued without host round trips.  The vector kernels read
// their coefficients from here and return at once when `done` is set.
struct MinresState {
  double beta1, beta, oldb, alfa;
  double dbar, epsln, phibar, rhs1, rhs2, tnorm2, gmax, gmin, cs, sn;
  double s;      // 1/beta: v = s * r2
  double s_prev; // 1/beta of the iteration whose w/x update is pending
  double c_r1;   // beta/oldb for itn >= 2, else 0: u = A v - c_r1 * r1
  double c_r2;   // alfa/beta: y = u - c_r2 * r2
  double oldeps, delta, denom, phi;
  double anorm, ynorm, rnorm, test1, test2, acond, root;
  double rtol;
  int maxiter;
  int itn;
  int istop;
  int done;
  int pending_xnorm;  // w/x updated, stopping tests still to run
  int pad;
};

}  // namespace mtl
