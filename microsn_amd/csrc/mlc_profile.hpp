// mlc_profile.hpp -- source brightness profiles on the reference's pixel grid
// and their derivative with respect to the half-light radius.  Host and
// device.  Follows FactoryProfile::createProfileFromHalfRadius
// (simulating_lcurves/include/profile.hpp:142-166) and the generateValues
// rules of profile.cpp:185-446, with incl = orient = 0 (the only values the
// driver uses, gpu_functions.cu:108-109).
#pragma once

#include <cmath>

#ifdef __HIPCC__
#include <hip/hip_runtime.h>
#define MLC_PHD __host__ __device__
#else
#define MLC_PHD
#endif

namespace mlc {

struct ProfileParams {
  int shape;
  int n;        // even grid size (pixels per side)
  double pix;   // pixel size (same unit as rhalf)
  double rhalf; // half-light radius the caller asked for
  double p0;    // R | sdev | sdev | Rin | b | sigma
  double p1;    // - | -    | Rin  | -   | a | -
  double rtr;   // truncation radius of wavy / exponential (2*Rhalf), else 0
  double dp0;   // d p0 / d rhalf
};

// Returns false for an unknown shape or an empty grid.
MLC_PHD inline bool profile_params(int shape, double rhalf, double pix, double rin, double a,
                                   ProfileParams* P) {
  P->shape = shape;
  P->pix = pix;
  P->rhalf = rhalf;
  P->p1 = 0.0;
  P->rtr = 0.0;
  int n = 0;
  switch (shape) {
    case 0: {  // UniformDisc: R = sqrt(2) rhalf; n = 2*(3+ceil(R/pix))  (profile.cpp:190-199)
      double R = sqrt(2.0) * rhalf;
      P->p0 = R;
      P->dp0 = sqrt(2.0);
      n = 2 * (3 + (int)ceil(R / pix));
      break;
    }
    case 1: {  // Gaussian: sdev = rhalf/1.18, Rhalf = 1.18 sdev (profile.cpp:233-244)
      double sdev = rhalf / 1.18;
      P->p0 = sdev;
      P->dp0 = 1.0 / 1.18;
      double Rh = 1.18 * sdev;
      n = (int)ceil(4.0 * Rh / pix);
      break;
    }
    case 2: {  // GaussianHole (profile.cpp:273-288)
      double dum = (pow(rhalf, 2.0) - pow(rin, 2.0)) / (2.0 * log(2.0));
      double sdev = sqrt(dum);
      P->p0 = sdev;
      P->p1 = rin;
      P->dp0 = rhalf / (2.0 * log(2.0) * sdev);
      double Rh = sqrt(2.0 * log(2.0) * pow(sdev, 2.0) + pow(rin, 2.0));
      n = (int)ceil(4.0 * Rh / pix);
      break;
    }
    case 3: {  // ThermalHole: Rin = rhalf/4, Rhalf = 4 Rin (profile.cpp:320-330)
      double Rin = rhalf / 4.0;
      P->p0 = Rin;
      P->dp0 = 0.25;
      double Rh = 4.0 * Rin;
      n = (int)ceil(4.0 * Rh / pix);
      break;
    }
    case 4: {  // Wavy, 3 nodes: b = 3 pi/(2 rhalf), Rhalf = 3 pi/(2 b) (profile.cpp:365-375)
      double b = 3 * M_PI / (2.0 * rhalf);
      P->p0 = b;
      P->p1 = a;
      P->dp0 = -b / rhalf;
      double Rh = 3 * M_PI / (2.0 * b);
      P->rtr = 2 * Rh;
      n = (int)ceil(4.0 * Rh / pix);
      break;
    }
    case 5: {  // Exponential thin disc: sigma = 0.4106 rhalf (profile.cpp:407-417)
      double sigma = 0.4106 * rhalf;
      P->p0 = sigma;
      P->dp0 = 0.4106;
      double Rh = sigma / 0.4106;
      P->rtr = 2 * Rh;
      n = (int)ceil(4.0 * Rh / pix);
      break;
    }
    default:
      return false;
  }
  if (n % 2 != 0) n += 1;  // BaseProfile::makeEven profile.cpp:175-179
  P->n = n;
  return n > 0;
}

// Unnormalised value f(r) and df/drhalf at radius r.
MLC_PHD inline void profile_value(const ProfileParams& P, double r, double* f, double* df) {
  double v = 0.0, dv = 0.0;
  switch (P.shape) {
    case 0: {  // profile.cpp:211-218.  The indicator has no derivative; the
               // gradient uses the edge of a one-pixel linear ramp instead.
      double R = P.p0;
      v = (r < R) ? 1.0 : 0.0;
      if (fabs(r - R) < 0.5 * P.pix) dv = P.dp0 / P.pix;
      break;
    }
    case 1: {  // profile.cpp:254-259
      double s = 2 * pow(P.p0, 2);
      v = exp(-r * r / s);
      dv = v * r * r / (P.p0 * P.p0 * P.p0) * P.dp0;
      break;
    }
    case 2: {  // profile.cpp:298-307
      double s = 2 * pow(P.p0, 2);
      if (r < P.p1) {
        v = 0.0;
      } else {
        v = exp(-r * r / s);
        dv = v * r * r / (P.p0 * P.p0 * P.p0) * P.dp0;
      }
      break;
    }
    case 3: {  // profile.cpp:342-350
      double Rin = P.p0;
      double fac = 3.0 * Rin / (2.0 * M_PI);
      if (r > Rin) {
        v = (1.0 - sqrt(Rin / r)) * fac / pow(r, 3.0);
        double dRin = (-0.5 / sqrt(Rin * r) * fac + (1.0 - sqrt(Rin / r)) * (3.0 / (2.0 * M_PI))) /
                      pow(r, 3.0);
        dv = dRin * P.dp0;
      }
      break;
    }
    case 4: {  // profile.cpp:387-393
      if (r < P.rtr) {
        v = pow(sin(P.p0 * r), 2) / (P.p1 * r);
        dv = sin(2.0 * P.p0 * r) / P.p1 * P.dp0;
      }
      break;
    }
    case 5: {  // profile.cpp:433-442: the r < r_inner branch is always
               // overwritten by the following if/else, so it has no effect.
      if (r < P.rtr) {
        double u = pow(r / P.p0, 0.75);
        double eu = exp(u);
        v = 1.0 / (eu - 1.0);
        dv = 0.75 * u * eu / (P.p0 * (eu - 1.0) * (eu - 1.0)) * P.dp0;
      }
      break;
    }
  }
  *f = v;
  *df = dv;
}

// Pixel-centre coordinate of grid index j (createGrid profile.cpp:143-157).
MLC_PHD inline double profile_coord(int j, int n, double pix) { return (j - n / 2) * pix + pix / 2.0; }

}  // namespace mlc
