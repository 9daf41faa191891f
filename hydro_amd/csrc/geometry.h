// geometry.h -- the per-direction / per-side scalars of the uniform mesh
// that the kernels take by value (plain C++: the host code builds them, see
// host_geometry.h, and the HIP kernels read them through device_common.h).
#pragma once

namespace hg {

enum Side { kLeft = 0, kRight = 1, kBottom = 2, kTop = 3 };

struct Geo {
  int nx, ny;
  int nfx;            // number of x-faces, (nx+1)*ny; y-faces follow
  double V;           // cell volume (CalcVolume)
  double sx;          // x-face surface, x component (= area of an x-face)
  double sy;          // y-face surface, y component
  double ax, ay;      // x-face / y-face areas (CalcArea)
  double tx0, tx1;    // x-face: x component of vect to cell 0 / cell 1
  double ty0, ty1;    // y-face: y component of vect to cell 0 / cell 1
  double dnx, dny;    // |to1 - to0| for x- / y-faces
  double bal[4];      // boundary derivative factor (1/|to_valid|)*(+-1) per side
  double ex_dist[4];  // extrapolation: |to_valid| per side
  double ex_dot[4];   // extrapolation: S_out(opposite face) . n_outward / per side
  double ex_den[4];   // extrapolation: 1/dist - S_out(self) . n_outward / V
  double rdist_x, rdist_y;  // |x_cell - x_face| for x- / y-faces
  // 1 / V, and whether V is a power of two: then x / V and x * (1 / V) are
  // the same real number, so both round to the same double (a division
  // costs about ten FP64 instructions, a product one)
  double invV;
  int pow2V;
};

}  // namespace hg
