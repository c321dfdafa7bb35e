// MC3D drop-in: tetrahedral engine with the reference class's public API
// (cpp/3d/MC3D.hpp:51-173); transport runs on the GPU through valomc.h.
// The geometry queries keep the reference's expressions
// (MC3D.hpp:306-457): Heron's formula for face areas and the edge cross
// product, in face vertex order, for normals.
#ifndef VMC_HOST_MC3D_HPP
#define VMC_HOST_MC3D_HPP
#include "MCEngine.hpp"

class MC3D : public MCEngine<3> {
public:
  MC3D &operator=(const MC3D &ref) {
    MCEngine<3>::operator=(ref);
    return *this;
  }

  // volume of tetrahedron el (MC3D.hpp:306-332)
  double ElementVolume(int_fast64_t el) {
    double ax = r(H(el, 0), 0), ay = r(H(el, 0), 1), az = r(H(el, 0), 2);
    double bx = r(H(el, 1), 0), by = r(H(el, 1), 1), bz = r(H(el, 1), 2);
    double cx = r(H(el, 2), 0), cy = r(H(el, 2), 1), cz = r(H(el, 2), 2);
    double dx = r(H(el, 3), 0), dy = r(H(el, 3), 1), dz = r(H(el, 3), 2);
    double v = (ax - dx) * ((by - dy) * (cz - dz) - (bz - dz) * (cy - dy)) -
               (ay - dy) * ((bx - dx) * (cz - dz) - (bz - dz) * (cx - dx)) +
               (az - dz) * ((bx - dx) * (cy - dy) - (by - dy) * (cx - dx));
    return fabs(v) / 6.0;
  }
  // area of boundary triangle ib (MC3D.hpp:335-343)
  double ElementArea(int_fast64_t ib) { return heron(BH(ib, 0), BH(ib, 1), BH(ib, 2)); }
  // area of face f of tetrahedron el, faces (0,1,2) (0,1,3) (0,2,3) (1,2,3);
  // 0 for another f (MC3D.hpp:346-383)
  double ElementArea(int_fast64_t el, long f) {
    int v[3];
    if (!face(f, v)) return 0.0;
    return heron(H(el, v[0]), H(el, v[1]), H(el, v[2]));
  }
  // unit normal of boundary triangle ib (MC3D.hpp:386-405)
  void Normal(int_fast64_t ib, double *normal) { cross_normal(BH(ib, 0), BH(ib, 1), BH(ib, 2), normal); }
  // unit normal of face f of tetrahedron el; zero for another f (MC3D.hpp:408-457)
  void Normal(int_fast64_t el, long f, double *normal) {
    int v[3];
    if (!face(f, v)) {
      normal[0] = normal[1] = normal[2] = 0.0;
      return;
    }
    cross_normal(H(el, v[0]), H(el, v[1]), H(el, v[2]), normal);
  }

private:
  static bool face(long f, int v[3]) {
    static const int fv[4][3] = {{0, 1, 2}, {0, 1, 3}, {0, 2, 3}, {1, 2, 3}};   // MC3D.hpp:1013
    if (f < 0 || f > 3) return false;
    for (int k = 0; k < 3; ++k) v[k] = fv[f][k];
    return true;
  }
  double dist(int_fast64_t a, int_fast64_t b) {
    return sqrt(pow(r(b, 0) - r(a, 0), 2) + pow(r(b, 1) - r(a, 1), 2) + pow(r(b, 2) - r(a, 2), 2));
  }
  double heron(int_fast64_t i0, int_fast64_t i1, int_fast64_t i2) {
    const double a = dist(i0, i1), b = dist(i0, i2), c = dist(i2, i1);
    return sqrt((a + b + c) * (b + c - a) * (c + a - b) * (a + b - c)) / 4.0;
  }
  void cross_normal(int_fast64_t i0, int_fast64_t i1, int_fast64_t i2, double *nrm) {
    const double x1 = r(i1, 0) - r(i0, 0), y1 = r(i1, 1) - r(i0, 1), z1 = r(i1, 2) - r(i0, 2);
    const double x2 = r(i2, 0) - r(i0, 0), y2 = r(i2, 1) - r(i0, 1), z2 = r(i2, 2) - r(i0, 2);
    const double nx = y1 * z2 - z1 * y2, ny = z1 * x2 - x1 * z2, nz = x1 * y2 - y1 * x2;
    const double len = sqrt(nx * nx + ny * ny + nz * nz);
    nrm[0] = nx / len;
    nrm[1] = ny / len;
    nrm[2] = nz / len;
  }
};
#endif
