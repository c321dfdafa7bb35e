// MC2D drop-in: triangular engine with the reference class's public API
// (cpp/2d/MC2D.hpp:59-180); transport runs on the GPU through valomc.h.
// The geometry queries keep the reference's expressions (MC2D.hpp:303-390):
// a triangle's area from the edge cross product, lengths of edges, and the
// normal (dy, -dx) / |d| of an edge d = v1 - v0.
#ifndef VMC_HOST_MC2D_HPP
#define VMC_HOST_MC2D_HPP
#include "MCEngine.hpp"

class MC2D : public MCEngine<2> {
public:
  MC2D &operator=(const MC2D &ref) {
    MCEngine<2>::operator=(ref);
    return *this;
  }

  // area of triangle el (MC2D.hpp:303-310)
  double ElementArea(int_fast64_t el) {
    const double ax = r(H(el, 1), 0) - r(H(el, 0), 0), ay = r(H(el, 1), 1) - r(H(el, 0), 1);
    const double bx = r(H(el, 2), 0) - r(H(el, 0), 0), by = r(H(el, 2), 1) - r(H(el, 0), 1);
    return 0.5 * fabs(ax * by - ay * bx);
  }
  // length of boundary edge ib (MC2D.hpp:313-318)
  double ElementLength(int_fast64_t ib) { return len(BH(ib, 0), BH(ib, 1)); }
  // length of edge f of triangle el, edges (0,1) (1,2) (2,0); 0 for another f
  // (MC2D.hpp:321-345)
  double ElementLength(int_fast64_t el, int_fast64_t f) {
    int v[2];
    if (!edge(f, v)) return 0.0;
    return len(H(el, v[0]), H(el, v[1]));
  }
  // unit normal of boundary edge ib (MC2D.hpp:348-356)
  void Normal(int_fast64_t ib, double *normal) { edge_normal(BH(ib, 0), BH(ib, 1), normal); }
  // unit normal of edge f of triangle el; zero for another f (MC2D.hpp:359-390)
  void Normal(int_fast64_t el, int_fast64_t f, double *normal) {
    int v[2];
    if (!edge(f, v)) {
      normal[0] = normal[1] = 0.0;
      return;
    }
    edge_normal(H(el, v[0]), H(el, v[1]), normal);
  }

private:
  static bool edge(int_fast64_t f, int v[2]) {
    static const int ev[3][2] = {{0, 1}, {1, 2}, {2, 0}};   // MC2D.hpp:921
    if (f < 0 || f > 2) return false;
    v[0] = ev[f][0];
    v[1] = ev[f][1];
    return true;
  }
  double len(int_fast64_t a, int_fast64_t b) { return sqrt(pow(r(b, 0) - r(a, 0), 2) + pow(r(b, 1) - r(a, 1), 2)); }
  void edge_normal(int_fast64_t a, int_fast64_t b, double *nrm) {
    const double dx = r(b, 0) - r(a, 0), dy = r(b, 1) - r(a, 1);
    const double l = sqrt(dx * dx + dy * dy);
    nrm[0] = dy / l;
    nrm[1] = -dx / l;
  }
};
#endif
