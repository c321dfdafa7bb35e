// Fortran-order array used by the MC3D/MC2D drop-in classes.
//
// Same interface and element layout as the reference container
// (cpp/3d/Array.hpp:9-85) so code written against it compiles unchanged:
// a(i, j) is data[i + Nx*j]; copies and assignments share storage
// (IsRef = 1) and only the owner frees it; resize() allocates fresh storage.
#ifndef VMC_HOST_ARRAY_HPP
#define VMC_HOST_ARRAY_HPP

#include <cstdint>
#include <cstring>

template <class T> class Array {
public:
  Array() : IsRef(0), rank(0), Nx(0), Ny(0), Nz(0), Nxy(0), N(0), data(nullptr) {}
  Array(const Array<T> &ref) { share(ref); }
  Array &operator=(const Array &ref) {
    if (this != &ref) {
      destroy();
      share(ref);
    }
    return *this;
  }
  ~Array() { destroy(); }

  T *resize(int_fast64_t nx, int_fast64_t ny, int_fast64_t nz) { return alloc(3, nx, ny, nz); }
  T *resize(int_fast64_t nx, int_fast64_t ny) { return alloc(2, nx, ny, 1); }
  T *resize(int_fast64_t nx) { return alloc(1, nx, 1, 1); }

  T &operator()(int_fast64_t ix, int_fast64_t iy, int_fast64_t iz) { return data[ix + Nx * iy + Nxy * iz]; }
  T &operator()(int_fast64_t ix, int_fast64_t iy) { return data[ix + Nx * iy]; }
  T &operator()(int_fast64_t ix) { return data[ix]; }
  T &operator[](int_fast64_t i) { return data[i]; }
  const T &operator[](int_fast64_t i) const { return data[i]; }

  void destroy() {
    if (!IsRef && data) delete[] data;
    data = nullptr;
    rank = Nx = Ny = Nz = Nxy = N = 0;
    IsRef = 0;
  }

public:
  int IsRef;
  int_fast64_t rank, Nx, Ny, Nz, Nxy, N;
  T *data;

private:
  void share(const Array &ref) {
    IsRef = 1;
    rank = ref.rank;
    Nx = ref.Nx; Ny = ref.Ny; Nz = ref.Nz; Nxy = ref.Nxy; N = ref.N;
    data = ref.data;
  }
  T *alloc(int rk, int_fast64_t nx, int_fast64_t ny, int_fast64_t nz) {
    destroy();
    rank = rk;
    Nx = nx; Ny = ny; Nz = nz;
    Nxy = nx * ny;
    N = nx * ny * nz;
    data = new T[N > 0 ? N : 1];
    std::memset(static_cast<void *>(data), 0, sizeof(T) * (N > 0 ? N : 1));
    return data;
  }
};

#endif
