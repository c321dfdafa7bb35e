// MT19937 random stream with the interface of the reference's mt_rng
// (cpp/3d/mt_rng.hpp:76-110): Seed, init_by_array, ulrand, lrand and the
// four uniform maps.  The generator is Matsumoto and Nishimura's MT19937
// ("Mersenne Twister: a 623-dimensionally equidistributed uniform
// pseudorandom number generator", ACM TOMACS 8, 1998) with the 2002
// initialisation by a seed or an array; seeded with s it yields the same
// sequence as std::mt19937(s mod 2^32).  The GPU transport draws from its
// own per-photon streams; this class serves host code written against the
// reference's members (MC3D::rng, UnifClosed, ...).
#ifndef VMC_HOST_MT_RNG_HPP
#define VMC_HOST_MT_RNG_HPP

#include <cstdint>

class mt_rng {
public:
  explicit mt_rng(unsigned long s = 5489UL) { Seed(s); }

  // state from a 32-bit seed: x[i] = 1812433253 (x[i-1] xor x[i-1] >> 30) + i
  void Seed(unsigned long s) {
    x_[0] = (uint32_t)s;
    for (int i = 1; i < kN; ++i) x_[i] = 1812433253u * (x_[i - 1] ^ (x_[i - 1] >> 30)) + (uint32_t)i;
    i_ = kN;
  }

  // state from a key array (the 2002 init_by_array)
  void init_by_array(unsigned long *key, int len) {
    Seed(19650218UL);
    int i = 1, j = 0;
    for (int k = (kN > len ? kN : len); k; --k) {
      x_[i] = (x_[i] ^ ((x_[i - 1] ^ (x_[i - 1] >> 30)) * 1664525u)) + (uint32_t)key[j] + (uint32_t)j;
      if (++i >= kN) {
        x_[0] = x_[kN - 1];
        i = 1;
      }
      if (++j >= len) j = 0;
    }
    for (int k = kN - 1; k; --k) {
      x_[i] = (x_[i] ^ ((x_[i - 1] ^ (x_[i - 1] >> 30)) * 1566083941u)) - (uint32_t)i;
      if (++i >= kN) {
        x_[0] = x_[kN - 1];
        i = 1;
      }
    }
    x_[0] = 0x80000000u;
    i_ = kN;
  }

  // on [0, 0xffffffff]
  unsigned long ulrand() {
    if (i_ >= kN) twist();
    uint32_t y = x_[i_++];
    y ^= y >> 11;
    y ^= (y << 7) & 0x9d2c5680u;
    y ^= (y << 15) & 0xefc60000u;
    y ^= y >> 18;
    return y;
  }
  long lrand() { return (long)(ulrand() >> 1); }                                          // [0, 0x7fffffff]
  double drand_closed() { return (double)ulrand() * (1.0 / 4294967295.0); }               // [0, 1]
  double drand_open() { return ((double)ulrand() + 0.5) * (1.0 / 4294967296.0); }         // ]0, 1[
  double drand_open_up() { return (double)ulrand() * (1.0 / 4294967296.0); }             // [0, 1[
  double drand_open_down() { return ((double)ulrand() + 1.0) * (1.0 / 4294967296.0); }   // ]0, 1]

private:
  static constexpr int kN = 624, kM = 397;
  uint32_t x_[kN];
  int i_ = kN;
  void twist() {
    for (int k = 0; k < kN; ++k) {
      const uint32_t y = (x_[k] & 0x80000000u) | (x_[(k + 1) % kN] & 0x7fffffffu);
      x_[k] = x_[(k + kM) % kN] ^ (y >> 1) ^ ((y & 1u) ? 0x9908b0dfu : 0u);
    }
    i_ = 0;
  }
};

#endif
