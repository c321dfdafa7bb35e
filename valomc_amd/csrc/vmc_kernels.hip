// Photon-packet transport kernels for gfx950 (MI355X).
//
// One lane carries one photon packet from creation at a light source to its
// end (exit, roulette, loss).  Waves are persistent: a wave whose lanes have
// finished takes the next photon ids from a wave-private pool that is filled
// from one global counter in chunks, so the walk keeps all 64 lanes busy
// until the photon budget runs out.  Photon i draws its random numbers from
// its own stream keyed by (seed, tag, i), so a result does not depend on the
// launch shape or on which lane ran which photon.
//
// Transport follows ValoMC's PropagatePhoton (cpp/3d/MC3D.hpp:1450-1638,
// cpp/2d/MC2D.hpp:1307-1511) step for step and consumes random numbers in
// the same order.  The exit face is found from precomputed outward face
// planes (argmin of the positive plane distances) instead of the reference's
// Moeller-Trumbore test on gathered vertices (MC3D.hpp:1006-1057,
// 1825-1869); for a photon inside a convex element both pick the same face.
//
// Tallies: the deposits of one element visit are summed in registers; a
// workgroup's LDS sector cache (sectors of 8 consecutive elements) takes
// the visits of the sectors that claimed its slots; other sectors are
// summed in per-lane LDS buffers that leave as coalesced atomics when the
// photon moves on.  Those go to one of `nrep` replicas, the XCD's own
// (workgroup-scope atomics) or, for an explicit count, one picked by
// workgroup index (device scope); the replicas are summed after the walk.
//
// Small meshes: when a single problem's element records and one cache slot
// per sector fit in a workgroup's LDS, walk_small_kernel copies the records
// into LDS at the start and every sector owns its slot, so the walk reads
// no element record from memory and adds no tally word to it before the
// workgroup's final flush.
#include "vmc_types.h"

#include <type_traits>

namespace vmc {

// ---------------------------------------------------------------------------
// Random numbers
// ---------------------------------------------------------------------------

// Philox4x32-10 (Salmon et al., SC'11) with the published constants.
__device__ __forceinline__ void philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
                                              uint32_t k0, uint32_t k1, uint32_t &o0,
                                              uint32_t &o1, uint32_t &o2, uint32_t &o3) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    if (r) {
      k0 += 0x9E3779B9u;
      k1 += 0xBB67AE85u;
    }
    const uint64_t p0 = (uint64_t)c0 * 0xD2511F53u;   // one v_mad_u64_u32 each
    const uint64_t p1 = (uint64_t)c2 * 0xCD9E8D57u;
    const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0, n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
    c0 = n0;
    c1 = (uint32_t)p1;
    c2 = n2;
    c3 = (uint32_t)p0;
  }
  o0 = c0;
  o1 = c1;
  o2 = c2;
  o3 = c3;
}

// Stream of photon p (64-bit id) under key (seed) and tag: a xoshiro128**
// generator (Blackman & Vigna, "Scrambled linear pseudorandom number
// generators", 2018) whose 128-bit state is the Philox4x32-10 block
// (0x80000000, tag, p lo, p hi) under key (seed lo, seed hi).  Creation and
// walk draw from it in the reference's order.  A draw costs a few full-rate
// integer instructions and is made only in the branch that needs it, so
// lanes that cross a face without scattering pay nothing; random 128-bit
// starting points in a 2^128 - 1 period make overlap of two photons'
// sequences negligible.
struct WalkRng {
  uint32_t s0, s1, s2, s3;
  __device__ __forceinline__ void seed(uint32_t plo, uint32_t phi, const ItemRec &it) {
    philox4x32_10(0x80000000u, it.tag, plo, phi, it.key0, it.key1, s0, s1, s2, s3);
    if ((s0 | s1 | s2 | s3) == 0u) s0 = 1u;   // the all-zero state is a fixed point
  }
  __device__ __forceinline__ uint32_t next(const WalkParams &) {
    const uint32_t r = __builtin_rotateleft32(s1 * 5u, 7) * 9u;
    const uint32_t t = s1 << 9;
    s2 ^= s0;
    s3 ^= s1;
    s1 ^= s2;
    s0 ^= s3;
    s2 ^= t;
    s3 = __builtin_rotateleft32(s3, 11);
    return r;
  }
};

// Uniform maps.  FP64 uses the reference's maps on the full 32-bit draw
// (cpp/3d/mt_rng.hpp:186-204); FP32 uses the top 24 bits so every value is
// exactly representable.
template <class R> struct Unif;
template <> struct Unif<double> {
  static __device__ __forceinline__ double closed(uint32_t u) { return (double)u * (1.0 / 4294967295.0); }
  static __device__ __forceinline__ double open(uint32_t u) { return ((double)u + 0.5) * (1.0 / 4294967296.0); }
  static __device__ __forceinline__ double halfup(uint32_t u) { return (double)u * (1.0 / 4294967296.0); }
};
template <> struct Unif<float> {
  static __device__ __forceinline__ float closed(uint32_t u) {
    return fminf((float)(u >> 8) * (1.0f / 16777215.0f), 1.0f);
  }
  static __device__ __forceinline__ float open(uint32_t u) { return ((float)(u >> 8) + 0.5f) * 5.9604644775390625e-8f; }
  static __device__ __forceinline__ float halfup(uint32_t u) { return (float)(u >> 8) * 5.9604644775390625e-8f; }
};

// ---------------------------------------------------------------------------
// Precision-generic math
// ---------------------------------------------------------------------------

__device__ __forceinline__ float m_sqrt(float x) { return __builtin_sqrtf(x); }
__device__ __forceinline__ double m_sqrt(double x) { return sqrt(x); }
__device__ __forceinline__ float m_log(float x) { return logf(x); }
__device__ __forceinline__ double m_log(double x) { return log(x); }
__device__ __forceinline__ float m_exp(float x) { return expf(x); }
__device__ __forceinline__ double m_exp(double x) { return exp(x); }
__device__ __forceinline__ float m_expm1(float x) { return expm1f(x); }
__device__ __forceinline__ double m_expm1(double x) { return expm1(x); }
__device__ __forceinline__ float m_asin(float x) { return asinf(x); }
__device__ __forceinline__ double m_asin(double x) { return asin(x); }
__device__ __forceinline__ float m_erf(float x) { return erff(x); }
__device__ __forceinline__ double m_erf(double x) { return erf(x); }
__device__ __forceinline__ void m_sincos(float x, float *s, float *c) { sincosf(x, s, c); }
__device__ __forceinline__ void m_sincos(double x, double *s, double *c) { sincos(x, s, c); }
__device__ __forceinline__ void m_sincospi(float x, float *s, float *c) { sincospif(x, s, c); }
__device__ __forceinline__ void m_sincospi(double x, double *s, double *c) { sincospi(x, s, c); }

// Walk-loop math.  FP64 keeps the libm functions of the reference's
// expressions.  FP32 uses the hardware transcendental instructions, whose
// errors (about an ulp for log2/exp2, a few 1e-7 absolute for sin/cos of an
// argument in revolutions) sit far below the FP32 transport's 1% budget:
//   -log(u) = -log2(u) ln 2            (u in (0, 1])
//   sin, cos of 2 pi u                 (u in revolutions, |u| < 256)
//   expm1(x), x <= 0                   Taylor series to x^5 on (-1/16, 0]
//                                      (relative error < 2e-9), else exp2 - 1
__device__ __forceinline__ float w_neglog(float u) { return -__builtin_amdgcn_logf(u) * 0.693147180559945309f; }
__device__ __forceinline__ double w_neglog(double u) { return -log(u); }
__device__ __forceinline__ void w_sincos_rev(float u, float *s, float *c) {
  *s = __builtin_amdgcn_sinf(u);
  *c = __builtin_amdgcn_cosf(u);
}
__device__ __forceinline__ void w_sincos_rev(double u, double *s, double *c) { sincospi(2.0 * u, s, c); }
__device__ __forceinline__ float w_expm1_neg(float x) {
  const float p = x * (1.0f + x * (0.5f + x * (1.0f / 6.0f + x * (1.0f / 24.0f + x * (1.0f / 120.0f)))));
  const float e = __builtin_amdgcn_exp2f(x * 1.44269504088896341f) - 1.0f;
  return x > -0.0625f ? p : e;
}
__device__ __forceinline__ double w_expm1_neg(double x) { return expm1(x); }
// sin and cos of pi x, x in [0, 1]
__device__ __forceinline__ void w_sincospi(float x, float *s, float *c) { w_sincos_rev(0.5f * x, s, c); }
__device__ __forceinline__ void w_sincospi(double x, double *s, double *c) { sincospi(x, s, c); }
// sin and cos of a small angle x in radians
__device__ __forceinline__ void w_sincos_small(float x, float *s, float *c) {
  w_sincos_rev(x * 0.159154943091895336f, s, c);
}
__device__ __forceinline__ void w_sincos_small(double x, double *s, double *c) { sincos(x, s, c); }
// a / b and 1 / sqrt(x): hardware reciprocal (about an ulp) in FP32,
// the reference's correctly rounded expressions in FP64
__device__ __forceinline__ float w_div(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
__device__ __forceinline__ double w_div(double a, double b) { return a / b; }
__device__ __forceinline__ float w_rsqrt(float x) { return __builtin_amdgcn_rsqf(x); }
__device__ __forceinline__ double w_rsqrt(double x) { return 1.0 / sqrt(x); }
// sqrt(x) of a normal or zero x (no denormal-range scaling in FP32)
__device__ __forceinline__ float w_sqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
__device__ __forceinline__ double w_sqrt(double x) { return sqrt(x); }
// sqrt(x) and 1 / sqrt(x) of x >= 1e-3: one reciprocal square root in FP32
__device__ __forceinline__ void w_sqrt_pair(float x, float *s, float *is) {
  *is = __builtin_amdgcn_rsqf(x);
  *s = x * *is;
}
__device__ __forceinline__ void w_sqrt_pair(double x, double *s, double *is) {
  *s = sqrt(x);
  *is = 1.0 / *s;
}

template <class R> struct Tol;
template <> struct Tol<float> { static constexpr float out = 1e-6f; };
template <> struct Tol<double> { static constexpr double out = 1e-12; };

// ---------------------------------------------------------------------------
// Tally adds
// ---------------------------------------------------------------------------

// Per-workgroup LDS cache of sector tallies.  A slot is claimed by the
// first sector that hashes to it and keeps that sector for the rest of the
// launch; flushes of a claimed sector stay in LDS and reach the global
// replica once, when the workgroup finishes.  Every photon starts at a light
// source, so the slots fill with the sectors next to the sources, which
// take the most flushes.  IDENT (small meshes): there is a slot per sector,
// slot s holds sector s from the start and nothing is claimed.
template <bool FIXED, int WS, bool IDENT> struct SectorCache {
  int *tag;                  // [S] sector id or -1
  unsigned long long *val;   // [S][WS] FP64 bits or fixed point
  int nslots;                // 0: cache off
  __device__ __forceinline__ int slot_of(int sect) const {
    return (int)(((uint64_t)((uint32_t)sect * 0x9E3779B1u) * (uint32_t)nslots) >> 32);
  }
  // the sector's slot, claimed if still free; -1 when the slot belongs to
  // another sector or the cache is off
  __device__ __forceinline__ int claim(int sect) {
    if constexpr (IDENT) return sect;
    if (nslots == 0) return -1;
    const int sl = slot_of(sect);
    int t = __hip_atomic_load(&tag[sl], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    if (t == -1) {
      const int old = atomicCAS(&tag[sl], -1, sect);
      t = old == -1 ? sect : old;
    }
    return t == sect ? sl : -1;
  }
  // moves the lane buffer words buf[j * stride] of `sect` into the sector's
  // slot (clearing them); false, leaving them, when the sector has no slot
  // (fixed point: `range_bad` is raised for a word at or beyond 2^62)
  template <class R>
  __device__ __forceinline__ bool add_from(int sect, R *buf, int stride, double scale, bool &range_bad) {
    const int sl = claim(sect);
    if (sl < 0) return false;
#pragma unroll 4
    for (int j = 0; j < WS; ++j) {
      const R v = buf[j * stride];
      buf[j * stride] = R(0);
      if (v == R(0)) continue;
      if constexpr (FIXED) {
        const double s = (double)v * scale;
        range_bad = range_bad || !(fabs(s) < 0x1p62);
        atomicAdd(val + (size_t)sl * WS + j, (unsigned long long)__double2ll_rn(s));
      } else {
        atomicAdd(reinterpret_cast<double *>(val) + (size_t)sl * WS + j, (double)v);
      }
    }
    return true;
  }
};

// Index of the XCD (accelerator complex die) running this wave:
// hardware register XCC_ID (id 20, bits 3:0) on gfx940 and later.
__device__ __forceinline__ int xcc_id() { return (int)(__builtin_amdgcn_s_getreg((3 << 11) | 20) & 15u); }

// ---------------------------------------------------------------------------
// Photon state and helpers
// ---------------------------------------------------------------------------

template <class R> struct Photon {
  R x[3], d[3];
  R w, prop;
  R cph, sph;   // cos/sin of the phase (frequency domain)
  R acc0, acc1; // deposits of the current element visit (ER, EI), flushed on leaving
  int el, face, nsteps;
};

// Element records, light sources and boundary records live in device
// global memory.  Their pointers are read from an item record, so without
// this address space the compiler would access them with flat instructions,
// which also wait on the LDS counter.
template <int NF, class R> using ElemPtr = const __attribute__((address_space(1))) Elem<NF, R> *;
template <class T> __device__ __forceinline__ const __attribute__((address_space(1))) T *global_ptr(const T *p) {
  return (const __attribute__((address_space(1))) T *)p;
}

// Element records staged in a workgroup's LDS (small meshes).
template <int NF, class R> using ElemLds = const __attribute__((address_space(3))) Elem<NF, R> *;

// Load the record of element `el` into registers (from global memory or,
// for a small mesh, from LDS).
template <int NF, class R, class EP>
__device__ __forceinline__ void load_elem(EP E, int el,
                                          typename Vec4T<R>::type (&pl)[NF],
                                          typename Vec4T<R>::type &opt, i4 &nb,
                                          typename Vec2T<R>::type (&nbo)[NF]) {
  const auto e = E + el;
#pragma unroll
  for (int f = 0; f < NF; ++f) pl[f] = e->pl[f];
  opt = e->opt;
  nb = e->nb;
#pragma unroll
  for (int f = 0; f < NF; ++f) nbo[f] = e->nbo[f];
}

// Fresnel transmission/reflection at a face with outward unit normal n
// (MC3D::FresnelPhoton, MC3D.hpp:1365-1447).  The reflection probability is
// the unpolarised Fresnel formula in cosine form, which equals the
// reference's sin/tan form given Snell's law, including its normal-incidence
// limit.  Returns true when the packet is reflected.
template <int D, class R, class S>
__device__ __forceinline__ bool fresnel(Photon<R> &p, const R n[3], R n1, R n2, bool refract,
                                        S &rs, const WalkParams &P) {
  const R eta = w_div(n1, n2);
  R ci = p.d[0] * n[0] + p.d[1] * n[1];
  if (D == 3) ci += p.d[2] * n[2];
  const R s2t = eta * eta * (R(1) - ci * ci);
  bool refl = true;
  if (s2t < R(1)) {
    const R ct = w_sqrt(R(1) - s2t);   // 1 - s2t >= 2^-24 in FP32: a normal number
    const R rs_ = w_div(eta * ci - ct, eta * ci + ct);
    const R rp = w_div(ci - eta * ct, ci + eta * ct);
    R Rf = R(0.5) * (rs_ * rs_ + rp * rp);
    // a direction longer than unit length (an unnormalised absolute light
    // direction before its first scattering) makes the reference's
    // acos(costhi) NaN, and its sin(thi + tht) > eps test then selects the
    // normal-incidence value (MC3D.hpp:1409-1419)
    if (ci > R(1)) Rf = w_div(eta - R(1), eta + R(1)) * w_div(eta - R(1), eta + R(1));
    refl = Unif<R>::closed(rs.next(P)) <= Rf;
    if (!refl && refract) {
      const R s = ct - eta * ci;
#pragma unroll
      for (int k = 0; k < D; ++k) p.d[k] = eta * p.d[k] + s * n[k];
    }
  }
  if (refl) {
#pragma unroll
    for (int k = 0; k < D; ++k) p.d[k] -= R(2) * ci * n[k];
  }
  return refl;
}

// Henyey-Greenstein scattering, MC3D::ScatterPhoton (MC3D.hpp:1297-1340).
template <class R, class S>
__device__ __forceinline__ void scatter3(Photon<R> &p, R g, S &rs, const WalkParams &P) {
  R cost;
  if (g != R(0)) {
    const R xi = Unif<R>::closed(rs.next(P));
    if (R(0) < xi && xi < R(1)) {
      const R g2 = g * g;
      const R t = w_div(R(1) - g2, R(1) - g * (R(1) - R(2) * xi));
      cost = w_div(R(1) + g2 - t * t, R(2) * g);
    } else {
      cost = xi >= R(1) ? R(1) : R(-1);
    }
  } else {
    cost = R(2) * Unif<R>::closed(rs.next(P)) - R(1);
  }
  cost = fmin(fmax(cost, R(-1)), R(1));
  const R sint = w_sqrt(fmax(R(1) - cost * cost, R(0)));
  R sp, cp;
  w_sincos_rev(Unif<R>::closed(rs.next(P)), &sp, &cp);
  R nx, ny, nz;
  if (fabs(p.d[2]) > R(0.999)) {
    nx = sint * cp;
    ny = sint * sp;
    nz = p.d[2] >= R(0) ? cost : -cost;
  } else {
    R tmp, itmp;   // |dz| <= 0.999 here
    w_sqrt_pair(R(1) - p.d[2] * p.d[2], &tmp, &itmp);
    nx = sint * (p.d[0] * p.d[2] * cp - p.d[1] * sp) * itmp + p.d[0] * cost;
    ny = sint * (p.d[1] * p.d[2] * cp + p.d[0] * sp) * itmp + p.d[1] * cost;
    nz = -sint * cp * tmp + p.d[2] * cost;
  }
  const R inv = w_rsqrt(nx * nx + ny * ny + nz * nz);
  p.d[0] = nx * inv;
  p.d[1] = ny * inv;
  p.d[2] = nz * inv;
  p.face = -1;
}

// 2D Henyey-Greenstein, MC2D::ScatterPhoton (MC2D.hpp:1135-1155):
// theta = 2 atan(tan(pi xi) (1-g)/(1+g)), applied as a rotation.
template <class R, class S>
__device__ __forceinline__ void scatter2(Photon<R> &p, R g, S &rs, const WalkParams &P) {
  const R xi = Unif<R>::halfup(rs.next(P));
  R s, c;
  if (g != R(0)) {
    w_sincospi(xi, &s, &c);
    const R q = w_div(R(1) - g, R(1) + g);
    const R a = q * s;
    const R den = w_div(R(1), c * c + a * a);
    const R ct = (c * c - a * a) * den, st = R(2) * a * c * den;
    // the reference rebuilds the direction from its angle, so the result
    // is a unit vector whatever the input length
    const R il = w_rsqrt(p.d[0] * p.d[0] + p.d[1] * p.d[1]);
    const R dx = p.d[0] * il, dy = p.d[1] * il;
    p.d[0] = dx * ct - dy * st;
    p.d[1] = dx * st + dy * ct;
  } else {
    w_sincospi(R(2) * xi, &s, &c);
    p.d[0] = c;
    p.d[1] = s;
  }
  p.face = -1;
}

// Truncated-normal launch angle by the reference's 16-step bisection of the
// CDF (MC2D.hpp:1044-1056).
template <class R> __device__ __forceinline__ R gauss_angle(R xi, R sig, R h) {
  const R pi = R(3.14159265358979323846);
  R a = -pi / R(2), b = pi / R(2), phi = R(0);
  const R inv = R(1) / (sig * R(1.41421356237309504880));
  for (int it = 0; it < 16; ++it) {
    phi = (a + b) / R(2);
    const R cdf = (R(0.5) + R(0.5) * m_erf(phi * inv) - h) / (R(1) - R(2) * h);
    if (cdf - xi < R(0)) a = phi; else b = phi;
  }
  return phi;
}

// CreatePhoton (MC3D.hpp:1060-1294, MC2D.hpp:958-1132).  Returns the source's
// mother element and initial face.
template <int D, class R>
__device__ __forceinline__ void create_photon(Photon<R> &p, WalkRng &rs, const ItemRec &it, const WalkParams &P) {
  // source entry: first CDF value above xi, clamped (MC2D.hpp:966-969)
  const double xi = Unif<double>::closed(rs.next(P));
  const int nsrc = it.nsrc;
  int lo = 0, hi = nsrc > 1 ? nsrc : 0;   // one source: nothing to search
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (xi < global_ptr(it.cdf)[mid]) hi = mid; else lo = mid + 1;
  }
  const auto &s = global_ptr(it.src)[lo < nsrc ? lo : nsrc - 1];
  p.el = s.mother;
  p.face = s.face;
  const int kind = s.kind;
  if constexpr (D == 3) {
    const R r1 = Unif<R>::open(rs.next(P)), r2 = Unif<R>::open(rs.next(P));
    const R sr = m_sqrt(r1);
    const R c0 = R(1) - sr, c1 = sr * (R(1) - r2), c2 = sr * r2;
#pragma unroll
    for (int k = 0; k < 3; ++k) p.x[k] = c0 * R(s.v0[k]) + c1 * R(s.v1[k]) + c2 * R(s.v2[k]);
    const R n0 = R(s.nin[0]), n1 = R(s.nin[1]), n2 = R(s.nin[2]);
    if (kind == kSrcGiven) {
#pragma unroll
      for (int k = 0; k < 3; ++k) p.d[k] = R(s.dir[k]);
    } else if (kind == kSrcPencil) {
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        p.d[k] = R(s.nin[k]);
        p.x[k] = R(s.dir[k]);
      }
    } else if (kind == kSrcIso3) {
      R v0 = n0, v1 = n1, v2 = n2;
      for (int it = 0; it < 256; ++it) {
        const R th = Unif<R>::halfup(rs.next(P));
        const R u = R(2) * Unif<R>::closed(rs.next(P)) - R(1);
        R st, ct;
        m_sincospi(R(2) * th, &st, &ct);
        const R q = m_sqrt(fmax(R(1) - u * u, R(0)));
        v0 = q * ct;
        v1 = q * st;
        v2 = u;
        if (n0 * v0 + n1 * v1 + n2 * v2 > R(0)) break;
      }
      p.d[0] = v0;
      p.d[1] = v1;
      p.d[2] = v2;
    } else if (kind == kSrcCos3) {
      const R ph = m_asin(R(2) * Unif<R>::open(rs.next(P)) - R(1));
      const R th = Unif<R>::closed(rs.next(P));
      R sp, cp, st, ct;
      m_sincos(ph, &sp, &cp);
      m_sincospi(R(2) * th, &st, &ct);
      const R f1 = ct * sp, f2 = st * sp;
#pragma unroll
      for (int k = 0; k < 3; ++k) p.d[k] = cp * R(s.nin[k]) + f1 * R(s.a1[k]) + f2 * R(s.a2[k]);
    } else {
      p.d[0] = n0;
      p.d[1] = n1;
      p.d[2] = n2;
    }
  } else {
    const R w = Unif<R>::open(rs.next(P));
#pragma unroll
    for (int k = 0; k < 2; ++k) p.x[k] = R(s.v0[k]) + w * (R(s.v1[k]) - R(s.v0[k]));
    p.x[2] = R(0);
    p.d[2] = R(0);
    const R n0 = R(s.nin[0]), n1 = R(s.nin[1]);
    const R phi0 = R(s.phi0);
    if (kind == kSrcNormal) {
      p.d[0] = n0;
      p.d[1] = n1;
    } else if (kind == kSrcGiven) {
      p.d[0] = R(s.dir[0]);
      p.d[1] = R(s.dir[1]);
    } else {
      for (int it = 0; it < 256; ++it) {
        R phi;
        if (kind == kSrcIso2) {
          phi = R(3.14159265358979323846) * (Unif<R>::open(rs.next(P)) - R(0.5));
        } else if (kind == kSrcCos2 || kind == kSrcCos2Rej) {
          phi = m_asin(R(2) * Unif<R>::open(rs.next(P)) - R(1));
        } else {
          phi = gauss_angle<R>(Unif<R>::open(rs.next(P)), R(s.gsig), R(s.gh));
        }
        R sn, cs;
        m_sincos(phi0 + phi, &sn, &cs);
        p.d[0] = cs;
        p.d[1] = sn;
        if (!(kind == kSrcCos2Rej || kind == kSrcGauss2Rej)) break;
        if (n0 * cs + n1 * sn > R(0)) break;
      }
    }
  }
  p.w = R(1);
  p.acc0 = R(0);
  p.acc1 = R(0);
  p.cph = R(it.cph0);
  p.sph = R(it.sph0);
  p.nsteps = 0;
}

// ---------------------------------------------------------------------------
// The walk
// ---------------------------------------------------------------------------

// Per-wave LDS ledger of the walk, words in the ledger's order (0 absorbed,
// 1 exited, 2 killed, 3 gained, 4 lost) plus the wave's loss count.  One
// problem: a single slot whose absorbed word stays in the lanes (summed
// across the wave at the end).  A batch: kBatchLedgerSlots slots
// direct-mapped by item, each with the item it sums for (-1: none).
template <bool BATCH> struct WaveLedger;
template <> struct WaveLedger<false> {
  static constexpr int slots = 1, first = 1;
  double v[5];   // exited, killed, gained, lost, loss count
  __device__ double &word(int, int q) { return v[q - 1]; }
  __device__ double &nloss() { return v[4]; }
  __device__ int item_of(int) const { return 0; }
  __device__ void set_item(int, int) {}
  __device__ void clear() {
    for (int q = 0; q < 5; ++q) v[q] = 0.0;
  }
};
template <> struct WaveLedger<true> {
  static constexpr int slots = kBatchLedgerSlots, first = 0;
  double v[slots][5];
  double nl;
  int it[slots];
  __device__ double &word(int sl, int q) { return v[sl][q]; }
  __device__ double &nloss() { return nl; }
  __device__ int item_of(int sl) const { return it[sl]; }
  __device__ void set_item(int sl, int item) { it[sl] = item; }
  __device__ void clear() {
    for (int sl = 0; sl < slots; ++sl) {
      for (int q = 0; q < 5; ++q) v[sl][q] = 0.0;
      it[sl] = -1;
    }
    nl = 0.0;
  }
};

// FP32 kernels fit 16 waves per CU in 128 VGPRs, so they take workgroups of
// up to 1024 lanes (one large LDS tally cache per CU); FP64 kernels need more
// registers per lane and stay at 256.
template <class R> struct MaxBlock { static constexpr int value = std::is_same<R, float>::value ? 1024 : 256; };
// waves per SIMD the register allocation must leave room for: 4 (128 VGPRs)
// for FP32, 2 (256) for FP64
template <class R> struct MinWaves { static constexpr int value = std::is_same<R, float>::value ? 4 : 2; };

// Boundary tally words written when a photon leaves through face ib, as
// offsets into its [EBR, EBI, DEBR, DEBI] slots.
template <int D, bool FD> struct ExitWords;
template <> struct ExitWords<3, false> { static constexpr int n = 1; static constexpr int off[1] = {0}; };
template <> struct ExitWords<3, true> { static constexpr int n = 2; static constexpr int off[2] = {0, 1}; };
template <> struct ExitWords<2, false> { static constexpr int n = 2; static constexpr int off[2] = {0, 2}; };
template <> struct ExitWords<2, true> { static constexpr int n = 4; static constexpr int off[4] = {0, 1, 2, 3}; };

// Tally path.  Elements are numbered along a space-filling curve, so a
// photon usually makes several visits in a row inside one sector of eight
// consecutive elements.  Each lane sums its visits into a private LDS
// buffer for its current sector; when the photon's visits move to another
// sector, the buffer is flushed as one run of 8W words: into the
// workgroup's sector cache if the sector holds a slot there, else to the
// replica with eight lanes per line, so each run costs one coalesced atomic
// request per 64-byte line.  (FP64 tallies add a cached sector's visits to
// its slot directly.)  Device
// atomics are served per line request (about 24 G requests/s on MI355X
// whatever the word type), which makes requests, not words, the budget.
// BATCH: items of different problems share the launch, so each lane carries
// its photon's element records, tally base and wave-number factor; without
// it every photon belongs to item 0 and these stay in scalar registers.
// SMALL (one item only): the item's P.mesh_elems element records are read
// from a copy in LDS and the sector cache has one slot per sector.
template <int D, class R, bool FD, bool FIXED, bool BATCH, bool SMALL>
__device__ __forceinline__ void walk_body(const WalkParams &P) {
  static_assert(!(SMALL && BATCH), "a small-mesh launch walks one item");
  constexpr int NF = D + 1;
  constexpr int W = FD ? 2 : 1;
  typedef ExitWords<D, FD> XW;
  typedef typename Vec4T<R>::type V4;
  const bool local = P.xcc_replicas != 0;
  const int rep = local ? (xcc_id() % P.nrep) : (int)(blockIdx.x % (unsigned)P.nrep);
  void *const etally = reinterpret_cast<char *>(P.etally) + (size_t)rep * P.ne * W * 8;
  void *const btally = reinterpret_cast<char *>(P.btally) + (size_t)rep * P.nb * 4 * 8;
  const int lane = (int)__lane_id();

  constexpr bool kF64 = std::is_same<R, double>::value;
  Photon<R> p;
  WalkRng rs{0u, 0u, 0u, 0u};   // the walk's random stream of the lane's photon
  uint32_t nrecheck = 0;        // FP32 exit-face choices decided in FP64
  bool alive = false;
  // Weight ledger: absorbed weight accumulates per lane; the rare
  // end-of-photon and roulette sums go to the wave's LDS ledger (WaveLedger).
  // A batch keeps one ledger per item (P.ledger[item * 5 + q]) and each wave
  // kLed slots direct-mapped by item: a wave's pool takes the slot of its
  // item when it moves to it, handing the sums of the item that held it to
  // that item's ledger, so the photons still alive from the wave's last few
  // pools (other items) keep adding in LDS; an event whose item lost its
  // slot goes straight to the item's ledger.  A lane's absorbed weight
  // leaves for its photon's item when the lane takes a photon of another
  // item, and at the end.
  double l_abs = 0;
  constexpr int kLed = WaveLedger<BATCH>::slots;
  __shared__ WaveLedger<BATCH> s_wl[MaxBlock<R>::value / 64];
  const int wid = threadIdx.x >> 6;
  WaveLedger<BATCH> &wl = s_wl[wid];
  if (lane == 0) wl.clear();
  constexpr int WS = kSectorElems * W;   // tally words per sector
  extern __shared__ __align__(16) unsigned char s_dyn[];
  const int S = P.lds_slots > 0 ? P.lds_slots : 0;
  const WalkLds lay(W, (int)sizeof(R), S, (int)blockDim.x);
  SectorCache<FIXED, WS, SMALL> cache{reinterpret_cast<int *>(s_dyn + lay.ctag),
                                      reinterpret_cast<unsigned long long *>(s_dyn + lay.cval), S};
  R *const sbuf = reinterpret_cast<R *>(s_dyn + lay.sbuf) + threadIdx.x;   // word j at [j * blockDim]
  const int bstride = (int)blockDim.x;
  for (int i = threadIdx.x; i < S; i += blockDim.x) cache.tag[i] = SMALL ? i : -1;
  for (int i = threadIdx.x; i < S * WS; i += blockDim.x) cache.val[i] = 0ull;
  // small mesh: the item's element records, copied after the lane buffers
  // and cache (16-byte words; the records are multiples of 16 bytes)
  const ElemLds<NF, R> EL = (ElemLds<NF, R>)(s_dyn + lay.total);
  if constexpr (SMALL) {
    static_assert(sizeof(Elem<NF, R>) % 16 == 0, "records are copied in 16-byte words");
    typedef typename Vec4T<float>::type W16;
    const W16 *src = reinterpret_cast<const W16 *>(P.items[0].elems);
    auto *dst = (__attribute__((address_space(3))) W16 *)(s_dyn + lay.total);
    const int n16 = P.mesh_elems * (int)(sizeof(Elem<NF, R>) / 16);
    for (int i = threadIdx.x; i < n16; i += blockDim.x) dst[i] = src[i];
  }
#pragma unroll
  for (int j = 0; j < WS; ++j) sbuf[j * bstride] = R(0);
  // cache epochs: waves that have finished, counted in LDS
  __shared__ int s_done;
  if (threadIdx.x == 0) s_done = 0;
  __syncthreads();
  uint64_t steps = 0;   // wave-uniform
  // The lane's current sector.  FP64 tallies: a sector with a cache slot
  // (cslot >= 0) takes each visit's deposits as LDS atomics straight into
  // the slot; one without sums them in the lane buffer, which goes to the
  // replica when the lane moves on.  Fixed point: every sector sums in the
  // lane buffer, so each rounding covers one photon's run of visits
  // whichever sectors hold slots.
  int sect = -1;
  int cslot = -1;
  const int epoch = S > 0 && !SMALL ? P.cache_epoch : 0;   // wave-uniform
  const int nwaves = (int)(blockDim.x >> 6);
  int pass = 0;            // wave-uniform: passes since the last epoch
  bool wave_done = false;  // wave-uniform: no photons left and none to come

  // Add to the replica: element word el*W + w (`etally`) or boundary word
  // ib*4 + j (`btally`).  With one replica per XCD (`local`), every add to a
  // replica comes from the CUs behind one L2, so a workgroup-scope atomic is
  // exact.  Kernel completion writes the L2s back before the reduction
  // kernel reads the replicas.
  auto replica_add = [&](void *base, int64_t i, unsigned long long bits) {
    if constexpr (FIXED) {
      unsigned long long *a = reinterpret_cast<unsigned long long *>(base) + i;
      if (local) __hip_atomic_fetch_add(a, bits, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
      else __hip_atomic_fetch_add(a, bits, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    } else {
      double *a = reinterpret_cast<double *>(base) + i;
      const double v = __longlong_as_double((long long)bits);
      if (local) __hip_atomic_fetch_add(a, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
      else __hip_atomic_fetch_add(a, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
  };
  // Fixed point: a word is v * 2^32 rounded to int64; a value at or beyond
  // 2^30 (or not finite) would leave the range in which the int64 sums stay
  // exact, so it raises the lane's range flag, reported after the launch.
  bool range_bad = false;
  auto to_bits = [&](double v) -> unsigned long long {
    if constexpr (FIXED) {
      const double s = v * P.fix_scale;
      range_bad = range_bad || !(fabs(s) < 0x1p62);
      return (unsigned long long)__double2ll_rn(s);
    } else {
      return (unsigned long long)__double_as_longlong(v);
    }
  };
  // Hands the buffers of the lanes with `go` (sector `sect`) to the cache
  // (fixed point) or to the replica, and clears them.  Converged: every lane
  // of the wave calls it together.  The wave moves 64 / (8W) flushing lanes'
  // buffers per instruction, a group of 8W lanes per buffer and one word per
  // lane, so a run costs one coalesced atomic request per 64-byte line and
  // the words go from the lane buffers to the replica without a copy.
  auto flush_sector = [&](bool go) {
    const bool need = go && (!FIXED ? true : !cache.add_from(sect, sbuf, bstride, P.fix_scale, range_bad));
    uint64_t m = __ballot(need);
    if (m == 0) return;
    if (P.direct_flush) {
      if (need) {
#pragma unroll 4
        for (int j = 0; j < WS; ++j) {
          const R w = sbuf[j * bstride];
          sbuf[j * bstride] = R(0);
          if (w != R(0)) replica_add(etally, sect * WS + j, to_bits((double)w));
        }
      }
      return;
    }
    constexpr int per = 64 / WS;   // buffers per wave instruction
    const int grp = lane / WS, jw = lane % WS;
    R *const wave_buf = sbuf - lane;   // word j of the wave's lane l at wave_buf[j * bstride + l]
    __builtin_amdgcn_wave_barrier();
    while (m) {
      int src = -1;   // the lane whose buffer this lane's group moves
#pragma unroll
      for (int k = 0; k < per; ++k) {
        const int b = m ? (int)__builtin_ctzll(m) : -1;
        src = grp == k ? b : src;
        m &= m - 1;
      }
      const int ss = __shfl(sect, src < 0 ? 0 : src);
      if (src >= 0) {
        R *const w = wave_buf + jw * bstride + src;
        const R v = *w;
        *w = R(0);
        if (v != R(0)) replica_add(etally, ss * WS + jw, to_bits((double)v));
      }
    }
    __builtin_amdgcn_wave_barrier();
  };

  // parked tally work of the last step
  int pe_el = -1;         // element whose visit ended, with its deposits
  R pe0 = R(0), pe1 = R(0);
  int pb_ib = -1;         // boundary face the photon left through, with its words
  R pb[XW::n];
#pragma unroll
  for (int j = 0; j < XW::n; ++j) pb[j] = R(0);

  // the lane's photon: its item's element records, tally base and (FD)
  // wave-number factor
  ElemPtr<NF, R> E = BATCH ? nullptr : (ElemPtr<NF, R>)(P.items[0].elems);
  int item = 0, tbase = BATCH ? 0 : P.items[0].tbase;
  R kf = BATCH ? R(0) : R(P.items[0].kfac);

  // whether the lane's photon's item holds its ledger slot (updated where
  // slots and photons change hands: at refills; a per-lane flag, which the
  // compiler keeps as a lane mask in scalar registers)
  bool own_slot = false;
  // ledger word q (0 absorbed, 1 exited, 2 killed, 3 gained, 4 lost) of
  // the lane's photon's item
  // (address-space-typed pointers: the LDS and global adds stay ds_ and
  // global_ atomics instead of merging into one flat atomic)
  auto led_add = [&](int q, double v) {
    if constexpr (BATCH) {
      if (!own_slot) {
        auto *g = (__attribute__((address_space(1))) double *)(P.ledger + (size_t)item * 5 + q);
        __hip_atomic_fetch_add(g, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        return;
      }
      auto *l = (__attribute__((address_space(3))) double *)&wl.word(item & (kLed - 1), q);
      __hip_atomic_fetch_add(l, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    } else {
      atomicAdd(&wl.word(0, q), v);
    }
  };
  // lane 0: moves slot sl's sums to its item's ledger and clears them
  auto flush_led_slot = [&](int sl) {
    const int it = wl.item_of(sl);
    if (it < 0) return;
    double *const L = P.ledger + (size_t)it * 5;
#pragma unroll
    for (int q = WaveLedger<BATCH>::first; q < 5; ++q) {
      const double v = wl.word(sl, q);
      wl.word(sl, q) = 0.0;
      if (v != 0.0) unsafeAtomicAdd(&L[q], v);
    }
  };

  // wave-private pool of virtual photon ids [pool_next, pool_end), all of
  // item pool_item; a chunk never straddles two items
  uint64_t pool_next = 0, pool_end = 0;
  int pool_item = 0;
  bool exhausted = false;
  const int refill_min = P.refill_min;
  // item owning virtual id v (items sorted by vstart; v < vtotal)
  auto find_item = [&](uint64_t v) {
    int lo = 0, hi = P.nitems;
    while (hi - lo > 1) {
      const int mid = (lo + hi) >> 1;
      if (P.items[mid].vstart <= v) lo = mid; else hi = mid;
    }
    return lo;
  };

  for (;;) {
    if (!exhausted) {
      const uint64_t dead = __ballot(!alive);
      const int ndead = __popcll(dead);
      if (ndead >= refill_min) {
        const uint64_t avail = pool_end - pool_next;
        uint64_t fresh = 0;
        int fresh_item = 0;
        if (avail < (uint64_t)ndead) {
          unsigned long long got = 0;
          if (lane == 0) got = atomicAdd(&P.counters[0], (unsigned long long)P.chunk);
          const uint32_t glo = __builtin_amdgcn_readfirstlane((uint32_t)got);
          const uint32_t ghi = __builtin_amdgcn_readfirstlane((uint32_t)(got >> 32));
          fresh = ((uint64_t)ghi << 32) | glo;
          if (BATCH && fresh < P.vtotal) fresh_item = find_item(fresh);
        }
        const int rank = (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(dead >> 32),
                                                        __builtin_amdgcn_mbcnt_lo((uint32_t)dead, 0u));
        uint64_t pid = 0;
        int pit = pool_item;
        if (!alive && (uint64_t)rank >= avail) {
          pid = fresh + (rank - avail);
          pit = fresh_item;
        } else if (!alive) {
          pid = pool_next + rank;
        }
        if constexpr (!BATCH) pit = 0;
        if (avail < (uint64_t)ndead) {
          pool_next = fresh + (ndead - avail);
          pool_end = fresh + (uint64_t)P.chunk;
          pool_item = fresh_item;
          if constexpr (BATCH) {
            // the pool's item takes its ledger slot; the photons of the item
            // that held it lose theirs
            const int sl = fresh_item & (kLed - 1);
            const int held = fresh < P.vtotal ? wl.item_of(sl) : fresh_item;
            if (held != fresh_item) {
              if (lane == 0) {
                flush_led_slot(sl);
                wl.set_item(sl, fresh_item);
              }
              own_slot = own_slot && item != held;
            }
          }
        } else {
          pool_next += ndead;
        }
        const bool invalid = !alive && pid >= P.vtotal;
        if (__ballot(invalid)) exhausted = true;
        const bool was_alive = alive;
        if (!alive && !invalid) {
          const ItemRec &it = P.items[pit];
          const uint64_t k = pid - it.vstart;
          if (k < it.count) {   // ids past the item's count pad its last chunk
            const uint64_t sid = it.id0 + k;
            rs.seed((uint32_t)sid, (uint32_t)(sid >> 32), it);
            if constexpr (BATCH) {
              if (pit != item) {   // the absorbed weight so far is the old item's
                if (l_abs != 0.0) led_add(0, l_abs);
                l_abs = 0.0;
              }
              item = pit;
              E = (ElemPtr<NF, R>)(it.elems);
              tbase = it.tbase;
              kf = R(it.kfac);
            }
            create_photon<D, R>(p, rs, it, P);
            p.prop = w_div(w_neglog(Unif<R>::open(rs.next(P))), SMALL ? EL[p.el].opt.y : E[p.el].opt.y);
            alive = true;
          }
        }
        if constexpr (BATCH) {
          // a new photon adds in LDS when its item holds the slot
          if (alive && !was_alive) own_slot = wl.item_of(item & (kLed - 1)) == item;
        }
      }
    }
    const uint64_t live = __ballot(alive);
    if (live == 0) {
      // a wave whose refill drew only the padding of an item's last chunk
      // is not finished while the dispenser still has ids: it refills again
      // at the next pass.  With cache epochs a finished wave keeps passing
      // through the loop (doing nothing) until the whole workgroup leaves it
      // together at an epoch.
      if (epoch == 0) {
        if (exhausted) break;
        continue;
      }
      if (exhausted && !wave_done) {
        wave_done = true;
        if (lane == 0) atomicAdd(&s_done, 1);
      }
      // a finished wave goes straight to the next epoch and waits in its
      // barrier instead of taking issue slots with empty passes
      if (wave_done) pass = epoch - 1;
    }
    steps += __popcll(live);

    // ---- one element step (returns early where the reference's loop does) ----
    auto step = [&]() {
      // the element record is fetched per step rather than carried across
      // iterations: it keeps the loop state small (more waves per SIMD) and a
      // repeated fetch of the same record hits in the caches
      V4 pl[NF];
      V4 opt;
      i4 nb;
      typename Vec2T<R>::type nbo[NF];
      if constexpr (SMALL) load_elem<NF, R>(EL, p.el, pl, opt, nb, nbo);
      else load_elem<NF, R>(E, p.el, pl, opt, nb, nbo);
      ++p.nsteps;
      // one element visit (entry to exit, scatterings inside included) adds
      // its deposits once
      auto end_visit = [&]() {
        if (FD || p.acc0 != R(0)) {
          pe_el = p.el;
          pe0 = p.acc0;
          pe1 = p.acc1;
        }
        p.acc0 = R(0);
        p.acc1 = R(0);
      };
      auto lose = [&]() {
        end_visit();
        led_add(4, (double)p.w);
        atomicAdd(&wl.nloss(), 1.0);
        atomicAdd(&P.item_loss[item], 1ull);
        alive = false;
      };
      // exit face: smallest positive plane distance over faces moving outward,
      // the current face excluded (MC3D.hpp:1020-1056); distances num/c are
      // compared as products, without dividing
      R bn = R(1), bc = R(0);   // best distance bn/bc
      R sn = R(1), sc = R(0);   // second best (FP32 only, for the tie test)
      int bf = -1, code = kNbNone;
      R mus_n = R(0), n_n = R(0);
      bool out_cur = false;
#pragma unroll
      for (int f = 0; f < NF; ++f) {
        R c = pl[f].x * p.d[0] + pl[f].y * p.d[1];
        R q = pl[f].x * p.x[0] + pl[f].y * p.x[1];
        if (D == 3) {
          c += pl[f].z * p.d[2];
          q += pl[f].z * p.x[2];
        }
        const R num = fmax(pl[f].w - q, R(0));
        const bool cur = f == p.face;
        out_cur = cur ? c > Tol<R>::out : out_cur;
        // selects, not branches: the faces are tested in every lane
        const bool valid = !cur && c > R(0);
        const bool best = valid && (f == 0 || num * bc < bn * c);
        if constexpr (!kF64) {
          const bool second = valid && !best && num * sc < sn * c;
          sn = best ? bn : (second ? num : sn);
          sc = best ? bc : (second ? c : sc);
        }
        bn = best ? num : bn;
        bc = best ? c : bc;
        bf = best ? f : bf;
      }
      // the record's second half (optical and neighbour words) is used only
      // past this branch; naming it here keeps its loads issued together
      // with the planes' instead of after the face test, and passing the
      // test's result through keeps the wait for them after the test (the
      // test needs the planes only, whose loads are issued first and return
      // first), so the second half arrives while the test runs
      asm volatile("" : "+v"(bf) : "v"(opt.x), "v"(opt.y), "v"(opt.z), "v"(opt.w), "v"(nb.x), "v"(nb.y), "v"(nb.z));
      if constexpr (D == 3) asm volatile("" : "+v"(bf) : "v"(nb.w));
#pragma unroll
      for (int f = 0; f < NF; ++f) asm volatile("" : "+v"(bf) : "v"(nbo[f].x), "v"(nbo[f].y));
      // the chosen face's neighbour words
#pragma unroll
      for (int f = 0; f < NF; ++f) {
        code = bf == f ? nb[f] : code;
        mus_n = bf == f ? nbo[f].x : mus_n;
        n_n = bf == f ? nbo[f].y : n_n;
      }
      if (out_cur || bf < 0 || (uint32_t)p.nsteps > P.step_cap) {
        // no face ahead: the reference's WhichFace miss (MC3D.hpp:1464-1468)
        lose();
        return;
      }
      if constexpr (!kF64) {
        // another outward face within 2^-16 of the best distance: decide in
        // FP64 from the same record and packet (the record is re-read from
        // the cache rather than kept live through the loop)
        if (sc > R(0) && sn * bc <= bn * R(1.0000152587890625) * sc) {
          ++nrecheck;
          V4 pl2[NF];
          V4 opt2;
          i4 nb2;
          typename Vec2T<R>::type nbo2[NF];
          if constexpr (SMALL) load_elem<NF, R>(EL, p.el, pl2, opt2, nb2, nbo2);
          else load_elem<NF, R>(E, p.el, pl2, opt2, nb2, nbo2);
          double bt = 0.0;
          int g = -1;
#pragma unroll
          for (int f = 0; f < NF; ++f) {
            double c = (double)pl2[f].x * p.d[0] + (double)pl2[f].y * p.d[1];
            double q = (double)pl2[f].x * p.x[0] + (double)pl2[f].y * p.x[1];
            if (D == 3) {
              c += (double)pl2[f].z * p.d[2];
              q += (double)pl2[f].z * p.x[2];
            }
            const double t = fmax((double)pl2[f].w - q, 0.0) / c;
            const bool take = f != p.face && c > 0.0 && (g < 0 || t < bt);
            bt = take ? t : bt;
            g = take ? f : g;
          }
#pragma unroll
          for (int f = 0; f < NF; ++f) {
            if (f != g || g == bf) continue;
            // the face's FP32 distance, as the first pass computed it
            R c = pl2[f].x * p.d[0] + pl2[f].y * p.d[1];
            R q = pl2[f].x * p.x[0] + pl2[f].y * p.x[1];
            if (D == 3) {
              c += pl2[f].z * p.d[2];
              q += pl2[f].z * p.x[2];
            }
            bn = fmax(pl2[f].w - q, R(0));
            bc = c;
            bf = f;
            code = nb2[f];
            mus_n = nbo2[f].x;
            n_n = nbo2[f].y;
          }
        }
      }
      // outward unit normal of the exit face, for the rarer boundary and
      // refraction branches (re-read rather than carried through the step)
      auto exit_normal = [&](R (&nrm)[3]) {
        const V4 q = SMALL ? (EL + p.el)->pl[bf] : (E + p.el)->pl[bf];
        nrm[0] = q.x;
        nrm[1] = q.y;
        nrm[2] = D == 3 ? q.z : R(0);
      };
      const R dist = w_div(bn, bc);
      const R ds = p.prop < dist ? p.prop : dist;   // fmin(prop, dist), NaN prop -> dist
#pragma unroll
      for (int k = 0; k < D; ++k) p.x[k] += p.d[k] * ds;

      const R mua = opt.x, mus = opt.y, nidx = opt.z;
      // element tally and attenuation (MC3D.hpp:1479-1526)
      // FP64: the reference's att = exp(-mua ds) and (1 - att) w.  FP32: the
      // deposit -w expm1(-mua ds) is subtracted from w exactly, so the weight
      // ledger closes to rounding of its FP64 sums.
      const R em = kF64 ? R(0) : w_expm1_neg(-mua * ds);
      const R att = kF64 ? m_exp(-mua * ds) : R(1) + em;
      const R dabs = kF64 ? p.w - p.w * att : -p.w * em;
      if constexpr (!FD) {
        const R dep = mua > R(0) ? (kF64 ? (R(1) - att) * p.w : dabs) : p.w * ds;
        p.acc0 += dep;
      } else {
        // w e^{-i phi} (1 - e^{-(mua + i k) ds}), MC3D.hpp:1519-1522, with the
        // phase advance b = k ds taken from half-angle terms
        const R b = kf * nidx * ds;
        R sh, ch;
        w_sincos_small(R(0.5) * b, &sh, &ch);
        const R sb = R(2) * sh * ch, ver = R(2) * sh * sh, cb = R(1) - ver;
        const R c1 = p.cph * cb - p.sph * sb, s1 = p.sph * cb + p.cph * sb;
        R er, ei;
        if constexpr (kF64) {
          er = p.w * (p.cph - c1 * att);
          ei = p.w * (-p.sph + s1 * att);
        } else {
          const R qq = -em + att * ver;   // 1 - att cos(b) without cancellation
          er = p.w * (p.cph * qq + att * p.sph * sb);
          ei = p.w * (-p.sph * qq + att * p.cph * sb);
        }
        p.acc0 += er;
        p.acc1 += ei;
        if constexpr (kF64) {
          p.cph = c1;
          p.sph = s1;
        } else {
          const R rn = R(1.5) - R(0.5) * (c1 * c1 + s1 * s1);   // keep |e^{i phi}| = 1
          p.cph = c1 * rn;
          p.sph = s1 * rn;
        }
      }
      l_abs += (double)dabs;
      p.w = kF64 ? p.w * att : p.w - dabs;
      p.prop -= ds;

      if (p.prop <= R(0)) {
        // scattering event (MC3D.hpp:1634-1636), then a new free path
        if (mus > R(0)) {
          if constexpr (D == 3) scatter3<R>(p, opt.w, rs, P);
          else scatter2<R>(p, opt.w, rs, P);
        }
        p.prop = w_div(w_neglog(Unif<R>::open(rs.next(P))), mus);
        return;
      }

      if (code < 0) {
        if (code == kNbNone) {
          lose();
          return;
        }
        // boundary (MC3D.hpp:1536-1567)
        const int ib = -1 - code;
        const auto &bref = global_ptr(P.items[item].bnd)[ib];
        const struct { double bcn; int32_t kind; } b = {bref.bcn, bref.kind};
        R nrm[3];
        exit_normal(nrm);
        if (b.kind == kBndMirror) {
          R cd = p.d[0] * nrm[0] + p.d[1] * nrm[1];
          if (D == 3) cd += p.d[2] * nrm[2];
#pragma unroll
          for (int k = 0; k < D; ++k) p.d[k] -= R(2) * cd * nrm[k];
          p.face = bf;
          return;
        }
        if (b.kind == kBndFresnel && fresnel<D, R>(p, nrm, nidx, R(b.bcn), false, rs, P)) {
          p.face = bf;
          return;
        }
        // exitance (MC3D.hpp:1556-1564); 2D adds it weighted by
        // 1/(dir . outward normal) (MC2D.hpp:1410-1437)
        pb_ib = P.items[item].bbase + ib;
        pb[0] = FD ? p.w * p.cph : p.w;
        if constexpr (FD) pb[1] = -p.w * p.sph;
        if constexpr (D == 2) {
          const R dn = p.d[0] * nrm[0] + p.d[1] * nrm[1];
          if constexpr (!FD) {
            pb[1] = p.w / dn;
          } else {
            pb[2] = p.w * p.cph / dn;
            pb[3] = -p.w * p.sph / dn;
          }
        }
        end_visit();
        led_add(1, (double)p.w);
        alive = false;
        return;
      }

      // interior face: enter the neighbour (MC3D.hpp:1570-1631)
      const int next = code >> 3;
      const bool back = (code >> 2) & 1;
      const int entry = code & 3;
      const R mus_c = mus, n_c = nidx;
      if (mus_c <= R(0) && mus_n > R(0)) {
        // vacuum -> scattering medium: fresh free path
        p.prop = w_div(w_neglog(Unif<R>::open(rs.next(P))), mus_n);
      }
      // the roulette constants come as FP32 copies for the FP32 walk, so
      // they stay in scalar registers
      const R weight0 = kF64 ? R(P.weight0) : R(P.weight0_f);
      const R chance = kF64 ? R(P.chance) : R(P.chance_f);
      if (p.w < weight0) {
        // Russian roulette at the crossing (MC3D.hpp:1597-1604)
        if (Unif<R>::closed(rs.next(P)) > chance) {
          end_visit();
          led_add(2, (double)p.w);
          alive = false;
          return;
        }
        const R wn = w_div(p.w, chance);
        led_add(3, (double)wn - (double)p.w);
        p.w = wn;
      }
      if (n_c != n_n) {
        R nrm[3];
        exit_normal(nrm);
        if (fresnel<D, R>(p, nrm, n_c, n_n, true, rs, P)) {
          p.face = bf;
          return;
        }
      }
      if (!(mus_c <= R(0) && mus_n > R(0))) p.prop *= w_div(mus_c, mus_n);
      if (!back) {
        // the neighbour does not list this element: back-lookup failure
        lose();
        return;
      }
      end_visit();
      p.el = next;
      p.face = entry;
    };
    const bool stepped = alive;
    if (alive) step();

    // ---- commit the parked tally work (converged) ----
    {
      const int ew = tbase + pe_el;   // arena element slot (>= 0 when pe_el is)
      const int es = pe_el >= 0 ? (int)((uint32_t)ew / kSectorElems) : -1;
      const bool chg = es >= 0 && es != sect;
      flush_sector(chg && sect >= 0 && cslot < 0);
      if constexpr (!FIXED) {
        if (chg) cslot = cache.claim(es);
      }
      if (es >= 0) {
        const int k = (int)((uint32_t)ew % kSectorElems);
        if (cslot >= 0) {
          double *cv = reinterpret_cast<double *>(cache.val) + (size_t)cslot * WS + k * W;
          atomicAdd(cv, (double)pe0);
          if constexpr (FD) atomicAdd(cv + 1, (double)pe1);
        } else {
          sbuf[(k * W) * bstride] += pe0;
          if constexpr (FD) sbuf[(k * W + 1) * bstride] += pe1;
        }
        sect = es;
      }
      if constexpr (FIXED) {
        // a buffer sum must hold one photon's deposits only, so fixed-point
        // results do not depend on which photons shared a lane
        const bool ended = stepped && !alive && sect >= 0;
        flush_sector(ended);
        if (ended) sect = -1;
      }
      if (pb_ib >= 0) {
#pragma unroll
        for (int j = 0; j < XW::n; ++j)
          replica_add(btally, pb_ib * 4 + XW::off[j], to_bits((double)pb[j]));
      }
      pe_el = -1;
      pb_ib = -1;
    }

    // ---- cache epoch: every `epoch` passes the workgroup moves the slots'
    // sums to its replica and frees the slots, so the sectors of the items
    // walked now (a batch's later items) can claim them.  Every wave makes
    // the same passes between two epochs (a finished wave skips to the next
    // epoch) and leaves the loop only here, all together, once every wave
    // has finished (the count is read between barriers, where no wave
    // writes it). ----
    if (epoch > 0 && ++pass == epoch) {
      pass = 0;
      __syncthreads();
      const bool all_done = *(volatile int *)&s_done == nwaves;
      if (!all_done) {
        for (int i = threadIdx.x; i < S * WS; i += blockDim.x) {
          const int sc = cache.tag[i / WS];
          const unsigned long long bits = cache.val[i];
          if (sc >= 0 && bits) replica_add(etally, (int64_t)sc * WS + i % WS, bits);
          cache.val[i] = 0ull;
        }
      }
      __syncthreads();
      if (all_done) break;
      for (int i = threadIdx.x; i < S; i += blockDim.x) cache.tag[i] = -1;
      __syncthreads();
      if (cslot >= 0) {   // the lane's sector sums went with its slot
        cslot = -1;
        sect = -1;
      }
    }
  }
  flush_sector(sect >= 0 && cslot < 0);

  // ---- flush the counters: one set of global atomics per wave ----
  if constexpr (BATCH) {
    if (l_abs != 0.0) led_add(0, l_abs);
    l_abs = 0.0;
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    if constexpr (!BATCH) l_abs += __shfl_xor(l_abs, off);
    nrecheck += __shfl_xor(nrecheck, off);
  }
  __syncthreads();
  // ---- flush the sector cache into this workgroup's replica: one sector
  // per WS consecutive lanes ----
  for (int i = threadIdx.x; i < S * WS; i += blockDim.x) {
    const int sc = cache.tag[i / WS];
    const unsigned long long bits = cache.val[i];
    if (sc >= 0 && bits) replica_add(etally, (int64_t)sc * WS + i % WS, bits);
  }
  if (lane == 0) {
    if constexpr (!BATCH) unsafeAtomicAdd(&P.ledger[0], l_abs);
#pragma unroll
    for (int sl = 0; sl < kLed; ++sl) flush_led_slot(sl);
    atomicAdd(&P.counters[1], (unsigned long long)steps);
    atomicAdd(&P.counters[2], (unsigned long long)wl.nloss());
    if (nrecheck) atomicAdd(&P.counters[3], (unsigned long long)nrecheck);
  }
  if constexpr (FIXED) {
    if (__ballot(range_bad) != 0 && lane == 0) atomicOr(&P.counters[4], 1ull);
  }
}

template <int D, class R, bool FD, bool FIXED, bool BATCH>
__global__ void __launch_bounds__(MaxBlock<R>::value)
    __attribute__((amdgpu_waves_per_eu(MinWaves<R>::value))) walk_kernel(const WalkParams P) {
  walk_body<D, R, FD, FIXED, BATCH, false>(P);
}

template <int D, class R, bool FD, bool FIXED>
__global__ void __launch_bounds__(MaxBlock<R>::value)
    __attribute__((amdgpu_waves_per_eu(MinWaves<R>::value))) walk_small_kernel(const WalkParams P) {
  walk_body<D, R, FD, FIXED, false, true>(P);
}

// ---------------------------------------------------------------------------
// Helper kernels
// ---------------------------------------------------------------------------

// p[r * stride + i] = v for r < nrep, i < n
__global__ void fill_u64_kernel(unsigned long long *p, size_t n, int nrep, size_t stride, unsigned long long v) {
  const size_t total = n * (size_t)nrep;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < total; i += (size_t)gridDim.x * blockDim.x)
    p[(i / n) * stride + i % n] = v;
}

// Sum the tally replicas in a fixed order: out[i] = sum_r rep[r * stride + i].
template <bool FIXED>
__global__ void reduce_replicas_kernel(const void *rep, double *out, size_t n, int nrep, size_t stride,
                                       double inv_scale) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    if (FIXED) {
      const long long *r = reinterpret_cast<const long long *>(rep);
      long long s = 0;
      for (int k = 0; k < nrep; ++k) s += r[(size_t)k * stride + i];
      out[i] = (double)s * inv_scale;
    } else {
      const double *r = reinterpret_cast<const double *>(rep);
      double s = 0.0;
      for (int k = 0; k < nrep; ++k) s += r[(size_t)k * stride + i];
      out[i] = s;
    }
  }
}

} // namespace vmc

// ---------------------------------------------------------------------------
// Scratch of the allocation-free entry point (vmc_problem_run_noalloc): the
// tally replicas, counters, ledger and item record of one run live in this
// static device array instead of library allocations.  Each GPU the library
// runs on holds its own copy, at its own address, until the process ends;
// every call on that GPU uses the same copy.
// ---------------------------------------------------------------------------

namespace vmc {

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

}  // namespace vmc

#include <mutex>

// ---------------------------------------------------------------------------
// Host-side launchers (called from the launch engine, vmc_engine.cpp)
// ---------------------------------------------------------------------------

namespace vmc {

// Device address of `scratch` on `device`, which must be the current device.
// The name `scratch` is only a host-side stand-in: its address is looked up
// once per device with hipGetSymbolAddress, and only that address is handed
// to kernels and copies.
hipError_t scratch_address(int device, void **ptr) {
  static std::mutex mu;
  static void *addr[kMaxDevices] = {};
  if (device < 0 || device >= kMaxDevices || !ptr) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lk(mu);
  if (!addr[device]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    addr[device] = p;
  }
  *ptr = addr[device];
  return hipSuccess;
}

typedef void (*WalkFn)(const WalkParams);

template <int D, class R>
static WalkFn walk_fn_dr(bool fd, bool fixed, bool batch, bool small) {
  if (small)
    return fd ? (fixed ? walk_small_kernel<D, R, true, true> : walk_small_kernel<D, R, true, false>)
              : (fixed ? walk_small_kernel<D, R, false, true> : walk_small_kernel<D, R, false, false>);
  if (batch)
    return fd ? (fixed ? walk_kernel<D, R, true, true, true> : walk_kernel<D, R, true, false, true>)
              : (fixed ? walk_kernel<D, R, false, true, true> : walk_kernel<D, R, false, false, true>);
  return fd ? (fixed ? walk_kernel<D, R, true, true, false> : walk_kernel<D, R, true, false, false>)
            : (fixed ? walk_kernel<D, R, false, true, false> : walk_kernel<D, R, false, false, false>);
}

static WalkFn walk_fn(int dim, bool f64, bool fd, bool fixed, bool batch, bool small) {
  if (dim == 3)
    return f64 ? walk_fn_dr<3, double>(fd, fixed, batch, small) : walk_fn_dr<3, float>(fd, fixed, batch, small);
  return f64 ? walk_fn_dr<2, double>(fd, fixed, batch, small) : walk_fn_dr<2, float>(fd, fixed, batch, small);
}

hipError_t walk_occupancy(int dim, bool f64, bool fd, bool fixed, bool batch, bool small, int block,
                          size_t lds_bytes, int *blocks_per_cu) {
  if (batch && small) return hipErrorInvalidValue;
  return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, walk_fn(dim, f64, fd, fixed, batch, small),
                                                      block, lds_bytes);
}

hipError_t launch_walk(int dim, bool f64, bool fd, bool fixed, bool batch, bool small, int grid, int block,
                       size_t lds_bytes, hipStream_t s, const WalkParams &P) {
  if (batch && small) return hipErrorInvalidValue;
  hipLaunchKernelGGL(walk_fn(dim, f64, fd, fixed, batch, small), dim3(grid), dim3(block), lds_bytes, s, P);
  return hipGetLastError();
}

hipError_t launch_fill_ranges(unsigned long long *p, size_t n, int nrep, size_t stride, unsigned long long v,
                              hipStream_t s) {
  if (n == 0 || nrep <= 0) return hipSuccess;
  const int block = 256;
  const size_t want = (n * (size_t)nrep + block - 1) / block;
  const int grid = (int)(want < 4096 ? want : 4096);
  hipLaunchKernelGGL(fill_u64_kernel, dim3(grid), dim3(block), 0, s, p, n, nrep, stride, v);
  return hipGetLastError();
}

hipError_t launch_fill_u64(unsigned long long *p, size_t n, unsigned long long v, hipStream_t s) {
  return launch_fill_ranges(p, n, 1, n, v, s);
}

hipError_t launch_reduce(bool fixed, const void *rep, double *out, size_t n, int nrep, size_t stride,
                         double inv_scale, hipStream_t s) {
  if (n == 0) return hipSuccess;
  const int block = 256;
  const size_t want = (n + block - 1) / block;
  const int grid = (int)(want < 8192 ? want : 8192);
  if (fixed)
    hipLaunchKernelGGL(reduce_replicas_kernel<true>, dim3(grid), dim3(block), 0, s, rep, out, n, nrep, stride,
                       inv_scale);
  else
    hipLaunchKernelGGL(reduce_replicas_kernel<false>, dim3(grid), dim3(block), 0, s, rep, out, n, nrep, stride,
                       inv_scale);
  return hipGetLastError();
}

} // namespace vmc
