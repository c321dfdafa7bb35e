// Internal host/device types of the transport engine.
//
// Element records are array-of-structs so one photon step touches one or two
// cache lines: the face planes, the per-face neighbour codes and the optical
// parameters of the element a photon is in are loaded together when it
// enters the element.  Coordinates are shifted by the mesh centre so FP32
// positions keep the most bits.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace vmc {

typedef float f4 __attribute__((ext_vector_type(4)));
typedef double d4 __attribute__((ext_vector_type(4)));
typedef int i4 __attribute__((ext_vector_type(4)));

typedef float f2 __attribute__((ext_vector_type(2)));
typedef double d2 __attribute__((ext_vector_type(2)));

template <class R> struct Vec4T;
template <> struct Vec4T<float> { typedef f4 type; };
template <> struct Vec4T<double> { typedef d4 type; };
template <class R> struct Vec2T;
template <> struct Vec2T<float> { typedef f2 type; };
template <> struct Vec2T<double> { typedef d2 type; };

// Neighbour code of one element face (reference HN, cpp/3d/MC3D.hpp:1025,
// 1616-1628, folded with the entry-face back-lookup):
//   interior: (next << 3) | (back << 2) | entry   back = 1 when next lists
//             this element among its own neighbours, entry = that face
//   boundary: -1 - ib
//   kNbNone:  face matched nothing (the reference leaves INT_FAST64_MAX)
constexpr int32_t kNbNone = INT32_MIN;
constexpr int64_t kMaxElements = (int64_t)1 << 28;

// Static device scratch of the allocation-free run entry point, per GPU
constexpr size_t kScratchBytes = (size_t)64 << 20;
constexpr int kMaxDevices = 64;

// Face f: outward unit normal n and offset d with n.x <= d inside.
// opt = (mua, mus, n, g).  nbo[f] = (mus, n) of the element across face f
// (zero on the boundary): a crossing needs them (free-path rescale and the
// Fresnel test, MC3D.hpp:1570-1613) and would otherwise gather them from the
// neighbour's record.  FP32 tetrahedra fill exactly one 128-byte line; the
// FP64 record is two lines.
template <int NF, class R> struct alignas(64) Elem {
  typename Vec4T<R>::type pl[NF];
  typename Vec4T<R>::type opt;
  i4 nb;
  typename Vec2T<R>::type nbo[NF];
};
static_assert(sizeof(Elem<4, float>) == 128, "FP32 tetrahedron record is one cache line");
static_assert(sizeof(Elem<4, double>) == 256, "FP64 tetrahedron record is two cache lines");

// Boundary element kinds for the walk (BCType, MC3D.hpp:1536-1567)
enum BndKind : int32_t {
  kBndAbsorb = 0,   // leave the mesh (no Fresnel: BCn <= 0)
  kBndFresnel = 1,  // Fresnel test against BCn, then leave
  kBndMirror = 2,   // m, L, I, C: specular reflection
};
struct alignas(16) BndRec {
  double bcn;
  int32_t kind;
  int32_t pad;
};

// How a light source launches photons (MC3D.hpp:1109-1287, MC2D.hpp:1004-1125)
enum SrcKind : int32_t {
  kSrcNormal = 0,    // along the inward normal
  kSrcGiven = 1,     // along BCLNormal as given
  kSrcPencil = 2,    // inward normal from the position in BCLNormal
  kSrcIso3 = 3,      // uniform on the inward hemisphere (rejection)
  kSrcCos3 = 4,      // phi = asin(2U-1) about frame (a0, a1, a2)
  kSrcIso2 = 5,      // phi0 + pi(U - 1/2)
  kSrcCos2 = 6,      // phi0 + asin(2U-1)
  kSrcCos2Rej = 7,   // as kSrcCos2, redrawn until it points inwards
  kSrcGauss2 = 8,    // phi0 + truncated-normal angle by 16-step bisection
  kSrcGauss2Rej = 9, // as kSrcGauss2, redrawn until it points inwards
};

struct alignas(16) SrcRec {
  double v0[3], v1[3], v2[3];  // triangle (2D: segment v0-v1), shifted
  double nin[3];               // inward unit normal (centroid test)
  double a1[3], a2[3];         // cosinic frame second and third axes
  double dir[3];               // given direction, or pencil position (shifted)
  double phi0, gsig, gh;       // 2D: base angle; Gaussian sigma and tail mass
  int32_t mother, face, kind, ib;
};

// Tally sectors: eight consecutive elements (device order), i.e. 8*W
// tally words, one 64-byte line of the replica in continuous-wave mode.
constexpr int kSectorElems = 8;

// Ledger slots per wave of a batch walk: the items of the photons a wave
// holds at once, direct-mapped (item & (slots - 1)).
constexpr int kBatchLedgerSlots = 4;
static_assert((kBatchLedgerSlots & (kBatchLedgerSlots - 1)) == 0, "ledger slots: a power of 2");

// Dynamic LDS of one walk workgroup, in bytes, in this order (each part
// 16-byte aligned):
//   sector cache values  [S][8W]  u64 (FP64 sums or fixed point)
//   lane sector buffers  [8W][block]  R
//   sector cache tags    [S]  int32
struct WalkLds {
  size_t cval, sbuf, ctag, total;
  __host__ __device__ WalkLds(int W, int rbytes, int S, int block) {
    const size_t ws = (size_t)kSectorElems * W;
    auto up = [](size_t x) { return (x + 15) & ~(size_t)15; };
    cval = 0;
    sbuf = up(cval + (size_t)S * ws * 8);
    ctag = up(sbuf + ws * (size_t)block * rbytes);
    total = up(ctag + (size_t)S * 4);
  }
};

// Tally layout: nrep replicas of [Ne rounded up to a sector][cw ? 1 : 2]
// element and [Nb][4] boundary slots (EBR, EBI, DEBR, DEBI).
// One instance of a launch: a resident problem (its element records, light
// sources and boundary), a photon range, a random-stream key and a place in
// the tally arena.  A launch walks the photons of all its items.
struct alignas(16) ItemRec {
  uint64_t vstart;          // first virtual photon id (a multiple of the chunk)
  uint64_t count;           // photons of this item in the launch
  uint64_t id0;             // stream id of its first photon
  const void *elems;        // element records
  const double *cdf;        // [nsrc] light-source CDF
  const SrcRec *src;        // [nsrc]
  const BndRec *bnd;        // [Nb]
  int32_t tbase;            // first element slot in the tally arena (whole sectors)
  int32_t bbase;            // first boundary slot in the tally arena
  int32_t nsrc;
  uint32_t key0, key1, tag; // stream key (seed) and tag
  double kfac;              // omega / c0: k = kfac * n
  double cph0, sph0;        // cos/sin of the initial phase
};

// Tally arena: nrep replicas of [ne][W] element words and [nb][4] boundary
// words (EBR, EBI, DEBR, DEBI).
struct WalkParams {
  const ItemRec *items;
  int32_t nitems;
  int32_t mesh_elems;           // small-mesh kernel: element records staged in LDS
  uint64_t vtotal;              // virtual photon ids [0, vtotal) of this launch
  void *etally;
  void *btally;
  double *ledger;               // [5] absorbed, exited, killed, gained, lost;
                                // a batch launch: [nitems][5], one per item
  unsigned long long *counters; // [8] photon dispenser, steps, loss, rechecks, fixed-point range flag
  unsigned long long *item_loss; // [nitems] photons lost per item (MC3D::loss)
  int64_t ne, nb;               // arena element slots (whole sectors) and boundary slots
  int32_t nrep;
  uint32_t step_cap;
  int32_t refill_min;
  int32_t chunk;
  int32_t xcc_replicas;         // 1: replica = XCD id, workgroup-scope (L2) atomics
  int32_t lds_slots;            // LDS sector-cache slots per workgroup (0 = none)
  int32_t direct_flush;         // 1: each lane adds its own sector words (no coalescing)
  int32_t cache_epoch;          // loop passes between sector-cache hand-overs (0: never)
  double weight0, chance;
  float weight0_f, chance_f;    // the same in FP32, for the FP32 walk
  double fix_scale;             // fixed-point scale (2^32)
};

} // namespace vmc
