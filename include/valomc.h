/*
 * valomc.h -- C ABI of the MI355X photon-transport engine.
 *
 * Drop-in boundary for ValoMC's MC3D/MC2D::MonteCarlo.  Inputs are the
 * arrays the reference engine holds as public members (cpp/3d/MC3D.hpp:108-172,
 * cpp/2d/MC2D.hpp:113-180) in the reference's own layout: column-major
 * (Fortran order, cpp/3d/Array.hpp:58-66), 0-based int64 node indices,
 * boundary neighbours encoded as -1-ib.  Outputs are the normalised
 * ER/EI/EBR/EBI/DEBR/DEBI the reference leaves in those members after
 * MonteCarlo() (MC3D.hpp:1793-1818, MC2D.hpp:1673-1703).
 *
 * No exceptions cross this boundary.  Status codes:
 *   0            success
 *   1 .. 16      1 + the reference's mcerror code (cpp/Errors.hpp:4-21)
 *   17           a light source whose direction the reference leaves
 *                undefined ('L', 'I' or 'p' with direction type 'n',
 *                MC3D.hpp:1112-1192 has no branch for them)
 *   18           invalid argument (NULL pointer, bad size or option)
 *   100 ..       device errors (see vmc_status_string / vmc_last_error);
 *                101 is also returned when a host allocation fails
 *
 * Threading: one problem handle per host thread; calls on one handle are
 * serialised by the caller.  Device memory and streams are owned by the
 * library and released by vmc_problem_destroy.
 */
#ifndef VALOMC_H
#define VALOMC_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define VMC_OK 0
#define VMC_ERR_UNDEFINED_DIRECTION 17
#define VMC_ERR_INVALID_ARGUMENT 18
#define VMC_ERR_HIP 100
#define VMC_ERR_OUT_OF_MEMORY 101
#define VMC_ERR_NO_DEVICE 102
#define VMC_ERR_ABORTED 103
#define VMC_ERR_TALLY_RANGE 104 /* VMC_TALLY_FIXED64: a deposit left the exact range */

/* transport arithmetic */
#define VMC_FP64 0 /* plain double-precision transport (the GPU FP64 oracle)   */
#define VMC_FP32 1 /* single-precision transport, FP64 tallies                 */

/* tally accumulation */
#define VMC_TALLY_FP64 0    /* double-precision atomics (order-dependent rounding) */
#define VMC_TALLY_FIXED64 1 /* 64-bit fixed point, 2^-32 resolution: bit-reproducible.
                               Each deposit is rounded to a multiple of 2^-32
                               and summed in int64, so every tally word's sum
                               over a run must stay below 2^31 in raw units
                               (weight; weight x mm in non-absorbing
                               elements; weight / exit cosine for 2D DEBR).
                               Runs are limited to fewer than 2^31 photons
                               (VMC_ERR_INVALID_ARGUMENT), and a single
                               deposit at or beyond 2^30, or not finite,
                               ends the run with VMC_ERR_TALLY_RANGE
                               (outputs are written but not exact).        */

/* Geometry, medium and boundary conditions: everything MC3D/MC2D holds
   before Init().  dim = 3 for tetrahedra (MC3D), 2 for triangles (MC2D). */
typedef struct vmc_problem_desc {
  int32_t dim;
  int64_t Ne, Nb, Nr;
  const int64_t *H;                  /* [Ne*(dim+1)]                        */
  const int64_t *HN;                 /* optional [Ne*(dim+1)], NULL = build  */
  const int64_t *BH;                 /* [Nb*dim]                             */
  const double *r;                   /* [Nr*dim]                             */
  const double *mua, *mus, *g, *n;   /* [Ne]                                 */
  const char *BCType;                /* [Nb] a m l L i I c C p (3D), a m l i c g L I C (2D) */
  const double *BCn;                 /* [Nb] exterior index; <= 0: no Fresnel */
  const double *BCLNormal;           /* optional [Nb*dim]                    */
  const char *BCLightDirectionType;  /* optional [Nb] n a r p                */
  const double *BCIntensity;         /* optional [Nb]                        */
  const double *GaussianSigma;       /* optional [Nb], 2D 'g' sources        */
} vmc_problem_desc;

/* One forward solve. */
typedef struct vmc_run_desc {
  double f;              /* modulation frequency [Hz]; 0 = CW                  */
  double phase0;         /* initial phase; negatives wrap into [0, 2pi)        */
  int64_t Nphoton;
  uint64_t seed;         /* photon i of a run draws from the stream (seed, i)  */
  int32_t precision;     /* VMC_FP64 | VMC_FP32                                */
  int32_t tally;         /* VMC_TALLY_FP64 | VMC_TALLY_FIXED64                 */
  int32_t block_size;    /* 0 = default (FP32 1024, or 512 for a launch of
                            several items; FP64 256); multiple of 64     */
  int32_t blocks_per_cu; /* resident workgroups per CU: 0 = as many as fit,
                            else at most this many                          */
  int32_t replicas;      /* tally copies: > 0 that many (device-scope atomics);
                            0 = automatic: one per XCD (8, workgroup-scope
                            atomics), or a
                            single one when a copy is 4..64 MiB             */
  int32_t refill_min;    /* dead lanes a wave collects before refilling,
                            0 = default (10 CW, 16 frequency domain)        */
  uint32_t stream_tag;   /* extra stream word: runs with different tags are
                            independent (batch instances use their index)      */
  int32_t lds_slots;     /* per-workgroup LDS tally cache entries (sectors of
                            8 elements); 0 = default (the LDS the lane buffers
                            leave free, about 1900 for FP32 CW),
                            -1 = off, at most 2^20 (must fit the CU's LDS)   */
  int32_t direct_flush;  /* 0 = default: a lane's sector sums leave LDS as
                            coalesced atomics, the wave's lanes together;
                            1 = each lane adds its own words (for testing)   */
  int32_t chunk;         /* photon ids a wave takes from the dispenser at a
                            time: 0 = 128, else a power of 2 in [64, 4096]   */
  int32_t cache_epoch;   /* loop passes after which a workgroup hands its LDS
                            sector cache to the replica and lets sectors claim
                            slots again: 0 = default (1024 for batches of
                            several items, never for one), -1 = never       */
  int32_t model;         /* 1: weight0, chance and c0 below are the caller's;
                            0: the reference's defaults 0.001, 0.1 and
                            2.99792458e11 (MC3D.hpp:178-183), whatever the
                            fields hold                                      */
  uint64_t first_photon; /* stream id of the run's first photon: the run walks
                            photons first_photon .. first_photon + Nphoton - 1
                            of the stream (seed, stream_tag), so disjoint
                            ranges split one solve over launches or GPUs
                            (raw sums add up exactly with fixed-point
                            tallies); 0 = from the start                   */
  /* Model constants, the reference's public members MC3D::weight0, chance
     and c0 (MC3D.hpp:146,160; MC2D.hpp:153,167), read when model == 1.  A
     packet whose weight is below weight0 at an interior face crossing plays
     roulette and survives with probability chance, its weight divided by
     chance (MC3D.hpp:1578-1604); c0 is the speed of light in mm/s, so the
     wave number is k = 2 pi f n / c0 (MC3D.hpp:662).  weight0 >= 0,
     0 < chance <= 1, c0 > 0, all finite.                                   */
  double weight0;
  double chance;
  double c0;
} vmc_run_desc;

typedef struct vmc_output {
  double *ER, *EI;       /* [Ne]; may be NULL                                   */
  double *EBR, *EBI;     /* [Nb]                                                */
  double *DEBR, *DEBI;   /* [Nb] (2D only; zero in 3D like the reference)       */
} vmc_output;

typedef struct vmc_stats {
  int64_t photons;       /* photons launched                                    */
  int64_t loss;          /* MC3D::loss: photons that left the mesh topology     */
  int64_t steps;         /* element steps executed on the device                */
  double kernel_ms;      /* transport kernels, HIP events on their stream       */
  double total_ms;       /* whole vmc_problem_run call                          */
  /* raw weight ledger: absorbed, exited, roulette killed, roulette gained,
     lost.  absorbed + exited + killed - gained + lost == photons.          */
  double ledger[5];
  int64_t rechecks;      /* FP32 exit-face choices re-evaluated in FP64          */
  int32_t grid_blocks, block_size;
  int32_t lds_mesh;      /* 1: a small mesh whose element records every
                            workgroup held in LDS (default cache knobs, the
                            records and a cache slot per 8 elements fit the
                            workgroup's LDS share); 0: records read from
                            device memory                                     */
} vmc_stats;

/* progress callback: percent done; return 0 to abort (MC3D.hpp:1688-1695) */
typedef int (*vmc_progress_fn)(double percent, void *user);

typedef struct vmc_problem vmc_problem;

/* ErrorChecks() + Init() of the reference (MC3D.hpp:460-738): validates,
   builds the neighbour table and light sources, uploads element records to
   the device.  The problem stays resident until destroyed. */
int vmc_problem_create(const vmc_problem_desc *desc, vmc_problem **out);

/* MonteCarlo() (MC3D.hpp:1641-1822): transport Nphoton photons and write
   normalised tallies.  If progress returns 0 the run stops after the
   current chunk and, like the reference, still normalises by the planned
   Nphoton; the status is then VMC_ERR_ABORTED. */
int vmc_problem_run(vmc_problem *prob, const vmc_run_desc *run, vmc_output *out,
                    vmc_stats *stats, vmc_progress_fn progress, void *user);

/* vmc_problem_run that allocates no device memory.  No counterpart in the
   reference: for callers that cannot call hipMalloc (tight serving loops,
   harnesses that forbid device allocation).  Same arguments, results and
   status codes as vmc_problem_run; the tally replicas, counters, ledger and
   launch record live in the library's static device scratch of 64 MiB per
   GPU (held until the process ends) instead of the problem's own arena.
   - The element records of run->precision must already be on the device:
     the first vmc_problem_run (or batch) of that precision on the problem
     allocates them; after that, new media are uploaded in place.
     Otherwise VMC_ERR_INVALID_ARGUMENT.
   - A run whose tallies do not fit the scratch (one replica per XCD is
     reduced to a single one first) returns VMC_ERR_OUT_OF_MEMORY: up to
     about 4M device element slots in FD mode and 8M in CW (a grid mesh
     takes 4 slots per 3 tetrahedra: about 3M and 6M tetrahedra).
   - Every call on a GPU shares that GPU's scratch and clears what it uses
     at its start; overlapping calls on one GPU would race on it, so the
     library serialises the calls of a process with a per-GPU lock. */
int vmc_problem_run_noalloc(vmc_problem *prob, const vmc_run_desc *run, vmc_output *out,
                            vmc_stats *stats, vmc_progress_fn progress, void *user);

/* Autotuner for a problem's launches (no counterpart in the reference).
   Times short runs of *run on prob -- trial_photons photons each, 0 = 2 %
   of run->Nphoton within [1e5, 1e7] -- over the launch shape (block_size,
   refill_min, chunk: photons per wave fetch) and the tally strategy
   (lds_slots cache on/off, cache_epoch, replicas), one knob at a time from
   the defaults,
   and writes *run with the fastest values into *tuned.  A knob moves only
   for a gain above 2 %.  Results do not depend on these knobs (bit for bit
   with VMC_TALLY_FIXED64).  trials (optional) receives the number of timed
   runs. */
int vmc_problem_autotune(vmc_problem *prob, const vmc_run_desc *run, vmc_run_desc *tuned,
                         int64_t trial_photons, int32_t *trials);

/* Host-only ErrorChecks() + the host part of Init(): validation, neighbour
   table and light-source list, without touching a device.  Replaces the
   reference's use of MC3D::BuildNeighbourhoods by createBH3mex
   (cpp/3d/createBH3mex.cpp:26-32).  HN_out (optional) receives the
   neighbour table [Ne*(dim+1)], nsources (optional) the number of
   light-source faces. */
int vmc_problem_check(const vmc_problem_desc *desc, int64_t *HN_out, int64_t *nsources);

/* BuildNeighbourhoods() alone (MC3D.hpp:835-871, MC2D.hpp:724-773), as
   createBH3mex (cpp/3d/createBH3mex.cpp:26-32) and the classes' public
   BuildNeighbourhoods() use it: the neighbour table of topology H
   [Ne*(dim+1)] (0-based node ids below Nr) with the optional boundary
   entries BH [Nb*dim] (Nb may be 0, BH then NULL), nothing else of
   ErrorChecks() or Init().  Face f of element e gets the element sharing
   it, -1-ib for boundary entry ib, or INT64_MAX (the reference's
   INT_FAST64_MAX) when nothing matches -- in 3D; in 2D an unmatched edge is
   MISSING_BOUNDARY, as in the reference.  Writes HN_out [Ne*(dim+1)],
   column-major.  Status: 0, 1 + mcerror (INCONSISTENT_H / INCONSISTENT_BH
   for a node id outside [0, Nr), the reference's duplicate-neighbour
   errors), VMC_ERR_INVALID_ARGUMENT for sizes. */
int vmc_build_neighbours(int32_t dim, int64_t Ne, int64_t Nb, int64_t Nr, const int64_t *H, const int64_t *BH,
                         int64_t *HN_out);

/* Host-only light-source list of Init() (MC3D::BuildLightSource,
   MC3D.hpp:941-1002; MC2D.hpp:851-911): the boundary faces that emit
   light in the reference's element-major order, the element each belongs
   to and the normalised cumulative distribution (face area, 2D edge
   length, times BCIntensity) photons are drawn from.  Writes up to
   `capacity` entries to each non-NULL array and the number of light-source
   faces to *nsources (so a first call with capacity 0 sizes the arrays). */
int vmc_problem_sources(const vmc_problem_desc *desc, int64_t *faces, int64_t *mothers, double *cdf,
                        int64_t capacity, int64_t *nsources);

/* Overrides for a variant of a resident problem; NULL keeps the base's. */
typedef struct vmc_variant_desc {
  const double *mua, *mus, *g, *n;   /* [Ne]                                      */
  const char *BCType;                /* [Nb]: light sources and boundary kinds   */
  const double *BCLNormal;           /* [Nb*dim], as given to vmc_problem_create  */
  const char *BCLightDirectionType;  /* [Nb]                                     */
  const double *BCIntensity;         /* [Nb]                                     */
  const double *GaussianSigma;       /* [Nb], 2D                                 */
} vmc_variant_desc;

/* A problem that shares the base's mesh (and, without medium overrides, its
   device element records) and differs in light sources and/or optical
   parameters: the many source positions and parameter sets an inverse
   solver evaluates on one mesh.  No counterpart in the reference, which
   rebuilds a whole MC3D object per configuration; the checks are those of
   ErrorChecks() and Init() that the overridden inputs can fail.  Destroy
   variants and base in any order. */
int vmc_problem_create_variant(const vmc_problem *base, const vmc_variant_desc *d, vmc_problem **out);

/* The neighbour table used (reference HN convention), [Ne*(dim+1)]. */
int vmc_problem_get_hn(const vmc_problem *prob, int64_t *HN);

/* Device element slots of a problem: whole tally sectors of 8, padding
   slots included (a sector holds whole cells of the element order, so a
   3D grid mesh takes 4 slots per 3 tetrahedra).  The element records (128
   bytes per slot in FP32, 256 in FP64) and every tally replica (8 bytes per
   slot, twice that in the frequency domain) are sized by it.  No
   counterpart in the reference: for callers that budget device memory,
   such as a cache of resident meshes. */
int vmc_problem_slots(const vmc_problem *prob, int64_t *slots);

/* Replace the per-element optical parameters of a resident problem
   (any pointer may be NULL to keep the current values). */
int vmc_problem_set_medium(vmc_problem *prob, const double *mua, const double *mus,
                           const double *g, const double *n);

void vmc_problem_destroy(vmc_problem *prob);

/* One instance of a batch: a resident problem, its run and its outputs. */
typedef struct vmc_batch_item {
  vmc_problem *prob;
  vmc_run_desc run;      /* Nphoton, seed, stream_tag, f, phase0 per item;
                            precision and tally mode must agree across the
                            batch, the tuning fields of item 0 apply to all */
  vmc_output out;        /* normalised like vmc_problem_run                 */
  int64_t loss;          /* out: MC3D::loss of this item                     */
  int32_t status;        /* out                                              */
  int32_t reserved;
  double ledger[5];      /* out: this item's raw weight ledger, as
                            vmc_stats::ledger of its run alone              */
} vmc_batch_item;

/* Many forward solves in one persistent launch per (dimension, CW/FD)
   group: the photons of all items share the GPU, each item keeps its own
   random streams (seed, stream_tag, photon index) and its own tallies, so
   every result equals the one vmc_problem_run gives for that item alone (bit
   for bit in VMC_TALLY_FIXED64 mode).  No counterpart in the reference, which
   runs one MonteCarlo() per object; this is the batched path the inverse
   solvers that call MC3Dmex in a loop (ValoMC.m) would use.  stats (optional)
   sums the whole batch. */
int vmc_batch_run(vmc_batch_item *items, int32_t n, vmc_stats *stats);

/* vmc_problem_autotune for a batch: times vmc_batch_run on the same items
   with trial_fraction of their photons (0 = 2 %) over the same knobs and
   writes items[0].run with the fastest values into *tuned; apply its
   block_size, refill_min, chunk, replicas, lds_slots and cache_epoch to
   every item.  No
   counterpart in the reference. */
int vmc_batch_autotune(const vmc_batch_item *items, int32_t n, double trial_fraction, vmc_run_desc *tuned,
                       int32_t *trials);

/* A batch recorded once as a HIP graph and replayed; no counterpart in the
   reference.  For loops that run the same batch again and again with new
   seeds or media (an inverse solver's iterations).  Creation checks the
   items as vmc_batch_run does, allocates every device buffer, and records
   the batch's device work in one graph: clearing of the tallies, the
   persistent walk launch of each (dimension, CW/FD) group and the replica
   sums.  The launch shape and tally strategy come from items[0].run.
   A run uploads the items' current records (seed, stream_tag, f and phase0
   with the same CW/FD split, and the problems' current media), replays the
   graph and writes outputs and stats as vmc_batch_run does.  Results equal
   vmc_batch_run's, bit for bit with VMC_TALLY_FIXED64.  The items of a run
   must name the plan's problems, photon counts, precision and tally mode,
   in the plan's order.  Destroy the plan before its problems. */
typedef struct vmc_batch_plan vmc_batch_plan;
int vmc_batch_plan_create(const vmc_batch_item *items, int32_t n, vmc_batch_plan **out);
int vmc_batch_plan_run(vmc_batch_plan *plan, vmc_batch_item *items, int32_t n, vmc_stats *stats);
void vmc_batch_plan_destroy(vmc_batch_plan *plan);

/* One-shot call with the argument list of the MEX gateways
   (cpp/3d/MC3Dmex.cpp:106-110, cpp/2d/MC2Dmex.cpp:105-109): create, run,
   destroy.  HN may be NULL; HN_out (optional) receives the table used.
   GaussianSigma is only read for dim == 2. */
int vmc_monte_carlo(int32_t dim, int64_t Ne, int64_t Nb, int64_t Nr,
                    const int64_t *H, const int64_t *HN, const int64_t *BH, const double *r,
                    const char *BCType, const double *BCIntensity,
                    const char *BCLightDirectionType, const double *BCLNormal,
                    const double *BCn, const double *mua, const double *mus,
                    const double *g, const double *n, double f, double phase0,
                    int64_t Nphoton, const double *GaussianSigma, uint64_t seed,
                    int32_t precision, double *ER, double *EI, double *EBR, double *EBI,
                    double *DEBR, double *DEBI, int64_t *HN_out, int64_t *loss);

/* ------------------------------------------------------------------------
   Serving: continuous batching of forward solves.  No counterpart in the
   reference; for callers that issue many MonteCarlo() solves as they come
   (an inverse-problem loop, a request queue).  A worker thread owns the GPU
   work.  Requests are admitted in arrival order when their tallies fit in the
   server's device arena; every launch advances admitted requests of one
   (dimension, CW/FD) group by parts of `launch_photons` (the scheduling
   policy below decides which and how much), so requests that arrive while
   others run join at the next launch.  Each request's photons
   keep their stream ids (seed, stream_tag, photon index), so its result equals
   vmc_problem_run's for the same inputs (bit for bit with
   VMC_TALLY_FIXED64).  A problem and its output buffers belong to the server
   from submit until wait returns for that ticket.
   ------------------------------------------------------------------------ */
typedef struct vmc_server vmc_server;

#define VMC_SERVE_SHORTEST 0 /* least estimated remaining device time first:
                                the launch's group is that request's, and the
                                budget goes to the group's requests in order
                                of remaining time (time per photon measured
                                per problem from earlier launches)           */
#define VMC_SERVE_FAIR 1     /* groups in turn; equal shares of a launch for
                                the group's admitted requests               */

typedef struct vmc_server_desc {
  int32_t precision;      /* VMC_FP64 | VMC_FP32, for every request              */
  int32_t tally;          /* VMC_TALLY_FP64 | VMC_TALLY_FIXED64                  */
  int64_t arena_bytes;    /* device memory for admitted requests' tallies, all
                             replicas; 0 = 4 GiB                                  */
  int64_t launch_photons; /* photons per launch, shared by the requests in it;
                             0 = 2e7 (about 0.15 s of FP32 work)                  */
  int64_t min_slice;      /* least photons a request gets in a launch; 0 = 262144 */
  int32_t policy;         /* VMC_SERVE_SHORTEST (0) | VMC_SERVE_FAIR             */
  int32_t reserved[7];
} vmc_server_desc;

int vmc_server_create(const vmc_server_desc *desc, vmc_server **out);
/* Queue one solve; run->precision and run->tally are the server's.  out's
   buffers are written before vmc_server_wait returns for the ticket.
   VMC_ERR_OUT_OF_MEMORY if the request could never fit in the arena. */
int vmc_server_submit(vmc_server *srv, vmc_problem *prob, const vmc_run_desc *run, const vmc_output *out,
                      uint64_t *ticket);
/* done = 1 once the ticket's outputs are written */
int vmc_server_poll(vmc_server *srv, uint64_t ticket, int32_t *done);
/* Blocks until the ticket is done and releases it; stats (optional) gets
   photons, loss, the request's own weight ledger, kernel_ms of the launches
   it took part in and total_ms from submit to completion.  Returns the
   request's status. */
int vmc_server_wait(vmc_server *srv, uint64_t ticket, vmc_stats *stats);
/* Launches so far, photons launched, summed kernel time, queue lengths. */
int vmc_server_counters(vmc_server *srv, int64_t *launches, int64_t *photons, double *busy_ms,
                        int32_t *waiting, int32_t *active);
/* Finishes the queued and running requests, then frees the server. */
void vmc_server_destroy(vmc_server *srv);

const char *vmc_status_string(int status);
/* last device error text of the calling thread */
const char *vmc_last_error(void);
/* device name, CU count, HBM size; returns VMC_ERR_NO_DEVICE without a GPU */
int vmc_device_info(char *buf, int32_t buflen, int32_t *num_cus);
/* library version string */
const char *vmc_version(void);

#ifdef __cplusplus
}
#endif
#endif
