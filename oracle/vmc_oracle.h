/*
 * oracle/vmc_oracle.h -- TEST INFRASTRUCTURE ONLY.
 *
 * The oracle is a CPU restatement of ValoMC's photon-transport path
 * (MC3D/MC2D::MonteCarlo and everything it calls).  Only tests/, the smoke
 * check in __graft_entry__.py and the cpu_baseline leg of bench.py may load
 * it, and only as the checker: the product path (valomc_amd/) never links or
 * calls it.
 *
 * The same problem/result structs are used by the reference harness
 * (ref_driver.cpp, built once per dimension), which runs the reference's
 * own engine, compiled from the reference's sources by oracle/Makefile into
 * oracle/_ref/.
 *
 * Array layout is the reference's: column-major (Fortran order), element
 * index fastest (cpp/3d/Array.hpp:58-66).  H[el + Ne*c], BH[ib + Nb*c],
 * r[node + Nr*d], BCLNormal[ib + Nb*d].
 */
#ifndef VMC_ORACLE_H
#define VMC_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct oracle_problem {
  int64_t Ne, Nb, Nr;
  const int64_t *H;          /* [Ne*4] (3D) / [Ne*3] (2D), 0-based node ids     */
  const int64_t *HN;         /* optional neighbour table, NULL = build it       */
  const int64_t *BH;         /* [Nb*3] (3D) / [Nb*2] (2D)                       */
  const double *r;           /* [Nr*3] / [Nr*2]                                 */
  const double *mua, *mus, *g, *n; /* [Ne]                                     */
  const char *BCType;        /* [Nb]                                            */
  const double *BCn;         /* [Nb]                                            */
  const double *BCLNormal;   /* optional [Nb*3] / [Nb*2]                        */
  const char *BCLDirType;    /* optional [Nb]                                   */
  const double *BCIntensity; /* optional [Nb]                                   */
  const double *GaussianSigma; /* optional [Nb], 2D only                        */
  double f, phase0;
  int64_t Nphoton;
  uint64_t seed;
  /* model constants, the reference's public members MC3D::weight0, chance,
     c0 (MC3D.hpp:146,160; defaults 0.001, 0.1, 2.99792458e11 at :178-183) */
  double weight0, chance, c0;
  /* Philox mode only: the stream tag and the stream id of the first photon
     (photon i of the run draws from stream (seed, tag, first_photon + i)) */
  uint32_t stream_tag;
  uint32_t pad;
  uint64_t first_photon;
} oracle_problem;

typedef struct oracle_result {
  double *ER, *EI;           /* [Ne] normalised, as MC3D::ER/EI after MonteCarlo */
  double *EBR, *EBI;         /* [Nb]                                            */
  double *DEBR, *DEBI;       /* [Nb]                                            */
  int64_t *HN;               /* optional out [Ne*4] / [Ne*3]                    */
  int64_t loss;              /* photons lost (MC3D::loss)                       */
  /* raw weight ledger (restatement only; zero from the reference harness):
     [0] absorbed, [1] exited, [2] roulette killed, [3] roulette gained,
     [4] lost.  absorbed + exited + killed - gained + lost == Nphoton.      */
  double ledger[5];
  double seconds;            /* wall time of the MonteCarlo call only          */
} oracle_result;

/* rng_mode */
#define ORACLE_RNG_MT19937 0 /* reference streams: thread t uses mt19937(seed+t),
                                photons split N/T, remainder to the last thread */
#define ORACLE_RNG_PHILOX 1  /* per-photon Philox4x32-10 + xoshiro128** streams */

/* Return 0 on success, 1 + mcerror code on an input error
   (cpp/Errors.hpp:4-21 order). */
int oracle_run3d(const oracle_problem *p, int nthreads, int rng_mode, oracle_result *out);
int oracle_run2d(const oracle_problem *p, int nthreads, int rng_mode, oracle_result *out);

/* Philox4x32-10 block function (Salmon et al., SC'11), for known-answer tests. */
void oracle_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);

/* n outputs of the per-photon xoshiro128** step started from state[4] (the
   walk's stream after seeding), for known-answer tests. */
void oracle_xoshiro128ss(const uint32_t state[4], int n, uint32_t *out);

/* Reference harness entry points (oracle/_ref/libref3d.so, libref2d.so). */
int ref3d_run(const oracle_problem *p, int nthreads, oracle_result *out);
int ref2d_run(const oracle_problem *p, int nthreads, oracle_result *out);

#ifdef __cplusplus
}
#endif
#endif
