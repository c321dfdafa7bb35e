/*
 * oracle/ref_driver.cpp -- TEST INFRASTRUCTURE ONLY.
 *
 * Harness that runs the reference engine itself (cpp/3d/MC3D.hpp or
 * cpp/2d/MC2D.hpp, compiled from the reference tree by oracle/Makefile) on
 * caller-supplied arrays, the way the reference's callers drive it
 * (cpp/3d/MC3D.cpp:56-91, cpp/2d/MC2D.cpp:79-110): fill the public members,
 * ErrorChecks(), Init(), MonteCarlo(progress, finalchecks), read the
 * normalised tallies.  Build with -DREF_DIM=3 against cpp/3d or -DREF_DIM=2
 * against cpp/2d; the result lands in oracle/_ref/ and is never committed.
 */
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <iostream>
#include <vector>
#include <omp.h>

#include "../oracle/vmc_oracle.h"

#if REF_DIM == 3
#include "MC3D.hpp"
typedef MC3D Engine;
#define REF_ENTRY ref3d_run
#else
#include "MC2D.hpp"
typedef MC2D Engine;
#define REF_ENTRY ref2d_run
#endif

namespace {

bool keep_going(double) { return true; }
void no_final_checks(int, int) {}

template <class T>
void fill(Array<T> &a, const T *src, int64_t rows, int64_t cols) {
  if (!src) return;
  if (cols == 1) a.resize(rows); else a.resize(rows, cols);
  std::memcpy(a.data, src, sizeof(T) * rows * cols);
}

} // namespace

extern "C" int REF_ENTRY(const oracle_problem *p, int nthreads, oracle_result *out) {
  const int dim = REF_DIM;
  omp_set_num_threads(nthreads < 1 ? 1 : nthreads);
  Engine MC;
  fill(MC.H, p->H, p->Ne, dim + 1);
  fill(MC.HN, p->HN, p->Ne, dim + 1);
  fill(MC.BH, p->BH, p->Nb, dim);
  fill(MC.r, p->r, p->Nr, dim);
  fill(MC.mua, p->mua, p->Ne, 1);
  fill(MC.mus, p->mus, p->Ne, 1);
  fill(MC.g, p->g, p->Ne, 1);
  fill(MC.n, p->n, p->Ne, 1);
  fill(MC.BCType, p->BCType, p->Nb, 1);
  fill(MC.BCn, p->BCn, p->Nb, 1);
  fill(MC.BCLNormal, p->BCLNormal, p->Nb, dim);
  fill(MC.BCLightDirectionType, p->BCLDirType, p->Nb, 1);
  fill(MC.BCIntensity, p->BCIntensity, p->Nb, 1);
#if REF_DIM == 2
  fill(MC.GaussianSigma, p->GaussianSigma, p->Nb, 1);
#endif
  MC.f = p->f;
  MC.phase0 = p->phase0;
  if (MC.phase0 < 0) MC.phase0 += 2 * M_PI * ceil(-MC.phase0 / (2 * M_PI));
  MC.Nphoton = p->Nphoton;
  MC.seed = (unsigned long)p->seed;
  MC.weight0 = p->weight0;
  MC.chance = p->chance;
  MC.c0 = p->c0;
  try {
    MC.ErrorChecks();
    MC.Init();
  } catch (mcerror e) {
    return 1 + (int)e;
  }
  auto t0 = std::chrono::steady_clock::now();
  MC.MonteCarlo(keep_going, no_final_checks);
  auto t1 = std::chrono::steady_clock::now();
  const int64_t Ne = MC.ER.N, Nb = MC.EBR.N;
  if (out->ER) std::memcpy(out->ER, MC.ER.data, sizeof(double) * Ne);
  if (out->EI) std::memcpy(out->EI, MC.EI.data, sizeof(double) * Ne);
  if (out->EBR) std::memcpy(out->EBR, MC.EBR.data, sizeof(double) * Nb);
  if (out->EBI) std::memcpy(out->EBI, MC.EBI.data, sizeof(double) * Nb);
  if (out->DEBR) std::memcpy(out->DEBR, MC.DEBR.data, sizeof(double) * Nb);
  if (out->DEBI) std::memcpy(out->DEBI, MC.DEBI.data, sizeof(double) * Nb);
  if (out->HN) std::memcpy(out->HN, MC.HN.data, sizeof(int64_t) * MC.HN.N);
  out->loss = MC.loss;
  for (int q = 0; q < 5; ++q) out->ledger[q] = 0.0;
  out->seconds = std::chrono::duration<double>(t1 - t0).count();
  return 0;
}
