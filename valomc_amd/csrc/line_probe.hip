// Random-line probe: the throughput ceilings the transport walk runs into.
//
// The walk reads one 128-byte element record per element step from a
// random-looking place in HBM (one 16-byte load per lane, eight per record)
// and adds its tallies with 64-byte atomic line requests.  This program
// measures, on the same launch shape as the walk (one 1024-lane workgroup
// per CU), how many such random 128-byte lines per second the memory system
// delivers for footprints from L2 size to beyond the Infinity Cache, and how
// many 64-byte FP64 atomic line requests (eight lanes per line, as the
// coalesced tally flush issues them) it accepts, alone and mixed with the
// line reads in the walk's proportions (the reads over a record-sized
// footprint, the atomics over a tally-sized one).  It prints one JSON line
// per (mode, footprint).  Not part of the library; built beside the command
// line executables and run by `bench.py --line-probe`.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

typedef float f4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint32_t mix32(uint32_t x) {
  x ^= x >> 16;
  x *= 0x7feb352dU;
  x ^= x >> 15;
  x *= 0x846ca68bU;
  x ^= x >> 16;
  return x;
}

// every lane reads `iters` random 128-byte lines, eight 16-byte loads each
__global__ void read_lines(const f4 *buf, uint32_t nlines, int iters, float *sink) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  f4 acc = {0, 0, 0, 0};
  uint32_t s = mix32(t * 2654435761u + 1u);
  for (int i = 0; i < iters; ++i) {
    s = mix32(s + (uint32_t)i);
    const f4 *p = buf + (size_t)(s % nlines) * 8;
#pragma unroll
    for (int k = 0; k < 8; ++k) acc += p[k];
  }
  if (acc.x == -1.0f) sink[t] = acc.y + acc.z + acc.w;   // never true: keeps the loads
}

// groups of eight lanes add to one random 64-byte line (eight doubles), as
// the tally flush does: one atomic line request per group and iteration
__global__ void atomic_lines(double *buf, uint32_t nlines, int iters) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t grp = t >> 3, j = t & 7u;
  uint32_t s = mix32(grp * 2654435761u + 7u);
  for (int i = 0; i < iters; ++i) {
    s = mix32(s + (uint32_t)i);
    __hip_atomic_fetch_add(buf + (size_t)(s % nlines) * 8 + j, 1.0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

// both at once: every lane reads one random 128-byte line per iteration
// and its group of eight lanes adds to `k` random 64-byte lines of a
// second buffer, i.e. k/8 atomic line requests per line read
__global__ void mixed_lines(const f4 *rbuf, uint32_t rlines, double *abuf, uint32_t alines, int iters, int k,
                            float *sink) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t grp = t >> 3, j = t & 7u;
  f4 acc = {0, 0, 0, 0};
  uint32_t s = mix32(t * 2654435761u + 1u), u = mix32(grp * 2654435761u + 7u);
  for (int i = 0; i < iters; ++i) {
    s = mix32(s + (uint32_t)i);
    const f4 *p = rbuf + (size_t)(s % rlines) * 8;
#pragma unroll
    for (int q = 0; q < 8; ++q) acc += p[q];
    for (int a = 0; a < k; ++a) {
      u = mix32(u + (uint32_t)(i * k + a));
      __hip_atomic_fetch_add(abuf + (size_t)(u % alines) * 8 + j, 1.0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
  }
  if (acc.x == -1.0f) sink[t] = acc.y + acc.z + acc.w;   // never true: keeps the loads
}

__global__ void fill(f4 *buf, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    buf[i] = f4{(float)(i & 1023), 1.f, 2.f, 3.f};
}

static int check(hipError_t e, const char *what) {
  if (e != hipSuccess) {
    fprintf(stderr, "%s: %s\n", what, hipGetErrorString(e));
    exit(1);
  }
  return 0;
}

int main() {
  const size_t cap = (size_t)1 << 30;   // largest footprint: 1 GiB
  const size_t sizes[] = {(size_t)2 << 20, (size_t)16 << 20, (size_t)64 << 20, (size_t)192 << 20, cap};
  int dev = 0, ncu = 0;
  check(hipGetDevice(&dev), "hipGetDevice");
  check(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev), "attribute");
  f4 *buf = nullptr;
  float *sink = nullptr;
  const int block = 1024, grid = ncu;   // one 16-wave workgroup per CU, as the FP32 walk
  check(hipMalloc((void **)&buf, cap), "hipMalloc");
  check(hipMalloc((void **)&sink, (size_t)grid * block * sizeof(float)), "hipMalloc");
  hipLaunchKernelGGL(fill, dim3(4096), dim3(256), 0, 0, buf, cap / sizeof(f4));
  check(hipDeviceSynchronize(), "fill");
  hipEvent_t e0, e1;
  check(hipEventCreate(&e0), "event");
  check(hipEventCreate(&e1), "event");
  for (int mode = 0; mode < 2; ++mode) {
    for (size_t sz : sizes) {
      // lines of 128 B (reads) or 64 B (atomics) inside the footprint
      const uint32_t nlines = (uint32_t)(mode == 0 ? sz / 128 : sz / 64);
      const int iters = mode == 0 ? 1600 : 400;
      float ms = 0.f;
      for (int rep = 0; rep < 2; ++rep) {   // the first launch warms the caches up
        if (mode == 1) hipLaunchKernelGGL(fill, dim3(4096), dim3(256), 0, 0, buf, sz / sizeof(f4));
        check(hipEventRecord(e0), "event");
        if (mode == 0)
          hipLaunchKernelGGL(read_lines, dim3(grid), dim3(block), 0, 0, buf, nlines, iters, sink);
        else
          hipLaunchKernelGGL(atomic_lines, dim3(grid), dim3(block), 0, 0, reinterpret_cast<double *>(buf), nlines,
                             iters);
        check(hipEventRecord(e1), "event");
        check(hipEventSynchronize(e1), "kernel");
        check(hipEventElapsedTime(&ms, e0, e1), "event");
      }
      const double lanes = (double)grid * block;
      const double lines = mode == 0 ? lanes * iters : lanes / 8 * iters;
      const double bytes = lines * (mode == 0 ? 128.0 : 64.0);
      printf("{\"probe\": \"%s\", \"footprint_mib\": %zu, \"ms\": %.3f, \"g_lines_per_s\": %.2f, "
             "\"tb_per_s\": %.3f}\n",
             mode == 0 ? "random 128-byte line reads" : "random 64-byte FP64 atomic lines", sz >> 20, ms,
             lines / ms / 1e6, bytes / ms / 1e9);
      fflush(stdout);
    }
  }
  // mixed: the box walk's proportions (0.651 line reads and 0.248 atomic
  // lines per step: k = 3, 0.375 per read) over its record and tally
  // footprints, and the FD box walk's (0.792 and 0.436: k = 4 and 5, 0.5
  // and 0.625 per read) over 256 MiB of records and a 48 MiB tally
  struct Mix { size_t rsz, asz; int k; };
  const Mix mixes[] = {{(size_t)64 << 20, (size_t)32 << 20, 3}, {(size_t)256 << 20, (size_t)48 << 20, 4},
                       {(size_t)256 << 20, (size_t)48 << 20, 5}};
  double *abuf = nullptr;
  const size_t acap = (size_t)48 << 20;
  check(hipMalloc((void **)&abuf, acap), "hipMalloc");
  for (const Mix &m : mixes) {
    const uint32_t rlines = (uint32_t)(m.rsz / 128), alines = (uint32_t)(m.asz / 64);
    const int iters = 400;
    float ms = 0.f;
    for (int rep = 0; rep < 2; ++rep) {
      hipLaunchKernelGGL(fill, dim3(4096), dim3(256), 0, 0, reinterpret_cast<f4 *>(abuf), m.asz / sizeof(f4));
      check(hipEventRecord(e0), "event");
      hipLaunchKernelGGL(mixed_lines, dim3(grid), dim3(block), 0, 0, buf, rlines, abuf, alines, iters, m.k, sink);
      check(hipEventRecord(e1), "event");
      check(hipEventSynchronize(e1), "kernel");
      check(hipEventElapsedTime(&ms, e0, e1), "event");
    }
    const double lanes = (double)grid * block;
    const double rl = lanes * iters, al = lanes / 8 * iters * m.k;
    printf("{\"probe\": \"mixed: random 128-byte line reads over %zu MiB + %d/8 random 64-byte FP64 atomic lines "
           "per read over %zu MiB\", \"ms\": %.3f, \"g_read_lines_per_s\": %.2f, \"g_atomic_lines_per_s\": %.2f}\n",
           m.rsz >> 20, m.k, m.asz >> 20, ms, rl / ms / 1e6, al / ms / 1e6);
    fflush(stdout);
  }
  check(hipFree(abuf), "hipFree");
  check(hipFree(buf), "hipFree");
  check(hipFree(sink), "hipFree");
  return 0;
}
