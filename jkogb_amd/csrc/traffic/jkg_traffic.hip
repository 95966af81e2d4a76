// jkg_traffic -- the workload of the HBM-traffic counter passes, as a plain program
// (no Python, no torch in the process):
//
//   jkg_traffic FN N [STEPS]
//
// First a plain device copy of 512 MiB, three times: the counters' scale (the bytes
// it reads and writes are known).  Then STEPS (default 2) stage-profiled steps of the
// mixed solver (jkg_dprofile_mixed) and of the plain solver (jkg_dprofile) on FN.Y
// (N x N FP64, column-major) and FN.J (N int64 signs), N_2 = N/2, unfused FMAD.
// Under rocprofv3 --pmc FETCH_SIZE (then WRITE_SIZE) --kernel-trace each dispatch gets
// its counter; the per-stage device times are printed for reference.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "jkg.h"

__global__ void __launch_bounds__(256) k_copy(const double4 *__restrict__ src, double4 *__restrict__ dst,
                                               size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

static bool read_file(const char *name, void *p, size_t bytes) {
  FILE *f = fopen(name, "rb");
  if (!f) return false;
  const size_t got = fread(p, 1, bytes, f);
  fclose(f);
  return got == bytes;
}

int main(int argc, char **argv) {
  if (argc < 3) {
    fprintf(stderr, "usage: %s FN N [STEPS]\n", argv[0]);
    return 2;
  }
  const int n = atoi(argv[2]);
  const int steps = argc > 3 ? atoi(argv[3]) : 2;
  if (n < 3 || n > JKG_NMAX_REAL || steps < 1) {
    fprintf(stderr, "N in [3, %d], STEPS >= 1\n", JKG_NMAX_REAL);
    return 2;
  }
  std::vector<double> y((size_t)n * n);
  std::vector<int64_t> j(n);
  char name[4096];
  snprintf(name, sizeof name, "%s.Y", argv[1]);
  if (!read_file(name, y.data(), sizeof(double) * y.size())) {
    fprintf(stderr, "cannot read %s\n", name);
    return 1;
  }
  snprintf(name, sizeof name, "%s.J", argv[1]);
  if (!read_file(name, j.data(), sizeof(int64_t) * j.size())) {
    fprintf(stderr, "cannot read %s\n", name);
    return 1;
  }

  // the counters' scale: 512 MiB read and written by one kernel, three times
  const size_t bytes = (size_t)512 << 20, count = bytes / sizeof(double4);
  double4 *src = nullptr, *dst = nullptr;
  if (hipMalloc(&src, bytes) != hipSuccess || hipMalloc(&dst, bytes) != hipSuccess) {
    fprintf(stderr, "hipMalloc failed\n");
    return 1;
  }
  if (hipMemset(src, 0, bytes) != hipSuccess) return 1;
  for (int k = 0; k < 3; ++k) hipLaunchKernelGGL(k_copy, dim3(8192), dim3(256), 0, 0, src, dst, count);
  if (hipDeviceSynchronize() != hipSuccess || hipGetLastError() != hipSuccess) {
    fprintf(stderr, "copy failed\n");
    return 1;
  }
  (void)hipFree(src);
  (void)hipFree(dst);
  printf("calibration: 3 copies of %zu bytes\n", bytes);
  fflush(stdout);

  jkg_ctx *ctx = nullptr;
  int rc = jkg_create(&ctx, 0);
  if (rc != JKG_OK) {
    fprintf(stderr, "jkg_create: %s\n", jkg_strerror(rc));
    return 1;
  }
  static const char *stage[7] = {"begin", "weights", "sort", "select", "pivots", "rotate", "end"};
  double ms[7];
  jkg_result res;
  jkg_mixed_opts m;
  jkg_mixed_opts_default(&m);
  for (int solver = 0; solver < 2; ++solver) {
    rc = solver == 0 ? jkg_dprofile_mixed(ctx, n, n / 2, y.data(), j.data(), 0, steps, &m, ms, &res)
                     : jkg_dprofile(ctx, n, n / 2, y.data(), j.data(), 0, steps, ms, &res);
    if (rc != JKG_OK) {
      fprintf(stderr, "%s: %s\n", solver == 0 ? "jkg_dprofile_mixed" : "jkg_dprofile", jkg_strerror(rc));
      jkg_destroy(ctx);
      return 1;
    }
    printf("%s: %d steps, IT %lld, SL %lld, b2 %lld, b4 %lld;", solver == 0 ? "mixed" : "plain", res.steps,
           (long long)res.sum_it, (long long)res.sum_sl, (long long)res.sum_b2, (long long)res.sum_b4);
    for (int k = 0; k < 7; ++k) printf(" %s %.4f", stage[k], ms[k] / steps);
    printf(" ms per step\n");
    fflush(stdout);
  }
  jkg_destroy(ctx);
  return 0;
}
