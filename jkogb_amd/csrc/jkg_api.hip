// jkg_api.hip -- the C ABI of libjkg.so (include/jkg.h): contexts, device
// buffers, the graph-captured step loop, and the step-level entry points.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <type_traits>
#include <vector>

#include "../../include/jkg.h"
#include "jkg_hsvd2.hpp"
#include "jkg_step.hpp"
#include "jkg_mlp.hpp"
#include "jkg_batch.hpp"

using namespace jkg;

static thread_local int g_last_hip = 0;

#define HIPCHK(call)                        \
  do {                                      \
    hipError_t e_ = (call);                 \
    if (e_ != hipSuccess) {                 \
      g_last_hip = (int)e_;                 \
      return JKG_EHIP;                      \
    }                                       \
  } while (0)

struct jkg_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  // buffers of the current problem size
  int n = 0, cplx = -1, trace_cap = 0;
  int streamed = 0;  // trace records of the running solve already copied to the caller's buffer
  size_t sort_n = 0;
  std::vector<void *> allocs;
  DeviceStep d{};
  double *sigma = nullptr;
  // captured steps: [0] FP64, [1] the FP32 warm start (each for the options it was captured with)
  hipGraph_t graph[2] = {nullptr, nullptr};
  hipGraphExec_t exec[2] = {nullptr, nullptr};
  // What a captured graph was captured from: every launch shape and kernel argument of the step
  // comes from the DeviceStep (sizes, options and every device pointer, such as the frame buffer of
  // jkg_set_frames), so a graph is reused only while the DeviceStep is byte for byte the same.
  struct GraphKey {
    bool valid = false;
    DeviceStep d{};
  } gkey[2];
  StepState st_init{};  // the loop state of the current problem before its first step
  // mixed solver buffers (allocated on first use)
  std::vector<void *> mallocs;
  int mixed_n = 0;
  // FP32 warm start: FP32 copies (in d), Newton-Schulz products M, T, P and block partials
  int warm_n = 0;
  void *wm_M = nullptr, *wm_T = nullptr, *wm_P = nullptr;
  double *wm_part = nullptr;
  int ncp3_n = 0;  // ID_NCP=3: per-offset totals allocated for this N
  // selection schedule: walk rounds per step (0: by size) and entries walked before each filter
  int sel_rounds = 0, sel_limit = 65536;
  // per-step frames of |A| (jkg_set_frames): device buffer, schedule, caller's buffer and count
  uint8_t *frames_dev = nullptr, *frames_host = nullptr;
  int32_t *frames_count = nullptr;
  int frame_every = 0, frame_size = 0, frame_max = 0;
  // step controller and its training (allocated on first use, kept across problems)
  float *ctl_master = nullptr, *ctl_table = nullptr, *ctl_trace = nullptr;
  uint16_t *ctl_q = nullptr;
  int ctl_trace_cap = 0;
  float *tr_buf = nullptr;  // training scratch
  size_t tr_floats = 0;
  // misc test buffers
  void *tmp = nullptr;
  size_t tmp_bytes = 0;
};

static void free_all(jkg_ctx *c) {
  for (int k = 0; k < 2; ++k) {
    if (c->exec[k]) (void)hipGraphExecDestroy(c->exec[k]);
    if (c->graph[k]) (void)hipGraphDestroy(c->graph[k]);
    c->exec[k] = nullptr;
    c->graph[k] = nullptr;
    c->gkey[k] = jkg_ctx::GraphKey{};
  }
  c->warm_n = 0;
  c->wm_M = c->wm_T = c->wm_P = nullptr;
  c->wm_part = nullptr;
  c->ncp3_n = 0;
  for (void *p : c->allocs) (void)hipFree(p);
  c->allocs.clear();
  for (void *p : c->mallocs) (void)hipFree(p);
  c->mallocs.clear();
  c->mixed_n = 0;
  c->n = 0;
  c->cplx = -1;
  c->trace_cap = 0;
  memset(&c->d, 0, sizeof(c->d));
  c->sigma = nullptr;
}

template <class T>
static int dalloc(jkg_ctx *c, T **p, size_t count) {
  void *q = nullptr;
  if (count == 0) count = 1;
  if (hipMalloc(&q, count * sizeof(T)) != hipSuccess) {
    (void)hipGetLastError();
    return JKG_ENOMEM;
  }
  c->allocs.push_back(q);
  *p = (T *)q;
  return JKG_OK;
}

#define ACHK(x)                 \
  do {                          \
    int r_ = (x);               \
    if (r_ != JKG_OK) {         \
      free_all(c);              \
      return r_;                \
    }                           \
  } while (0)

// Walk rounds per step by list length: about one round (65536 entries walked by one
// workgroup, then a whole-GPU filter) per 768K candidate pairs, 2 to 12 (measured on
// MI355X at N=4096, select stage per step: 10 x 65536 0.86 ms, 12 x 49152 0.89 ms,
// 8 x 98304 0.91 ms, 16 x 32768 0.95 ms, 4 x 262144 1.26 ms, 2 x 1048576 1.73 ms).
static int select_rounds_for(size_t nn) {
  const size_t r = nn / 786432 + 1;
  return (int)(r < 2 ? 2 : (r > 12 ? 12 : r));
}

// (Re)allocates the buffers of an N x N problem (real or complex).
static int ensure_buffers(jkg_ctx *c, int n, int cplx, int trace_cap) {
  if (trace_cap < 1) trace_cap = 1;
  if (n > 65535) return JKG_ENOTSUP;  // pairs are packed as two 16-bit indices
  if (c->n == n && c->cplx == cplx && c->trace_cap >= trace_cap) return JKG_OK;
  free_all(c);
  DeviceStep &d = c->d;
  const size_t nn = (size_t)n * (n - 1) / 2;
  const size_t esz = cplx ? 16 : 8;
  const int n2max = n / 2 > 0 ? n / 2 : 1;
  d.n = n;
  d.lda = n;
  d.cplx = cplx;
  d.nn = (int)nn;
  d.n2max = n2max;
  ACHK(dalloc(c, (char **)&d.A, esz * n * n));
  ACHK(dalloc(c, (char **)&d.U, esz * n * n));
  ACHK(dalloc(c, (char **)&d.Z, esz * n * n));
  ACHK(dalloc(c, &d.J, n));
  ACHK(dalloc(c, &d.plus, n));
  ACHK(dalloc(c, &d.minus, n));
  ACHK(dalloc(c, &d.sort.k0, nn));
  ACHK(dalloc(c, &d.sort.v0, nn + LIST_PAD));
  ACHK(dalloc(c, &d.sort.k1, 2 * nn));
  ACHK(dalloc(c, &d.sort.k2, 2 * nn));
  ACHK(dalloc(c, &d.sort.v1, nn));
  ACHK(dalloc(c, &d.sort.v2, nn));
  ACHK(dalloc(c, &d.sort.d1, sort_digit_bytes((int)nn)));
  ACHK(dalloc(c, &d.sort.d2, sort_digit_bytes((int)nn)));
  ACHK(dalloc(c, (char **)&d.sort.scratch, sort_scratch_bytes((int)nn)));
  for (int k = 0; k < 2; ++k) {
    ACHK(dalloc(c, &d.lists.val[k], nn + LIST_PAD));
    ACHK(dalloc(c, &d.lists.pos[k], nn + LIST_PAD));
  }
  ACHK(dalloc(c, &d.fcount, select_count_ints((int)nn)));
  ACHK(dalloc(c, &d.used, n));
  ACHK(dalloc(c, &d.piv_p, n2max));
  ACHK(dalloc(c, &d.piv_q, n2max));
  ACHK(dalloc(c, &d.piv_pos, n2max));
  ACHK(dalloc(c, &d.piv_info, n2max));
  ACHK(dalloc(c, &d.piv_w, n2max));
  ACHK(dalloc(c, (char **)&d.piv_V, esz * 4 * n2max));
  ACHK(dalloc(c, (char **)&d.piv_W, esz * 4 * n2max));
  ACHK(dalloc(c, &d.piv_d, 2 * n2max));
  ACHK(dalloc(c, &d.rot_row, n2max));
  ACHK(dalloc(c, &d.rot_col, n2max));
  if (cplx) ACHK(dalloc(c, (char **)&d.zrot, 3 * zrot_bytes() * n2max));
  ACHK(dalloc(c, &d.colpiv, n));
  ACHK(dalloc(c, &d.st, 1));
  ACHK(dalloc(c, &d.trace, trace_cap));
  ACHK(dalloc(c, &c->sigma, n));
  d.select_limit = c->sel_limit;  // walked before each filter of the list
  d.sel_rounds = c->sel_rounds > 0 ? c->sel_rounds : select_rounds_for(nn);
  if (step_prepare(d) != 0) {
    free_all(c);
    return JKG_ENOTSUP;
  }
  c->n = n;
  c->cplx = cplx;
  c->trace_cap = trace_cap;
  return JKG_OK;
}

// Buffers of the Gram-form weights: (O^H O)(plus, minus) (complex: real parts,
// then imaginary parts) and the column norms, then the diagonal.
static int ensure_mixed(jkg_ctx *c) {
  DeviceStep &d = c->d;
  if (c->mixed_n == d.n) return JKG_OK;  // the buffers are freed with the problem's (size or type change)
  const size_t n = (size_t)d.n, parts = d.cplx ? 2 : 1;
  void *g = nullptr, *cn = nullptr;
  if (hipMalloc(&g, sizeof(double) * parts * gram_plane_doubles(d.n)) != hipSuccess ||
      hipMalloc(&cn, sizeof(double) * (1 + parts) * n) != hipSuccess) {
    (void)hipGetLastError();
    if (g) (void)hipFree(g);
    return JKG_ENOMEM;
  }
  c->mallocs.push_back(g);
  c->mallocs.push_back(cn);
  d.gram = (double *)g;
  d.coln = (double *)cn;
  c->mixed_n = d.n;
  return JKG_OK;
}

// Buffers of the FP32 warm start: A32, U32, Z32 and the hand-over's products.
static int ensure_warm(jkg_ctx *c) {
  DeviceStep &d = c->d;
  if (c->warm_n == d.n) return JKG_OK;  // freed with the problem's buffers
  const size_t n2 = (size_t)d.n * d.n;
  void *p[7] = {};
  const size_t bytes[7] = {8 * n2, 8 * n2, 8 * n2, 16 * n2, 16 * n2, 16 * n2,
                           sizeof(double) * (size_t)ns_blocks(d.n)};
  for (int k = 0; k < 7; ++k) {
    if (hipMalloc(&p[k], bytes[k]) != hipSuccess) {
      (void)hipGetLastError();
      for (int i = 0; i < k; ++i) (void)hipFree(p[i]);
      return JKG_ENOMEM;
    }
  }
  for (int k = 0; k < 7; ++k) c->mallocs.push_back(p[k]);
  d.A32 = p[0];
  d.U32 = p[1];
  d.Z32 = p[2];
  c->wm_M = p[3];
  c->wm_T = p[4];
  c->wm_P = p[5];
  c->wm_part = (double *)p[6];
  c->warm_n = d.n;
  return JKG_OK;
}

// ID_NCP=3: the per-offset totals (NN doubles) and the walk kernel's LDS size.
static int ensure_ncp3(jkg_ctx *c) {
  DeviceStep &d = c->d;
  if (d.n > NCP3_NMAX) return JKG_ENOTSUP;
  if (c->ncp3_n != d.n) {
    void *w = nullptr;
    if (hipMalloc(&w, sizeof(double) * (size_t)(d.nn > 0 ? d.nn : 1)) != hipSuccess) {
      (void)hipGetLastError();
      return JKG_ENOMEM;
    }
    c->mallocs.push_back(w);
    d.ncp3_wsum = (double *)w;
    c->ncp3_n = d.n;
  }
  return ncp3_prepare(d.n, d.n2max) ? JKG_ENOTSUP : JKG_OK;
}

// Controller parameters (FP32 master + BF16 copy), phase table and per-step trace.
static int ensure_ctl(jkg_ctx *c, int trace_cap) {
  auto grab = [&](void **p, size_t bytes) -> int {
    if (*p) return JKG_OK;
    if (hipMalloc(p, bytes) != hipSuccess) {
      (void)hipGetLastError();
      *p = nullptr;
      return JKG_ENOMEM;
    }
    return JKG_OK;
  };
  int rc = grab((void **)&c->ctl_master, sizeof(float) * MLP_NPARAM);
  if (!rc) rc = grab((void **)&c->ctl_q, sizeof(uint16_t) * MLP_NPARAM);
  if (!rc) rc = grab((void **)&c->ctl_table, sizeof(float) * 8);
  if (rc) return rc;
  if (trace_cap > c->ctl_trace_cap) {
    if (c->ctl_trace) (void)hipFree(c->ctl_trace);
    c->ctl_trace = nullptr;
    c->ctl_trace_cap = 0;
    if (hipMalloc(&c->ctl_trace, sizeof(float) * CTL_REC * (size_t)trace_cap) != hipSuccess) {
      (void)hipGetLastError();
      c->ctl_trace = nullptr;
      return JKG_ENOMEM;
    }
    c->ctl_trace_cap = trace_cap;
  }
  return JKG_OK;
}

// Training scratch of at least `floats` floats.
static int ensure_train(jkg_ctx *c, size_t floats) {
  if (c->tr_floats >= floats) return JKG_OK;
  if (c->tr_buf) (void)hipFree(c->tr_buf);
  c->tr_buf = nullptr;
  c->tr_floats = 0;
  if (hipMalloc(&c->tr_buf, sizeof(float) * floats) != hipSuccess) {
    (void)hipGetLastError();
    return JKG_ENOMEM;
  }
  c->tr_floats = floats;
  return JKG_OK;
}

static int tmp_bytes(jkg_ctx *c, size_t bytes) {
  if (c->tmp_bytes >= bytes) return JKG_OK;
  if (c->tmp) (void)hipFree(c->tmp);
  c->tmp = nullptr;
  c->tmp_bytes = 0;
  if (hipMalloc(&c->tmp, bytes) != hipSuccess) {
    (void)hipGetLastError();
    return JKG_ENOMEM;
  }
  c->tmp_bytes = bytes;
  return JKG_OK;
}

extern "C" {

void jkg_opts_default(jkg_opts *o) {
  memset(o, 0, sizeof(*o));
  o->max_steps = -1;
}

const char *jkg_strerror(int code) {
  switch (code) {
    case JKG_OK: return "success";
    case JKG_EHIP: return "HIP runtime error";
    case JKG_ENOMEM: return "device allocation failed";
    case JKG_ENOTSUP: return "option not supported on the GPU";
    case JKG_EJ: return "J has an entry outside {-1,+1} (R%TRU)";
    case JKG_ENOCONV: return "the step loop did not end within 1000 N + 1000 steps";
    default: return (code < 0) ? "invalid argument" : "unknown";
  }
}

int jkg_last_hip_error(void) { return g_last_hip; }

static int noalloc_setup(int device);

int jkg_create(jkg_ctx **out, int device) {
  if (!out) return -1;
  *out = nullptr;
  int count = 0;
  HIPCHK(hipGetDeviceCount(&count));
  if (device < 0 || device >= count) return -2;
  HIPCHK(hipSetDevice(device));
  {
    const int rc = noalloc_setup(device);  // the device address of this GPU's scratch array, once
    if (rc) return rc;
  }
  jkg_ctx *c = new jkg_ctx();
  c->device = device;
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess) {
    g_last_hip = (int)hipGetLastError();
    delete c;
    return JKG_EHIP;
  }
  *out = c;
  return JKG_OK;
}

void jkg_destroy(jkg_ctx *c) {
  if (!c) return;
  (void)hipSetDevice(c->device);
  (void)hipStreamSynchronize(c->stream);
  free_all(c);
  if (c->tmp) (void)hipFree(c->tmp);
  for (void *p : {(void *)c->ctl_master, (void *)c->ctl_q, (void *)c->ctl_table, (void *)c->ctl_trace,
                  (void *)c->tr_buf, (void *)c->frames_dev})
    if (p) (void)hipFree(p);
  (void)hipEventDestroy(c->ev0);
  (void)hipEventDestroy(c->ev1);
  (void)hipStreamDestroy(c->stream);
  delete c;
}

int jkg_jstep_len(int n, int n2) {  // src/shared/jstep.F90:124-133
  if (n2 == 0) return (n / 2 > 0) ? n / 2 : 0;
  int m = (n2 > 0) ? n2 : 0;
  return (m < n / 2) ? m : n / 2;
}

int jkg_tru(int id_tru, int n, const int32_t *j, int nn, int32_t *p, int32_t *q, int32_t pn[2]) {
  // src/shared/jstep.F90:9-120: trigonometric pairs from the front,
  // hyperbolic pairs from the back; column-cyclic (1) or row-cyclic (2)
  pn[0] = pn[1] = 0;
  if (n < 0) return -1;
  if (nn < 0) return -3;
  if (n <= 1) return 0;
  for (int i = 0; i < n; ++i) {
    if (j[i] == 1) ++pn[0];
    else if (j[i] == -1) ++pn[1];
    else return -2;
  }
  int it = 1, ih = nn;
  auto put = [&](int ip, int iq) {
    if (j[ip - 1] == j[iq - 1]) { p[it - 1] = ip; q[it - 1] = iq; ++it; }
    else { p[ih - 1] = ip; q[ih - 1] = iq; --ih; }
  };
  if (id_tru == 2) {
    for (int ip = 1; ip <= n - 1; ++ip)
      for (int iq = ip + 1; iq <= n; ++iq) {
        if (it > ih) return 0;
        put(ip, iq);
      }
  } else {
    for (int iq = 2; iq <= n; ++iq)
      for (int ip = 1; ip <= iq - 1; ++ip) {
        if (it > ih) return 0;
        put(ip, iq);
      }
  }
  return it - 1;
}

uint32_t jkg_pair_index(int n, int p, int q) { return pair_index(n, p, q); }

double jkg_key_weight(uint64_t key) {
  if (key == KEY_NAN) return NAN;
  uint64_t ord = ~key;
  uint64_t u = (ord & 0x8000000000000000ull) ? (ord & 0x7FFFFFFFFFFFFFFFull) : ~ord;
  double w;
  memcpy(&w, &u, 8);
  return w;
}

}  // extern "C"

// ---------------------------------------------------------------- problem setup
static int upload_problem(jkg_ctx *c, int n, int n2, const void *y, const int64_t *j, int *npos, int *nneg,
                          int max_steps, int trace_cap, double tau = 0.0, double tau_cap = 1e-6,
                          float handoff = -10.f, int step0 = 0) {
  // (the element size follows the buffers: c->d.cplx)
  DeviceStep &d = c->d;
  std::vector<int> J32(n), pl, mi;
  for (int i = 0; i < n; ++i) {
    if (j[i] != 1 && j[i] != -1) return JKG_EJ;
    J32[i] = (int)j[i];
    (j[i] == 1 ? pl : mi).push_back(i);
  }
  *npos = (int)pl.size();
  *nneg = (int)mi.size();
  d.npl = (int)pl.size();
  d.nmi = (int)mi.size();
  const size_t esz = d.cplx ? 16 : 8;
  HIPCHK(hipMemcpyAsync(d.A, y, esz * n * n, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(d.J, J32.data(), sizeof(int) * n, hipMemcpyHostToDevice, c->stream));
  if (!pl.empty()) HIPCHK(hipMemcpyAsync(d.plus, pl.data(), sizeof(int) * pl.size(), hipMemcpyHostToDevice, c->stream));
  if (!mi.empty()) HIPCHK(hipMemcpyAsync(d.minus, mi.data(), sizeof(int) * mi.size(), hipMemcpyHostToDevice, c->stream));
  StepState s{};
  s.step = step0;  // a resumed loop continues the step count (the trace's STEP, INFO)
  s.step0 = step0;
  s.max_steps = (max_steps >= 0) ? step0 + max_steps : -1;
  s.n2 = n2;
  s.trace_cap = trace_cap;
  s.tau = tau;
  s.n2max = n2;
  s.n2f = 1.f;
  s.ltau = (tau > 0.0) ? (float)std::log10(tau) : -16.f;
  s.off_best = HUGE_VAL;
  s.ltau_cap = (float)std::log10(tau_cap);
  s.lhandoff = handoff;
  {
    double y2 = 0.0;  // ||Y||_F^2 (the controller's scale)
    const double *yd = (const double *)y;
    const size_t cnt = (size_t)n * n * (d.cplx ? 2 : 1);
    for (size_t e = 0; e < cnt; ++e) y2 += yd[e] * yd[e];
    s.ynorm2 = y2 > 0.0 ? y2 : 1.0;
  }
  c->st_init = s;
  HIPCHK(hipMemcpyAsync(d.st, &c->st_init, sizeof(s), hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));  // the host vectors go out of scope
  return JKG_OK;
}

// Captures (or reuses) the step graph of the current precision: slot 1 while d.prec32.
static int capture_step(jkg_ctx *c, int fused) {
  const int slot = c->d.prec32 ? 1 : 0;
  c->d.fused = fused;
  if (c->exec[slot] && c->gkey[slot].valid && memcmp(&c->gkey[slot].d, &c->d, sizeof(DeviceStep)) == 0)
    return JKG_OK;
  if (c->exec[slot]) (void)hipGraphExecDestroy(c->exec[slot]);
  if (c->graph[slot]) (void)hipGraphDestroy(c->graph[slot]);
  c->exec[slot] = nullptr;
  c->graph[slot] = nullptr;
  c->gkey[slot].valid = false;
  HIPCHK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
  step_launch(c->d, c->stream);
  HIPCHK(hipStreamEndCapture(c->stream, &c->graph[slot]));
  HIPCHK(hipGraphInstantiate(&c->exec[slot], c->graph[slot], nullptr, nullptr, 0));
  memcpy(&c->gkey[slot].d, &c->d, sizeof(DeviceStep));
  c->gkey[slot].valid = true;
  return JKG_OK;
}

// Launches the captured step of the current precision until the device reports
// done, or until `remaining` (if >= 0) steps are used up.  The reference needs
// about 10 N steps (10 NN with the one-pivot cyclic sweeps); a run that has not
// ended after 1000 times that + 1000 steps is stopped and reported (a guard,
// never reached by a converging run).
static int launch_until_done(jkg_ctx *c, const jkg_opts *o, long &remaining, long &launched) {
  DeviceStep &d = c->d;
  hipGraphExec_t exec = c->exec[d.prec32 ? 1 : 0];
  int batch = o->steps_per_sync > 0 ? o->steps_per_sync : 16;
  const long cap = 1000L * (d.cyclic >= 2 ? (long)d.nn : (long)d.n) + 1000;
  for (;;) {
    const long nb = (remaining >= 0 && remaining < batch) ? remaining : batch;
    for (long k = 0; k < nb; ++k) HIPCHK(hipGraphLaunch(exec, c->stream));
    launched += nb;
    if (remaining >= 0) remaining -= nb;
    if (remaining == 0) return JKG_OK;
    StepState hs;
    HIPCHK(hipMemcpyAsync(&hs, d.st, sizeof(hs), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    const int done = hs.done;
    // the trace records of the steps done so far, for a caller that prints them as they come
    const int avail = std::min(hs.nrec, o->trace ? std::min(o->trace_cap, c->trace_cap) : 0);
    if (avail > c->streamed) {
      HIPCHK(hipMemcpy(o->trace + c->streamed, d.trace + c->streamed, sizeof(jkg_steprec) * (avail - c->streamed),
                       hipMemcpyDeviceToHost));
      c->streamed = avail;
    }
    if (o->progress) o->progress(o->user, hs.nrec, done);
    if (done) return JKG_OK;
    if (launched > cap) return JKG_ENOCONV;
    if (o->steps_per_sync <= 0 && batch < 64) batch *= 2;
  }
}

// The loop's own state (A, U, Z before the finalisation), in and out of a resumed loop.
struct LoopState {
  void *a, *u, *z;
  int step0;
};

// After the last step: the end event, the loop's state to the caller of a resumed
// loop, finalisation and the result record.
static int finish_run(jkg_ctx *c, jkg_result *res, const LoopState *ls = nullptr) {
  DeviceStep &d = c->d;
  HIPCHK(hipEventRecord(c->ev1, c->stream));
  if (ls) {
    const size_t bytes = (d.cplx ? 16 : 8) * (size_t)d.n * d.n;
    HIPCHK(hipMemcpyAsync(ls->a, d.A, bytes, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipMemcpyAsync(ls->u, d.U, bytes, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipMemcpyAsync(ls->z, d.Z, bytes, hipMemcpyDeviceToHost, c->stream));
  }
  finalize_launch(d, c->stream, c->sigma);
  HIPCHK(hipGetLastError());
  StepState s;
  HIPCHK(hipMemcpyAsync(&s, d.st, sizeof(s), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  float ms = 0.f;
  HIPCHK(hipEventElapsedTime(&ms, c->ev0, c->ev1));
  res->info = s.info;
  res->steps = s.nrec;
  res->seconds = ms * 1e-3;
  res->sum_sl = (int64_t)s.acc_sl;
  res->sum_it = (int64_t)s.acc_it;
  res->sum_b2 = (int64_t)s.acc_b2;
  res->sum_b4 = (int64_t)s.acc_b4;
  res->sum_abod = (int64_t)acc_abod_total(s);
  return JKG_OK;
}

// With a step limit K, exactly K step graphs are launched and then the
// step-begin kernel alone, which records the end (no idle graphs).
static int run_loop(jkg_ctx *c, const jkg_opts *o, jkg_result *res, const LoopState *ls = nullptr) {
  DeviceStep &d = c->d;
  d.prec32 = 0;
  int rc = capture_step(c, o->fused ? 1 : 0);
  if (rc) return rc;
  if (ls) {  // U and Z of the resumed loop (A came with the problem)
    const size_t bytes = (d.cplx ? 16 : 8) * (size_t)d.n * d.n;
    HIPCHK(hipMemcpyAsync(d.U, ls->u, bytes, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(d.Z, ls->z, bytes, hipMemcpyHostToDevice, c->stream));
  } else {
    init_identity_launch(d, c->stream);
  }
  HIPCHK(hipGetLastError());
  HIPCHK(hipEventRecord(c->ev0, c->stream));
  long launched = 0, remaining = (o->max_steps >= 0) ? (long)o->max_steps : -1;
  rc = launch_until_done(c, o, remaining, launched);
  if (rc) return rc;
  if (remaining == 0) {
    step_begin_launch(d, c->stream);
    HIPCHK(hipGetLastError());
  }
  return finish_run(c, res, ls);
}

// Newton-Schulz iteration for the polar factor of X in the unitary group (J
// null) or the J-unitary group: X <- X (3 I - X* X) / 2 until ||X* X - I||_F
// (measured before each update) is below 1e-12 N or stops halving.  Returns 0
// if the start is too far from the group (||.||_F >= 1/2) or the iteration
// grows.
static int polar_ns(jkg_ctx *c, void *X, const int *J, int *iters, double *res) {
  const int n = c->d.n, cx = c->d.cplx;
  hipStream_t s = c->stream;
  const int nb = ns_blocks(n);
  std::vector<double> part(nb);
  double prev = HUGE_VAL;
  *iters = 0;
  *res = HUGE_VAL;
  for (int it = 0; it < 12; ++it) {
    if (J) {
      jrows_launch(n, cx, X, J, c->wm_T, s);
      gemm_h_launch(n, cx, X, c->wm_T, c->wm_M, s);  // X^H J X
    } else {
      gemm_h_launch(n, cx, X, X, c->wm_M, s);  // X^H X
    }
    ns_step_launch(n, cx, c->wm_M, J, c->wm_T, c->wm_part, s);  // T = (3 I - S M) / 2
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(part.data(), c->wm_part, sizeof(double) * nb, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    double r2 = 0.0;
    for (double v : part) r2 += v;
    const double r = std::sqrt(r2);
    *res = r;
    if (!(r < 0.5) || (it > 0 && r > prev)) return 0;
    if (r < 1e-14) break;  // already in the group to rounding
    adj_launch(n, cx, X, c->wm_P, s);
    gemm_h_launch(n, cx, c->wm_P, c->wm_T, X, s);  // (X^H)^H T = X T
    HIPCHK(hipGetLastError());
    ++*iters;
    if (r < 1e-12 * n || r > 0.5 * prev) break;
    prev = r;
  }
  return 1;
}

struct MixedCfg;
static int run_warm(jkg_ctx *c, const jkg_opts *o, jkg_result *res, const MixedCfg &mc);

// What the mixed solver adds to a solve (all zero: the plain solver).
struct MixedCfg {
  int wmode = 0;
  double tol = 0.0, tau = 0.0;
  int ctl_mode = 0;
  const float *ctl = nullptr;
  float *ctl_trace = nullptr;
  int ctl_trace_cap = 0;
  double tau_cap = 1e-6;
  float handoff = -10.f;  // log10(TW / TW_0) of the controller's hand-back to the full ordering
  int cyclic = 0;  // round-robin sweeps in place of the dynamic ordering
  // FP32 warm start
  int warm = 0, cut32 = 0;
  double switch_tol = 0.0;
  jkg_warm_info *wi = nullptr;
};

// The FP32 warm start, its hand-over to FP64 (polar factors of U and Z in
// their groups, A = U^H Y Z) and the FP64 steps.  If a polar iteration cannot
// start (the FP32 factors are too far from their groups), the solve restarts
// in FP64 from Y (reported as a fallback).
static int run_warm(jkg_ctx *c, const jkg_opts *o, jkg_result *res, const MixedCfg &mc) {
  DeviceStep &d = c->d;
  const int fused = o->fused ? 1 : 0;
  jkg_warm_info wi;
  memset(&wi, 0, sizeof(wi));
  d.cut32 = mc.cut32;
  d.switch_tol = mc.switch_tol;
  d.prec32 = 1;
  int rc = capture_step(c, fused);
  d.prec32 = 0;
  if (!rc) rc = capture_step(c, fused);
  if (rc) return rc;
  hipEvent_t em = nullptr, es = nullptr;
  HIPCHK(hipEventCreate(&em));
  HIPCHK(hipEventCreate(&es));
  auto done_events = [&]() {
    (void)hipEventDestroy(em);
    (void)hipEventDestroy(es);
  };
  HIPCHK(hipEventRecord(c->ev0, c->stream));
  warm_begin_launch(d, c->stream);
  long launched = 0, remaining = (o->max_steps >= 0) ? (long)o->max_steps : -1;
  d.prec32 = 1;
  rc = launch_until_done(c, o, remaining, launched);
  d.prec32 = 0;
  if (rc) {
    done_events();
    return rc;
  }
  HIPCHK(hipEventRecord(em, c->stream));
  StepState s;
  HIPCHK(hipMemcpyAsync(&s, d.st, sizeof(s), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  wi.steps32 = s.nrec;
  wi.warm_end = s.warm_end ? s.warm_end : (s.done ? 3 : 4);  // 3: converged in FP32, 4: step limit
  wi.off_rel = s.off_rel;
  warm_promote_launch(d, c->stream);
  const int ok = polar_ns(c, d.U, nullptr, &wi.ns_u, &wi.res_u) && polar_ns(c, d.Z, d.J, &wi.ns_z, &wi.res_z);
  if (ok) {
    gemm_h_launch(d.n, d.cplx, d.U, d.A, c->wm_M, c->stream);  // U^H Y
    adj_launch(d.n, d.cplx, c->wm_M, c->wm_P, c->stream);
    gemm_h_launch(d.n, d.cplx, c->wm_P, d.Z, d.A, c->stream);  // (U^H Y) Z
    s.step = s.nrec;
    s.done = 0;
    s.info = 0;
    s.active = 0;
  } else {  // restart from Y in FP64
    wi.fallback = 1;
    init_identity_launch(d, c->stream);
    s = c->st_init;
    remaining = (o->max_steps >= 0) ? (long)o->max_steps : -1;
  }
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(d.st, &s, sizeof(s), hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipEventRecord(es, c->stream));
  if (remaining != 0) rc = launch_until_done(c, o, remaining, launched);
  if (rc) {
    done_events();
    return rc;
  }
  if (remaining == 0) {
    step_begin_launch(d, c->stream);
    HIPCHK(hipGetLastError());
  }
  rc = finish_run(c, res);
  float m1 = 0.f, m2 = 0.f;
  if (!rc) {
    (void)hipEventElapsedTime(&m1, c->ev0, em);
    (void)hipEventElapsedTime(&m2, em, es);
  }
  wi.seconds32 = m1 * 1e-3;
  wi.seconds_switch = m2 * 1e-3;
  done_events();
  if (mc.wi) *mc.wi = wi;
  return rc;
}

static int solve_common(jkg_ctx *c, int cplx, int n, int n2, const void *y, const int64_t *j, void *yu, void *zz,
                        double *sy, const jkg_opts *opts, jkg_result *res, const MixedCfg &mc = MixedCfg{},
                        const LoopState *ls = nullptr) {
  const int wmode = mc.wmode;
  const double tol = mc.tol, tau = mc.tau;
  if (!c) return -1;
  if (n < 3) return -2;
  if (!y) return -4;
  if (!j) return -5;
  if (!yu) return -6;
  if (!zz) return -7;
  if (!sy) return -8;
  if (!res) return -10;
  jkg_opts o;
  if (opts) o = *opts;
  else jkg_opts_default(&o);
  if (o.id_ncp < 0 || o.id_ncp > 3) return -9;
  if (o.id_tru < 0 || o.id_tru > 2) return -9;
  memset(res, 0, sizeof(*res));
  HIPCHK(hipSetDevice(c->device));
  n2 = jkg_jstep_len(n, n2);
  res->n2 = n2;
  if (n2 <= 0) return -3;
  int cap = std::max(o.trace_cap, 1);
  int rc = ensure_buffers(c, n, cplx, cap);
  if (rc) return rc;
  c->d.wmode = wmode;
  c->d.neg_tol = tol;
  if (wmode) {
    rc = ensure_mixed(c);
    if (rc) return rc;
  }
  if (mc.warm) {
    rc = ensure_warm(c);
    if (rc) return rc;
  }
  c->d.cyclic = mc.cyclic;
  c->d.frames = (c->frame_every > 0) ? c->frames_dev : nullptr;
  c->d.frame_every = c->frame_every;
  c->d.frame_size = c->frame_size;
  c->d.frame_max = c->frame_max;
  c->d.ncp3 = 0;
  if (o.id_ncp == 3) {
    rc = ensure_ncp3(c);
    if (rc) return rc;
    c->d.ncp3 = 1;
  }
  c->d.ctl_mode = mc.ctl_mode;
  c->d.ctl_trace = nullptr;
  c->d.ctl_trace_cap = 0;
  if (mc.ctl_mode) {
    rc = ensure_ctl(c, std::max(mc.ctl_trace_cap, 1));
    if (rc) return rc;
    c->d.ctl_master = c->ctl_master;
    c->d.ctl_q = c->ctl_q;
    c->d.ctl_table = c->ctl_table;
    if (mc.ctl_trace) {
      c->d.ctl_trace = c->ctl_trace;
      c->d.ctl_trace_cap = mc.ctl_trace_cap;
    }
    if (mc.ctl_mode == 1) {
      HIPCHK(hipMemcpyAsync(c->ctl_master, mc.ctl, sizeof(float) * MLP_NPARAM, hipMemcpyHostToDevice, c->stream));
      mlp_quantize_launch(c->ctl_master, c->ctl_q, c->stream);
      HIPCHK(hipGetLastError());
    } else {
      HIPCHK(hipMemcpyAsync(c->ctl_table, mc.ctl, sizeof(float) * 8, hipMemcpyHostToDevice, c->stream));
    }
    HIPCHK(hipStreamSynchronize(c->stream));
  }
  // TT: the number of trigonometric pairs (the reference's R%TRU result)
  int npos = 0, nneg = 0;
  rc = upload_problem(c, n, n2, y, j, &npos, &nneg, o.max_steps, c->trace_cap, tau, mc.tau_cap, mc.handoff,
                      ls ? ls->step0 : 0);
  if (rc) return rc;
  res->npos = npos;
  res->nneg = nneg;
  res->tt = (int)((int64_t)npos * (npos - 1) / 2 + (int64_t)nneg * (nneg - 1) / 2);
  c->streamed = 0;
  rc = mc.warm ? run_warm(c, &o, res, mc) : run_loop(c, &o, res, ls);
  if (rc) return rc;
  DeviceStep &d = c->d;
  if (o.trace && o.trace_cap > 0) {
    int cnt = std::min(res->steps, o.trace_cap);
    if (cnt > 0) HIPCHK(hipMemcpy(o.trace, d.trace, sizeof(jkg_steprec) * cnt, hipMemcpyDeviceToHost));
  }
  if (c->d.frames && c->frames_host) {
    // the steps s = 0 mod every among this run's step0 .. step0 + steps - 1
    const int s0 = ls ? ls->step0 : 0, ev = c->frame_every;
    const int first = (s0 + ev - 1) / ev, last = (s0 + res->steps - 1) / ev;
    const int cnt = res->steps > 0 ? std::min(std::max(last - first + 1, 0), c->frame_max) : 0;
    if (cnt > 0)
      HIPCHK(hipMemcpy(c->frames_host, c->frames_dev, (size_t)cnt * c->frame_size * c->frame_size,
                       hipMemcpyDeviceToHost));
    if (c->frames_count) *c->frames_count = cnt;
  }
  if (mc.ctl_mode && mc.ctl_trace && mc.ctl_trace_cap > 0) {
    int cnt = std::min(res->steps, mc.ctl_trace_cap);
    if (cnt > 0)
      HIPCHK(hipMemcpy(mc.ctl_trace, d.ctl_trace, sizeof(float) * CTL_REC * cnt, hipMemcpyDeviceToHost));
  }
  // written also after a failed step (res->info < 0): the loop's finalisation ran, and the
  // reference writes U, Z and S after its ERROR line (src/D/djk.F90:86-114, dstep.F90:594-603)
  const size_t esz = cplx ? 16 : 8;
  HIPCHK(hipMemcpy(yu, d.U, esz * n * n, hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(zz, d.Z, esz * n * n, hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(sy, c->sigma, sizeof(double) * n, hipMemcpyDeviceToHost));
  return JKG_OK;
}

extern "C" int jkg_dsolve(jkg_ctx *c, int n, int n2, const double *y, const int64_t *j, double *yu, double *zz,
                          double *sy, const jkg_opts *opts, jkg_result *res) {
  return solve_common(c, 0, n, n2, y, j, yu, zz, sy, opts, res);
}

// DSTEP_LOOP resumed from its own state (src/D/dstep.F90:472-604 and src/Z/zstep.F90
// after the identity set-up): (a, u, z) in place of (Y, I, I), steps counted from step0.
static int solve_from(jkg_ctx *c, int cplx, int n, int n2, void *a, void *u, void *z, const int64_t *j, int step0,
                      void *yu, void *zz, double *sy, const jkg_opts *opts, jkg_result *res) {
  // negative argument positions, as for jkg_dsolve
  if (!c) return -1;
  if (n < 3) return -2;
  if (!a) return -4;
  if (!u) return -5;
  if (!z) return -6;
  if (!j) return -7;
  if (step0 < 0) return -8;
  if (!yu) return -9;
  if (!zz) return -10;
  if (!sy) return -11;
  if (!res) return -13;
  LoopState ls{a, u, z, step0};
  return solve_common(c, cplx, n, n2, a, j, yu, zz, sy, opts, res, MixedCfg{}, &ls);
}

extern "C" int jkg_dsolve_from(jkg_ctx *c, int n, int n2, double *a, double *u, double *z, const int64_t *j,
                               int step0, double *yu, double *zz, double *sy, const jkg_opts *opts,
                               jkg_result *res) {
  return solve_from(c, 0, n, n2, a, u, z, j, step0, yu, zz, sy, opts, res);
}

extern "C" int jkg_zsolve_from(jkg_ctx *c, int n, int n2, double *a, double *u, double *z, const int64_t *j,
                               int step0, double *yu, double *zz, double *sy, const jkg_opts *opts,
                               jkg_result *res) {
  return solve_from(c, 1, n, n2, a, u, z, j, step0, yu, zz, sy, opts, res);
}

extern "C" void jkg_mixed_opts_default(jkg_mixed_opts *m) {
  memset(m, 0, sizeof(*m));
  m->weights = JKG_W_GRAM;
  m->switch_tol = 1e-5;
  m->key_bits = 64;
  m->tol = 0x1p-52;  // double epsilon
}

extern "C" int jkg_dsolve_mixed(jkg_ctx *c, int n, int n2, const double *y, const int64_t *j, double *yu,
                                double *zz, double *sy, const jkg_opts *opts, const jkg_mixed_opts *mopts,
                                jkg_result *res) {
  jkg_mixed_opts m;
  if (mopts) m = *mopts;
  else jkg_mixed_opts_default(&m);
  if (m.weights != JKG_W_PAIR && m.weights != JKG_W_GRAM) return -10;
  if (!(m.tol >= 0.0) || !(m.threshold >= 0.0 && m.threshold < 1.0)) return -10;
  if (m.controller < 0 || m.controller > 2 || (m.controller && !m.ctl) || m.ctl_trace_cap < 0) return -10;
  if (!(m.ctl_tau_max >= 0.0 && m.ctl_tau_max <= 1e-6)) return -10;
  if (!(m.ctl_handoff >= -16.f && m.ctl_handoff <= 0.f)) return -10;
  if (m.ordering < 0 || m.ordering > 3) return -10;
  if (m.fp32_warm < 0 || m.fp32_warm > 1 || (m.key_bits != 32 && m.key_bits != 64)) return -10;
  if (m.fp32_warm && !(m.switch_tol > 0.0 && m.switch_tol < 1.0)) return -10;
  if (m.key_bits == 32 && !m.fp32_warm) return JKG_ENOTSUP;  // 32-bit keys: the warm start only
  if (m.weights == JKG_W_PAIR && (m.tol != 0.0 || m.controller || m.fp32_warm)) return JKG_ENOTSUP;
  if (m.ordering != 0 && (m.controller || m.threshold != 0.0 || m.fp32_warm)) return JKG_ENOTSUP;
  MixedCfg mc;
  mc.warm = m.fp32_warm;
  mc.cut32 = m.key_bits == 32;
  mc.switch_tol = m.switch_tol;
  mc.wi = m.warm;
  mc.cyclic = m.ordering;
  mc.wmode = m.weights == JKG_W_GRAM ? 1 : 0;
  mc.tol = m.tol;
  mc.tau = m.controller ? 0.0 : m.threshold;
  mc.ctl_mode = m.controller;
  mc.ctl = m.ctl;
  mc.ctl_trace = m.ctl_trace;
  mc.ctl_trace_cap = m.ctl_trace ? m.ctl_trace_cap : 0;
  mc.tau_cap = m.ctl_tau_max > 0.0 ? m.ctl_tau_max : 1e-6;
  mc.handoff = m.ctl_handoff != 0.f ? m.ctl_handoff : -10.f;
  return solve_common(c, 0, n, n2, y, j, yu, zz, sy, opts, res, mc);
}

extern "C" int jkg_zsolve(jkg_ctx *c, int n, int n2, const double *y, const int64_t *j, double *yu, double *zz,
                          double *sy, const jkg_opts *opts, jkg_result *res) {
  return solve_common(c, 1, n, n2, y, j, yu, zz, sy, opts, res);
}

extern "C" int jkg_zsolve_mixed(jkg_ctx *c, int n, int n2, const double *y, const int64_t *j, double *yu,
                                double *zz, double *sy, const jkg_opts *opts, const jkg_mixed_opts *mopts,
                                jkg_result *res) {
  jkg_mixed_opts m;
  if (mopts) m = *mopts;
  else jkg_mixed_opts_default(&m);
  if (m.weights != JKG_W_PAIR && m.weights != JKG_W_GRAM) return -10;
  if (!(m.tol >= 0.0) || !(m.threshold >= 0.0 && m.threshold < 1.0)) return -10;
  if (m.controller < 0 || m.controller > 2 || m.ctl_trace_cap < 0) return -10;
  if (m.fp32_warm < 0 || m.fp32_warm > 1 || (m.key_bits != 32 && m.key_bits != 64)) return -10;
  if (m.fp32_warm && !(m.switch_tol > 0.0 && m.switch_tol < 1.0)) return -10;
  if (m.controller || m.ordering != 0) return JKG_ENOTSUP;  // real-only options
  if (m.key_bits == 32 && !m.fp32_warm) return JKG_ENOTSUP;  // 32-bit keys: the warm start only
  if (m.weights == JKG_W_PAIR && (m.tol != 0.0 || m.fp32_warm)) return JKG_ENOTSUP;
  MixedCfg mc;
  mc.wmode = m.weights == JKG_W_GRAM ? 1 : 0;
  mc.tol = m.tol;
  mc.tau = m.threshold;
  mc.warm = m.fp32_warm;
  mc.cut32 = m.key_bits == 32;
  mc.switch_tol = m.switch_tol;
  mc.wi = m.warm;
  return solve_common(c, 1, n, n2, y, j, yu, zz, sy, opts, res, mc);
}

// ---------------------------------------------------------------- allocation-free solve
// jkg_dsolve_noalloc: jkg_dsolve for callers that cannot allocate device memory
// (code under HIP graph capture, serving loops, harnesses that forbid it).  Every
// buffer of the step -- A, U, Z, keys, lists, pivots, loop state and trace -- lives
// in `scratch`, a file-scope device array of NOALLOC_BYTES held by each GPU until
// the process ends; the caller passes Y, YU, ZZ and SY as device buffers.  All calls
// on one GPU share the array (it keeps the last call's contents; every part a step
// reads is written first in the same call), so calls that overlap on one GPU race.
constexpr int NOALLOC_NMAX = 256;
constexpr int NOALLOC_TRACE = 64 * NOALLOC_NMAX + 64;
constexpr size_t NOALLOC_BYTES = (size_t)6 << 20;
static __device__ __align__(256) unsigned char scratch[NOALLOC_BYTES];
static void *g_scratch_ptr[64];  // per device: the array's device address (hipGetSymbolAddress)
static bool g_scratch_set[64];

// Once per device, with that device current (called by jkg_create).
static int noalloc_setup(int device) {
  if (device < 0 || device >= 64) return JKG_ENOTSUP;
  if (g_scratch_set[device]) return JKG_OK;
  void *p = nullptr;
  HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  g_scratch_ptr[device] = p;
  g_scratch_set[device] = true;
  return JKG_OK;
}

// The step's buffers for an n x n real problem laid out from base (256-byte aligned
// pieces); returns the bytes used.  base == nullptr: the size only.
static size_t noalloc_layout(DeviceStep &d, double **sigma, char *base, int n, int trace_cap) {
  size_t off = 0;
  auto take = [&](auto **p, size_t count) {
    using T = typename std::remove_reference<decltype(**p)>::type;
    off = (off + 255) & ~(size_t)255;
    if (base) *p = (T *)(base + off);
    off += sizeof(T) * (count ? count : 1);
  };
  const size_t nn = (size_t)n * (n - 1) / 2;
  const int n2max = n / 2 > 0 ? n / 2 : 1;
  memset(&d, 0, sizeof(d));
  d.n = n;
  d.lda = n;
  d.cplx = 0;
  d.nn = (int)nn;
  d.n2max = n2max;
  take((double **)&d.A, (size_t)n * n);
  take((double **)&d.U, (size_t)n * n);
  take((double **)&d.Z, (size_t)n * n);
  take(&d.J, n);
  take(&d.plus, n);
  take(&d.minus, n);
  take(&d.sort.k0, nn);
  take(&d.sort.v0, nn + LIST_PAD);
  take(&d.sort.k1, 2 * nn);
  take(&d.sort.k2, 2 * nn);
  take(&d.sort.v1, nn);
  take(&d.sort.v2, nn);
  take(&d.sort.d1, sort_digit_bytes((int)nn));
  take(&d.sort.d2, sort_digit_bytes((int)nn));
  take((char **)&d.sort.scratch, sort_scratch_bytes((int)nn));
  for (int k = 0; k < 2; ++k) {
    take(&d.lists.val[k], nn + LIST_PAD);
    take(&d.lists.pos[k], nn + LIST_PAD);
  }
  take(&d.fcount, select_count_ints((int)nn));
  take(&d.used, n);
  take(&d.piv_p, n2max);
  take(&d.piv_q, n2max);
  take(&d.piv_pos, n2max);
  take(&d.piv_info, n2max);
  take(&d.piv_w, n2max);
  take((double **)&d.piv_V, 4 * (size_t)n2max);
  take((double **)&d.piv_W, 4 * (size_t)n2max);
  take(&d.piv_d, 2 * (size_t)n2max);
  take(&d.rot_row, n2max);
  take(&d.rot_col, n2max);
  take(&d.colpiv, n);
  take(&d.st, 1);
  take(&d.trace, trace_cap);
  take(sigma, n);
  return off;
}

extern "C" int jkg_dsolve_noalloc(jkg_ctx *c, int n, int n2, const double *y, const int64_t *j, double *yu,
                                  double *zz, double *sy, const jkg_opts *opts, jkg_result *res) {
  if (!c) return -1;
  if (n < 3) return -2;
  if (!y) return -4;
  if (!j) return -5;
  if (!yu) return -6;
  if (!zz) return -7;
  if (!sy) return -8;
  if (!res) return -10;
  if (n > NOALLOC_NMAX) return JKG_ENOTSUP;
  jkg_opts o;
  if (opts) o = *opts;
  else jkg_opts_default(&o);
  if (o.id_ncp < 0 || o.id_ncp > 2 || o.id_tru < 0 || o.id_tru > 2) return -9;
  memset(res, 0, sizeof(*res));
  n2 = jkg_jstep_len(n, n2);
  res->n2 = n2;
  if (n2 <= 0) return -3;
  if (!g_scratch_set[c->device]) return JKG_ENOTSUP;
  HIPCHK(hipSetDevice(c->device));
  const int trace_cap = std::min(std::max(o.trace_cap, 1), NOALLOC_TRACE);
  DeviceStep d;
  double *sigma = nullptr;
  if (noalloc_layout(d, &sigma, nullptr, n, trace_cap) > NOALLOC_BYTES) return JKG_ENOTSUP;
  noalloc_layout(d, &sigma, (char *)g_scratch_ptr[c->device], n, trace_cap);
  d.fused = o.fused ? 1 : 0;
  d.select_limit = c->sel_limit;
  d.sel_rounds = c->sel_rounds > 0 ? c->sel_rounds : select_rounds_for((size_t)d.nn);
  // the problem: J and its classes from fixed host arrays (no allocation), Y copied on the device
  int J32[NOALLOC_NMAX], pl[NOALLOC_NMAX], mi[NOALLOC_NMAX];
  int npl = 0, nmi = 0;
  for (int i = 0; i < n; ++i) {
    if (j[i] != 1 && j[i] != -1) return JKG_EJ;
    J32[i] = (int)j[i];
    if (j[i] == 1) pl[npl++] = i;
    else mi[nmi++] = i;
  }
  d.npl = npl;
  d.nmi = nmi;
  hipStream_t s = c->stream;
  HIPCHK(hipMemcpyAsync(d.A, y, sizeof(double) * n * n, hipMemcpyDeviceToDevice, s));
  HIPCHK(hipMemcpyAsync(d.J, J32, sizeof(int) * n, hipMemcpyHostToDevice, s));
  if (npl) HIPCHK(hipMemcpyAsync(d.plus, pl, sizeof(int) * npl, hipMemcpyHostToDevice, s));
  if (nmi) HIPCHK(hipMemcpyAsync(d.minus, mi, sizeof(int) * nmi, hipMemcpyHostToDevice, s));
  StepState st{};
  st.max_steps = o.max_steps;
  st.n2 = n2;
  st.n2max = n2;
  st.trace_cap = trace_cap;
  st.n2f = 1.f;
  st.ltau = -16.f;
  st.off_best = HUGE_VAL;
  st.ltau_cap = -6.f;
  st.lhandoff = -10.f;
  st.ynorm2 = 1.0;
  HIPCHK(hipMemcpyAsync(d.st, &st, sizeof(st), hipMemcpyHostToDevice, s));
  init_identity_launch(d, s);
  HIPCHK(hipGetLastError());
  res->npos = npl;
  res->nneg = nmi;
  res->tt = (int)((int64_t)npl * (npl - 1) / 2 + (int64_t)nmi * (nmi - 1) / 2);
  // the steps, launched one by one (no graph: instantiating one may allocate), the done flag
  // polled every `batch` steps
  const int batch = o.steps_per_sync > 0 ? o.steps_per_sync : 32;
  const long cap = 1000L * n + 1000;
  long launched = 0, remaining = (o.max_steps >= 0) ? (long)o.max_steps : -1;
  const auto t0 = std::chrono::steady_clock::now();
  for (;;) {
    const long nb = (remaining >= 0 && remaining < batch) ? remaining : batch;
    for (long k = 0; k < nb; ++k) step_launch(d, s);
    HIPCHK(hipGetLastError());
    launched += nb;
    if (remaining >= 0) remaining -= nb;
    if (remaining == 0) {
      step_begin_launch(d, s);  // records the end of a step-limited run, as run_loop does
      break;
    }
    HIPCHK(hipMemcpyAsync(&st, d.st, sizeof(st), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    if (o.progress) o.progress(o.user, st.nrec, st.done);
    if (st.done) break;
    if (launched > cap) return JKG_ENOCONV;
  }
  finalize_launch(d, s, sigma);
  HIPCHK(hipMemcpyAsync(yu, d.U, sizeof(double) * n * n, hipMemcpyDeviceToDevice, s));
  HIPCHK(hipMemcpyAsync(zz, d.Z, sizeof(double) * n * n, hipMemcpyDeviceToDevice, s));
  HIPCHK(hipMemcpyAsync(sy, sigma, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
  HIPCHK(hipMemcpyAsync(&st, d.st, sizeof(st), hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  if (o.trace && o.trace_cap > 0) {
    const int cnt = std::min(st.nrec, std::min(o.trace_cap, trace_cap));
    if (cnt > 0) HIPCHK(hipMemcpy(o.trace, d.trace, sizeof(jkg_steprec) * cnt, hipMemcpyDeviceToHost));
  }
  res->info = st.info;
  res->steps = st.nrec;
  res->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  res->sum_sl = (int64_t)st.acc_sl;
  res->sum_it = (int64_t)st.acc_it;
  res->sum_b2 = (int64_t)st.acc_b2;
  res->sum_b4 = (int64_t)st.acc_b4;
  res->sum_abod = (int64_t)acc_abod_total(st);
  return JKG_OK;
}

// ---------------------------------------------------------------- step level
template <bool F>
__global__ void k_dhsvd2_batch(int count, const double *a, const int *j, double th1, double *u, double *z,
                               double *dg, int *info) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  double A[4], U[4] = {0, 0, 0, 0}, Z[4] = {0, 0, 0, 0};
  for (int k = 0; k < 4; ++k) A[k] = a[4 * i + k];
  int J[2] = {j[2 * i], j[2 * i + 1]};
  int inf = dhsvd2<F>(A, J, U, Z, th1);
  for (int k = 0; k < 4; ++k) {
    u[4 * i + k] = U[k];
    z[4 * i + k] = Z[k];
  }
  dg[2 * i] = A[0];
  dg[2 * i + 1] = A[3];
  info[i] = inf;
}

template <bool F>
__global__ void k_zhsvd2_batch(int count, const dz *a, const int *j, double th1, dz *u, dz *z, double *dg,
                               int *info) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  dz A[4], U[4], Z[4];
  for (int k = 0; k < 4; ++k) {
    A[k] = a[4 * i + k];
    U[k] = zmk(0, 0);
    Z[k] = zmk(0, 0);
  }
  int J[2] = {j[2 * i], j[2 * i + 1]};
  int inf = zhsvd2<F>(A, J, U, Z, th1);
  for (int k = 0; k < 4; ++k) {
    u[4 * i + k] = U[k];
    z[4 * i + k] = Z[k];
  }
  dg[2 * i] = A[0].re;
  dg[2 * i + 1] = A[3].re;
  info[i] = inf;
}

template <class T, bool CPLX>
static int hsvd2_batch(jkg_ctx *c, int count, const double *a, const int32_t *j, double th1, int fused, double *u,
                       double *z, double *dg, int32_t *info) {
  if (!c) return -1;
  if (count < 0) return -2;
  if (count == 0) return JKG_OK;
  HIPCHK(hipSetDevice(c->device));
  const size_t esz = CPLX ? 16 : 8;
  const size_t ba = esz * 4 * count, bj = sizeof(int) * 2 * count, bd = sizeof(double) * 2 * count,
               bi = sizeof(int) * count;
  int rc = tmp_bytes(c, 3 * ba + bj + bd + bi + 256);
  if (rc) return rc;
  char *base = (char *)c->tmp;
  char *da = base, *du = da + ba, *dzp = du + ba, *dj = dzp + ba, *dd = dj + bj, *di = dd + bd;
  HIPCHK(hipMemcpyAsync(da, a, ba, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(dj, j, bj, hipMemcpyHostToDevice, c->stream));
  const int g = (count + 127) / 128;
  if (CPLX) {
    if (fused)
      hipLaunchKernelGGL(k_zhsvd2_batch<true>, dim3(g), dim3(128), 0, c->stream, count, (const dz *)da,
                         (const int *)dj, th1, (dz *)du, (dz *)dzp, (double *)dd, (int *)di);
    else
      hipLaunchKernelGGL(k_zhsvd2_batch<false>, dim3(g), dim3(128), 0, c->stream, count, (const dz *)da,
                         (const int *)dj, th1, (dz *)du, (dz *)dzp, (double *)dd, (int *)di);
  } else {
    if (fused)
      hipLaunchKernelGGL(k_dhsvd2_batch<true>, dim3(g), dim3(128), 0, c->stream, count, (const double *)da,
                         (const int *)dj, th1, (double *)du, (double *)dzp, (double *)dd, (int *)di);
    else
      hipLaunchKernelGGL(k_dhsvd2_batch<false>, dim3(g), dim3(128), 0, c->stream, count, (const double *)da,
                         (const int *)dj, th1, (double *)du, (double *)dzp, (double *)dd, (int *)di);
  }
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(u, du, ba, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipMemcpyAsync(z, dzp, ba, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipMemcpyAsync(dg, dd, bd, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipMemcpyAsync(info, di, bi, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return JKG_OK;
}

extern "C" int jkg_dhsvd2(jkg_ctx *c, int count, const double *a, const int32_t *j, double th1, int fused, double *u,
                          double *z, double *d, int32_t *info) {
  return hsvd2_batch<double, false>(c, count, a, j, th1, fused, u, z, d, info);
}
extern "C" int jkg_zhsvd2(jkg_ctx *c, int count, const double *a, const int32_t *j, double th1, int fused, double *u,
                          double *z, double *d, int32_t *info) {
  return hsvd2_batch<dz, true>(c, count, a, j, th1, fused, u, z, d, info);
}

// Prepares the device state of one step on (a, j) without running it.
static int prepare_single(jkg_ctx *c, int n, int n2, const void *a, const int64_t *j, int fused, int cplx = 0) {
  if (n < 3) return -2;
  HIPCHK(hipSetDevice(c->device));
  int rc = ensure_buffers(c, n, cplx, 1);
  if (rc) return rc;
  c->d.fused = fused;
  c->d.wmode = 0;
  c->d.neg_tol = 0.0;
  c->d.ctl_mode = 0;
  c->d.ncp3 = 0;
  c->d.cyclic = 0;
  c->d.frames = nullptr;
  int npos, nneg;
  return upload_problem(c, n, jkg_jstep_len(n, n2), a, j, &npos, &nneg, -1, c->trace_cap);
}

extern "C" int jkg_dweights(jkg_ctx *c, int n, const double *a, const int64_t *j, int fused, uint64_t *keys) {
  if (!c) return -1;
  int rc = prepare_single(c, n, 0, a, j, fused);
  if (rc) return rc;
  step_begin_launch(c->d, c->stream);
  weights_launch(c->d, c->stream);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(keys, c->d.sort.k0, sizeof(uint64_t) * c->d.nn, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return JKG_OK;
}

extern "C" int jkg_dweights_mixed(jkg_ctx *c, int n, const double *a, const int64_t *j, int fused,
                                  uint64_t *keys) {
  if (!c) return -1;
  int rc = prepare_single(c, n, 0, a, j, fused);
  if (rc) return rc;
  rc = ensure_mixed(c);
  if (rc) return rc;
  c->d.wmode = 1;
  step_begin_launch(c->d, c->stream);
  weights_launch(c->d, c->stream);
  c->d.wmode = 0;
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(keys, c->d.sort.k0, sizeof(uint64_t) * c->d.nn, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return JKG_OK;
}

extern "C" int jkg_zweights(jkg_ctx *c, int n, const double *a, const int64_t *j, int fused, uint64_t *keys) {
  if (!c) return -1;
  int rc = prepare_single(c, n, 0, a, j, fused, 1);
  if (rc) return rc;
  step_begin_launch(c->d, c->stream);
  weights_launch(c->d, c->stream);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(keys, c->d.sort.k0, sizeof(uint64_t) * c->d.nn, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return JKG_OK;
}

extern "C" int jkg_zweights_mixed(jkg_ctx *c, int n, const double *a, const int64_t *j, int fused,
                                  uint64_t *keys) {
  if (!c) return -1;
  int rc = prepare_single(c, n, 0, a, j, fused, 1);
  if (rc) return rc;
  rc = ensure_mixed(c);
  if (rc) return rc;
  c->d.wmode = 1;
  step_begin_launch(c->d, c->stream);
  weights_launch(c->d, c->stream);
  c->d.wmode = 0;
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(keys, c->d.sort.k0, sizeof(uint64_t) * c->d.nn, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return JKG_OK;
}


extern "C" int jkg_dbuild(jkg_ctx *c, int n, int n2, const double *a, const int64_t *j, int fused, int32_t *it,
                          jkg_aw *sorted, int32_t *sl, int32_t *step) {
  if (!c) return -1;
  int rc = prepare_single(c, n, n2, a, j, fused);
  if (rc) return rc;
  DeviceStep &d = c->d;
  step_begin_launch(d, c->stream);
  weights_launch(d, c->stream);
  sort_pairs_launch(d.sort, &d.st->nn_sort, d.nn, c->stream, d.n);
  select_launch(d, c->stream);
  HIPCHK(hipGetLastError());
  StepState s;
  HIPCHK(hipMemcpyAsync(&s, d.st, sizeof(s), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  // the sorted packed pairs; the keys stay in pair-index order
  std::vector<uint64_t> keys(d.nn);
  std::vector<uint32_t> vals(s.it);
  HIPCHK(hipMemcpy(keys.data(), d.sort.k0, sizeof(uint64_t) * d.nn, hipMemcpyDeviceToHost));
  if (s.it > 0) HIPCHK(hipMemcpy(vals.data(), d.sort.v0, sizeof(uint32_t) * s.it, hipMemcpyDeviceToHost));
  for (int i = 0; i < s.it; ++i) {
    const int p = pq_p(vals[i]) + 1, q = pq_q(vals[i]) + 1;
    sorted[i].w = jkg_key_weight(keys[pair_index(n, p, q)]);
    sorted[i].p = p;
    sorted[i].q = q;
    sorted[i].b = q - p;
    sorted[i].i = (int32_t)pair_index(n, p, q);
  }
  *it = s.it;
  *sl = s.sl;
  if (s.sl > 0) HIPCHK(hipMemcpy(step, d.piv_pos, sizeof(int) * s.sl, hipMemcpyDeviceToHost));
  return JKG_OK;
}

extern "C" int jkg_dstep(jkg_ctx *c, int n, int n2, double *a, const int64_t *j, double *u, double *z, int fused,
                         jkg_steprec *rec) {
  if (!c) return -1;
  int rc = prepare_single(c, n, n2, a, j, fused);
  if (rc) return rc;
  DeviceStep &d = c->d;
  HIPCHK(hipMemcpyAsync(d.U, u, sizeof(double) * n * n, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(d.Z, z, sizeof(double) * n * n, hipMemcpyHostToDevice, c->stream));
  step_launch(d, c->stream);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(a, d.A, sizeof(double) * n * n, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipMemcpyAsync(u, d.U, sizeof(double) * n * n, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipMemcpyAsync(z, d.Z, sizeof(double) * n * n, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipMemcpyAsync(rec, d.trace, sizeof(jkg_steprec), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return JKG_OK;
}

// ---------------------------------------------------------------- DPROC plug-ins
__global__ void k_plugin_pairkeys(int n, int nn, const int *pq, uint64_t *k0) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nn) k0[i] = pair_index(n, pq[2 * i], pq[2 * i + 1]);
}
__global__ void k_plugin_wkeys(int nn, const double *w, const uint32_t *perm, uint64_t *k0) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nn) k0[i] = weight_key(w[perm[i]]);
}

// ASRT: sort by (pair index) then stably by weight key == AW_CMP order.
extern "C" int jkg_srt(jkg_ctx *c, int nn, int nm, jkg_aw *dz, int64_t *info) {
  if (!c) return -1;
  if (nn < 0) { *info = -1; return -2; }
  if (nm < nn) { *info = -2; return -3; }
  *info = 0;
  if (nn == 0) return JKG_OK;
  HIPCHK(hipSetDevice(c->device));
  int n = 0;
  for (int i = 0; i < nn; ++i) n = std::max(n, (int)dz[i].q);
  std::vector<int> pq(2 * (size_t)nn);
  std::vector<double> w(nn);
  for (int i = 0; i < nn; ++i) {
    pq[2 * i] = dz[i].p;
    pq[2 * i + 1] = dz[i].q;
    w[i] = dz[i].w;
    if (dz[i].p < 1 || dz[i].q <= dz[i].p) return -4;
  }
  // buffers: pair keys, weight keys, the first and the final permutation, the sort's records and
  // scratch, pq, w, n
  const size_t bk = sizeof(uint64_t) * nn, bv = sizeof(uint32_t) * nn;
  const size_t bs = sort_scratch_bytes(nn), bpq = sizeof(int) * 2 * nn, bw = sizeof(double) * nn;
  const size_t bdg = sort_digit_bytes(nn);
  int rc = tmp_bytes(c, 6 * bk + 4 * bv + 2 * bdg + bs + bpq + bw + 256 * 15);  // 14 pieces, each 256-byte aligned
  if (rc) return rc;
  char *p = (char *)c->tmp;
  auto take = [&](size_t b) {
    char *r = p;
    p += (b + 255) & ~(size_t)255;
    return r;
  };
  SortBuffers sb;
  sb.k0 = (uint64_t *)take(bk);
  uint64_t *wk = (uint64_t *)take(bk);
  sb.v0 = (uint32_t *)take(bv);
  uint32_t *perm2 = (uint32_t *)take(bv);
  sb.k1 = (uint64_t *)take(2 * bk);
  sb.k2 = (uint64_t *)take(2 * bk);
  sb.v1 = (uint32_t *)take(bv);
  sb.v2 = (uint32_t *)take(bv);
  sb.d1 = (uint8_t *)take(bdg);
  sb.d2 = (uint8_t *)take(bdg);
  sb.vin = nullptr;
  sb.scratch = (int *)take(bs);
  int *dpq = (int *)take(bpq);
  double *dw = (double *)take(bw);
  int *dn = (int *)take(sizeof(int));
  HIPCHK(hipMemcpyAsync(dpq, pq.data(), bpq, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(dw, w.data(), bw, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(dn, &nn, sizeof(int), hipMemcpyHostToDevice, c->stream));
  const int g = (nn + 255) / 256;
  // first by pair index, then stably by weight key over that permutation
  hipLaunchKernelGGL(k_plugin_pairkeys, dim3(g), dim3(256), 0, c->stream, n, nn, dpq, sb.k0);
  sort_pairs_launch(sb, dn, nn, c->stream, 0);
  hipLaunchKernelGGL(k_plugin_wkeys, dim3(g), dim3(256), 0, c->stream, nn, dw, sb.v0, wk);
  SortBuffers s2 = sb;
  s2.k0 = wk;
  s2.vin = sb.v0;
  s2.v0 = perm2;
  sort_pairs_launch(s2, dn, nn, c->stream, 0);
  HIPCHK(hipGetLastError());
  std::vector<uint32_t> perm(nn);
  HIPCHK(hipMemcpyAsync(perm.data(), perm2, bv, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  std::vector<jkg_aw> out(nn);
  for (int i = 0; i < nn; ++i) out[i] = dz[perm[i]];
  memcpy(dz, out.data(), sizeof(jkg_aw) * nn);
  return JKG_OK;
}

// ANCP (AW_NCP1/AW_NCP2): the device's greedy walk over a given sorted list.
// weight_key() on the host: IEEE bits mapped to an unsigned key that grows as w decreases (-0 -> +0)
static uint64_t host_weight_key(double w) {
  uint64_t u;
  memcpy(&u, &w, 8);
  if (w == 0.0) u = 0;
  const uint64_t ord = (u & 0x8000000000000000ull) ? ~u : (u | 0x8000000000000000ull);
  return ~ord;
}

extern "C" int jkg_ncp(jkg_ctx *c, int n, int nn, int nm, jkg_aw *dz, int n2, int32_t *sl, int32_t *step,
                       int64_t *info) {
  if (!c) return -1;
  *sl = 0;
  *info = 0;
  if (n < 0) { *info = -1; return -2; }
  if (nn < 0) { *info = -2; return -3; }
  if (nm < nn) { *info = -3; return -4; }
  if (n2 < 0) { *info = -5; return -7; }
  if (n == 0 || nn == 0 || n2 == 0) return JKG_OK;
  for (int i = 0; i < nn; ++i)
    if (!(dz[i].w == dz[i].w) || dz[i].p < 1 || dz[i].q <= dz[i].p || dz[i].q > n) return -5;
  HIPCHK(hipSetDevice(c->device));
  int nb = std::max(n, 3);
  int rc = ensure_buffers(c, nb, 0, 1);
  if (rc) return rc;
  if ((size_t)nn > (size_t)c->d.nn || n2 > c->d.n2max) {
    // the candidate list holds at most n(n-1)/2 distinct pairs
    return -3;
  }
  DeviceStep &d = c->d;
  std::vector<uint32_t> val(nn), pos(nn);
  for (int i = 0; i < nn; ++i) {
    val[i] = pack_pq(dz[i].p - 1, dz[i].q - 1);
    pos[i] = (uint32_t)(i + 1);
  }
  HIPCHK(hipMemcpyAsync(d.lists.val[1], val.data(), 4 * (size_t)nn, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(d.lists.pos[1], pos.data(), 4 * (size_t)nn, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemsetAsync(d.used, 0, nb, c->stream));
  StepState s{};
  s.n2 = n2;
  s.sel_round = 1;  // the walk starts with round 1 over the given list
  s.next_len = nn;
  s.active = 1;
  HIPCHK(hipMemcpyAsync(d.st, &s, sizeof(s), hipMemcpyHostToDevice, c->stream));
  // the lists of round 1 are read from index (1-1)&1 = 0: swap the pair; round 0 (the sorted
  // list) is skipped, so at least two rounds
  DeviceStep dd = d;
  if (dd.sel_rounds < 2) dd.sel_rounds = 2;
  std::swap(dd.lists.val[0], dd.lists.val[1]);
  std::swap(dd.lists.pos[0], dd.lists.pos[1]);
  select_launch(dd, c->stream);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(&s, d.st, sizeof(s), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  *sl = s.sl;
  if (s.sl > 0) HIPCHK(hipMemcpy(step, d.piv_pos, sizeof(int) * s.sl, hipMemcpyDeviceToHost));
  return JKG_OK;
}

// Frames of |A| for later solves on this context (see k_frame): every `every`-th step
// (<= 0: off), size x size bytes each, at most max_frames; after each solve they are
// copied to frames (max_frames * size^2 bytes) and their number to *count.
extern "C" int jkg_set_frames(jkg_ctx *c, int every, int size, int max_frames, uint8_t *frames,
                              int32_t *count) {
  if (!c) return -1;
  HIPCHK(hipSetDevice(c->device));
  if (c->frames_dev) (void)hipFree(c->frames_dev);
  c->frames_dev = nullptr;
  c->frame_every = 0;
  c->frames_host = nullptr;
  c->frames_count = nullptr;
  if (every <= 0) return JKG_OK;
  if (size < 1 || size > 1024) return -3;
  if (max_frames < 1) return -4;
  if (!frames) return -5;
  if (hipMalloc(&c->frames_dev, (size_t)max_frames * size * size) != hipSuccess) {
    (void)hipGetLastError();
    c->frames_dev = nullptr;
    return JKG_ENOMEM;
  }
  c->frame_every = every;
  c->frame_size = size;
  c->frame_max = max_frames;
  c->frames_host = frames;
  c->frames_count = count;
  return JKG_OK;
}

// The selection schedule of later solves on this context: `rounds` walks per step (0: by size),
// each but the last stopping after `limit` entries of its list, followed by a
// whole-GPU pass that drops the entries with a used index.  The pivots chosen do
// not depend on it (the filter removes only entries the walk would skip).
extern "C" int jkg_set_select(jkg_ctx *c, int rounds, int limit) {
  if (!c) return -1;
  if (rounds < 0 || rounds > 16) return -2;
  if (limit < 2048) return -3;
  c->sel_rounds = rounds;
  c->sel_limit = limit;
  c->d.sel_rounds = rounds > 0 ? rounds : select_rounds_for((size_t)c->d.nn);
  c->d.select_limit = limit;
  return JKG_OK;
}

// AW_NCP3 (src/shared/atypes.F90:488-592) behind the ANCP interface: the same
// arguments and checks as jkg_ncp; the selection is the best total weight over
// every start offset of the sorted list.
extern "C" int jkg_ncp3(jkg_ctx *c, int n, int nn, int nm, jkg_aw *dz, int n2, int32_t *sl, int32_t *step,
                        int64_t *info) {
  if (!c) return -1;
  *sl = 0;
  *info = 0;
  if (n < 0) { *info = -1; return -2; }
  if (nn < 0) { *info = -2; return -3; }
  if (nm < nn) { *info = -3; return -4; }
  if (n2 < 0) { *info = -5; return -7; }
  if (n == 0 || nn == 0 || n2 == 0) return JKG_OK;
  for (int i = 0; i < nn; ++i)
    if (!(dz[i].w == dz[i].w) || dz[i].p < 1 || dz[i].q <= dz[i].p || dz[i].q > n) return -5;
  HIPCHK(hipSetDevice(c->device));
  int nb = std::max(n, 3);
  int rc = ensure_buffers(c, nb, 0, 1);
  if (rc) return rc;
  if ((size_t)nn > (size_t)c->d.nn || n2 > c->d.n2max) return -3;
  rc = ensure_ncp3(c);
  if (rc) return rc;
  DeviceStep &d = c->d;
  // the list as a finished sort: the packed pairs in list order, each key at its pair index (of
  // the buffers' order nb), and the sort's figures (no NaN weights: they were refused above)
  std::vector<uint64_t> key(d.nn, ~0ull);
  std::vector<uint32_t> val(nn);
  for (int i = 0; i < nn; ++i) {
    uint64_t &k = key[pair_index(nb, dz[i].p, dz[i].q)];
    if (k != ~0ull) return -5;  // a pair listed twice
    k = host_weight_key(dz[i].w);
    val[i] = pack_pq(dz[i].p - 1, dz[i].q - 1);
  }
  SortStats ss{};
  ss.n = nn;
  ss.nan_count = 0;
  ss.minkey = ~0ull;
  for (int i = 0; i < nn; ++i) ss.minkey = std::min<unsigned long long>(ss.minkey, host_weight_key(dz[i].w));
  HIPCHK(hipMemcpyAsync(d.sort.k0, key.data(), 8 * (size_t)d.nn, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(d.sort.v0, val.data(), 4 * (size_t)nn, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(sort_stats(d.sort), &ss, sizeof(ss), hipMemcpyHostToDevice, c->stream));
  StepState s{};
  s.n2 = n2;
  s.active = 1;
  HIPCHK(hipMemcpyAsync(d.st, &s, sizeof(s), hipMemcpyHostToDevice, c->stream));
  ncp3_launch(d, c->stream);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(&s, d.st, sizeof(s), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  *sl = s.sl;
  if (s.sl > 0) HIPCHK(hipMemcpy(step, d.piv_pos, sizeof(int) * s.sl, hipMemcpyDeviceToHost));
  return JKG_OK;
}

// ---------------------------------------------------------------- test matrices
namespace {
uint64_t splitmix64(uint64_t &x) {
  uint64_t z = (x += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
double unit(uint64_t &x) { return (double)(splitmix64(x) >> 11) * 0x1p-53; }  // [0, 1)
}  // namespace

// G = Q diag(sqrt|lam|) H with Q orthogonal, H J-orthogonal, J = sign(lam): the
// hyperbolic singular values of (G, J) are sqrt|lam| (see jkg_gen.hip).
extern "C" int jkg_dgen(jkg_ctx *c, int n, const double *lam, int rounds, double tmax, uint64_t seed, double *g,
                        int64_t *j) {
  if (!c) return -1;
  if (n < 1) return -2;
  if (!lam) return -3;
  if (!(tmax >= 0.0) || tmax > 20.0) return -5;
  if (!g) return -7;
  if (!j) return -8;
  for (int i = 0; i < n; ++i)
    if (!(lam[i] == lam[i]) || lam[i] == 0.0 || __builtin_isinf(lam[i])) return -3;
  if (rounds <= 0) {
    int lg = 0;
    while ((1 << lg) < n) ++lg;
    rounds = 2 * lg + 2;
  }
  HIPCHK(hipSetDevice(c->device));
  if (gen_rows_prepare(n)) return JKG_ENOTSUP;
  const int np = n / 2;
  const size_t bG = sizeof(double) * (size_t)n * n, bd = sizeof(double) * n, bJ = sizeof(int) * n,
               bp = sizeof(int) * 2 * (size_t)np * 2 * rounds, bc = sizeof(double) * 2 * (size_t)np * 2 * rounds;
  int rc = tmp_bytes(c, bG + bd + bJ + bp + bc + 256);
  if (rc) return rc;
  char *base = (char *)c->tmp;
  double *dG = (double *)base, *dd = (double *)(base + bG), *dcs = (double *)(base + bG + bd);
  int *dJ = (int *)(base + bG + bd + bc), *dp = (int *)(base + bG + bd + bc + bJ);
  std::vector<double> d(n), cs(2 * (size_t)np * 2 * rounds);
  std::vector<int> J(n), pairs(2 * (size_t)np * 2 * rounds), perm(n);
  for (int i = 0; i < n; ++i) {
    d[i] = std::sqrt(std::fabs(lam[i]));
    J[i] = lam[i] > 0.0 ? 1 : -1;
    j[i] = J[i];
  }
  // 2 * rounds matchings: rounds for H (columns), then rounds for Q (rows); pairs as
  // [round][a[0..np), b[0..np)], angles as [round][(c, s) x np]
  uint64_t x = seed;
  for (int r = 0; r < 2 * rounds; ++r) {
    for (int i = 0; i < n; ++i) perm[i] = i;
    for (int i = n - 1; i > 0; --i) std::swap(perm[i], perm[(int)(splitmix64(x) % (uint64_t)(i + 1))]);
    int *pa = &pairs[(size_t)r * 2 * np], *pb = pa + np;
    double *rcs = &cs[(size_t)r * 2 * np];
    for (int k = 0; k < np; ++k) {
      pa[k] = perm[2 * k];
      pb[k] = perm[2 * k + 1];
      const bool hyp = r < rounds && J[pa[k]] != J[pb[k]];
      if (hyp) {
        const double t = tmax * (2.0 * unit(x) - 1.0);
        rcs[2 * k] = std::cosh(t);
        rcs[2 * k + 1] = std::sinh(t);
      } else {
        const double t = M_PI * (2.0 * unit(x) - 1.0);
        rcs[2 * k] = std::cos(t);
        rcs[2 * k + 1] = std::sin(t);
      }
    }
  }
  HIPCHK(hipMemcpyAsync(dd, d.data(), bd, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(dJ, J.data(), bJ, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(dp, pairs.data(), bp, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(dcs, cs.data(), bc, hipMemcpyHostToDevice, c->stream));
  gen_diag_launch(n, dd, dG, c->stream);
  for (int r = 0; r < rounds; ++r)
    gen_cols_launch(n, np, dG, dp + (size_t)r * 2 * np, dp + (size_t)r * 2 * np + np, dcs + (size_t)r * 2 * np, dJ,
                    c->stream);
  for (int r = rounds; r < 2 * rounds; ++r)
    gen_rows_launch(n, np, dG, dp + (size_t)r * 2 * np, dp + (size_t)r * 2 * np + np, dcs + (size_t)r * 2 * np,
                    c->stream);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(g, dG, bG, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return JKG_OK;
}

// ---------------------------------------------------------------- profiling
static int profile_common(jkg_ctx *c, int n, int n2, const double *y, const int64_t *j, int fused, int steps,
                          double *stage_ms, jkg_result *res, int wmode, double tol = 0.0, int cplx = 0) {
  if (!c) return -1;
  if (n < 3) return -2;
  if (steps < 1) return -7;
  memset(res, 0, sizeof(*res));
  HIPCHK(hipSetDevice(c->device));
  n2 = jkg_jstep_len(n, n2);
  int rc = ensure_buffers(c, n, cplx, steps + 1);
  if (rc) return rc;
  c->d.wmode = wmode;
  c->d.neg_tol = tol;
  c->d.ctl_mode = 0;
  c->d.ncp3 = 0;
  c->d.cyclic = 0;
  c->d.frames = nullptr;
  if (wmode) {
    rc = ensure_mixed(c);
    if (rc) return rc;
  }
  int npos, nneg;
  rc = upload_problem(c, n, n2, y, j, &npos, &nneg, steps, c->trace_cap);
  if (rc) return rc;
  DeviceStep &d = c->d;
  d.fused = fused ? 1 : 0;
  init_identity_launch(d, c->stream);
  const int NST = 7;
  hipEvent_t ev[NST + 1];
  for (int k = 0; k <= NST; ++k) HIPCHK(hipEventCreate(&ev[k]));
  for (int k = 0; k < NST; ++k) stage_ms[k] = 0.0;
  for (int s = 0; s < steps; ++s) {
    HIPCHK(hipEventRecord(ev[0], c->stream));
    step_begin_launch(d, c->stream);
    HIPCHK(hipEventRecord(ev[1], c->stream));
    weights_launch(d, c->stream);
    HIPCHK(hipEventRecord(ev[2], c->stream));
    sort_pairs_launch(d.sort, &d.st->nn_sort, d.nn, c->stream, d.n, (d.prec32 && d.cut32) ? 32 : 0);
    HIPCHK(hipEventRecord(ev[3], c->stream));
    select_launch(d, c->stream);
    HIPCHK(hipEventRecord(ev[4], c->stream));
    pivots_launch(d, c->stream);
    HIPCHK(hipEventRecord(ev[5], c->stream));
    rotate_launch(d, c->stream);
    HIPCHK(hipEventRecord(ev[6], c->stream));
    step_end_launch(d, c->stream);
    HIPCHK(hipEventRecord(ev[7], c->stream));
    HIPCHK(hipEventSynchronize(ev[7]));
    for (int k = 0; k < NST; ++k) {
      float ms = 0.f;
      HIPCHK(hipEventElapsedTime(&ms, ev[k], ev[k + 1]));
      stage_ms[k] += ms;
    }
  }
  for (int k = 0; k <= NST; ++k) (void)hipEventDestroy(ev[k]);
  StepState st;
  HIPCHK(hipMemcpy(&st, d.st, sizeof(st), hipMemcpyDeviceToHost));
  res->info = st.done ? st.info : st.step;
  res->steps = st.nrec;
  res->n2 = n2;
  res->npos = npos;
  res->nneg = nneg;
  res->sum_sl = (int64_t)st.acc_sl;
  res->sum_it = (int64_t)st.acc_it;
  res->sum_b2 = (int64_t)st.acc_b2;
  res->sum_b4 = (int64_t)st.acc_b4;
  res->sum_abod = (int64_t)acc_abod_total(st);
  double tot = 0;
  for (int k = 0; k < NST; ++k) tot += stage_ms[k];
  res->seconds = tot * 1e-3;
  return JKG_OK;
}

extern "C" int jkg_dprofile(jkg_ctx *c, int n, int n2, const double *y, const int64_t *j, int fused, int steps,
                            double *stage_ms, jkg_result *res) {
  return profile_common(c, n, n2, y, j, fused, steps, stage_ms, res, 0);
}

extern "C" int jkg_dprofile_mixed(jkg_ctx *c, int n, int n2, const double *y, const int64_t *j, int fused,
                                  int steps, const jkg_mixed_opts *mopts, double *stage_ms, jkg_result *res) {
  jkg_mixed_opts m;
  if (mopts) m = *mopts;
  else jkg_mixed_opts_default(&m);
  if (m.weights != JKG_W_PAIR && m.weights != JKG_W_GRAM) return -8;
  if (!(m.tol >= 0.0)) return -8;
  if (m.fp32_warm != 0 || m.key_bits != 64 || m.controller != 0 || m.threshold != 0.0) return JKG_ENOTSUP;
  if (m.weights == JKG_W_PAIR && m.tol != 0.0) return JKG_ENOTSUP;
  return profile_common(c, n, n2, y, j, fused, steps, stage_ms, res, m.weights == JKG_W_GRAM ? 1 : 0, m.tol);
}

extern "C" int jkg_zprofile(jkg_ctx *c, int n, int n2, const double *y, const int64_t *j, int fused, int steps,
                            double *stage_ms, jkg_result *res) {
  return profile_common(c, n, n2, y, j, fused, steps, stage_ms, res, 0, 0.0, 1);
}

extern "C" int jkg_zprofile_mixed(jkg_ctx *c, int n, int n2, const double *y, const int64_t *j, int fused,
                                  int steps, const jkg_mixed_opts *mopts, double *stage_ms, jkg_result *res) {
  jkg_mixed_opts m;
  if (mopts) m = *mopts;
  else jkg_mixed_opts_default(&m);
  if (m.weights != JKG_W_PAIR && m.weights != JKG_W_GRAM) return -8;
  if (!(m.tol >= 0.0)) return -8;
  if (m.fp32_warm != 0 || m.key_bits != 64 || m.controller != 0 || m.threshold != 0.0) return JKG_ENOTSUP;
  if (m.weights == JKG_W_PAIR && m.tol != 0.0) return JKG_ENOTSUP;
  return profile_common(c, n, n2, y, j, fused, steps, stage_ms, res, m.weights == JKG_W_GRAM ? 1 : 0, m.tol, 1);
}

// ---------------------------------------------------------------- controller training
// Layout of the training scratch (floats): master | Adam m1 | Adam m2 | grad |
// q (BF16, NPARAM/2 + 1 floats) | loss | X | T | W | Y | block partials | block losses.
namespace {
struct TrainBufs {
  float *master, *m1, *m2, *grad, *loss, *X, *T, *W, *Y, *part, *lpart;
  uint16_t *q;
};
size_t train_floats(int ns) {
  const size_t np = MLP_NPARAM, nb = (size_t)mlp_blocks(ns);
  return 4 * np + (np / 2 + 2) + 1 + (size_t)ns * (MLP_IN + 2 * MLP_OUT + 1) + nb * np + nb + 64;
}
TrainBufs train_bufs(float *b, int ns) {
  const size_t np = MLP_NPARAM, nb = (size_t)mlp_blocks(ns);
  TrainBufs t;
  t.master = b;
  t.m1 = b + np;
  t.m2 = b + 2 * np;
  t.grad = b + 3 * np;
  t.q = (uint16_t *)(b + 4 * np);
  float *p = b + 4 * np + (np / 2 + 2);
  t.loss = p++;
  t.X = p;
  p += (size_t)ns * MLP_IN;
  t.T = p;
  p += (size_t)ns * MLP_OUT;
  t.W = p;
  p += ns;
  t.Y = p;
  p += (size_t)ns * MLP_OUT;
  t.part = p;
  p += nb * np;
  t.lpart = p;
  return t;
}
}  // namespace

extern "C" int jkg_ctl_grad(jkg_ctx *c, int ns, const float *x, const float *t, const float *w,
                            const float *params, float *grad, float *loss, float *y) {
  if (!c) return -1;
  if (ns < 1) return -2;
  if (!x) return -3;
  if (!t) return -4;
  if (!params) return -6;
  HIPCHK(hipSetDevice(c->device));
  int rc = ensure_train(c, train_floats(ns));
  if (rc) return rc;
  TrainBufs b = train_bufs(c->tr_buf, ns);
  double wsum = ns;
  if (w) {
    wsum = 0.0;
    for (int i = 0; i < ns; ++i) wsum += w[i];
    if (!(wsum > 0.0)) return -5;
  }
  HIPCHK(hipMemcpyAsync(b.master, params, sizeof(float) * MLP_NPARAM, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(b.X, x, sizeof(float) * ns * MLP_IN, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(b.T, t, sizeof(float) * ns * MLP_OUT, hipMemcpyHostToDevice, c->stream));
  if (w) HIPCHK(hipMemcpyAsync(b.W, w, sizeof(float) * ns, hipMemcpyHostToDevice, c->stream));
  mlp_quantize_launch(b.master, b.q, c->stream);
  mlp_grad_launch(b.q, b.master, b.X, b.T, w ? b.W : nullptr, ns, (float)(1.0 / wsum), b.part, b.lpart, b.grad,
                  b.loss, y ? b.Y : nullptr, c->stream);
  HIPCHK(hipGetLastError());
  if (grad) HIPCHK(hipMemcpyAsync(grad, b.grad, sizeof(float) * MLP_NPARAM, hipMemcpyDeviceToHost, c->stream));
  if (loss) HIPCHK(hipMemcpyAsync(loss, b.loss, sizeof(float), hipMemcpyDeviceToHost, c->stream));
  if (y) HIPCHK(hipMemcpyAsync(y, b.Y, sizeof(float) * ns * MLP_OUT, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return JKG_OK;
}

extern "C" int jkg_ctl_fit(jkg_ctx *c, int ns, const float *x, const float *t, const float *w, int epochs,
                           int batch, float lr, float weight_decay, uint64_t seed, float *params, float *loss) {
  if (!c) return -1;
  if (ns < 1) return -2;
  if (!x) return -3;
  if (!t) return -4;
  if (epochs < 1) return -6;
  if (batch < 1) return -7;
  if (!(lr > 0.f)) return -8;
  if (!params) return -11;
  HIPCHK(hipSetDevice(c->device));
  batch = std::min(batch, ns);
  // scratch sized for the full set (the per-epoch loss) and for one batch
  int rc = ensure_train(c, train_floats(ns) + train_floats(batch));
  if (rc) return rc;
  TrainBufs full = train_bufs(c->tr_buf, ns);
  TrainBufs bt = train_bufs(c->tr_buf + train_floats(ns), batch);
  std::vector<float> zeros(MLP_NPARAM, 0.f);
  HIPCHK(hipMemcpyAsync(full.master, params, sizeof(float) * MLP_NPARAM, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(full.m1, zeros.data(), sizeof(float) * MLP_NPARAM, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(full.m2, zeros.data(), sizeof(float) * MLP_NPARAM, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(full.X, x, sizeof(float) * ns * MLP_IN, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(full.T, t, sizeof(float) * ns * MLP_OUT, hipMemcpyHostToDevice, c->stream));
  std::vector<float> wv(ns, 1.f);
  if (w) for (int i = 0; i < ns; ++i) wv[i] = w[i];
  double wsum = 0.0;
  for (int i = 0; i < ns; ++i) wsum += wv[i];
  if (!(wsum > 0.0)) return -5;
  HIPCHK(hipMemcpyAsync(full.W, wv.data(), sizeof(float) * ns, hipMemcpyHostToDevice, c->stream));
  mlp_quantize_launch(full.master, full.q, c->stream);
  // host-side permutation per epoch (splitmix64 + Fisher-Yates), batches copied in order
  std::vector<int> perm(ns);
  std::vector<float> bx((size_t)batch * MLP_IN), btg((size_t)batch * MLP_OUT), bw(batch);
  uint64_t st = seed;
  auto rnd = [&]() {
    st += 0x9E3779B97F4A7C15ull;
    uint64_t z = st;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  };
  int stepno = 0;
  for (int e = 0; e < epochs; ++e) {
    for (int i = 0; i < ns; ++i) perm[i] = i;
    for (int i = ns - 1; i > 0; --i) std::swap(perm[i], perm[rnd() % (uint64_t)(i + 1)]);
    for (int b0 = 0; b0 + batch <= ns; b0 += batch) {
      double bws = 0.0;
      for (int k = 0; k < batch; ++k) {
        const int i = perm[b0 + k];
        for (int f = 0; f < MLP_IN; ++f) bx[(size_t)k * MLP_IN + f] = x[(size_t)i * MLP_IN + f];
        for (int o = 0; o < MLP_OUT; ++o) btg[(size_t)k * MLP_OUT + o] = t[(size_t)i * MLP_OUT + o];
        bw[k] = wv[i];
        bws += wv[i];
      }
      if (!(bws > 0.0)) continue;
      HIPCHK(hipMemcpyAsync(bt.X, bx.data(), sizeof(float) * batch * MLP_IN, hipMemcpyHostToDevice, c->stream));
      HIPCHK(hipMemcpyAsync(bt.T, btg.data(), sizeof(float) * batch * MLP_OUT, hipMemcpyHostToDevice, c->stream));
      HIPCHK(hipMemcpyAsync(bt.W, bw.data(), sizeof(float) * batch, hipMemcpyHostToDevice, c->stream));
      mlp_grad_launch(full.q, full.master, bt.X, bt.T, bt.W, batch, (float)(1.0 / bws), bt.part, bt.lpart, bt.grad,
                      nullptr, nullptr, c->stream);
      mlp_adam_launch(full.master, full.q, bt.grad, full.m1, full.m2, lr, ++stepno, weight_decay, c->stream);
      HIPCHK(hipGetLastError());
      HIPCHK(hipStreamSynchronize(c->stream));  // the host batch buffers are reused
    }
    if (loss) {
      mlp_grad_launch(full.q, full.master, full.X, full.T, full.W, ns, (float)(1.0 / wsum), full.part, full.lpart,
                      full.grad, full.loss, nullptr, c->stream);
      HIPCHK(hipGetLastError());
      HIPCHK(hipMemcpyAsync(&loss[e], full.loss, sizeof(float), hipMemcpyDeviceToHost, c->stream));
    }
  }
  HIPCHK(hipMemcpyAsync(params, full.master, sizeof(float) * MLP_NPARAM, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return JKG_OK;
}

// ---------------------------------------------------------------- batched solver
extern "C" void jkg_batch_opts_default(jkg_batch_opts *o) {
  memset(o, 0, sizeof(*o));
  o->max_steps = -1;
  o->variant = -1;
}

namespace {
// Device memory of one batch, freed by the destructor.
struct BatchMem {
  std::vector<void *> ptrs;
  ~BatchMem() {
    for (void *p : ptrs) (void)hipFree(p);
  }
  template <class T>
  int get(T **p, size_t count) {
    void *q = nullptr;
    if (hipMalloc(&q, sizeof(T) * (count ? count : 1)) != hipSuccess) {
      (void)hipGetLastError();
      return JKG_ENOMEM;
    }
    ptrs.push_back(q);
    *p = (T *)q;
    return JKG_OK;
  }
};
}  // namespace

extern "C" int jkg_dsolve_batched(jkg_ctx *c, int count, const int32_t *ns, const int32_t *n2s, const double *y,
                                  const int64_t *j, double *yu, double *zz, double *sy, int32_t *info,
                                  jkg_batch_opts *opts) {
  if (!c) return -1;
  if (count < 1) return -2;
  if (!ns) return -3;
  if (!y) return -5;
  if (!j) return -6;
  if (!yu) return -7;
  if (!zz) return -8;
  if (!sy) return -9;
  if (!info) return -10;
  jkg_batch_opts o;
  if (opts) o = *opts;
  else jkg_batch_opts_default(&o);
  if (o.grid < 0 || o.grid > JKG_BATCH_GRID_MAX || o.variant < -1 || o.variant >= batch_variants() ||
      !(o.tol >= 0.0))
    return -11;
  HIPCHK(hipSetDevice(c->device));
  // layout of the pools
  std::vector<BInst> inst(count);
  long long mat = 0, pairs = 0, ints = 0, dbl = 0, rot = 0, syo = 0;
  std::vector<int> J32;
  std::vector<int> lists;
  for (int i = 0; i < count; ++i) {
    const int n = ns[i];
    if (n < 3 || n > 1024) return -3;
    BInst &I = inst[i];
    I.n = n;
    I.n2 = jkg_jstep_len(n, n2s ? n2s[i] : 0);
    if (I.n2 <= 0) return -4;
    I.a = mat;
    I.u = mat + (long long)n * n;
    I.z = mat + 2LL * n * n;
    mat += 3LL * n * n;
    I.k = pairs;
    pairs += (long long)n * (n - 1) / 2;
    I.j = ints;
    I.pm = ints + n;
    I.cp = ints + 2 * n;
    I.pv = ints + 3 * n;
    ints += 3LL * n + 4LL * I.n2;
    I.pw = dbl;
    dbl += 3LL * I.n2;
    I.pr = rot;
    rot += 2LL * I.n2;
    I.sy = syo;
    syo += n;
  }
  // J, plus/minus lists (host) into the int pool image
  std::vector<int> ipool((size_t)ints, 0);
  {
    long long jo = 0;
    for (int i = 0; i < count; ++i) {
      BInst &I = inst[i];
      int np = 0, nm = 0;
      std::vector<int> pl, mi;
      for (int r = 0; r < I.n; ++r) {
        const int64_t v = j[jo + r];
        if (v != 1 && v != -1) return JKG_EJ;
        ipool[I.j + r] = (int)v;
        (v == 1 ? pl : mi).push_back(r);
      }
      np = (int)pl.size();
      nm = (int)mi.size();
      I.npl = np;
      I.nmi = nm;
      for (int r = 0; r < np; ++r) ipool[I.pm + r] = pl[r];
      for (int r = 0; r < nm; ++r) ipool[I.pm + np + r] = mi[r];
      jo += I.n;
    }
  }
  // schedule: instances by decreasing cost (N^3) onto the least-loaded workgroup
  std::vector<int> order(count);
  for (int i = 0; i < count; ++i) order[i] = i;
  std::sort(order.begin(), order.end(), [&](int x, int y2) { return inst[x].n > inst[y2].n; });
  auto make_sched = [&](int grid, std::vector<int> &sched, std::vector<int> &off) {
    std::vector<std::vector<int>> lists2(grid);
    std::vector<double> load(grid, 0.0);
    for (int i : order) {
      int b = (int)(std::min_element(load.begin(), load.end()) - load.begin());
      lists2[b].push_back(i);
      load[b] += (double)inst[i].n * inst[i].n * inst[i].n;
    }
    sched.clear();
    off.assign(grid + 1, 0);
    for (int b = 0; b < grid; ++b) {
      off[b] = (int)sched.size();
      for (int i : lists2[b]) sched.push_back(i);
    }
    off[grid] = (int)sched.size();
  };
  BatchMem m;
  BatchArgs a{};
  a.count = count;
  a.tol = o.tol;
  a.pair_total = pairs;
  BInst *dinst;
  int rc = JKG_OK;
  int *dsched, *doff;
  const int maxgrid = JKG_BATCH_GRID_MAX;
  if ((rc = m.get(&dinst, count)) || (rc = m.get(&a.state, count)) || (rc = m.get(&a.mat, mat)) ||
      (rc = m.get(&a.dbl, dbl)) || (rc = m.get(&a.ints, ints)) || (rc = m.get(&a.keys, 2 * pairs)) ||
      (rc = m.get(&a.vals, 2 * pairs)) || (rc = m.get(&a.rot, rot)) || (rc = m.get(&dsched, count)) ||
      (rc = m.get(&doff, maxgrid + 1)) || (rc = m.get(&a.ndone, 1)))
    return rc;
  double *dsy;
  if ((rc = m.get(&dsy, syo))) return rc;
  a.inst = dinst;
  a.sched = dsched;
  a.sched_off = doff;
  HIPCHK(hipMemcpy(dinst, inst.data(), sizeof(BInst) * count, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(a.ints, ipool.data(), sizeof(int) * ints, hipMemcpyHostToDevice));
  // (re)starts every instance from its Y
  auto reset = [&](int max_steps) -> int {
    long long yo = 0;
    for (int i = 0; i < count; ++i) {
      const long long nn2 = (long long)inst[i].n * inst[i].n;
      HIPCHK(hipMemcpyAsync(a.mat + inst[i].a, y + yo, sizeof(double) * nn2, hipMemcpyHostToDevice, c->stream));
      yo += nn2;
    }
    std::vector<BState> st0(count);
    for (int i = 0; i < count; ++i) {
      memset(&st0[i], 0, sizeof(BState));
      st0[i].max_steps = max_steps;
    }
    HIPCHK(hipMemcpyAsync(a.state, st0.data(), sizeof(BState) * count, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemsetAsync(a.ndone, 0, sizeof(int), c->stream));
    batch_identity_launch(a, c->stream);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(c->stream));
    return JKG_OK;
  };
  std::vector<int> sched, off;
  auto set_grid = [&](int grid) -> int {
    make_sched(grid, sched, off);
    HIPCHK(hipMemcpy(dsched, sched.data(), sizeof(int) * count, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(doff, off.data(), sizeof(int) * (grid + 1), hipMemcpyHostToDevice));
    return JKG_OK;
  };
  // autotuner: two steps of every candidate (grid x variant) from the start
  int grid = o.grid, variant = o.variant;
  const double t_tune0 = (double)clock() / CLOCKS_PER_SEC;
  if (grid == 0 || variant < 0) {
    std::vector<int> grids;
    for (int g : {256, 512, 1024, 2048})
      if (o.grid == 0 ? (g <= std::max(count, 256)) : g == o.grid) grids.push_back(g);
    if (o.grid != 0 && grids.empty()) grids.push_back(o.grid);
    float best = 1e30f;
    for (int g : grids) {
      if ((rc = set_grid(g))) return rc;
      for (int v = 0; v < batch_variants(); ++v) {
        if (o.variant >= 0 && v != o.variant) continue;
        if ((rc = reset(-1))) return rc;
        HIPCHK(hipEventRecord(c->ev0, c->stream));
        for (int k = 0; k < 2; ++k) batch_step_launch(a, g, v, o.fused, c->stream);
        HIPCHK(hipEventRecord(c->ev1, c->stream));
        HIPCHK(hipEventSynchronize(c->ev1));
        float ms = 0.f;
        HIPCHK(hipEventElapsedTime(&ms, c->ev0, c->ev1));
        if (ms < best) {
          best = ms;
          grid = g;
          variant = v;
        }
      }
    }
  }
  o.tune_seconds = (double)clock() / CLOCKS_PER_SEC - t_tune0;
  if ((rc = set_grid(grid))) return rc;
  if ((rc = reset(o.max_steps))) return rc;
  // the step loop: a graph of 8 batched steps, replayed until every instance is done
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  const int per_graph = 8;
  HIPCHK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
  for (int k = 0; k < per_graph; ++k) batch_step_launch(a, grid, variant, o.fused, c->stream);
  HIPCHK(hipStreamEndCapture(c->stream, &graph));
  if (hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) != hipSuccess) {
    (void)hipGraphDestroy(graph);
    return JKG_EHIP;
  }
  int maxn = 0;
  for (int i = 0; i < count; ++i) maxn = std::max(maxn, inst[i].n);
  const long cap = (o.max_steps >= 0) ? (long)o.max_steps + per_graph : 1000L * maxn + 1000;
  long launched = 0;
  int ndone = 0;
  rc = JKG_OK;
  (void)hipEventRecord(c->ev0, c->stream);
  for (;;) {
    if (hipGraphLaunch(exec, c->stream) != hipSuccess) {
      rc = JKG_EHIP;
      break;
    }
    launched += per_graph;
    if (hipMemcpyAsync(&ndone, a.ndone, sizeof(int), hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
        hipStreamSynchronize(c->stream) != hipSuccess) {
      rc = JKG_EHIP;
      break;
    }
    if (ndone >= count) break;
    if (launched > cap) {
      rc = JKG_ENOCONV;
      break;
    }
  }
  (void)hipEventRecord(c->ev1, c->stream);
  (void)hipGraphExecDestroy(exec);
  (void)hipGraphDestroy(graph);
  if (rc) return rc;
  HIPCHK(hipEventSynchronize(c->ev1));
  float ms = 0.f;
  HIPCHK(hipEventElapsedTime(&ms, c->ev0, c->ev1));
  batch_finalize_launch(a, dsy, c->stream);
  HIPCHK(hipGetLastError());
  std::vector<BState> st(count);
  HIPCHK(hipMemcpyAsync(st.data(), a.state, sizeof(BState) * count, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipMemcpyAsync(sy, dsy, sizeof(double) * syo, hipMemcpyDeviceToHost, c->stream));
  {
    long long yo = 0;
    for (int i = 0; i < count; ++i) {
      const long long nn2 = (long long)inst[i].n * inst[i].n;
      HIPCHK(hipMemcpyAsync(yu + yo, a.mat + inst[i].u, sizeof(double) * nn2, hipMemcpyDeviceToHost, c->stream));
      HIPCHK(hipMemcpyAsync(zz + yo, a.mat + inst[i].z, sizeof(double) * nn2, hipMemcpyDeviceToHost, c->stream));
      yo += nn2;
    }
  }
  HIPCHK(hipStreamSynchronize(c->stream));
  for (int i = 0; i < count; ++i) info[i] = st[i].info;
  o.steps = (int)launched;
  o.tuned_grid = grid;
  o.tuned_variant = variant;
  o.seconds = ms * 1e-3;
  if (opts) *opts = o;
  return JKG_OK;
}
