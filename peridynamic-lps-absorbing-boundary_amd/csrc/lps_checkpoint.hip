// Checkpoint and restart of the time loop (lps_checkpoint_save / lps_checkpoint_load). The reference
// has none: a run of src/main.cpp:716-1166 starts from the initial conditions every time.
//
// A checkpoint holds everything the next step reads (lps_ctx::state): U^n, U^{n-1}, V^n, V^{n-1},
// A^n, the prediction P^n, theta^n and theta^{n-1}, per slot, with the step count and E0 (one per
// instance). Restoring it into a context of the same problem and options and stepping on gives
// the same bits as the run that was never interrupted.
//
// File layout (little-endian):
//   char[8]   "LPSCKPT1"
//   int64     step
//   int64     n_slots
//   int32     n_inst, n_nodes, n_abc, 0
//   uint64    fingerprint of the context (slot layout, kernels' path, per-instance constants and
//             ABC inputs)
//   double    E0[n_inst]
//   double    14 arrays of n_slots, in slot order: U^n x, y; U^{n-1} x, y; V^n x, y;
//             V^{n-1} x, y; A^n x, y; P^n x, y; theta^n; theta^{n-1}
#include <cstdio>
#include <cstring>
#include <memory>

#include "lps_ctx.h"

namespace {

constexpr char kMagic[8] = {'L', 'P', 'S', 'C', 'K', 'P', 'T', '1'};
constexpr int kArrays = 14;

struct Fnv {
  uint64_t h = 1469598103934665603ull;
  void bytes(const void *p, size_t n) {
    const unsigned char *b = (const unsigned char *)p;
    for (size_t i = 0; i < n; i++) h = (h ^ b[i]) * 1099511628211ull;
  }
  template <class T>
  void value(const T &v) { bytes(&v, sizeof(T)); }
};

// What a checkpoint may only be restored into: the same slots (node order and layout), the same
// arithmetic path and ABC method, the same per-instance constants and the same ABC nodes, clouds
// and operators (a context built with other G_u / G_v / G_theta continues differently).
uint64_t fingerprint(const lps_ctx *c) {
  Fnv f;
  f.value(c->n_slots);
  f.value(c->n);
  f.value(c->n_abc);
  f.value(c->n_inst);
  f.value(c->fast);
  f.value(c->opt.exact);
  f.value(c->opt.theta_method);
  f.value(c->opt.abc_enabled);
  f.value(c->hb.nnz);  // the bonds, after any prenotch cut
  f.bytes(c->perm_h.data(), c->perm_h.size() * sizeof(int32_t));
  for (const lps::InstParams &p : c->inst_h) {
    f.value(p.dt);
    f.value(p.rho);
    f.value(p.lam);
    f.value(p.mu);
  }
  for (uint64_t h : c->abc_hash) f.value(h);
  return f.h;
}

// The 14 device arrays of the state at step s, in file order (see lps_ctx::state).
void state_arrays(const lps_ctx *c, int64_t s, double *arr[kArrays]) {
  const lps::Dev2 &Uc = c->U[s % 3], &Um = c->U[(s + 2) % 3];
  const lps::Dev2 &Vc = c->V[s % 3], &Vm = c->V[(s + 2) % 3];
  const lps::Dev2 &Ac = c->A[s % 2], &Pc = c->P[s % 6];
  double *a[kArrays] = {Uc.x, Uc.y, Um.x, Um.y, Vc.x, Vc.y, Vm.x, Vm.y,
                        Ac.x, Ac.y, Pc.x, Pc.y, c->TH[s % 3], c->TH[(s + 2) % 3]};
  for (int k = 0; k < kArrays; k++) arr[k] = a[k];
}

struct FileCloser {
  void operator()(FILE *f) const {
    if (f) fclose(f);
  }
};
using File = std::unique_ptr<FILE, FileCloser>;

struct Header {
  char magic[8];
  int64_t step, n_slots;
  int32_t n_inst, n_nodes, n_abc, pad;
  uint64_t fp;
};
static_assert(sizeof(Header) == 48, "checkpoint header layout");

}  // namespace

extern "C" {

int32_t lps_checkpoint_save(lps_ctx *c, const char *path) {
  if (!c) return LPS_ERR_ARG;
  if (!path) return c->fail(LPS_ERR_ARG, "checkpoint path is NULL");
  if (c->clo)
    return c->fail(LPS_ERR_ARG, "a context with a learned closure attached cannot be checkpointed");
  CTX_TRY(c, hipSetDevice(c->opt.device));
  int32_t rc = lps::ctx_sync_u(c, 2);  // U^n, U^{n-1} of every node
  if (rc != LPS_OK) return rc;
  CTX_TRY(c, hipStreamSynchronize(c->stream));
  if (c->stream2) CTX_TRY(c, hipStreamSynchronize(c->stream2));
  if (c->stream3) CTX_TRY(c, hipStreamSynchronize(c->stream3));
  File f(fopen(path, "wb"));
  if (!f) return c->fail(LPS_ERR_ARG, std::string("cannot open ") + path + " for writing");
  Header h{};
  memcpy(h.magic, kMagic, 8);
  h.step = c->step;
  h.n_slots = c->n_slots;
  h.n_inst = c->n_inst;
  h.n_nodes = c->n;
  h.n_abc = c->n_abc;
  h.fp = fingerprint(c);
  std::vector<double> e0(c->n_inst);
  for (int32_t i = 0; i < c->n_inst; i++) e0[i] = c->batch ? c->E0_inst[i] : c->E0;
  bool ok = fwrite(&h, sizeof(h), 1, f.get()) == 1 &&
            fwrite(e0.data(), sizeof(double), e0.size(), f.get()) == e0.size();
  double *arr[kArrays];
  state_arrays(c, c->step, arr);
  std::vector<double> buf(c->n_slots);
  for (int k = 0; ok && k < kArrays; k++) {
    CTX_TRY(c, hipMemcpy(buf.data(), arr[k], sizeof(double) * buf.size(), hipMemcpyDeviceToHost));
    ok = fwrite(buf.data(), sizeof(double), buf.size(), f.get()) == buf.size();
  }
  if (!ok || fflush(f.get()) != 0)
    return c->fail(LPS_ERR_ARG, std::string("write to ") + path + " failed");
  return LPS_OK;
}

int32_t lps_checkpoint_load(lps_ctx *c, const char *path) {
  if (!c) return LPS_ERR_ARG;
  if (!path) return c->fail(LPS_ERR_ARG, "checkpoint path is NULL");
  if (c->clo)
    return c->fail(LPS_ERR_ARG, "a context with a learned closure attached cannot be restored");
  File f(fopen(path, "rb"));
  if (!f) return c->fail(LPS_ERR_ARG, std::string("cannot open ") + path);
  Header h{};
  if (fread(&h, sizeof(h), 1, f.get()) != 1 || memcmp(h.magic, kMagic, 8) != 0)
    return c->fail(LPS_ERR_ARG, std::string(path) + " is not a checkpoint");
  if (h.n_slots != c->n_slots || h.n_inst != c->n_inst || h.n_nodes != c->n ||
      h.n_abc != c->n_abc || h.fp != fingerprint(c))
    return c->fail(LPS_ERR_ARG, std::string(path) +
                                    " was written by a context of another problem or options");
  if (h.step < 0) return c->fail(LPS_ERR_ARG, std::string(path) + ": negative step");
  std::vector<double> e0(c->n_inst);
  if (fread(e0.data(), sizeof(double), e0.size(), f.get()) != e0.size())
    return c->fail(LPS_ERR_ARG, std::string(path) + " is truncated");
  // read everything before touching the device state, so a bad file leaves the context as it was
  std::vector<double> buf((size_t)kArrays * c->n_slots);
  if (fread(buf.data(), sizeof(double), buf.size(), f.get()) != buf.size())
    return c->fail(LPS_ERR_ARG, std::string(path) + " is truncated");
  CTX_TRY(c, hipSetDevice(c->opt.device));
  CTX_TRY(c, hipDeviceSynchronize());
  double *arr[kArrays];
  state_arrays(c, h.step, arr);
  for (int k = 0; k < kArrays; k++)
    CTX_TRY(c, hipMemcpy(arr[k], buf.data() + (size_t)k * c->n_slots,
                         sizeof(double) * c->n_slots, hipMemcpyHostToDevice));
  // the interior nodes' U^n, U^{n-1} back into the prediction ring
  for (int j = 0; j < 2; j++) {
    const lps::Dev2 &R = c->ring_u(h.step, j), &U = c->U[(h.step + 3 - j) % 3];
    CTX_TRY(c, hipMemcpy(R.x, U.x, sizeof(double) * c->n_slots, hipMemcpyDeviceToDevice));
    CTX_TRY(c, hipMemcpy(R.y, U.y, sizeof(double) * c->n_slots, hipMemcpyDeviceToDevice));
  }
  c->u_synced = 2;
  c->step = h.step;
  c->pk_last_step = -1;  // the ABC members' records hold another state
  if (c->batch) {
    c->E0_inst = e0;
    c->E0 = -1.0;
  } else {
    c->E0 = e0[0];
  }
  return LPS_OK;
}

}  // extern "C"
