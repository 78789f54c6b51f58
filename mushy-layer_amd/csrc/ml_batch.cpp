// ml_batch.cpp — the batched form of the (H, Theta) backward-Euler step:
// many independent instances (own grid size, parameters and BCs, each an
// ml_hc_op) advance one FP32-refined implicit step together, every kernel
// launch covering all instances that take part in it.
//
// Instance i's step is exactly the single-instance sequence
//   porosity -> bCoef -> ml_hc_set_coefs(NULL, bCoef, 1, -dt) ->
//   ml_hc_solve_be(phi, copy of phi, NULL, dt, FP32-refined)
// (the same kernels' tile bodies, the same pass splits, the same stopping
// rule and step-length backtracking), so its result is bit-identical to it.
//
// Launch sequence of one outer iteration: the tiled levels of the instances
// are aligned by their distance from the single-workgroup level lc_i, so
// that down-sweep launch k (k = 0 .. K-1, K = max lc_i) relaxes and
// restricts level k - (K - lc_i) of every instance with lc_i > K - 1 - k;
// one launch runs every instance's coarse V-cycle (one workgroup each); the
// up-sweep mirrors the down-sweep; one launch computes every instance's
// FP64 residual, Jacobian and stopping rule.  2K + 2 launches per iteration
// for the whole batch, replayed from a HIP graph.  An instance that has
// converged returns at once from every later launch (its done flag).
#include <algorithm>
#include <cstring>
#include <functional>
#include <vector>

#include "ml_internal.h"

namespace ml {
namespace {

constexpr int kBatchSlots = 128;   // norm slots per instance (max_iter + 1)
constexpr int kSmallCells = 1024;  // levels of at most this many cells go to the single-workgroup kernel
constexpr int kMaxHalvings = 6;    // step-length backtracking, as the single-instance solve

struct BatchInst {
  ml_hc_op op = nullptr;
  int L = 0, lc = 0;
  std::vector<LinLevel<float>> lin;
  std::vector<float*> ealt;   // partners of the tiled levels' corrections
  double* alt0 = nullptr;
  double* rhs = nullptr;
  int nres = 0, res_ntx = 0, res_nty = 0;   // residual launch tiles
};

// one tiled relaxation launch: the instances taking part and their levels
struct LinLaunch {
  std::vector<int> inst, level;
  std::vector<int> blocks_v;   // workgroups under each tile shape
  int blocks = 0;              // ... under the shape in use
  size_t job_off = 0, map_off = 0;
  double bytes = 0;   // algorithmic bytes (every input read once, every output written once)
};

size_t align256(size_t x) { return (x + 255) & ~size_t(255); }

}  // namespace
}  // namespace ml

struct ml_hc_batch_s {
  ml_ctx ctx = nullptr;
  std::vector<ml::BatchInst> inst;
  int K = 0;
  size_t coarse_lds = 0;
  // device memory: the buffers requested during set-up are carved out of
  // one allocation and zeroed by one fill launch (counts kept on the host)
  unsigned char* dev_block = nullptr;
  size_t dev_bytes = 0;
  std::vector<std::function<void(unsigned char*)>> dev_req;   // sets one buffer's pointer from the base
  int dev_allocs = 0, dev_fills = 0;
  ml::SolveCtl* ctl = nullptr;
  ml::SolveCtl* hctl = nullptr;           // pinned mirror
  unsigned long long* slots = nullptr;
  unsigned int* arrive = nullptr;
  // job tables and workgroup maps: one device arena, one pinned mirror
  unsigned char* arena = nullptr;
  unsigned char* harena = nullptr;
  size_t arena_bytes = 0;
  size_t setup_off = 0, coarsen_off = 0, res_off = 0, desc_off = 0;
  size_t setup_map_off = 0, coarsen_map_off = 0, res_map_off = 0;
  int setup_blocks = 0, coarsen_blocks = 0, res_blocks = 0;
  std::vector<ml::LinLaunch> down, up;
  // what the uploaded tables were built for
  bool tables_ok = false;
  std::vector<double*> key_phi;
  std::vector<const double*> key_src;
  double key_dt = 0;
  int key_nu1 = -1, key_nu2 = -1, key_bottom = -1;
  // workgroup maps over the instances still running (stragglers), built
  // when few remain: [down k | up k] per k, then the residual
  unsigned* cmap = nullptr;
  unsigned* hcmap = nullptr;
  size_t cmap_cap = 0;
  std::vector<int> cdown_blocks, cup_blocks;
  std::vector<double> cdown_bytes, cup_bytes;
  std::vector<size_t> cdown_off, cup_off;
  size_t cres_off = 0;
  int cres_blocks = 0;
  // backtracking trials: instance list, workgroup map, gathered norms
  int* tlist = nullptr;
  unsigned* tmap = nullptr;
  unsigned long long* tnorm = nullptr;
  int* htlist = nullptr;
  unsigned* htmap = nullptr;
  unsigned long long* htnorm = nullptr;
  // tile shape of the batched relaxation: picked by timing every shape on
  // the batch's own launches at the first step, unless set
  int variant = 0;
  bool tuned = false;
  std::vector<double> tune_ms;
  // one iteration's launches captured as a HIP graph
  hipGraphExec_t gexec = nullptr;
  int batch_iters = 4;   // iterations enqueued before the first host check

  template <typename T>
  T* at(size_t off) const {
    return reinterpret_cast<T*>(arena + off);
  }
  template <typename T>
  T* hat(size_t off) const {
    return reinterpret_cast<T*>(harena + off);
  }
};

namespace ml {
namespace {

int batch_fail(ml_hc_batch b, int code, const std::string& msg) { return fail(b->ctx, code, msg); }

// Requests `bytes` of zeroed device memory for *field; the pointer is set
// by commit_device() once every buffer of the batch has been requested.
template <typename T>
int dalloc(ml_hc_batch b, size_t bytes, T** field) {
  *field = nullptr;
  const size_t off = b->dev_bytes;
  b->dev_req.push_back([field, off](unsigned char* base) { *field = reinterpret_cast<T*>(base + off); });
  b->dev_bytes = align256(b->dev_bytes + bytes);
  return ML_OK;
}

// one hipMalloc for every requested buffer, one fill launch to zero it;
// then the pointers that set-up derived from them
int commit_device(ml_hc_batch b) {
  ml_ctx ctx = b->ctx;
  void* v = nullptr;
  ML_HIP(ctx, hipMalloc(&v, b->dev_bytes));
  b->dev_block = static_cast<unsigned char*>(v);
  ++b->dev_allocs;
  for (const auto& set : b->dev_req) set(b->dev_block);
  ML_HIP(ctx, vec_set((long long)(b->dev_bytes / sizeof(double)), reinterpret_cast<double*>(b->dev_block), 0.0,
                      ctx->stream));
  ++b->dev_fills;
  for (size_t i = 0; i < b->inst.size(); ++i)
    for (LinLevel<float>& Lv : b->inst[i].lin) Lv.skip = &b->ctl[i].done;
  return ML_OK;
}

void batch_free(ml_hc_batch b) {
  if (!b) return;
  if (b->ctx) (void)hipStreamSynchronize(b->ctx->stream);
  if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
  if (b->dev_block) (void)hipFree(b->dev_block);
  if (b->hctl) (void)hipHostFree(b->hctl);
  if (b->harena) (void)hipHostFree(b->harena);
  if (b->hcmap) (void)hipHostFree(b->hcmap);
  if (b->htlist) (void)hipHostFree(b->htlist);
  if (b->htmap) (void)hipHostFree(b->htmap);
  if (b->htnorm) (void)hipHostFree(b->htnorm);
  delete b;
}

// levels, LDS and allocations of one instance (the checks of the
// single-instance FP32-refined form)
int inst_init(ml_hc_batch b, int i, ml_hc_op op, size_t maxlds) {
  BatchInst& I = b->inst[i];
  I.op = op;
  I.L = int(op->lev.size());
  I.lc = I.L;
  for (int l = 0; l < I.L; ++l)
    if (op->lev[l].g.nvalid() <= kSmallCells) {
      I.lc = l;
      break;
    }
  if (op->dim != 2) return batch_fail(b, ML_ENOTSUP, "batched form: 2D instances only");
  if (I.lc >= I.L) return batch_fail(b, ML_ENOTSUP, "batched form: an instance has no MG level of at most 1024 cells");
  if (op->lev[I.L - 1].g.nvalid() > 512)
    return batch_fail(b, ML_ENOTSUP, "batched form: an instance's coarsest level exceeds 512 cells");
  if (I.L - I.lc > lin_coarse_levels_max(2))
    return batch_fail(b, ML_ENOTSUP, "batched form: too many levels below the single-workgroup level");
  {
    int ntx, nty;
    // a single solve switches to wide tiles above this count; the batch
    // keeps the deep-halo tiles, so it stays below it
    if (lin_residual_tiles2(op->lev[0].g, ntx, nty) > lin_deep_tiles_max2())
      return batch_fail(b, ML_ENOTSUP, "batched form: an instance's finest level is too large");
  }
  for (int l = 0; l < I.L; ++l)
    if (op->lev[l].a) return batch_fail(b, ML_ENOTSUP, "batched form supports aCoef == 1 only");
  std::vector<Geom> gs(I.L);
  for (int l = 0; l < I.L; ++l) gs[l] = op->lev[l].g;
  const size_t lds = lin_coarse_lds_bytes(gs.data(), I.lc, I.L);
  if (lds + lin_coarse_static_lds() > maxlds)
    return batch_fail(b, ML_ENOTSUP, "batched form: an instance's coarse levels exceed the workgroup's LDS");
  b->coarse_lds = std::max(b->coarse_lds, lds);
  I.lin.assign(I.L, LinLevel<float>{});
  I.ealt.assign(I.L, nullptr);
  auto fl = [&](size_t n, auto** field) -> int { return dalloc(b, n * sizeof(float), field); };
  int rc;
  for (int l = 0; l < I.L; ++l) {
    LinLevel<float>& Lv = I.lin[l];
    const Geom& g = op->lev[l].g;
    Lv.g = g;
    if ((rc = fl(2 * g.ncell, &Lv.e))) return rc;
    if ((rc = fl(2 * g.ncell, &Lv.r))) return rc;
    if ((rc = fl(4 * g.ncell, &Lv.M))) return rc;
    for (int d = 0; d < 3; ++d) {
      Lv.b[d] = nullptr;
      if (d < g.dim && (rc = fl(2 * g.nface[d], &Lv.b[d]))) return rc;
    }
    Lv.a = nullptr;
    Lv.skip = nullptr;   // &ctl[i].done, set once the device memory exists
    if (l < I.lc && (rc = fl(2 * g.ncell, &I.ealt[l]))) return rc;
  }
  const size_t bytes0 = sizeof(double) * 2 * op->lev[0].g.ncell;
  if ((rc = dalloc(b, bytes0, &I.alt0))) return rc;
  if ((rc = dalloc(b, bytes0, &I.rhs))) return rc;
  I.nres = lin_residual_tiles2(op->lev[0].g, I.res_ntx, I.res_nty);
  return ML_OK;
}

// FP32 algorithmic bytes of a relaxation launch over the jobs whose
// instance runs (run == nullptr: all): down-sweep (zero start) r, M, faces
// in, e out, restricted r and M out; up-sweep e, r, M, faces, coarse e in,
// e out
double launch_bytes(ml_hc_batch b, const LinLaunch& Lh, bool down, const std::vector<char>* run) {
  double cells = 0, faces = 0, coarse = 0;
  for (size_t j = 0; j < Lh.inst.size(); ++j) {
    if (run && !(*run)[Lh.inst[j]]) continue;
    const ml_hc_op op = b->inst[Lh.inst[j]].op;
    const int l = Lh.level[j];
    cells += double(op->lev[l].g.nvalid());
    for (int d = 0; d < op->dim; ++d) faces += double(op->lev[l].g.nface[d]);
    coarse += double(op->lev[l + 1].g.nvalid());
  }
  return down ? 4.0 * (cells * (2 + 2 + 4) + 2 * faces + coarse * (2 + 4))
              : 4.0 * (cells * (2 + 2 + 2 + 4) + 2 * faces + coarse * 2);
}

// arena layout and the workgroup maps (geometry only)
int layout(ml_hc_batch b) {
  const int n = int(b->inst.size());
  const long long chunk = batch_chunk_items();
  size_t off = 0;
  auto take = [&](size_t bytes) {
    const size_t o = off;
    off = align256(off + bytes);
    return o;
  };
  b->setup_off = take(sizeof(BatchSetupJob) * n);
  b->coarsen_off = take(sizeof(BatchCoarsenJob) * n);
  b->res_off = take(sizeof(BatchResJob) * n);
  b->desc_off = take(sizeof(LinDesc) * n);
  // block counts
  std::vector<int> setup_chunks(n), coarsen_chunks(n);
  b->setup_blocks = b->coarsen_blocks = b->res_blocks = 0;
  for (int i = 0; i < n; ++i) {
    const BatchInst& I = b->inst[i];
    const Geom& g = I.op->lev[0].g;
    long long items = 2 * g.ncell + kBatchSlots;
    for (int d = 0; d < g.dim; ++d) items += g.nface[d];
    setup_chunks[i] = int((items + chunk - 1) / chunk);
    long long citems = 0;
    for (int l = 1; l < I.L; ++l)
      for (int d = 0; d < g.dim; ++d) citems += 2 * I.op->lev[l].g.nface[d];
    coarsen_chunks[i] = int((citems + chunk - 1) / chunk);
    b->setup_blocks += setup_chunks[i];
    b->coarsen_blocks += coarsen_chunks[i];
    b->res_blocks += I.nres;
  }
  b->K = 0;
  for (const BatchInst& I : b->inst) b->K = std::max(b->K, I.lc);
  b->down.assign(b->K, LinLaunch{});
  b->up.assign(b->K, LinLaunch{});
  const int nv = lin_relax_batch_tiles();
  for (int k = 0; k < b->K; ++k) {
    for (LinLaunch* Lh : {&b->down[k], &b->up[k]}) Lh->blocks_v.assign(nv, 0);
    for (int i = 0; i < n; ++i) {
      const BatchInst& I = b->inst[i];
      const int l = k - (b->K - I.lc);
      if (l < 0) continue;
      for (LinLaunch* Lh : {&b->down[k], &b->up[k]}) {
        Lh->inst.push_back(i);
        Lh->level.push_back(l);
        for (int v = 0; v < nv; ++v) {
          int ntx, nty;
          Lh->blocks_v[v] += lin_relax_batch_tiles_of(v, I.op->lev[l].g, ntx, nty);
        }
      }
    }
    for (LinLaunch* Lh : {&b->down[k], &b->up[k]}) Lh->blocks = Lh->blocks_v[b->variant];
  }
  for (int k = 0; k < b->K; ++k) {
    b->down[k].bytes = launch_bytes(b, b->down[k], true, nullptr);
    b->up[k].bytes = launch_bytes(b, b->up[k], false, nullptr);
  }
  for (int k = 0; k < b->K; ++k) {
    b->down[k].job_off = take(sizeof(BatchLinJob) * b->down[k].inst.size());
    b->up[k].job_off = take(sizeof(BatchLinJob) * b->up[k].inst.size());
  }
  b->setup_map_off = take(sizeof(unsigned) * b->setup_blocks);
  b->coarsen_map_off = take(sizeof(unsigned) * b->coarsen_blocks);
  b->res_map_off = take(sizeof(unsigned) * b->res_blocks);
  auto most = [](const LinLaunch& Lh) { return *std::max_element(Lh.blocks_v.begin(), Lh.blocks_v.end()); };
  for (int k = 0; k < b->K; ++k) {
    b->down[k].map_off = take(sizeof(unsigned) * most(b->down[k]));
    b->up[k].map_off = take(sizeof(unsigned) * most(b->up[k]));
  }
  b->arena_bytes = off;
  ml_ctx ctx = b->ctx;
  void* v = nullptr;
  int rc;
  b->cmap_cap = size_t(b->res_blocks);
  for (int k = 0; k < b->K; ++k) b->cmap_cap += size_t(most(b->down[k])) + size_t(most(b->up[k]));
  if ((rc = dalloc(b, sizeof(unsigned) * b->cmap_cap, &b->cmap))) return rc;
  ML_HIP(ctx, hipHostMalloc(&v, sizeof(unsigned) * b->cmap_cap, hipHostMallocDefault));
  b->hcmap = static_cast<unsigned*>(v);
  if ((rc = dalloc(b, sizeof(int) * n, &b->tlist))) return rc;
  if ((rc = dalloc(b, sizeof(unsigned) * size_t(b->res_blocks), &b->tmap))) return rc;
  if ((rc = dalloc(b, sizeof(unsigned long long) * n, &b->tnorm))) return rc;
  ML_HIP(ctx, hipHostMalloc(&v, sizeof(int) * n, hipHostMallocDefault));
  b->htlist = static_cast<int*>(v);
  ML_HIP(ctx, hipHostMalloc(&v, sizeof(unsigned) * size_t(b->res_blocks), hipHostMallocDefault));
  b->htmap = static_cast<unsigned*>(v);
  ML_HIP(ctx, hipHostMalloc(&v, sizeof(unsigned long long) * n, hipHostMallocDefault));
  b->htnorm = static_cast<unsigned long long*>(v);
  if ((rc = dalloc(b, off, &b->arena))) return rc;
  ML_HIP(ctx, hipHostMalloc(&v, off, hipHostMallocDefault));
  b->harena = static_cast<unsigned char*>(v);
  std::memset(b->harena, 0, off);
  // maps (static)
  auto fill = [&](size_t moff, int job, int count, int& pos) {
    unsigned* m = b->hat<unsigned>(moff);
    for (int t = 0; t < count; ++t) m[pos++] = (unsigned(job) << kBatchTileBits) | unsigned(t);
  };
  for (int i = 0; i < n; ++i)
    if (unsigned(i) >= (1u << (32 - kBatchTileBits)) || setup_chunks[i] > int(kBatchTileMask))
      return batch_fail(b, ML_ENOTSUP, "batched form: too many instances or an instance too large");
  int p = 0;
  for (int i = 0; i < n; ++i) fill(b->setup_map_off, i, setup_chunks[i], p);
  p = 0;
  for (int i = 0; i < n; ++i) fill(b->coarsen_map_off, i, coarsen_chunks[i], p);
  p = 0;
  for (int i = 0; i < n; ++i) fill(b->res_map_off, i, b->inst[i].nres, p);
  return ML_OK;
}

// relaxation maps for the tile shape in use (into the pinned mirror)
void fill_lin_maps(ml_hc_batch b) {
  for (int k = 0; k < b->K; ++k)
    for (LinLaunch* Lh : {&b->down[k], &b->up[k]}) {
      Lh->blocks = Lh->blocks_v[b->variant];
      unsigned* m = b->hat<unsigned>(Lh->map_off);
      int p = 0;
      for (size_t j = 0; j < Lh->inst.size(); ++j) {
        int ntx, nty;
        const int nt = lin_relax_batch_tiles_of(b->variant, b->inst[Lh->inst[j]].op->lev[Lh->level[j]].g, ntx, nty);
        for (int t = 0; t < nt; ++t) m[p++] = (unsigned(j) << kBatchTileBits) | unsigned(t);
      }
    }
}

// job tables for the given states, dt and cycle parameters
int build_tables(ml_hc_batch b, double* const* phi, const double* const* src, double dt, const ml_solve_opts& o) {
  ml_ctx ctx = b->ctx;
  const int n = int(b->inst.size());
  // the pinned mirror may still be read by an earlier upload
  ML_HIP(ctx, hipStreamSynchronize(ctx->stream));
  if (b->gexec) {
    (void)hipGraphExecDestroy(b->gexec);
    b->gexec = nullptr;
  }
  for (int i = 0; i < n; ++i) {
    BatchInst& I = b->inst[i];
    ml_hc_op op = I.op;
    op->alpha = 1.0;
    op->beta = -dt;
    SolveCtl* ctl = b->ctl + i;
    unsigned long long* slots = b->slots + size_t(i) * kBatchSlots;
    // set-up
    BatchSetupJob& S = b->hat<BatchSetupJob>(b->setup_off)[i];
    S.g = op->lev[0].g;
    S.cp = op->cp;
    S.kcond = op->params.heatConductivityRatio;
    S.DH = op->params.heatDiffusionCoeff;
    S.DC = op->params.saltDiffusionCoeff;
    S.phi = phi[i];
    S.alt = I.alt0;
    S.rhs = I.rhs;
    S.src = src ? src[i] : nullptr;
    S.dt = dt;
    for (int d = 0; d < 3; ++d) {
      S.b64[d] = d < op->dim ? op->lev[0].b[d] : nullptr;
      S.b32[d] = d < op->dim ? const_cast<float*>(I.lin[0].b[d]) : nullptr;
    }
    S.slots = slots;
    S.nslots = kBatchSlots;
    S.ctl = ctl;
    // coefficient hierarchy
    BatchCoarsenJob& C = b->hat<BatchCoarsenJob>(b->coarsen_off)[i];
    std::memset(&C, 0, sizeof(C));
    C.g0 = op->lev[0].g;
    for (int d = 0; d < 3; ++d) C.b0[d] = d < op->dim ? op->lev[0].b[d] : nullptr;
    long long tot = 0;
    for (int l = 1; l < I.L; ++l)
      for (int d = 0; d < op->dim; ++d) {
        C.begin[C.njobs] = tot;
        C.level[C.njobs] = short(l);
        C.dir[C.njobs] = short(d);
        C.b64[C.njobs] = op->lev[l].b[d];
        C.b32[C.njobs] = const_cast<float*>(I.lin[l].b[d]);
        tot += 2 * op->lev[l].g.nface[d];
        ++C.njobs;
      }
    C.begin[C.njobs] = tot;
    // residual + Jacobian + stopping rule
    BatchResJob& R = b->hat<BatchResJob>(b->res_off)[i];
    R.A = op->args(0);
    R.A.skip = &ctl->done;
    R.pb[0] = phi[i];
    R.pb[1] = I.alt0;
    R.rhs = I.rhs;
    R.e = I.lin[0].e;
    R.r32 = I.lin[0].r;
    R.M32 = I.lin[0].M;
    R.slots = slots;
    R.ctl = ctl;
    R.arrive = b->arrive + i;
    R.ntiles = unsigned(I.nres);
    R.ntx = I.res_ntx;
    R.nty = I.res_nty;
    // coarse V-cycle
    LinDesc& D = b->hat<LinDesc>(b->desc_off)[i];
    std::memset(&D, 0, sizeof(D));
    D.l0 = I.lc;
    D.nlev = I.L;
    D.nu1 = o.nu1;
    D.nu2 = o.nu2;
    D.bottom = o.bottom;
    D.bc = op->bc;
    D.alpha = op->alpha;
    D.beta = op->beta;
    D.skip = &ctl->done;
    for (int l = 0; l < I.L; ++l) D.lev[l] = I.lin[l];
    for (int d = 0; d < 3; ++d) D.b64[d] = op->lev[I.L - 1].b[d];
    D.a64 = nullptr;
  }
  // tiled relaxations: down-sweep from a zero start with the restriction,
  // up-sweep with the prolongation of the coarser correction
  for (int k = 0; k < b->K; ++k) {
    for (int pass = 0; pass < 2; ++pass) {
      const LinLaunch& Lh = pass == 0 ? b->down[k] : b->up[k];
      BatchLinJob* jobs = b->hat<BatchLinJob>(Lh.job_off);
      for (size_t j = 0; j < Lh.inst.size(); ++j) {
        const BatchInst& I = b->inst[Lh.inst[j]];
        const int l = Lh.level[j];
        BatchLinJob& J = jobs[j];
        std::memset(&J, 0, sizeof(J));
        J.L = I.lin[l];
        J.bc = I.op->bc;
        J.alpha = float(I.op->alpha);
        J.beta = float(I.op->beta);
        lin_relax_batch_tiles_of(b->variant, J.L.g, J.ntx, J.nty);
        if (pass == 0) {
          J.ein = I.lin[l].e;
          J.eout = I.ealt[l];
          J.gc = J.L.g;
          J.pc = nullptr;
          J.gr = I.lin[l + 1].g;
          J.rc = I.lin[l + 1].r;
          J.Mc = I.lin[l + 1].M;
        } else {
          J.ein = I.ealt[l];
          J.eout = I.lin[l].e;
          J.gc = I.lin[l + 1].g;
          J.pc = I.lin[l + 1].e;
          J.gr = J.L.g;
          J.rc = nullptr;
          J.Mc = nullptr;
        }
      }
    }
  }
  fill_lin_maps(b);
  ML_HIP(ctx, hipMemcpyAsync(b->arena, b->harena, b->arena_bytes, hipMemcpyHostToDevice, ctx->stream));
  b->key_phi.assign(phi, phi + n);
  b->key_src.assign(n, nullptr);
  if (src) b->key_src.assign(src, src + n);
  b->key_dt = dt;
  b->key_nu1 = o.nu1;
  b->key_nu2 = o.nu2;
  b->key_bottom = o.bottom;
  b->tables_ok = true;
  return ML_OK;
}

// workgroup maps restricted to the running instances (run[i] != 0)
int build_compact(ml_hc_batch b, const std::vector<char>& run) {
  ml_ctx ctx = b->ctx;
  size_t pos = 0;
  auto lin = [&](const LinLaunch& Lh, size_t& off, int& blocks) {
    off = pos;
    blocks = 0;
    for (size_t j = 0; j < Lh.inst.size(); ++j) {
      if (!run[Lh.inst[j]]) continue;
      int ntx, nty;
      const int nt = lin_relax_batch_tiles_of(b->variant, b->inst[Lh.inst[j]].op->lev[Lh.level[j]].g, ntx, nty);
      for (int t = 0; t < nt; ++t) b->hcmap[pos++] = (unsigned(j) << kBatchTileBits) | unsigned(t);
      blocks += nt;
    }
  };
  b->cdown_off.assign(b->K, 0);
  b->cup_off.assign(b->K, 0);
  b->cdown_blocks.assign(b->K, 0);
  b->cup_blocks.assign(b->K, 0);
  b->cdown_bytes.assign(b->K, 0);
  b->cup_bytes.assign(b->K, 0);
  for (int k = 0; k < b->K; ++k) {
    lin(b->down[k], b->cdown_off[k], b->cdown_blocks[k]);
    lin(b->up[k], b->cup_off[k], b->cup_blocks[k]);
    b->cdown_bytes[k] = launch_bytes(b, b->down[k], true, &run);
    b->cup_bytes[k] = launch_bytes(b, b->up[k], false, &run);
  }
  b->cres_off = pos;
  b->cres_blocks = 0;
  for (size_t i = 0; i < b->inst.size(); ++i) {
    if (!run[i]) continue;
    for (int t = 0; t < b->inst[i].nres; ++t) b->hcmap[pos++] = (unsigned(i) << kBatchTileBits) | unsigned(t);
    b->cres_blocks += b->inst[i].nres;
  }
  ML_HIP(ctx, hipMemcpyAsync(b->cmap, b->hcmap, sizeof(unsigned) * pos, hipMemcpyHostToDevice, ctx->stream));
  return ML_OK;
}

// one outer iteration over the compact maps (direct launches)
hipError_t enqueue_compact(ml_hc_batch b, const ml_solve_opts& o, hipStream_t s, const LaunchTimer* tm) {
  hipError_t e;
  for (int k = 0; k < b->K; ++k)
    if ((e = lin_relax_batch(b->at<BatchLinJob>(b->down[k].job_off), b->cmap + b->cdown_off[k], b->cdown_blocks[k],
                             2 * o.nu1, true, true, b->variant, s, tm, b->cdown_bytes[k])) != hipSuccess)
      return e;
  if ((e = lin_coarse_vcycle_batch(b->at<LinDesc>(b->desc_off), int(b->inst.size()), b->coarse_lds, s)) !=
      hipSuccess)
    return e;
  for (int k = b->K - 1; k >= 0; --k)
    if ((e = lin_relax_batch(b->at<BatchLinJob>(b->up[k].job_off), b->cmap + b->cup_off[k], b->cup_blocks[k],
                             2 * o.nu2, false, false, b->variant, s, tm, b->cup_bytes[k])) != hipSuccess)
      return e;
  return hc_residual_jacobian_batch(b->at<BatchResJob>(b->res_off), b->cmap + b->cres_off, b->cres_blocks, false, s);
}

// one outer iteration: linear V-cycle of every instance, then the FP64
// residual with the stopping rule
hipError_t enqueue_iteration(ml_hc_batch b, const ml_solve_opts& o, hipStream_t s, const LaunchTimer* tm) {
  hipError_t e;
  for (int k = 0; k < b->K; ++k) {
    const LinLaunch& Lh = b->down[k];
    if ((e = lin_relax_batch(b->at<BatchLinJob>(Lh.job_off), b->at<unsigned>(Lh.map_off), Lh.blocks, 2 * o.nu1, true,
                             true, b->variant, s, tm, Lh.bytes)) != hipSuccess)
      return e;
  }
  if ((e = lin_coarse_vcycle_batch(b->at<LinDesc>(b->desc_off), int(b->inst.size()), b->coarse_lds, s)) !=
      hipSuccess)
    return e;
  for (int k = b->K - 1; k >= 0; --k) {
    const LinLaunch& Lh = b->up[k];
    if ((e = lin_relax_batch(b->at<BatchLinJob>(Lh.job_off), b->at<unsigned>(Lh.map_off), Lh.blocks, 2 * o.nu2, false,
                             false, b->variant, s, tm, Lh.bytes)) != hipSuccess)
      return e;
  }
  return hc_residual_jacobian_batch(b->at<BatchResJob>(b->res_off), b->at<unsigned>(b->res_map_off), b->res_blocks,
                                    false, s);
}

// The autotuner: every tile shape of the batched relaxation runs the
// batch's own down- and up-sweep launches (one warm-up, then kTuneReps
// timed with HIP events); the fastest shape is kept.  The relaxations only
// read and write correction buffers that every step recomputes from
// scratch, so the trial runs leave no trace in the results.
constexpr int kTuneReps = 3;

int autotune(ml_hc_batch b, double* const* phi, const double* const* src, double dt, const ml_solve_opts& o) {
  ml_ctx ctx = b->ctx;
  hipStream_t s = ctx->stream;
  const int nv = lin_relax_batch_tiles();
  hipEvent_t e0, e1;
  ML_HIP(ctx, hipEventCreate(&e0));
  ML_HIP(ctx, hipEventCreate(&e1));
  b->tune_ms.assign(nv, 0.0);
  int best = 0, rc = ML_OK;
  for (int v = 0; v < nv && rc == ML_OK; ++v) {
    b->variant = v;
    if ((rc = build_tables(b, phi, src, dt, o))) break;
    for (int rep = 0; rep <= kTuneReps && rc == ML_OK; ++rep) {
      if (rep == 1 && hipEventRecord(e0, s) != hipSuccess) rc = fail(ctx, ML_EHIP, "autotune event");
      for (int k = 0; k < b->K && rc == ML_OK; ++k)
        if (lin_relax_batch(b->at<BatchLinJob>(b->down[k].job_off), b->at<unsigned>(b->down[k].map_off),
                            b->down[k].blocks, 2 * o.nu1, true, true, v, s) != hipSuccess)
          rc = fail(ctx, ML_EHIP, "autotune launch");
      for (int k = b->K - 1; k >= 0 && rc == ML_OK; --k)
        if (lin_relax_batch(b->at<BatchLinJob>(b->up[k].job_off), b->at<unsigned>(b->up[k].map_off), b->up[k].blocks,
                            2 * o.nu2, false, false, v, s) != hipSuccess)
          rc = fail(ctx, ML_EHIP, "autotune launch");
    }
    if (rc != ML_OK) break;
    float ms = 0;
    if (hipEventRecord(e1, s) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
        hipEventElapsedTime(&ms, e0, e1) != hipSuccess) {
      rc = fail(ctx, ML_EHIP, "autotune timing");
      break;
    }
    b->tune_ms[v] = ms / kTuneReps;
    if (b->tune_ms[v] < b->tune_ms[best]) best = v;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc != ML_OK) return rc;
  b->variant = best;
  b->tuned = true;
  return build_tables(b, phi, src, dt, o);
}

int capture(ml_hc_batch b, const ml_solve_opts& o) {
  ml_ctx ctx = b->ctx;
  hipStream_t s = ctx->stream;
  hipGraph_t graph = nullptr;
  ML_HIP(ctx, hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
  const hipError_t e = enqueue_iteration(b, o, s, nullptr);
  const hipError_t e2 = hipStreamEndCapture(s, &graph);
  if (e != hipSuccess || e2 != hipSuccess) {
    if (graph) (void)hipGraphDestroy(graph);
    return fail(ctx, ML_EHIP, std::string("batched iteration capture: ") + hipGetErrorString(e != hipSuccess ? e : e2));
  }
  const hipError_t e3 = hipGraphInstantiate(&b->gexec, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  if (e3 != hipSuccess) {
    b->gexec = nullptr;
    return fail(ctx, ML_EHIP, std::string("batched iteration graph: ") + hipGetErrorString(e3));
  }
  return ML_OK;
}

// host step-length backtracking after a rejected step, the single-instance
// rule for each instance in `idx`: each step length's trials run for all of
// them in one launch (the residual kernel's tile body at step length omega),
// then one synchronisation
int backtrack(ml_hc_batch b, const std::vector<int>& idx) {
  ml_ctx ctx = b->ctx;
  hipStream_t s = ctx->stream;
  const BatchResJob* jobs = b->at<BatchResJob>(b->res_off);
  std::vector<int> trying = idx;
  std::vector<double> norm(b->inst.size());
  for (int i : idx) norm[i] = b->hctl[i].norm;
  double omega = 1.0;
  for (int k = 1; k <= kMaxHalvings && !trying.empty(); ++k) {
    omega *= 0.5;
    const int nt = int(trying.size());
    int pos = 0;
    for (int j = 0; j < nt; ++j) {
      const int i = trying[j];
      b->htlist[j] = i;
      for (int t = 0; t < b->inst[i].nres; ++t) b->htmap[pos++] = (unsigned(i) << kBatchTileBits) | unsigned(t);
    }
    ML_HIP(ctx, hipMemcpyAsync(b->tlist, b->htlist, sizeof(int) * nt, hipMemcpyHostToDevice, s));
    ML_HIP(ctx, hipMemcpyAsync(b->tmap, b->htmap, sizeof(unsigned) * pos, hipMemcpyHostToDevice, s));
    ML_HIP(ctx, batch_trial_slots(jobs, b->tlist, nt, false, nullptr, s));
    ML_HIP(ctx, hc_residual_jacobian_batch_trial(jobs, b->tmap, pos, omega, s));
    ML_HIP(ctx, batch_trial_slots(jobs, b->tlist, nt, true, b->tnorm, s));
    ML_HIP(ctx, hipMemcpyAsync(b->htnorm, b->tnorm, sizeof(unsigned long long) * nt, hipMemcpyDeviceToHost, s));
    ML_HIP(ctx, hipStreamSynchronize(s));
    std::vector<int> left;
    for (int j = 0; j < nt; ++j) {
      const int i = trying[j];
      std::memcpy(&norm[i], b->htnorm + j, sizeof(double));
      if (!(norm[i] < b->hctl[i].prev)) left.push_back(i);
    }
    trying.swap(left);
  }
  for (int i : idx) {
    SolveCtl& c = b->hctl[i];
    const double nm = norm[i];
    c.iter = c.iter + 1;
    int status = 0;
    if (nm <= c.eps * c.r0) status |= 1;
    if (nm <= c.thresh) status |= 8;
    if (status == 0 && nm > (1.0 - c.hang) * c.prev) status |= 4;
    c.prev = nm;
    c.norm = nm;
    c.need_host = 0;
    c.done = 0;
    if (status != 0 || c.iter >= c.max_iter) {
      c.status = status != 0 ? status : 2;
      c.done = 1;
    }
  }
  // no kernel has changed a control block since the host read them
  ML_HIP(ctx, hipMemcpyAsync(b->ctl, b->hctl, sizeof(SolveCtl) * b->inst.size(), hipMemcpyHostToDevice, s));
  ML_HIP(ctx, hipStreamSynchronize(s));
  return ML_OK;
}

}  // namespace
}  // namespace ml

using namespace ml;

extern "C" {

int ml_hc_batch_create(ml_ctx ctx, int count, const ml_hc_op ops[], ml_hc_batch* out) {
  if (!ctx || !ops || !out || count < 1) return ML_EINVAL;
  *out = nullptr;
  ml_hc_batch b = new ml_hc_batch_s();
  b->ctx = ctx;
  b->inst.assign(count, BatchInst{});
  int rc = ML_OK;
  size_t maxlds = 0;
  void* v = nullptr;
  for (int i = 0; i < count && rc == ML_OK; ++i)
    if (!ops[i] || ops[i]->ctx != ctx) rc = fail(ctx, ML_EINVAL, "ml_hc_batch_create: null operator or another context");
  if (rc == ML_OK) rc = max_lds(ctx, &maxlds);
  if (rc == ML_OK) rc = dalloc(b, sizeof(SolveCtl) * count, &b->ctl);
  if (rc == ML_OK) rc = dalloc(b, sizeof(unsigned long long) * kBatchSlots * count, &b->slots);
  if (rc == ML_OK) rc = dalloc(b, sizeof(unsigned int) * count, &b->arrive);
  if (rc == ML_OK) {
    if (hipHostMalloc(&v, sizeof(SolveCtl) * count, hipHostMallocDefault) != hipSuccess)
      rc = fail(ctx, ML_ENOMEM, "hipHostMalloc");
    else
      b->hctl = static_cast<SolveCtl*>(v);
  }
  for (int i = 0; i < count && rc == ML_OK; ++i) rc = inst_init(b, i, ops[i], maxlds);
  if (rc == ML_OK) rc = layout(b);
  if (rc == ML_OK) rc = commit_device(b);
  if (rc == ML_OK && hipStreamSynchronize(ctx->stream) != hipSuccess) rc = fail(ctx, ML_EHIP, "batch set-up");
  if (rc != ML_OK) {
    batch_free(b);
    return rc;
  }
  *out = b;
  return ML_OK;
}

int ml_hc_batch_destroy(ml_hc_batch b) {
  if (!b) return ML_EINVAL;
  batch_free(b);
  return ML_OK;
}

int ml_hc_batch_size(ml_hc_batch b) { return b ? int(b->inst.size()) : ML_EINVAL; }

int ml_hc_batch_memory(ml_hc_batch b, size_t* bytes, int* allocs, int* fills) {
  if (!b) return ML_EINVAL;
  if (bytes) *bytes = b->dev_bytes;
  if (allocs) *allocs = b->dev_allocs;
  if (fills) *fills = b->dev_fills;
  return ML_OK;
}

int ml_hc_batch_set_tile(ml_hc_batch b, int variant) {
  if (!b || variant < -1 || variant >= lin_relax_batch_tiles()) return ML_EINVAL;
  if (variant < 0) {
    b->tuned = false;
  } else {
    b->variant = variant;
    b->tuned = true;
  }
  b->tables_ok = false;
  return ML_OK;
}

int ml_hc_batch_tile(ml_hc_batch b, int* variant, int* tile_x, int* tile_y, int* threads, float* tune_ms,
                     int max_shapes) {
  if (!b || !variant || !tile_x || !tile_y || !threads) return ML_EINVAL;
  *variant = b->variant;
  lin_relax_batch_tile_shape(b->variant, tile_x, tile_y, threads);
  if (tune_ms)
    for (int v = 0; v < max_shapes && v < lin_relax_batch_tiles(); ++v)
      tune_ms[v] = v < int(b->tune_ms.size()) ? float(b->tune_ms[v]) : 0.0f;
  return lin_relax_batch_tiles();
}

int ml_hc_batch_step(ml_hc_batch b, double* const phi[], double dt, const ml_solve_opts* opts, ml_solve_stats stats[]) {
  return ml_hc_batch_step_src(b, phi, nullptr, dt, opts, stats);
}

int ml_hc_batch_step_src(ml_hc_batch b, double* const phi[], const double* const src[], double dt,
                         const ml_solve_opts* opts, ml_solve_stats stats[]) {
  if (!b || !phi) return ML_EINVAL;
  ml_ctx ctx = b->ctx;
  hipStream_t s = ctx->stream;
  ml_solve_opts o;
  if (opts)
    o = *opts;
  else
    ml_default_opts(&o);
  const int n = int(b->inst.size());
  for (int i = 0; i < n; ++i)
    if (!phi[i]) return fail(ctx, ML_EINVAL, "ml_hc_batch_step: null state");
  if (o.precision != ML_PRECISION_FP32_REFINED)
    return fail(ctx, ML_ENOTSUP, "batched form: FP32-refined precision only");
  if (o.max_iter < 0 || o.max_iter + 1 > kBatchSlots) return fail(ctx, ML_EINVAL, "batched form: max_iter too large");
  if (o.nu1 < 1 || o.nu2 < 1 || o.bottom < 1)
    return fail(ctx, ML_EINVAL, "batched form needs nu1, nu2, bottom >= 1");
  // one relaxation launch per level and sweep direction, as a single solve
  // makes it on these levels (2 nu1 passes + the restriction, 2 nu2 passes)
  if (2 * o.nu1 + 1 > 8 || 2 * o.nu2 > 8) return fail(ctx, ML_ENOTSUP, "batched form needs nu1 <= 3 and nu2 <= 4");
  int rc;
  bool same = b->tables_ok && b->key_dt == dt && b->key_nu1 == o.nu1 && b->key_nu2 == o.nu2 &&
              b->key_bottom == o.bottom;
  for (int i = 0; same && i < n; ++i) same = b->key_phi[i] == phi[i] && b->key_src[i] == (src ? src[i] : nullptr);
  if (!b->tuned) {
    if ((rc = autotune(b, phi, src, dt, o))) return rc;
  } else if (!same && (rc = build_tables(b, phi, src, dt, o))) {
    return rc;
  }
  for (BatchInst& I : b->inst) {
    I.op->alpha = 1.0;
    I.op->beta = -dt;
  }
  const bool use_graph = !(o.flags & ML_SOLVE_NO_GRAPH);
  const LaunchTimer tm = ctx->timer(ML_KT_BATCH_RELAX);
  if (use_graph && !b->gexec && (rc = capture(b, o))) return rc;
  // set-up, initial residual, iteration control
  ML_HIP(ctx, batch_setup(b->at<BatchSetupJob>(b->setup_off), b->at<unsigned>(b->setup_map_off), b->setup_blocks, s));
  ML_HIP(ctx, batch_coarsen(b->at<BatchCoarsenJob>(b->coarsen_off), b->at<unsigned>(b->coarsen_map_off),
                            b->coarsen_blocks, s));
  ML_HIP(ctx, hc_residual_jacobian_batch(b->at<BatchResJob>(b->res_off), b->at<unsigned>(b->res_map_off),
                                         b->res_blocks, true, s));
  ML_HIP(ctx, batch_ctl_init(b->ctl, b->slots, kBatchSlots, n, o.max_iter, o.eps, o.hang, o.norm_thresh, s));
  const bool lookahead = !(o.flags & ML_SOLVE_NO_LOOKAHEAD);
  int batch = lookahead ? std::max(1, std::min(b->batch_iters, o.max_iter)) : 1;
  int low = 0;   // fewest accepted iterations among the running instances
  int fewest = o.max_iter;   // iterations of the instance that finished first
  int compacted = n;         // running instances when the maps were last compacted
  bool compact = false;
  std::vector<char> run(n, 1);
  for (;;) {
    const int todo = std::min(batch, o.max_iter - low);
    for (int k = 0; k < todo; ++k) {
      if (compact)
        ML_HIP(ctx, enqueue_compact(b, o, s, &tm));
      else if (use_graph)
        ML_HIP(ctx, hipGraphLaunch(b->gexec, s));
      else
        ML_HIP(ctx, enqueue_iteration(b, o, s, &tm));
    }
    ML_HIP(ctx, hipMemcpyAsync(b->hctl, b->ctl, sizeof(SolveCtl) * n, hipMemcpyDeviceToHost, s));
    ML_HIP(ctx, hipStreamSynchronize(s));
    std::vector<int> rejected;
    for (int i = 0; i < n; ++i)
      if (b->hctl[i].need_host) rejected.push_back(i);
    if (!rejected.empty() && (rc = backtrack(b, rejected))) return rc;
    int running = 0;
    low = o.max_iter;
    for (int i = 0; i < n; ++i) {
      const SolveCtl& c = b->hctl[i];
      run[i] = !c.done && c.iter < o.max_iter;
      if (run[i]) {
        ++running;
        low = std::min(low, c.iter);
      } else {
        fewest = std::min(fewest, c.iter);
      }
    }
    if (!running) break;
    // once the running instances have halved, later iterations launch only
    // their tiles (the finished ones would return at once from every
    // workgroup)
    if (2 * running <= compacted) {
      if ((rc = build_compact(b, run))) return rc;
      compact = true;
      compacted = running;
    }
    batch = lookahead ? 2 : 1;
  }
  // the next step's first host check comes when the first instance is
  // expected to finish
  if (lookahead) b->batch_iters = std::max(1, std::min(fewest, o.max_iter));
  ML_HIP(ctx, batch_finish(b->at<BatchSetupJob>(b->setup_off), b->at<unsigned>(b->setup_map_off), b->setup_blocks, s));
  if (stats) {
    ML_HIP(ctx, hipMemcpyAsync(b->hctl, b->ctl, sizeof(SolveCtl) * n, hipMemcpyDeviceToHost, s));
    ML_HIP(ctx, hipStreamSynchronize(s));
    for (int i = 0; i < n; ++i) {
      const SolveCtl& c = b->hctl[i];
      stats[i].exit_status = c.status != 0 ? c.status : 2;
      stats[i].iterations = c.iter;
      stats[i].initial_residual = c.r0;
      stats[i].final_residual = c.norm;
    }
  }
  return ML_OK;
}

}  // extern "C"
