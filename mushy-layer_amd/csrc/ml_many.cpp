// ml_many.cpp — batched forms of per-operator set-up calls for many
// independent operators (ml_vc_set_coefs_many, ml_hc_set_coefs_many): the
// level-0 coefficient copies of every operator in one launch and the
// coarsened hierarchies of every operator in two (coarsen_records /
// coarsen_recs, the arithmetic of coarsen_hierarchy), instead of several
// launches per operator.  Results equal the per-operator calls bit for bit.
#include "ml_internal.h"

namespace ml {

hipError_t TableStage::put(const void* src, size_t bytes, hipStream_t s, void** out) {
  const int i = next;
  next ^= 1;
  hipError_t e = hipSuccess;
  if (ev[i]) {
    if ((e = hipEventSynchronize(ev[i])) != hipSuccess) return e;
  } else if ((e = hipEventCreateWithFlags(&ev[i], hipEventDisableTiming)) != hipSuccess) {
    return e;
  }
  if (host_cap[i] < bytes) {
    if (host[i]) hipHostFree(host[i]);
    host[i] = nullptr;
    host_cap[i] = 0;
    if ((e = hipHostMalloc(&host[i], bytes, hipHostMallocDefault)) != hipSuccess) return e;
    host_cap[i] = bytes;
  }
  if (dev_cap < bytes) {
    // kernels enqueued earlier may still read the old table
    if ((e = hipStreamSynchronize(s)) != hipSuccess) return e;
    if (dev) hipFree(dev);
    dev = nullptr;
    dev_cap = 0;
    if ((e = hipMalloc(&dev, bytes)) != hipSuccess) return e;
    dev_cap = bytes;
  }
  std::memcpy(host[i], src, bytes);
  if ((e = hipMemcpyAsync(dev, host[i], bytes, hipMemcpyHostToDevice, s)) != hipSuccess) return e;
  if ((e = hipEventRecord(ev[i], s)) != hipSuccess) return e;
  *out = dev;
  return hipSuccess;
}

void TableStage::release() {
  for (int i = 0; i < 2; ++i) {
    if (ev[i]) hipEventDestroy(ev[i]);
    if (host[i]) hipHostFree(host[i]);
    ev[i] = nullptr;
    host[i] = nullptr;
    host_cap[i] = 0;
  }
  if (dev) hipFree(dev);
  dev = nullptr;
  dev_cap = 0;
}

}  // namespace ml

namespace {

using ml::fail;

size_t align256(size_t x) { return (x + 255) & ~size_t(255); }

// Level-0 copies and hierarchy averages of many operators, packed into one
// staged table: [copy jobs | copy map | light records | light map | heavy
// records | heavy map], then three launches.
struct ManyCoefs {
  std::vector<ml::WordsJob> words;
  std::vector<unsigned> wmap;
  std::vector<ml::CoarsenRec<double>> light, heavy;
  int dim = 2;

  void copy(const double* src, double* dst, long long n) {
    const unsigned j = unsigned(words.size());
    words.push_back(ml::WordsJob{reinterpret_cast<const unsigned long long*>(src),
                                 reinterpret_cast<unsigned long long*>(dst), n});
    for (long long c = 0; c * ml::kWordsChunk < n; ++c) wmap.push_back((j << ml::kBatchTileBits) | unsigned(c));
  }

  int run(ml_ctx ctx) {
    if (words.size() > (size_t(1) << (32 - ml::kBatchTileBits)))
      return fail(ctx, ML_ENOTSUP, "batched set-up: too many coefficient copies in one call");
    std::vector<unsigned long long> lmap, hmap;
    for (size_t r = 0; r < light.size(); ++r)
      for (long long c = 0; c * 256 < light[r].count; ++c) lmap.push_back((static_cast<unsigned long long>(r) << 32) | c);
    for (size_t r = 0; r < heavy.size(); ++r)
      for (long long c = 0; c < heavy[r].count; ++c) hmap.push_back((static_cast<unsigned long long>(r) << 32) | c);
    size_t off[7];
    off[0] = 0;
    off[1] = align256(off[0] + words.size() * sizeof(ml::WordsJob));
    off[2] = align256(off[1] + wmap.size() * sizeof(unsigned));
    off[3] = align256(off[2] + light.size() * sizeof(ml::CoarsenRec<double>));
    off[4] = align256(off[3] + lmap.size() * sizeof(unsigned long long));
    off[5] = align256(off[4] + heavy.size() * sizeof(ml::CoarsenRec<double>));
    off[6] = align256(off[5] + hmap.size() * sizeof(unsigned long long));
    std::vector<unsigned char> blob(off[6]);
    std::memcpy(blob.data() + off[0], words.data(), words.size() * sizeof(ml::WordsJob));
    std::memcpy(blob.data() + off[1], wmap.data(), wmap.size() * sizeof(unsigned));
    std::memcpy(blob.data() + off[2], light.data(), light.size() * sizeof(ml::CoarsenRec<double>));
    std::memcpy(blob.data() + off[3], lmap.data(), lmap.size() * sizeof(unsigned long long));
    std::memcpy(blob.data() + off[4], heavy.data(), heavy.size() * sizeof(ml::CoarsenRec<double>));
    std::memcpy(blob.data() + off[5], hmap.data(), hmap.size() * sizeof(unsigned long long));
    if (blob.empty()) return ML_OK;
    void* dv = nullptr;
    ML_HIP(ctx, ctx->tables.put(blob.data(), blob.size(), ctx->stream, &dv));
    unsigned char* d = static_cast<unsigned char*>(dv);
    // the copies first: the averages read the finest level's copies
    ML_HIP(ctx, ml::words_batch(reinterpret_cast<const ml::WordsJob*>(d + off[0]),
                                reinterpret_cast<const unsigned*>(d + off[1]), int(wmap.size()), ctx->stream));
    ML_HIP(ctx, ml::coarsen_recs<double>(dim, reinterpret_cast<const ml::CoarsenRec<double>*>(d + off[2]),
                                         reinterpret_cast<const unsigned long long*>(d + off[3]), int(lmap.size()),
                                         reinterpret_cast<const ml::CoarsenRec<double>*>(d + off[4]),
                                         reinterpret_cast<const unsigned long long*>(d + off[5]), int(hmap.size()),
                                         ctx->stream));
    return ML_OK;
  }
};

int vc_alloc_level(ml_ctx ctx, size_t n, double** p) {
  void* v = nullptr;
  ML_HIP(ctx, hipMalloc(&v, n * sizeof(double)));
  *p = static_cast<double*>(v);
  ML_HIP(ctx, ml::vec_set((long long)n, *p, 0.0, ctx->stream));
  return ML_OK;
}

}  // namespace

extern "C" {

int ml_vc_set_coefs_many(ml_ctx ctx, int n, const ml_vc_op ops[], const double* const aCoef[],
                         const double* const bCoef[], const double* const cCoef[], const double alpha[],
                         const double beta[]) {
  if (!ctx || n < 0 || (n > 0 && (!ops || !bCoef || !alpha || !beta))) return ML_EINVAL;
  ManyCoefs M;
  for (int i = 0; i < n; ++i) {
    ml_vc_op op = ops[i];
    if (!op || op->ctx != ctx) return fail(ctx, ML_EINVAL, "ml_vc_set_coefs_many: operator of another context");
    for (int j = 0; j < i; ++j)
      if (ops[j] == op) return fail(ctx, ML_EINVAL, "ml_vc_set_coefs_many: an operator appears twice");
    if (i > 0 && op->dim != ops[0]->dim) return fail(ctx, ML_EINVAL, "ml_vc_set_coefs_many: operators of mixed dimension");
    for (int d = 0; d < op->dim; ++d)
      if (!bCoef[i * op->dim + d]) return fail(ctx, ML_EINVAL, "ml_vc_set_coefs_many: null bCoef");
  }
  if (n == 0) return ML_OK;
  M.dim = ops[0]->dim;
  for (int i = 0; i < n; ++i) {
    ml_vc_op op = ops[i];
    const int dim = op->dim;
    const double* a = aCoef ? aCoef[i] : nullptr;
    const double* c = cCoef ? cCoef[i] : nullptr;
    op->alpha = alpha[i];
    op->beta = beta[i];
    op->has_a = a != nullptr;
    op->has_c = c != nullptr;
    for (auto& L : op->lev) {
      int rc;
      if (a && !L.a && (rc = vc_alloc_level(ctx, L.g.ncell, &L.a))) return rc;
      if (c && !L.c && (rc = vc_alloc_level(ctx, L.g.ncell, &L.c))) return rc;
    }
    const ml::Geom& g0 = op->lev[0].g;
    for (int d = 0; d < dim; ++d) M.copy(bCoef[i * dim + d], op->lev[0].b[d], g0.nface[d]);
    if (a) M.copy(a, op->lev[0].a, g0.ncell);
    if (c) M.copy(c, op->lev[0].c, g0.ncell);
    // the hierarchy of ml_vc_set_coefs (DarcyBrinkmanOpFactory::MGnewOp,
    // src/DarcyBrinkmanOp.cpp:977-1008): faces and a, then c
    const int L = int(op->lev.size());
    const double* bf[3] = {op->lev[0].b[0], op->lev[0].b[1], op->lev[0].b[2]};
    std::vector<double* const*> bc(L), nob(L, nullptr);
    std::vector<double*> ac(L, nullptr), cc(L, nullptr);
    for (int l = 0; l < L; ++l) {
      bc[l] = op->lev[l].b;
      ac[l] = op->lev[l].a;
      cc[l] = op->lev[l].c;
    }
    ml::coarsen_records<double>(g0, 1, L, bf, bc.data(), a ? op->lev[0].a : nullptr, ac.data(), true, M.light, M.heavy);
    if (c) {
      const double* none[3] = {nullptr, nullptr, nullptr};
      ml::coarsen_records<double>(g0, 1, L, none, nob.data(), op->lev[0].c, cc.data(), false, M.light, M.heavy);
    }
  }
  return M.run(ctx);
}

int ml_hc_set_coefs_many(ml_ctx ctx, int n, const ml_hc_op ops[], const double* const aCoef[],
                         const double* const bCoef[], const double alpha[], const double beta[]) {
  if (!ctx || n < 0 || (n > 0 && (!ops || !bCoef || !alpha || !beta))) return ML_EINVAL;
  ManyCoefs M;
  for (int i = 0; i < n; ++i) {
    ml_hc_op op = ops[i];
    if (!op || op->ctx != ctx) return fail(ctx, ML_EINVAL, "ml_hc_set_coefs_many: operator of another context");
    for (int j = 0; j < i; ++j)
      if (ops[j] == op) return fail(ctx, ML_EINVAL, "ml_hc_set_coefs_many: an operator appears twice");
    if (i > 0 && op->dim != ops[0]->dim) return fail(ctx, ML_EINVAL, "ml_hc_set_coefs_many: operators of mixed dimension");
    for (int d = 0; d < op->dim; ++d)
      if (!bCoef[i * op->dim + d]) return fail(ctx, ML_EINVAL, "ml_hc_set_coefs_many: null bCoef");
  }
  if (n == 0) return ML_OK;
  M.dim = ops[0]->dim;
  for (int i = 0; i < n; ++i) {
    ml_hc_op op = ops[i];
    const int dim = op->dim;
    const double* a = aCoef ? aCoef[i] : nullptr;
    op->alpha = alpha[i];
    op->beta = beta[i];
    for (auto& L : op->lev) {
      if (a && !L.a) {
        void* p = nullptr;
        if (hipMalloc(&p, sizeof(double) * 2 * L.g.ncell) != hipSuccess) return fail(ctx, ML_ENOMEM, "hipMalloc aCoef");
        L.a = static_cast<double*>(p);
        ML_HIP(ctx, ml::vec_set(2 * L.g.ncell, L.a, 0.0, ctx->stream));
      }
      if (!a && L.a) {
        hipStreamSynchronize(ctx->stream);
        hipFree(L.a);
        L.a = nullptr;
      }
    }
    const ml::Geom& g0 = op->lev[0].g;
    for (int d = 0; d < dim; ++d) M.copy(bCoef[i * dim + d], op->lev[0].b[d], 2 * g0.nface[d]);
    if (a) M.copy(a, op->lev[0].a, 2 * g0.ncell);
    // the hierarchy of ml_hc_set_coefs (MGnewOp,
    // src/AMRNonLinearMultiCompOp.cpp:1423-1494)
    const int L = int(op->lev.size());
    const double* bf[3] = {op->lev[0].b[0], op->lev[0].b[1], op->lev[0].b[2]};
    std::vector<double* const*> bc(L);
    std::vector<double*> ac(L, nullptr);
    for (int l = 0; l < L; ++l) {
      bc[l] = op->lev[l].b;
      ac[l] = op->lev[l].a;
    }
    ml::coarsen_records<double>(g0, 2, L, bf, bc.data(), a ? op->lev[0].a : nullptr, ac.data(), true, M.light, M.heavy);
  }
  return M.run(ctx);
}

}  // extern "C"
