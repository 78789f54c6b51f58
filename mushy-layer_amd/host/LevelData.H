// LevelData.H — device-resident single-box LevelData<FArrayBox> stand-in for
// host C++ callers of libmlhip.so.
//
// The reference's operators act on Chombo LevelData<FArrayBox> (one FAB per
// box, one ghost layer, component outermost).  Here a level is one box
// starting at the origin; the data live in HBM in exactly the FAB layout the
// C ABI (include/mlhip.h) expects, so a host LevelData can be uploaded or
// downloaded with one copy.
#pragma once

#include <stdexcept>
#include <string>
#include <utility>

#include "mlhip.h"

namespace mlhost {

inline void check(int rc, ml_ctx ctx, const char* what) {
  if (rc != ML_OK) {
    std::string msg = std::string(what) + " failed (" + std::to_string(rc) + ")";
    if (ctx) msg += ": " + std::string(ml_last_error(ctx));
    throw std::runtime_error(msg);
  }
}

class LevelData {
 public:
  LevelData() = default;
  LevelData(ml_ctx ctx, int dim, const int n[3], int ncomp) { define(ctx, dim, n, ncomp); }
  ~LevelData() { clear(); }
  LevelData(const LevelData&) = delete;
  LevelData& operator=(const LevelData&) = delete;
  LevelData(LevelData&& o) noexcept { swap(o); }
  LevelData& operator=(LevelData&& o) noexcept {
    if (this != &o) {
      clear();
      swap(o);
    }
    return *this;
  }

  void define(ml_ctx ctx, int dim, const int n[3], int ncomp) {
    clear();
    m_ctx = ctx;
    m_dim = dim;
    for (int d = 0; d < 3; ++d) m_n[d] = d < dim ? n[d] : 1;
    m_ncomp = ncomp;
    m_ncell = ml_cell_count(dim, m_n);
    void* p = nullptr;
    check(ml_buf_alloc(ctx, bytes(), &p), ctx, "ml_buf_alloc");
    m_data = static_cast<double*>(p);
    m_owned = true;
    check(ml_vec_set(ctx, size(), m_data, 0.0), ctx, "ml_vec_set");
  }

  // Wrap caller-owned device memory (no ownership transfer).
  void alias(ml_ctx ctx, int dim, const int n[3], int ncomp, double* dev) {
    clear();
    m_ctx = ctx;
    m_dim = dim;
    for (int d = 0; d < 3; ++d) m_n[d] = d < dim ? n[d] : 1;
    m_ncomp = ncomp;
    m_ncell = ml_cell_count(dim, m_n);
    m_data = dev;
    m_owned = false;
  }

  bool isDefined() const { return m_data != nullptr; }
  int nComp() const { return m_ncomp; }
  int dim() const { return m_dim; }
  const int* n() const { return m_n; }
  size_t ncell() const { return m_ncell; }
  size_t size() const { return m_ncell * m_ncomp; }
  size_t bytes() const { return size() * sizeof(double); }
  double* data() { return m_data; }
  const double* data() const { return m_data; }
  ml_ctx ctx() const { return m_ctx; }

  void upload(const double* host) { check(ml_buf_upload(m_ctx, m_data, host, bytes()), m_ctx, "upload"); }
  void download(double* host) const { check(ml_buf_download(m_ctx, host, m_data, bytes()), m_ctx, "download"); }
  void copyFrom(const LevelData& o) {
    if (o.size() != size()) throw std::runtime_error("LevelData::copyFrom size mismatch");
    check(ml_buf_copy(m_ctx, m_data, o.m_data, bytes()), m_ctx, "copy");
  }

 private:
  void clear() {
    if (m_owned && m_data) ml_buf_free(m_ctx, m_data);
    m_data = nullptr;
    m_owned = false;
  }
  void swap(LevelData& o) {
    std::swap(m_ctx, o.m_ctx);
    std::swap(m_dim, o.m_dim);
    std::swap(m_n, o.m_n);
    std::swap(m_ncomp, o.m_ncomp);
    std::swap(m_ncell, o.m_ncell);
    std::swap(m_data, o.m_data);
    std::swap(m_owned, o.m_owned);
  }

  ml_ctx m_ctx = nullptr;
  int m_dim = 0;
  int m_n[3] = {0, 0, 0};
  int m_ncomp = 0;
  size_t m_ncell = 0;
  double* m_data = nullptr;
  bool m_owned = false;
};

}  // namespace mlhost
