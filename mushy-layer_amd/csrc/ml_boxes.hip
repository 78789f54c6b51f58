// ml_boxes.hip — multi-box levels: copies between the boxes of a
// DisjointBoxLayout (each box its own FAB with one ghost layer, as
// LevelData<FArrayBox> holds them) and the single array on the problem
// domain that the solves run on, plus the face coefficients (FluxBox).
//
// The reference's levels are DisjointBoxLayouts (max_grid_size 128 cuts a
// 256^2 domain into four boxes, srcSubcycle/MushyLayerSubcycleUtils.cpp
// :1011-1012).  One GPU holds the whole level, so the solve itself needs no
// box structure: gather the boxes into the domain array once, solve, and
// scatter back, where the scatter also does what LevelData::exchange does
// for the boxes' ghost cells (values from the neighbouring box or the
// periodic image).  Each copy is one pass over the boxes' cells (one launch
// for all boxes, a workgroup map of (box, chunk) pairs).
#include <hip/hip_runtime.h>

#include <vector>

#include "ml_internal.h"

namespace ml {
namespace {

struct BoxJob {
  int lo[3], hi[3];   // inclusive valid cell bounds in the domain
  double* fab;        // the box's FAB (box grown by 1, component outermost)
  long long first;    // first item (cell of the box, grown for the scatter) of this box
};

constexpr int kMaxBoxes = 4096;

struct BoxJobs {
  const BoxJob* jobs;
  int nbox;
  long long total;
};

__device__ inline int find_box(const BoxJob* J, int nbox, long long t) {
  int a = 0, b = nbox - 1;
  while (a < b) {
    const int m = (a + b + 1) / 2;
    if (J[m].first <= t)
      a = m;
    else
      b = m - 1;
  }
  return a;
}

// gather (grow = 0) or scatter (grow = 1) one item: a cell of box b (grown by
// one for the scatter) and its FAB index
template <int DIM>
__global__ void box_copy_kernel(Geom g, BoxJobs B, int ncomp, int grow, int to_level, double* __restrict__ level) {
  for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < B.total;
       t += (long long)gridDim.x * blockDim.x) {
    const int b = find_box(B.jobs, B.nbox, t);
    const BoxJob& J = B.jobs[b];
    const long long r = t - J.first;
    const int bx = J.hi[0] - J.lo[0] + 1, by = J.hi[1] - J.lo[1] + 1;
    const int ex = bx + 2 * grow, ey = by + 2 * grow;
    const int i = int(r % ex) - grow;
    const long long rr = r / ex;
    const int j = int(rr % ey) - grow;
    const int k = DIM == 3 ? int(rr / ey) - grow : 0;
    // FAB layout of the box: grown by one in every direction
    const long long fsy = bx + 2, fsz = (long long)(bx + 2) * (by + 2);
    const long long fncell = fsz * (DIM == 3 ? (J.hi[2] - J.lo[2] + 3) : 1);
    const long long fq = (i + 1) + fsy * (j + 1) + fsz * (DIM == 3 ? k + 1 : 0);
    int gi[3] = {J.lo[0] + i, J.lo[1] + j, DIM == 3 ? J.lo[2] + k : 0};
    bool ok = true;
    for (int d = 0; d < DIM; ++d) {
      if (gi[d] >= 0 && gi[d] < g.n[d]) continue;
      if (g.periodic[d])
        gi[d] = gi[d] < 0 ? gi[d] + g.n[d] : gi[d] - g.n[d];
      else
        ok = false;   // a ghost on a physical wall: left to the caller's BCs
    }
    if (!ok) continue;
    const long long lq = g.ci(gi[0], gi[1], gi[2]);
    for (int c = 0; c < ncomp; ++c) {
      if (to_level)
        level[lq + c * g.ncell] = J.fab[fq + c * fncell];
      else
        J.fab[fq + c * fncell] = level[lq + c * g.ncell];
    }
  }
}

// faces of direction d: box b's faces on surroundingNodes(box, d) into the
// level's face array (boxes later in the list overwrite shared faces)
template <int DIM>
__global__ void box_faces_kernel(Geom g, int d, BoxJobs B, int ncomp, double* __restrict__ lf) {
  for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < B.total;
       t += (long long)gridDim.x * blockDim.x) {
    const int b = find_box(B.jobs, B.nbox, t);
    const BoxJob& J = B.jobs[b];
    const long long r = t - J.first;
    int fn[3];
    for (int e = 0; e < 3; ++e) fn[e] = e < DIM ? J.hi[e] - J.lo[e] + 1 + (e == d) : 1;
    const int i = int(r % fn[0]);
    const long long rr = r / fn[0];
    const int j = int(rr % fn[1]);
    const int k = DIM == 3 ? int(rr / fn[1]) : 0;
    const long long fq = i + (long long)fn[0] * (j + (long long)fn[1] * k);
    const long long fnf = (long long)fn[0] * fn[1] * fn[2];
    const long long lq = g.fi(d, J.lo[0] + i, J.lo[1] + j, DIM == 3 ? J.lo[2] + k : 0);
    for (int c = 0; c < ncomp; ++c) lf[lq + c * g.nface[d]] = J.fab[fq + c * fnf];
  }
}

int grid_for(long long n) {
  long long b = (n + 255) / 256;
  if (b > 65535 * 4) b = 65535 * 4;
  return b < 1 ? 1 : int(b);
}

// checks the boxes (inside the domain, disjoint, covering it) and builds
// the job table; `faces` >= 0: the items are the faces of that direction
int build_jobs(ml_ctx ctx, const Geom& g, int nbox, const int* lo, const int* hi, double* const* fabs, int grow,
               int faces, std::vector<BoxJob>& jobs, long long& total) {
  if (nbox < 1 || nbox > kMaxBoxes) return fail(ctx, ML_EINVAL, "box count out of range");
  jobs.assign(nbox, BoxJob{});
  long long vol = 0;
  total = 0;
  for (int b = 0; b < nbox; ++b) {
    BoxJob& J = jobs[b];
    if (!fabs[b]) return fail(ctx, ML_EINVAL, "null box FAB");
    long long v = 1, items = 1;
    for (int d = 0; d < 3; ++d) {
      J.lo[d] = d < g.dim ? lo[3 * b + d] : 0;
      J.hi[d] = d < g.dim ? hi[3 * b + d] : 0;
      if (d < g.dim && (J.lo[d] < 0 || J.hi[d] >= g.n[d] || J.hi[d] < J.lo[d]))
        return fail(ctx, ML_EINVAL, "box outside the domain");
      const long long len = J.hi[d] - J.lo[d] + 1;
      v *= len;
      items *= d < g.dim ? (faces >= 0 ? len + (d == faces) : len + 2 * grow) : 1;
    }
    J.fab = fabs[b];
    J.first = total;
    total += items;
    vol += v;
    for (int c = 0; c < b; ++c) {
      bool overlap = true;
      for (int d = 0; d < g.dim; ++d)
        if (jobs[c].hi[d] < J.lo[d] || J.hi[d] < jobs[c].lo[d]) overlap = false;
      if (overlap) return fail(ctx, ML_EINVAL, "boxes overlap");
    }
  }
  if (vol != g.nvalid()) return fail(ctx, ML_EINVAL, "the boxes do not cover the domain");
  return ML_OK;
}

int run(ml_ctx ctx, int dim, const int n[], const int periodic[], int ncomp, int nbox, const int box_lo[],
        const int box_hi[], double* const* fabs, double* level, int to_level, int face_dir) {
  if (!ctx || !n || !box_lo || !box_hi || !fabs || !level || ncomp < 1) return ML_EINVAL;
  int per[3] = {0, 0, 0};
  if (periodic)
    for (int d = 0; d < dim; ++d) per[d] = periodic[d];
  if (!valid_grid(dim, n, per, 1.0)) return fail(ctx, ML_EINVAL, "bad domain");
  const Geom g = make_geom(dim, n, per, 1.0);
  std::vector<BoxJob> jobs;
  long long total = 0;
  const int grow = (face_dir < 0 && !to_level) ? 1 : 0;
  int rc = build_jobs(ctx, g, nbox, box_lo, box_hi, fabs, grow, face_dir, jobs, total);
  if (rc) return rc;
  // the job table travels in a small device buffer owned by this call; it is
  // freed after the stream has consumed it
  void* dj = nullptr;
  ML_HIP(ctx, hipMalloc(&dj, sizeof(BoxJob) * jobs.size()));
  hipError_t e = hipMemcpyAsync(dj, jobs.data(), sizeof(BoxJob) * jobs.size(), hipMemcpyHostToDevice, ctx->stream);
  BoxJobs B{static_cast<const BoxJob*>(dj), nbox, total};
  if (e == hipSuccess) {
    if (face_dir >= 0) {
      if (dim == 2)
        hipLaunchKernelGGL((box_faces_kernel<2>), dim3(grid_for(total)), dim3(256), 0, ctx->stream, g, face_dir, B,
                           ncomp, level);
      else
        hipLaunchKernelGGL((box_faces_kernel<3>), dim3(grid_for(total)), dim3(256), 0, ctx->stream, g, face_dir, B,
                           ncomp, level);
    } else if (dim == 2) {
      hipLaunchKernelGGL((box_copy_kernel<2>), dim3(grid_for(total)), dim3(256), 0, ctx->stream, g, B, ncomp, grow,
                         to_level, level);
    } else {
      hipLaunchKernelGGL((box_copy_kernel<3>), dim3(grid_for(total)), dim3(256), 0, ctx->stream, g, B, ncomp, grow,
                         to_level, level);
    }
    e = hipGetLastError();
  }
  const hipError_t e2 = hipStreamSynchronize(ctx->stream);
  hipFree(dj);
  if (e != hipSuccess) return fail(ctx, ML_EHIP, std::string("box copy: ") + hipGetErrorString(e));
  if (e2 != hipSuccess) return fail(ctx, ML_EHIP, std::string("box copy: ") + hipGetErrorString(e2));
  return ML_OK;
}

}  // namespace
}  // namespace ml

extern "C" {

int ml_level_gather(ml_ctx ctx, int dim, const int n[], const int periodic[], int ncomp, int nbox,
                    const int box_lo[], const int box_hi[], const double* const fabs[], double* level) {
  return ml::run(ctx, dim, n, periodic, ncomp, nbox, box_lo, box_hi, const_cast<double* const*>(fabs), level, 1, -1);
}

int ml_level_scatter(ml_ctx ctx, int dim, const int n[], const int periodic[], int ncomp, int nbox,
                     const int box_lo[], const int box_hi[], const double* level, double* const fabs[]) {
  return ml::run(ctx, dim, n, periodic, ncomp, nbox, box_lo, box_hi, fabs, const_cast<double*>(level), 0, -1);
}

int ml_level_gather_faces(ml_ctx ctx, int dim, const int n[], int ncomp, int nbox, const int box_lo[],
                          const int box_hi[], int dir, const double* const fabs[], double* level_faces) {
  if (dir < 0 || dir >= dim) return ML_EINVAL;
  return ml::run(ctx, dim, n, nullptr, ncomp, nbox, box_lo, box_hi, const_cast<double* const*>(fabs), level_faces, 1,
                 dir);
}

}  // extern "C"
