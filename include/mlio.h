/*
 * mlio.h — the reference's on-disk format: Chombo HDF5 checkpoint and plot
 * files (srcSubcycle/AMRLevelMushyLayerIO.cpp:131-183 writeCheckpointHeader,
 * :184-241 writeCheckpointLevel, :243-420 readCheckpointLevel, :525-551
 * writePlotHeader, :554-622 writePlotLevel), written and read by
 * libmlio.so, a host-only library over the HDF5 C API (no GPU calls; the
 * caller copies device arrays to host memory first).
 *
 * File layout (Chombo's HDF5 AMR layout, which the reference writes through
 * Chombo's HDF5Handle / HDF5HeaderData / write(LevelData)):
 *   /                 attributes: num_levels, max_level, iteration, time,
 *                     num_components, component_<i>, ... (header data)
 *   /Chombo_global    attributes SpaceDim, testReal
 *   /level_<l>        attributes (ref_ratio, dx, dt, time, prob_domain, ...);
 *                     datasets "boxes" (compound lo_i,lo_j[,lo_k],hi_i,hi_j[,hi_k])
 *                     and "Processors" (int per box);
 *                     per variable: "<name>:datatype=0" (doubles, box after
 *                     box, each box's components one after another over the
 *                     box grown by outputGhost, i fastest),
 *                     "<name>:offsets=0" (int64, nbox+1 prefix offsets),
 *                     group "<name>_attributes" (comps, ghost, outputGhost,
 *                     objectType "FArrayBox"; "unknown" for face data).
 * Face data (LevelData<FluxBox>, e.g. the checkpoint's "advVel") stores per
 * box the face FAB of direction 0, then 1 (then 2), each over
 * surroundingNodes(box, d) with all components.
 *
 * In memory: a box's FAB is the layout of mlhip.h (box grown by `ghost`,
 * component outermost, i fastest); face FABs are surroundingNodes(box, d)
 * without ghosts.  Boxes are given by nbox*3 inclusive lo/hi index triples.
 *
 * Errors: ML_H5_OK (0) or a negative code; ml_h5_error() returns the last
 * message of the calling thread.
 */
#ifndef MLIO_H
#define MLIO_H

#ifdef __cplusplus
extern "C" {
#endif

enum { ML_H5_OK = 0, ML_H5_EINVAL = -1, ML_H5_EIO = -2, ML_H5_ENOTFOUND = -3 };

typedef struct ml_h5_s* ml_h5;

const char* ml_h5_error(void);

/* create (truncate) a file of space dimension dim; writes /Chombo_global
 * (HDF5Handle::CREATE) */
int ml_h5_create(const char* path, int dim, ml_h5* out);
/* open read-only; the dimension comes from /Chombo_global/SpaceDim */
int ml_h5_open(const char* path, ml_h5* out);
int ml_h5_close(ml_h5 f);
int ml_h5_dim(ml_h5 f);

/* header attributes (HDF5HeaderData m_int / m_real / m_string / m_box) of
 * a group: "" or "/" is the root, else e.g. "level_0".  get_* return
 * ML_H5_ENOTFOUND when the attribute is absent. */
int ml_h5_put_int(ml_h5 f, const char* group, const char* name, int v);
int ml_h5_put_real(ml_h5 f, const char* group, const char* name, double v);
int ml_h5_put_string(ml_h5 f, const char* group, const char* name, const char* v);
int ml_h5_put_box(ml_h5 f, const char* group, const char* name, const int lo[], const int hi[]);
int ml_h5_get_int(ml_h5 f, const char* group, const char* name, int* v);
int ml_h5_get_real(ml_h5 f, const char* group, const char* name, double* v);
int ml_h5_get_string(ml_h5 f, const char* group, const char* name, char* buf, int cap);
int ml_h5_get_box(ml_h5 f, const char* group, const char* name, int lo[], int hi[]);

/* the box layout of a level (write(HDF5Handle&, const BoxLayout&)):
 * "boxes" and "Processors" (all boxes on rank 0) */
int ml_h5_put_boxes(ml_h5 f, int level, int nbox, const int lo[], const int hi[]);
/* nbox on return; lo/hi receive min(nbox, cap) boxes (may be NULL with cap 0) */
int ml_h5_get_boxes(ml_h5 f, int level, int cap, int* nbox, int lo[], int hi[]);

/* LevelData<FArrayBox> (write(HDF5Handle&, const LevelData&, name,
 * outputGhost)): fabs[b] holds ncomp components over box b grown by
 * `ghost`; the cells of box b grown by out_ghost (<= ghost) are written */
int ml_h5_put_fabs(ml_h5 f, int level, const char* name, int nbox, const int lo[], const int hi[], int ncomp,
                   int ghost, int out_ghost, const double* const fabs[]);
/* components and outputGhost of a stored variable */
int ml_h5_fab_info(ml_h5 f, int level, const char* name, int* ncomp, int* out_ghost);
/* read components [c0, c0+ncomp) of a stored variable into fabs[b] (box b
 * grown by `ghost`): the valid cells, plus the stored ghost cells where the
 * file holds them (outputGhost) and the FAB has room; the boxes must be the
 * file's boxes of the level, in the file's order */
int ml_h5_get_fabs(ml_h5 f, int level, const char* name, int nbox, const int lo[], const int hi[], int c0,
                   int ncomp, int ghost, double* const fabs[]);

/* LevelData<FluxBox>: faces[b*dim + d] holds ncomp components over
 * surroundingNodes(box b, d) */
int ml_h5_put_faces(ml_h5 f, int level, const char* name, int nbox, const int lo[], const int hi[], int ncomp,
                    const double* const faces[]);
int ml_h5_get_faces(ml_h5 f, int level, const char* name, int nbox, const int lo[], const int hi[], int ncomp,
                    double* const faces[]);

/* ---- the reference's headers (single AMR level or several) ------------- */
/* a level's header values: writeCheckpointLevel (:184-241) writes
 * ref_ratio, dx, dt, time, prob_domain, tag_buffer_size, is_periodic_<d>;
 * writePlotLevel (:554-577) ref_ratio, dx, dt, time, prob_domain,
 * timestepFailed */
typedef struct {
  int lo[3], hi[3];       /* prob_domain */
  int periodic[3];
  int ref_ratio;
  int tag_buffer_size;    /* checkpoint */
  int timestep_failed;    /* plot */
  double dx, dt, time;
} ml_h5_level;
/* root of a checkpoint: AMR::writeCheckpointFile's max_level (num_levels-1),
 * num_levels, iteration, time, regrid_interval_<l> / steps_since_regrid_<l>
 * for l < max_level (NULL arrays when num_levels == 1), then
 * writeCheckpointHeader (:131-183): num_components, component_<i> */
int ml_h5_write_checkpoint_header(ml_h5 f, int num_levels, int iteration, double time, const int regrid_interval[],
                                  const int steps_since_regrid[], int ncomp, const char* const names[]);
int ml_h5_write_checkpoint_level_header(ml_h5 f, int level, const ml_h5_level* L);
/* root of a plot file: max_level, num_levels, iteration, time, then
 * writePlotHeader (:525-551): num_components, component_<i> and
 * MushyLayerParams::writeToHDF5 (src/MushyLayerParams.cpp:1214-1225) */
typedef struct {
  double compositionRatio, heatConductivityRatio, specificHeatRatio, waterDistributionCoeff, stefan, lewis, darcy;
  const char* timescale;       /* e.g. "Diffusive" */
  const char* velocity_scale;  /* e.g. "Advective" */
} ml_h5_plot_params;
int ml_h5_write_plot_header(ml_h5 f, int num_levels, int iteration, double time, int ncomp, const char* const names[],
                            const ml_h5_plot_params* p);
int ml_h5_write_plot_level_header(ml_h5 f, int level, const ml_h5_level* L);
/* readCheckpointLevel's header part (:243-330): ref_ratio, dx, dt, time and
 * prob_domain are required (ML_H5_ENOTFOUND otherwise), is_periodic_<d>,
 * tag_buffer_size and timestepFailed default to 0 */
int ml_h5_read_level_header(ml_h5 f, int level, ml_h5_level* L);

/* the file's structure as text, one line per object:
 *   "<path> group" | "<path> dataset <class> <n>" | "<path>@<attr> attr <class>"
 * with class int / float / string / compound(<member>,...); lines sorted.
 * Returns the length needed (excluding the terminator); writes at most cap
 * bytes. */
int ml_h5_describe(ml_h5 f, char* buf, int cap);

#ifdef __cplusplus
}
#endif

#endif
