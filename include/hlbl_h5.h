/* hlbl_amd result file: the kernel-sum rows of a run in HDF5, in the layout of the
 * reference's ResultWriter (src/output.cpp:131-247, include/hlbl/output.hpp:45-86;
 * format table in README.md:266-290), so result files are interchangeable:
 *
 *   kernel_sum/x[N][4], y[N][4], yv[N][4]   int     (x, y in [0, L); yv minimum image)
 *   kernel_sum/chain[N], window[N]          int
 *   kernel_sum/kernel_sum[N][2]             double  (flavour 0, flavour 1)
 *   chains[C][10]                           int     (start, direction, stride, steps)
 *   attributes lattice, procs, nconf, muon_mass_MeV, lattice_spacing_fm, xunit, kernel,
 *   flavour_order, coordinate_ranges, chains_columns; with one chain also direction,
 *   origin, stride;
 *   with with_2p2 the group 2p2: w[M][4], chain[M], window[M] int, P1[M][4][4][4][Lmax]
 *   double (one row per source), pair_w, pair_y, pair_yv [R][4] int, P234[R][5][4][4][4]
 *   double (one row per (source, target)), attributes rows, P1_index, P234_index, flavour.
 *
 * Row datasets are unlimited along N and chunked; every append is flushed, so a job
 * that is killed keeps the windows it finished.
 * Separate library (libhlbl_h5.so, host code, links HDF5) so that libhlbl_amd.so does
 * not depend on HDF5. Errors: NULL / nonzero return, message in hlbl_h5_last_error. */
#ifndef HLBL_H5_H
#define HLBL_H5_H

#include "hlbl_amd.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Chain (include/hlbl/input.hpp:25-30), steps resolved */
typedef struct {
  hlbl_coord start, direction;
  int stride, steps;
} hlbl_chain_info;

/* RunInfo (include/hlbl/output.hpp:37-47); kernel NULL: "LLambda0.4" */
typedef struct {
  hlbl_coord lattice, procs;
  int nconf;
  const hlbl_chain_info* chains;
  int n_chains;
  const char* kernel;
  double muon_mass_mev, lattice_spacing_fm, xunit;
  int with_2p2;  /* write the 2p2 group */
} hlbl_run_info;

/* The 2+2 row type hlbl_source_2p2, Result2p2 of include/hlbl/output.hpp:25-35, is declared in hlbl_amd.h */

typedef struct hlbl_result_file hlbl_result_file;

/* ResultWriter(path, info): creates (truncates) the file and writes everything but rows */
hlbl_result_file* hlbl_result_open(const char* path, const hlbl_run_info* info);
/* ResultWriter::append: n rows of 16 doubles as hlbl_run_chain writes them
 * (x[4], y[4], yv[4], window, chain, ks0, ks1), then a flush */
int hlbl_result_append(hlbl_result_file* f, const double* rows, int n);
/* the same with the argument types of hlbl_window_rows_fn, for hlbl_run_chain_each_window */
int hlbl_result_append_rows(void* f, const double* rows, int n);
/* ResultWriter::append_2p2 (src/output.cpp:249-297): n sources, then a flush; needs
 * with_2p2 */
int hlbl_result_append_2p2(hlbl_result_file* f, const hlbl_source_2p2* rows, int n);
/* the same with the argument types of hlbl_window_2p2_fn, for hlbl_run_chain_2p2 */
int hlbl_result_append_2p2_rows(void* f, const hlbl_source_2p2* rows, int n);
long long hlbl_result_rows(const hlbl_result_file* f);
int hlbl_result_close(hlbl_result_file* f);
const char* hlbl_h5_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* HLBL_H5_H */
