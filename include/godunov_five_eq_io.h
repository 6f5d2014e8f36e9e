/*
 * HDF5 + XDMF output and restart of the conservative state (libg5h5.so, over the HDF5 1.10 C
 * library).  Kept out of libg5.so so that the step library needs no HDF5.
 *
 * File layout (the writer of the reference, HDF5_Xdmf_Writer, lives in the kalypsso core, which
 * is not part of the reference tree; what the reference fixes is kept: one dataset per
 * conservative variable named as models/FiveEq.h:93-135, the scalar attributes "time" and
 * "iteration" of "run" (SolverGodunovFiveEq.cpp:804-806), every variable in a checkpoint
 * (SolverGodunovFiveEq.cpp:812-846), and the restart reading them back by name
 * (init/init.cpp:178-196)):
 *   /rho_mix /e_tot /rho_vx /rho_vy [/rho_vz] /alpha_0 /alpha_rho_0 /alpha_1 /alpha_rho_1
 *       FP64 (IEEE little endian), cell-centred, in global cell order: shape (ny, nx) in 2-D,
 *       (nz, ny, nx) in 3-D (x fastest), so that the .xmf describes them as a CoRectMesh
 *   /run   attributes time (f64), iteration (i64)
 *   /mesh  attributes dim (i32), block_size (i32), nb (i32[3]), dx (f64), xmin (f64[3])
 * and <basename>.xmf (XDMF 2) pointing at the datasets.
 */
#ifndef GODUNOV_FIVE_EQ_IO_H
#define GODUNOV_FIVE_EQ_IO_H

#include <stddef.h>
#include <stdint.h>

#include "godunov_five_eq.h"

#ifdef __cplusplus
extern "C"
{
#endif

  /* Write <dir>/<basename>.h5 and <dir>/<basename>.xmf from U[block][var][cell] (blocks in brick
   * order, n = nblocks * nvar * bsize^dim).  var_mask: bit v selects conservative variable v;
   * 0 selects all of them (a checkpoint).  Replaces save_solution_hdf5
   * (SolverGodunovFiveEq.cpp:773-917) on a fixed single-level mesh; derived quantities and leaf
   * data are not written. */
  int g5_h5_write(const char * dir, const char * basename, const g5_config * cfg, const double * U, size_t n, double t,
                  int64_t iteration, uint32_t var_mask);

  /* Read a checkpoint (every variable present) into U[block][var][cell]; t and iteration from
   * the "run" attributes.  The file's mesh must be cfg's (dim, block size, blocks, dx).
   * Replaces the HDF5 part of init_restart (init/init.cpp:132-207). */
  int g5_h5_read(const char * path, const g5_config * cfg, double * U, size_t n, double * t, int64_t * iteration);

  /* Name of conservative variable v of a dim-D state (models/FiveEq.h:93-135), NULL if out of
   * range. */
  const char * g5_h5_var_name(int dim, int v);

  /* Message of the last failed g5_h5_* call of this thread. */
  const char * g5_h5_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* GODUNOV_FIVE_EQ_IO_H */
