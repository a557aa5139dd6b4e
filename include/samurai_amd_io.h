/* Restart files of samurai's layout (io/restart.hpp) and samurai's HDF5/XDMF
 * output (io/hdf5.hpp), written and read with the HDF5 C library: a separate
 * shared library (libsamurai_amd_io.so) on top of the C ABI in samurai_amd.h,
 * so the compute library does not depend on HDF5.
 *
 * Restart file layout (the reference's, io/restart.hpp:52-187, single process):
 *   /n_process                 size_t (1)
 *   /mesh/dim                  size_t
 *   /mesh/min_level, max_level size_t (the mesh's configured levels)
 *   /mesh/origin_point         double[dim] (0)
 *   /mesh/scaling_factor       double (1)
 *   /mesh/level/{l}/dim/{d}/intervals/{partition,data}
 *        the LevelCellArray of the leaves of level l, dimension d: intervals
 *        {start, end, step, index} (int32 x3 + int64, 24 bytes);
 *        dimension 0 index = storage offset of the run (value = index + x)
 *   /mesh/level/{l}/dim/{d}/offsets/{partition,data}   d >= 1, size_t
 *   /fields/{name}/n_comp      size_t (1)
 *   /fields/{name}/data/{partition,data}  double, for_each_cell order
 * Levels without leaves and empty arrays are not written (io/restart.hpp:82-85).
 *
 * Return values and errors follow samurai_amd.h; the message of a failure is
 * available from smr_io_last_error().
 */
#ifndef SAMURAI_AMD_IO_H
#define SAMURAI_AMD_IO_H

#include "samurai_amd.h"

#ifdef __cplusplus
extern "C" {
#endif

/* samurai::dump(path, filename, mesh, fields...) (io/restart.hpp:189-200):
 * `file` is the full file name (the reference appends ".h5" to filename) */
int smr_restart_dump(const char* file, const smr_mesh* mesh, smr_field* const* fields, int nfields);

/* samurai::load(path, filename, mesh) (io/restart.hpp:423-462): a new mesh on
 * ctx from the file's leaves, with `cfg`'s stencil and graduation settings and
 * the file's min and max levels */
int smr_restart_load_mesh(const char* file, smr_ctx* ctx, const smr_mesh_config* cfg, smr_mesh** out);

/* load_field(file, mesh, field) (io/restart.hpp:341-386): the values of
 * /fields/{field name}, for_each_cell order, into a field living on `mesh`
 * (loaded from the same file) */
int smr_restart_load_field(const char* file, const smr_mesh* mesh, smr_field* field);

/* samurai::load(path, filename, mesh, fields...) (io/restart.hpp:423-462): the
 * mesh is replaced in place by the file's leaves (its configuration kept, the
 * levels taken from the file), then every field on it is resized and given
 * the values of /fields/{field name}.  Every field must be in the file; the
 * mesh is left unchanged when one is missing or the dimensions differ. */
int smr_restart_load(const char* file, smr_mesh* mesh, smr_field* const* fields, int nfields);

/* samurai::dump / load of a CellArray alone (io/restart.hpp:126-150, 313-337):
 * n leaves given as rows (level, x, y, z) of int32; load returns the leaf
 * count, writing at most `capacity` rows (call with capacity 0 to size) */
int smr_restart_dump_cells(const char* file, int dim, const int32_t* cells, int64_t n);
int64_t smr_restart_load_cells(const char* file, int* dim, int32_t* cells, int64_t capacity);

/* samurai::save(path, filename, options, mesh, fields...) (io/hdf5.hpp:1280-1329,
 * one process): {path}/{filename}.h5 with the leaves as an unstructured grid
 * (/mesh/points: double [n_points][3], /mesh/connectivity: uint64
 * [n_cells][2^dim], points numbered in the order the cells' corners are first
 * met, io/hdf5.hpp:84-145), one dataset per scalar field
 * (/mesh/fields/{name}), and {path}/{filename}.xdmf describing them.  `path`
 * is created if missing.  options: 0, or Hdf5Options's by_level
 * (SMR_SAVE_BY_LEVEL: one grid per level, /level/{l}/mesh, from min_level - 2)
 * and by_mesh_id (SMR_SAVE_BY_MESH_ID: one grid per mesh id, /mesh/{id name},
 * fields on every cell of the id, ghosts included) (io/hdf5.hpp:146-158, 428-522) */
#define SMR_SAVE_BY_LEVEL 1
#define SMR_SAVE_BY_MESH_ID 2
int smr_save(const char* path, const char* filename, const smr_mesh* mesh, smr_field* const* fields, int nfields, int options);

/* samurai::save(path, filename, options, cell_array, fields...) for a CellArray
 * (Hdf5_CellArray, io/hdf5.hpp:958-1000): n leaves as rows (level, x, y, z)
 * of int32 in for_each_cell order, and nfields host arrays of n values with
 * their names; the sub-mesh of by_mesh_id is "cell_array" */
int smr_save_cell_array(const char* path, const char* filename, int dim, const int32_t* cells, int64_t n, const double* const* values,
                        const char* const* names, int nfields, int options);

const char* smr_io_last_error(void);

#ifdef __cplusplus
}
#endif

#endif
