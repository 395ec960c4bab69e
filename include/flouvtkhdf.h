/*
 * flouvtkhdf.h -- C ABI of libflouvtkhdf.so, the host-side writer of Flou's VTKHDF
 * solution files (an HDF5 file whose /VTKHDF group describes a VTK unstructured grid of
 * Lagrange cells), built on the HDF5 C library.
 *
 * The device side of the output path, projecting a state held in HBM onto the equispaced
 * points of the cells, is flou_hip_equispaced() in flouhip.h; this library only lays the
 * host arrays out in the file.  It has no GPU dependency.
 *
 * File contents, as FlouBiz/FlouSpatial write them (src/FlouSpatial/IO.jl:16-76,
 * src/FlouBiz/FlouBiz.jl:38-103):
 *   /VTKHDF                          group
 *     attribute Version              int64[2] = {1, 0}
 *     attribute Type                 fixed-length ASCII string "UnstructuredGrid"
 *                                    (NUL-terminated padding, scalar dataspace)
 *     NumberOfPoints                 int64[1]
 *     Points                         real[npoints][3]   (x, y, z of each point; unused
 *                                                        coordinates are 0)
 *     NumberOfConnectivityIds        int64[1]
 *     Connectivity                   int64[nconn]
 *     NumberOfCells                  int64[1]
 *     Types                          uint8[ncells]      (VTK cell type ids)
 *     Offsets                        int64[ncells + 1]
 *     CellData/Region                int64[ncells]
 *     FieldData/<name>, CellData/<name>, PointData/<name>   added by flou_vtkhdf_write
 * "real" is float64 or float32 (the discretisation's precision).  Datasets are
 * contiguous and unfiltered; dimensions are listed slowest first (C order), so Flou's
 * column-major (3, npoints) Points matrix appears as [npoints][3].
 *
 * Return codes: 0 on success, negative FLOU_VTKHDF_E* on failure; the message of the last
 * failure on the calling thread is flou_vtkhdf_last_error().  A file handle is not
 * thread safe.
 */
#ifndef FLOUVTKHDF_H
#define FLOUVTKHDF_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct flou_vtkhdf_file flou_vtkhdf_file;

enum {
    FLOU_VTKHDF_OK = 0,
    FLOU_VTKHDF_EINVAL = -1, /* invalid argument */
    FLOU_VTKHDF_EHDF5 = -2   /* the HDF5 library reported an error */
};

/* Element types of the datasets */
enum { FLOU_VTKHDF_F64 = 0, FLOU_VTKHDF_F32 = 1, FLOU_VTKHDF_I64 = 2, FLOU_VTKHDF_U8 = 3 };

/* open_for_write(filename, disc) (src/FlouBiz/FlouBiz.jl:38-42) together with its
 * MultielementDisc method open_for_write! (src/FlouSpatial/IO.jl:16-76): create `filename`
 * (truncating an existing file), write the /VTKHDF group, its two attributes and the mesh
 * datasets listed above.
 *   points       npoints x 3 values of type points_dtype (F64 or F32), point-major
 *   connectivity nconn point ids, 0-based
 *   types        ncells VTK cell types (68 Lagrange curve, 70 Lagrange quadrilateral,
 *                72 Lagrange hexahedron)
 *   offsets      ncells + 1 offsets into connectivity, offsets[0] = 0,
 *                offsets[ncells] = nconn, non-decreasing
 *   regions      ncells region ids
 * On success *out owns the open file; release it with flou_vtkhdf_close. */
int flou_vtkhdf_open(const char* filename, int64_t npoints, const void* points, int points_dtype,
                     int64_t nconn, const int64_t* connectivity, int64_t ncells,
                     const uint8_t* types, const int64_t* offsets, const int64_t* regions,
                     flou_vtkhdf_file** out);

/* HDF5.write(file, "VTKHDF/" * path, data) as add_fielddata!, add_celldata!,
 * add_pointdata! and add_solution! call it (src/FlouBiz/FlouBiz.jl:49-103): a new
 * contiguous dataset at /VTKHDF/<path> (groups along the path are created when missing)
 * holding `data`, of element type `dtype`, rank 0..4 (0: a scalar dataspace, dims unused),
 * dimensions dims[0..rank-1] slowest first.  A path that already exists is an error, as it
 * is in HDF5. */
int flou_vtkhdf_write(flou_vtkhdf_file* f, const char* path, int dtype, int rank,
                      const int64_t* dims, const void* data);

/* close_file!(file) (src/FlouBiz/FlouBiz.jl:110-113): flush and close the file and free the
 * handle (which is freed also when the close fails). */
int flou_vtkhdf_close(flou_vtkhdf_file* f);

/* Message of the last failure of a call on this thread ("" if none). */
const char* flou_vtkhdf_last_error(void);

/* HDF5 library version of the build, "major.minor.release". */
const char* flou_vtkhdf_hdf5_version(void);

#ifdef __cplusplus
}
#endif

#endif /* FLOUVTKHDF_H */
