/* HDF5 project files (part of libdxmc_hip.so): the operations OpenDXMC's
 * HDF5Wrapper performs on its project files (hdf5wrapper.cpp) -- nested
 * groups, N-d numeric datasets stored whole or as one deflate-compressed
 * chunk, variable-length string lists, and double / uint64 attributes on
 * groups -- on the HDF5 library itself, loaded at first use
 * ($DXG_HDF5_LIBRARY, else /opt/conda/lib/libhdf5.so.103, else libhdf5.so).
 * The one change in how the bytes are produced: a compressed dataset's chunk
 * is deflated by several threads into one zlib stream and handed to HDF5 as
 * already-filtered data, so the file is the one the HDF5 deflate filter would
 * read and write, made faster.
 *
 * Paths are absolute or relative to the root ("beams/CTSpiralBeams/1").
 * Status codes and dxg_last_error() as in dxmc_hip.h.
 */
#ifndef DXMC_H5_H
#define DXMC_H5_H

#include "dxmc_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct dxg_h5 dxg_h5;

/* Element types (memory side; HDF5 converts from the file's numeric type on
 * reads as H5Dread does). */
enum {
    DXG_H5_F64 = 0,   /* NATIVE_DOUBLE (hdf5wrapper.cpp:134) */
    DXG_H5_U64 = 1,   /* NATIVE_UINT64 (:136) */
    DXG_H5_U8 = 2,    /* NATIVE_UINT8 (:138) */
    DXG_H5_F32 = 3,
    DXG_H5_I32 = 4,
    DXG_H5_I64 = 5
};

/* H5File(path, H5F_ACC_TRUNC) / H5File(path, H5F_ACC_RDONLY)
 * (hdf5wrapper.cpp:363-374); close flushes and frees. */
int dxg_h5_create(const char* path, dxg_h5** out);
int dxg_h5_open(const char* path, dxg_h5** out);
int dxg_h5_close(dxg_h5* f);

/* Creates the group and its missing parents (getGroup(..., create = true),
 * hdf5wrapper.cpp:73-91). */
int dxg_h5_create_group(dxg_h5* f, const char* path);
/* 0 when nothing is at path, 1 a group, 2 a dataset (nameExists, :259). */
int dxg_h5_kind(dxg_h5* f, const char* path);

/* saveArray (hdf5wrapper.cpp:99-177): rank dims in file order (slowest
 * first; a volume of nx x ny x nz voxels is { nz, ny, nx }).
 * deflate_level < 0 stores the data whole (contiguous); 0..9 stores it as one
 * chunk of the whole extent through the deflate filter at that level.
 * Missing parent groups are created. */
int dxg_h5_write_array(dxg_h5* f, const char* path, int type, int rank, const uint64_t* dims, const void* data,
                       int deflate_level);
/* saveArray of strings (hdf5wrapper.cpp:202-246): 1-d, variable-length C
 * strings. */
int dxg_h5_write_strings(dxg_h5* f, const char* path, const char* const* strings, uint64_t n);
/* saveAttribute (hdf5wrapper.cpp:299-326) on the group at path: type
 * DXG_H5_F64 or DXG_H5_U64; scalar != 0 writes a scalar dataspace (n must be
 * 1), else a 1-d one of n values. */
int dxg_h5_write_attribute(dxg_h5* f, const char* path, const char* name, int type, const void* data, uint64_t n,
                           int scalar);

/* Dataset extent (up to 32 dimensions) and the file's element class:
 * 0 integer, 1 float, 3 string, 9 variable-length (H5T_class_t). */
int dxg_h5_dataset_shape(dxg_h5* f, const char* path, int* rank, uint64_t* dims, int* type_class, uint64_t* type_size);
/* loadArray (hdf5wrapper.cpp:254-290): the whole dataset converted to type;
 * n must equal its element count. */
int dxg_h5_read_array(dxg_h5* f, const char* path, int type, void* out, uint64_t n);
/* loadArray<std::string>: the strings NUL-separated into buf (cap bytes);
 * *count strings, *bytes used (call with buf = NULL for the sizes). */
int dxg_h5_read_strings(dxg_h5* f, const char* path, char* buf, uint64_t cap, uint64_t* count, uint64_t* bytes);
/* loadAttribute (hdf5wrapper.cpp:327-360): *n values, *rank 0 for a scalar;
 * *n = 0 (status 0) when the attribute is absent, has rank > 1, or is not an
 * 8-byte float (F64) / 8-byte integer (U64) -- the reference's std::nullopt.
 * With out = NULL only *n and *rank are set. */
int dxg_h5_read_attribute(dxg_h5* f, const char* path, const char* name, int type, void* out, uint64_t cap, uint64_t* n,
                          int* rank);
/* Names of the members of a group, NUL-separated, in HDF5's name order. */
int dxg_h5_list(dxg_h5* f, const char* path, char* buf, uint64_t cap, uint64_t* count, uint64_t* bytes);

/* Milliseconds spent compressing in the last dxg_h5_write_array on this
 * thread, and the number of threads it used (1: the HDF5 library's own
 * filter, used when it lacks H5Dwrite_chunk or DXG_H5_LIBRARY_DEFLATE=1,
 * and then the time covers the library's whole write). */
double dxg_h5_last_deflate_ms(int* threads);

#ifdef __cplusplus
}
#endif

#endif /* DXMC_H5_H */
