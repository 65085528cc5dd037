/*
 * aprh_io.h -- APR files (HDF5 .apr) for libaprh: libaprh_io.so, a host library over the
 * system HDF5 and c-blosc libraries.  Same conventions as aprh.h: aprh_status returns,
 * aprh_io_last_error() for the message of the last failure of this library.
 *
 * Replaces APRFile (src/io/APRFile.hpp:19-141) for APRs in device memory: the file layout,
 * dataset and attribute names and the blosc filter parameters are the reference's
 * (APRWriter.hpp:163-215, :357-460, :595-740), so files are interchangeable.
 */
#ifndef APRH_IO_H
#define APRH_IO_H

#include <stddef.h>
#include <stdint.h>

#include "aprh.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct aprh_file aprh_file;

const char* aprh_io_last_error(void);

/* APRFile::open / close (APRFile.hpp:148-169): mode "READ", "WRITE" (truncates) or "READWRITE" */
int aprh_file_open(const char* path, const char* mode, aprh_file** out);
int aprh_file_close(aprh_file* f);

/* APRFile::write_apr (APRFile.hpp:178-263): image dims, level geometry, total particle count
   and the pipeline parameters (par may be NULL: defaults) as attributes of
   ParticleRepr/<channel><t>, the linear access (y_vec, xz_end_vec), and with_tree the tree's
   linear access under .../Tree.  channel NULL: "t". */
int aprh_file_write_apr(aprh_context* ctx, aprh_file* f, aprh_apr* apr, uint64_t t, const char* channel, int with_tree,
                        const aprh_params* par);
/* the same from a linear access in host memory (no device; no tree) */
int aprh_file_write_structure(aprh_file* f, uint64_t t, const char* channel, const int64_t* dims3 /* y, x, z */,
                              const uint64_t* xz_end_vec, uint64_t xz_end_size, const uint16_t* y_vec, uint64_t total_particles,
                              const aprh_params* par);
/* APRFile::read_apr (APRFile.hpp:299-356): a device APR from the file's linear access
   (validated like aprh_apr_from_structure) */
int aprh_file_read_apr(aprh_context* ctx, aprh_file* f, uint64_t t, const char* channel, aprh_apr** out);
/* APRFile::read_metadata (APRFile.hpp:358-388): image dims (y, x, z), particle count and
   the size of xz_end_vec */
int aprh_file_read_info(aprh_file* f, uint64_t t, const char* channel, int64_t* dims3, uint64_t* total_particles,
                        uint64_t* xz_end_size);
/* the linear access into host memory (sizes from aprh_file_read_info); either may be NULL */
int aprh_file_read_structure(aprh_file* f, uint64_t t, const char* channel, uint64_t* xz_end_vec, uint16_t* y_vec);
/* APRFile::write_particles (APRFile.hpp:211-270): n values of dtype (APRH_DTYPE_U16 / _F32) as
   dataset `name` of the time point (apr_or_tree 0: of its Tree group); device or host data
   (ctx NULL: host memory); an existing dataset of the same size and type is overwritten */
int aprh_file_write_particles(aprh_context* ctx, aprh_file* f, const char* name, const void* parts, int dtype, uint64_t n,
                              int apr_or_tree, uint64_t t, const char* channel);
/* APRFile::read_particles (APRFile.hpp:390-470): *dtype and *n first (out NULL), then the
   values into out (device or host memory; ctx NULL: host) of capacity values; values stored
   with APRCompress (compress_type != 0) are not supported */
int aprh_file_read_particles(aprh_context* ctx, aprh_file* f, const char* name, int apr_or_tree, uint64_t t, const char* channel,
                             void* out, uint64_t capacity, int* dtype, uint64_t* n);

#ifdef __cplusplus
}
#endif

#endif
