/*
 * usct_h5.h -- C ABI of the acquisition-dataset file reader and writer
 * (libusct_h5.so, host code).
 *
 * The acquisition dataset is an HDF5 file laid out as
 * software/schema/README.md:13-41 specifies and simulation/ringfwi/dataset.py:112-174
 * writes and reads through h5py: a root group with string attributes, groups
 * /geometry, /acquisition and optional /ground_truth holding numeric and string
 * attributes and N-dimensional numeric datasets, the channel data float32
 * gzip-compressed in chunks.  This library reads and writes that subset of the
 * HDF5 file format (version-0 superblock, version-1 object headers,
 * symbol-table groups, contiguous / compact / chunked layouts with a version-1
 * B-tree chunk index, the deflate, shuffle and Fletcher-32 filters, fixed and
 * variable-length strings) without libhdf5.  Chunks are inflated in parallel
 * (OpenMP) straight into the caller's buffer, which may be pinned host memory
 * that a device copy reads next.
 *
 * Reference interfaces these replace (paths relative to the reference root):
 *   usct_h5_open / usct_h5_read / usct_h5_attr_*   h5py.File(path, "r") and the
 *       reads of Dataset.load, simulation/ringfwi/dataset.py:142-174
 *   usct_h5_create / usct_h5_add_* / usct_h5_finish  h5py.File(path, "w") and the
 *       writes of Dataset.save, simulation/ringfwi/dataset.py:112-139
 *
 * Every entry point returns 0 or a usct_h5_status code with a message in
 * usct_h5_last_error() (thread-local).  A file handle may be read from several
 * threads at once; a writer belongs to one thread.
 */
#ifndef USCT_H5_H
#define USCT_H5_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    USCT_H5_OK = 0,
    USCT_H5_EINVAL = 1,   /* bad argument */
    USCT_H5_ENOMEM = 2,
    USCT_H5_EIO = 3,      /* the operating system refused an open / read / write */
    USCT_H5_ENOTSUP = 4,  /* valid HDF5 this library does not read (e.g. dense link storage) */
    USCT_H5_EFORMAT = 5,  /* not an HDF5 file, or a damaged / truncated one */
    USCT_H5_ENOENT = 6    /* no object or attribute of that name */
} usct_h5_status;

/* Element types (file types reported by the info calls, and output types). */
enum {
    USCT_H5_F32 = 1, USCT_H5_F64 = 2, USCT_H5_I32 = 3, USCT_H5_I64 = 4, USCT_H5_U8 = 5,
    USCT_H5_STR = 6,  /* fixed or variable-length string */
    USCT_H5_I8 = 7, USCT_H5_I16 = 8, USCT_H5_U16 = 9, USCT_H5_U32 = 10, USCT_H5_U64 = 11,
    USCT_H5_F16 = 12,
    USCT_H5_OTHER = 99  /* a type this library cannot convert (compound, enum, ...) */
};

/* Object kinds. */
enum { USCT_H5_ABSENT = 0, USCT_H5_GROUP = 1, USCT_H5_DATASET = 2 };

#define USCT_H5_MAX_RANK 32

typedef struct usct_h5_file usct_h5_file;
typedef struct usct_h5_writer usct_h5_writer;

/* ---- reading ---------------------------------------------------------- */

int usct_h5_open(usct_h5_file **file, const char *path);
int usct_h5_close(usct_h5_file *file);

/* Kind of the object at an absolute path ("/", "/acquisition/data", ...). */
int usct_h5_kind(usct_h5_file *file, const char *path, int *kind);

/* Names of a group's members, each followed by '\0', in the file's (sorted)
 * order.  *needed receives the bytes required; buf may be NULL. */
int usct_h5_list(usct_h5_file *file, const char *path, char *buf, int64_t cap,
                 int64_t *needed);

/* Shape and element type of a dataset.  dims has room for USCT_H5_MAX_RANK
 * entries; a scalar has ndim 0.  chunked: 1 when stored in chunks. */
int usct_h5_dataset_info(usct_h5_file *file, const char *path, int *ndim, int64_t *dims,
                         int *type, int *chunked);

/* Read rows [row0, row0 + nrows) along the first dimension (all of every other
 * dimension) into out, converted to out_type (any numeric type; integers and
 * floats convert as C casts do).  nrows < 0 reads to the end; a scalar
 * dataset takes row0 = 0, nrows = -1.  out_elems must equal the number of
 * elements read.  nthreads 0 = one per core of the process. */
int usct_h5_read(usct_h5_file *file, const char *path, int64_t row0, int64_t nrows,
                 int out_type, void *out, int64_t out_elems, int nthreads);

/* Attribute names of an object, as usct_h5_list. */
int usct_h5_attr_list(usct_h5_file *file, const char *path, char *buf, int64_t cap,
                      int64_t *needed);

/* Type, element count and rank (0 = scalar) of an attribute (strings: type
 * USCT_H5_STR and n = bytes of the first string, without a terminator). */
int usct_h5_attr_info(usct_h5_file *file, const char *path, const char *name, int *type,
                      int64_t *n, int *rank);

/* Read a numeric attribute (n elements of out_type) or a string attribute
 * (out_type USCT_H5_STR: cap bytes of buf, the string's bytes then '\0'). */
int usct_h5_attr_read(usct_h5_file *file, const char *path, const char *name, int out_type,
                      void *out, int64_t cap);

/* ---- writing ---------------------------------------------------------- */

/* Create (truncate) a file.  Data is written as datasets are added; the
 * metadata and the superblock are written by usct_h5_finish. */
int usct_h5_create(usct_h5_writer **writer, const char *path);

/* Add a group; its parent must exist ("/" always does). */
int usct_h5_add_group(usct_h5_writer *writer, const char *path);

/* Add an N-dimensional dataset of type (F32, F64, I32, I64, U8) from a
 * C-contiguous buffer.  chunk: NULL = contiguous storage, else ndim chunk
 * extents; gzip_level 0..9 compresses the chunks (deflate, as h5py's
 * compression="gzip"), -1 stores them unfiltered. */
int usct_h5_add_dataset(usct_h5_writer *writer, const char *path, int type, int ndim,
                        const int64_t *dims, const void *data, const int64_t *chunk,
                        int gzip_level, int nthreads);

/* Add an attribute to a group or dataset: n elements of a numeric type
 * (n < 0: a scalar), or a UTF-8 string (type USCT_H5_STR, value a
 * '\0'-terminated string; stored variable-length, as h5py stores str). */
int usct_h5_add_attr(usct_h5_writer *writer, const char *path, const char *name, int type,
                     int64_t n, const void *value);

/* Write the metadata and close.  commit 0 discards the file instead.  The
 * writer is freed either way. */
int usct_h5_finish(usct_h5_writer *writer, int commit);

const char *usct_h5_last_error(void);

#ifdef __cplusplus
}
#endif
#endif /* USCT_H5_H */
