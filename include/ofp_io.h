/*
 * ofp_io.h -- vector_dist save / load in the reference's HDF5 checkpoint layout
 * (libofp_io.so, host code over the HDF5 C library).
 *
 * Reference: vector_dist::save / load (src/Vector/vector_dist.hpp:2459-2475),
 * whose HDF5_writer / HDF5_reader live in openfpm_io (not in the reference
 * tree).  The layout is the one of the reference's checkpoint file
 * test_data/vector_dist_24.h5, read by vector_dist_hdf5_load_test
 * (src/Vector/tests/vector_dist_HDF5_chckpnt_restart_test.cpp:119-176):
 *   dataset "metadata":    int32[n_blocks], the byte size of each block;
 *   dataset "vector_dist": int8[sum of the sizes], the blocks back to back.
 * A block (one per writing processor in the reference) is
 *   uint64 n | n positions (3 values of the position type each) |
 *   uint64 n | n property objects (obj_bytes each),
 * everything little-endian and unpadded between the parts; a property object is
 * the aggregate's members in declaration order at their natural alignment
 * (ofp_h5_field gives each scalar plane's offset inside it).  Loading
 * concatenates every block in file order.  Saving writes one block, or several
 * when one would exceed the int32 size field.
 *
 * Host memory only: the caller copies device data to the host first.  Every
 * function returns 0 or a negative OFP_E_* code (ofp_nn.h) and sets
 * ofp_io_last_error().
 */
#ifndef OFP_IO_H
#define OFP_IO_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* One scalar plane of the properties: value of particle p at
 * (char*)ptr + p * elem_bytes, stored at byte `offset` of the packed object. */
typedef struct ofp_h5_field {
  void* ptr;
  uint32_t elem_bytes; /* 1, 2, 4 or 8 */
  uint32_t offset;
} ofp_h5_field;

const char* ofp_io_last_error(void);

/* save (vector_dist.hpp:2459-2463): n particles, positions as 3 planes of
 * `dtype` (OFP_F32 / OFP_F64) `pos_stride` elements apart. */
int ofp_h5_save(const char* filename, int dtype, const void* pos, size_t pos_stride, size_t n,
                const ofp_h5_field* fields, int n_fields, uint32_t obj_bytes);

/* The particle count of a checkpoint whose blocks match (dtype, obj_bytes);
 * OFP_E_ARG if the file's layout does not. */
int ofp_h5_load_count(const char* filename, int dtype, uint32_t obj_bytes, size_t* n);

/* load (vector_dist.hpp:2471-2475): every block's particles, in file order,
 * into positions (3 planes, pos_stride apart) and the property planes; at most
 * n_cap particles (OFP_E_CAPACITY otherwise); *n = particles loaded. */
int ofp_h5_load(const char* filename, int dtype, void* pos, size_t pos_stride, size_t n_cap,
                const ofp_h5_field* fields, int n_fields, uint32_t obj_bytes, size_t* n);

#ifdef __cplusplus
}
#endif
#endif
