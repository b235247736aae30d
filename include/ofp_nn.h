/*
 * ofp_nn.h -- C ABI of libofp_nn.so, the MI355X (gfx950) particle
 * neighbour-search library behind the vector_dist GPU surface.
 *
 * The reference (openfpm_pdata) has no C ABI: its GPU path is C++ templates that
 * call CUDA kernels directly.  Each entry point below replaces one step of that
 * path; the comment names the reference function (file:line, relative to the
 * openfpm_pdata source tree) whose work it does.  The C++ facade in
 * include/openfpm_amd/ (vector_dist_gpu, CellList_gpu, ...) is the user-facing
 * layer and calls only these functions.
 *
 * Conventions
 *  - dtype: OFP_F32 or OFP_F64, the position type St of vector_dist<3,St,...>.
 *  - Positions are SoA "planes": coordinate k of particle i lives at
 *    pos[k * stride + i] (stride in elements).  This is the layout of
 *    openfpm's memory_traits_inte for Point<3,St>.
 *  - A property component is an ofp_plane: element i at ptr + i * elem_bytes.
 *    elem_bytes is 4 or 8.
 *  - All device pointers are HIP device memory.  `stream` is a hipStream_t
 *    (NULL = the default stream).  Calls are stream-ordered; only the calls
 *    documented as "synchronous" wait for the device.
 *  - Every function returns OFP_OK (0) or a negative OFP_E_* code; the last
 *    error message of the calling thread is available from ofp_last_error().
 *  - Handles are not thread-safe (one host thread per device, as the
 *    reference's one-rank-one-thread model).
 */
#ifndef OFP_NN_H
#define OFP_NN_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OFP_ABI_VERSION 1

enum ofp_status {
  OFP_OK = 0,
  OFP_E_ARG = -1,      /* invalid argument (null pointer, bad dtype, size) */
  OFP_E_HIP = -2,      /* a HIP runtime call failed */
  OFP_E_CAPACITY = -3, /* an output buffer is smaller than required */
  OFP_E_LIMIT = -4,    /* a documented implementation limit was exceeded */
  OFP_E_STATE = -5     /* handle used before the step it depends on */
};

enum ofp_dtype { OFP_F32 = 0, OFP_F64 = 1 };

/* Pair arithmetic of the interaction kernels. */
enum ofp_precision {
  OFP_PREC_F64 = 0, /* plain FP64 (positions promoted to double) */
  OFP_PREC_F32C = 1 /* FP32 storage: quantised cell-relative records, FP32 P/rho^2 and forces; pairs evaluated in FP64 */
};

/* Cell-list options: the reference's CL_* flags (vector_dist.hpp:1340-1372) keep
 * their meaning; OFP_CL_QUANT asks the build for the cell-relative stream of the
 * FP32 pair kernels. */
#define OFP_CL_NON_SYMMETRIC 0x1u
#define OFP_CL_GPU_REORDER_POSITION 0x100u
#define OFP_CL_GPU_REORDER_PROPERTY 0x200u
#define OFP_CL_QUANT 0x10000u

/* Reduction operators of reduce_local (vector_dist_cuda_funcs.cuh:247-270). */
enum ofp_reduce_op { OFP_RED_SUM = 0, OFP_RED_MAX = 1, OFP_RED_MIN = 2 };

/* Value type of a plane that ofp_ghost_put / ofp_ghost_put_sorted add into. */
enum ofp_put_type { OFP_PUT_F32 = 0, OFP_PUT_F64 = 1, OFP_PUT_U32 = 2, OFP_PUT_I64 = 3 };

typedef struct ofp_plane {
  void* ptr;
  uint32_t elem_bytes;
  uint32_t pad_;
} ofp_plane;

/* Opaque handles: each owns the device buffers of one structure. */
typedef struct ofp_ghost ofp_ghost;
typedef struct ofp_ghost_remote ofp_ghost_remote;
typedef struct ofp_map ofp_map;
typedef struct ofp_cl ofp_cl;
typedef struct ofp_vl ofp_vl;

/* Non-owning device view of a built cell list: what CellList_gpu::toKernel()
 * hands to user kernels.  Sorted index j = cell_start[c] + rank of the particle
 * inside cell c; inside a cell particles are in ascending original index. */
typedef struct ofp_cl_view {
  /* C + 1 packed entries: low 32 bits = first sorted index of cell c, high 32
   * bits = position of the cell's first domain particle in dom_sort_ids. */
  const uint64_t* cell_start;
  const uint32_t* sort_to_nonsort;
  const uint32_t* nonsort_to_sort;
  const uint32_t* dom_sort_ids; /* sorted ids of domain particles, ascending */
  const int32_t* quant;         /* 4 x int32 per sorted particle (OFP_CL_QUANT) or NULL */
  double lo[3];                 /* low corner of the padded-cell origin (cell 1 starts here) */
  double spacing[3];
  uint32_t div[3];  /* cells per dim without padding */
  uint32_t divc[3]; /* div + 2 */
  uint64_t n_cells; /* divc[0]*divc[1]*divc[2] */
  uint64_t n_part;  /* particles binned (domain + ghost) */
  uint64_t ghost_marker;
  int dtype;
  int pad_;
} ofp_cl_view;

/* Non-owning device view of a Verlet list (CSR).  Row k belongs to the domain
 * particle dom_sort_ids[k] of the cell list it was built from; entries are
 * sorted indices in traversal order (self included). */
typedef struct ofp_vl_view {
  const uint64_t* offsets; /* n_rows + 1 */
  const uint32_t* nbr;
  uint64_t n_rows;
  uint64_t n_entries;
} ofp_vl_view;

/* Non-owning device view of the sparse form of a built cell list
 * (CELLLIST_GPU_SPARSE, vector_dist.hpp:49): only the non-empty cells, in
 * ascending cell id.  Entry m: cell id cell_key[m], particles (sorted index)
 * [cell_start[m], cell_start[m + 1]), and the nn_per_cell = (2B + 1)^3 cells of
 * its stencil (B = box_nn of the list, ofp_cl_set_box_nn) cell_nn[nn_per_cell m
 * + k], k = (dx+B) + W (dy+B) + W^2 (dz+B), W = 2B + 1, as entries of this view,
 * -1 where the cell is empty or off the grid (B = 1: the 27-cell box). */
typedef struct ofp_cl_sparse_view {
  const uint32_t* cell_key;
  const uint32_t* cell_start; /* n_cells + 1 */
  const int32_t* cell_nn;     /* nn_per_cell per cell */
  uint64_t n_cells;           /* non-empty cells */
  uint32_t box_nn;            /* stencil half-width B */
  uint32_t nn_per_cell;       /* (2B + 1)^3 */
} ofp_cl_sparse_view;

/* Lennard-Jones parameters (truncated at r_cut). */
typedef struct ofp_lj_params {
  double r_cut, sigma, eps;
} ofp_lj_params;

/* SPH parameters: cubic spline of smoothing length h (support 2h), particle mass,
 * Tait EOS (rho0, c0) used by ofp_sph_eos. */
typedef struct ofp_sph_params {
  double h, mass, rho0, c0;
} ofp_sph_params;

/* ---------------------------------------------------------------- status */
const char* ofp_last_error(void);
int ofp_abi_version(void);
/* Allocates nothing; reports the device the library will use (hipGetDevice). */
int ofp_device_info(int* device, int* n_cu, char* arch, size_t arch_len);

/* -------------------------------------------------------------- geometry
 * Host-only; write results into host arrays (values of type dtype). */

/* shift_vect_converter::generateShiftVectors_ld,
 * Decomposition/shift_vect_converter.hpp:36-66.  out: 27 x 3 values, index =
 * comb lin = sum_k (c_k + 1) 3^k. */
int ofp_shift_vectors(int dtype, const void* dom_lo, const void* dom_hi, void* out);

/* createShiftBox (Vector/vector_dist_comm.hpp:588-668) over the local ghost boxes
 * of ie_loc_ghost::create (Decomposition/ie_loc_ghost.hpp:102-271) for one
 * sub-domain equal to the domain.  ghost_lo/hi: the Ghost box (ghost_lo < 0).
 * bc[k] != 0 means periodic.  Writes up to 26 boxes (lo, hi: 3 values each),
 * their shift ids, and *n_boxes. */
int ofp_local_ghost_boxes(int dtype, const void* dom_lo, const void* dom_hi,
                          const void* ghost_lo, const void* ghost_hi, const int* bc,
                          void* box_lo, void* box_hi, uint32_t* shift_id, int* n_boxes);

/* getCellListGPU geometry (vector_dist.hpp:1222-1260): ghost.magnify(factor),
 * cl_param_calculate(pbox, div, r_cut, ghostEnlarge).  Outputs the cell-list box
 * and div. */
int ofp_cl_param(int dtype, const void* pbox_lo, const void* pbox_hi, double r_cut,
                 const void* ghost_lo, const void* ghost_hi, double factor, void* box_lo,
                 void* box_hi, uint32_t* div);

/* ----------------------------------------------------------------- ghost */

/* Uploads the local ghost boxes and the 27 shift vectors (createShiftBox's
 * host-to-device step, vector_dist_comm.hpp:657-663). */
int ofp_ghost_create(ofp_ghost** h, int dtype, int n_boxes, const void* box_lo,
                     const void* box_hi, const uint32_t* shift_id, const void* shifts27);
int ofp_ghost_destroy(ofp_ghost* h);

/* local_ghost_from_dec_impl first half (Vector/cuda/vector_dist_comm_util_funcs.cuh:
 * 247-280): num_shift_ghost_each_part + scan + read back of the total.
 * Synchronous (it returns the ghost count the caller must make room for). */
int ofp_ghost_label(ofp_ghost* h, const void* pos, size_t stride, size_t n_dom,
                    size_t* n_ghost, void* stream);

/* local_ghost_from_dec_impl second half (vector_dist_comm_util_funcs.cuh:281-300)
 * = shift_ghost_each_part (vector_dist_cuda_funcs.cuh:184-233): writes ghost
 * positions at [old_size, old_size + n_ghost), copies every plane, records
 * o_part_loc.  pos must have room for old_size + n_ghost particles. */
int ofp_ghost_emit(ofp_ghost* h, void* pos, size_t stride, size_t n_dom, size_t old_size,
                   const ofp_plane* planes, int n_planes, void* stream);

/* local_ghost_from_opart (vector_dist_comm.hpp:677-724) =
 * process_ghost_particles_local (vector_dist_cuda_funcs.cuh:133-155): the
 * SKIP_LABELLING refresh from the cached (p, shift) list.  with_pos = 0 is
 * NO_POSITION. */
int ofp_ghost_refresh(ofp_ghost* h, void* pos, size_t stride, size_t old_size, int with_pos,
                      const ofp_plane* planes, int n_planes, void* stream);

/* Device pointers to the cached o_part_loc list (particle id, shift id). */
int ofp_ghost_opart(const ofp_ghost* h, const uint32_t** pid, const uint32_t** sid,
                    size_t* n);

/* ghost_get<prp...>(SKIP_LABELLING | NO_POSITION) while the vector is in the
 * sorted order of `cl` (between updateCellListGPU and restoreOrder,
 * vector_dist.hpp:1334-1402 and :3042-3049): each ghost's planes take its source
 * particle's values, both located through the cell list's permutation.  `cl`
 * must have been built on the particles as laid out by the last emit
 * (ghost_marker = n_dom of the label). */
int ofp_ghost_sync_sorted(const ofp_ghost* h, const ofp_cl* cl, const ofp_plane* planes,
                          int n_planes, void* stream);

/* ghost_put<add_, prp...> for the local ghosts (Vector/vector_dist_comm.hpp:
 * 2086-2185, its o_part_loc merge at :2158-2180, merge_prp_v with add_) while
 * the vector is in the sorted order of `cl`: every ghost's value is added to its
 * source particle's, in ghost order, so that the sums of a symmetric pass reach
 * the particles that own them.  planes[q] holds n_part values of type types[q]
 * (OFP_PUT_*), sorted index; the ghosts' values are left as they were.  The
 * ghost list must be the emit's (source-major). */
int ofp_ghost_put_sorted(const ofp_ghost* h, const ofp_cl* cl, void* const* planes, const int* types,
                         int n_planes, void* stream);

/* ghost_put<add_, prp...> (vector_dist.hpp:2304-2310 -> vector_dist_comm.hpp:
 * 2086-2185, merge_prp_v at :2158-2180) in storage order: ghost k of the last
 * emit sits at n_dom + k; its value is added to its source particle's, the
 * ghosts of one source in ghost order.  planes[q] holds n_dom + n_ghost values
 * of type types[q] (OFP_PUT_*); the ghosts' values are left as they were. */
int ofp_ghost_put(const ofp_ghost* h, void* const* planes, const int* types, int n_planes, void* stream);

/* map(RUN_ON_DEVICE) on one rank: apply_bc_each_part (vector_dist_cuda_funcs.cuh:
 * 45-53) -> applyPointBC_no_dec (Decomposition/cuda/CartDecomposition_gpu.cuh:67-74). */
int ofp_apply_bc(int dtype, void* pos, size_t stride, size_t n, const void* dom_lo,
                 const void* dom_hi, const int* bc, void* stream);

/* ------------------------------------------------ remote ghosts (send side) */

/* The internal ghost boxes of this rank that other ranks need (ie_ghost's
 * vb_int_box with its (processor, shift id) pairs vb_int, as ie_ghost_gpu holds
 * them for the kernels, Decomposition/cuda/ie_ghost_gpu.cuh:117-160), and the
 * shift vectors (3 per shift id).  The geometric cell list over the boxes
 * (ie_ghost_gpu's geo_cell) is built on the host by the rule of
 * ofp_ghost_remote_geo and uploaded with them. */
int ofp_ghost_remote_create(ofp_ghost_remote** h, int dtype, int n_boxes, const void* box_lo,
                            const void* box_hi, const uint32_t* proc, const uint32_t* shift_id,
                            int n_shifts, const void* shifts);
int ofp_ghost_remote_destroy(ofp_ghost_remote* h);

/* Host only: the geometric cell list of the boxes (box corners in double).  Grid
 * over the boxes' bounding box with div[k] = clamp(floor(extent / narrowest box
 * width), 1, 64) cells; cell c lists, ascending, the boxes whose closed extent
 * meets the cell widened by 1e-9 of the extent.  cell_start (n_cells + 1) and
 * cell_boxes (*n_entries) may be NULL to query the sizes. */
int ofp_ghost_remote_geo(int n_boxes, const double* box_lo, const double* box_hi, double grid_lo[3],
                         double cell_size[3], uint32_t div[3], uint32_t* cell_start,
                         uint32_t* cell_boxes, size_t* n_entries);

/* labelParticlesGhost_impl (Vector/cuda/vector_dist_comm_util_funcs.cuh:53-177):
 * num_proc_ghost_each_part, scan, proc_label_id_ghost
 * (Vector/cuda/vector_dist_cuda_funcs.cuh:19-43, by the walk of
 * ghost_processorID_general_impl, Decomposition/cuda/ie_ghost_gpu.cuh:40-96),
 * then the outputs ordered by destination processor (stable: processor
 * ascending, then particle-major as labelled) and counted per processor (the
 * sort + find_buffer_offsets + prc/prc_sz of :110-170).  Synchronous: returns
 * the number of outputs and of destination processors with at least one. */
int ofp_ghost_remote_label(ofp_ghost_remote* h, const void* pos, size_t stride, size_t n,
                           size_t* n_out, int* n_procs, void* stream);

/* The last label's destination processors (ascending) and their output counts
 * (prc, prc_sz); arrays of *n_procs entries. */
int ofp_ghost_remote_procs(const ofp_ghost_remote* h, uint32_t* procs, uint64_t* counts);

/* Device view of the last label's ordered outputs (g_opart_device): destination
 * processor and particle | shift id << 32 per output. */
int ofp_ghost_remote_opart(const ofp_ghost_remote* h, const uint32_t** proc, const uint64_t** id,
                           size_t* n);

/* process_ghost_particles_pos / _prp (Vector/cuda/vector_dist_cuda_funcs.cuh:90-131):
 * send entry i (in the ordered outputs) = position of its particle minus its
 * shift vector (3 planes of send_stride >= n_out), and each plane's value of the
 * particle. */
int ofp_ghost_remote_pack(const ofp_ghost_remote* h, const void* pos, size_t stride, void* send_pos,
                          size_t send_stride, const ofp_plane* planes_in, const ofp_plane* planes_out,
                          int n_planes, void* stream);

/* ------------------------------------------------------------------- map */

/* The sub-domains of every rank (CartDecomposition's sub_domains_global:
 * box + processor, as CartDecomposition_gpu holds them,
 * Decomposition/cuda/CartDecomposition_gpu.cuh:76-112), the domain and its
 * periodicity, and this rank.  The geometric cell list over the sub-domains
 * (the role of its fine_s cell list) is built on the host by the rule of
 * ofp_ghost_remote_geo.  At most 255 distinct processors. */
int ofp_map_create(ofp_map** h, int dtype, int n_sub, const void* sub_lo, const void* sub_hi,
                   const uint32_t* sub_proc, const void* dom_lo, const void* dom_hi, const int* bc,
                   uint32_t rank);
int ofp_map_destroy(ofp_map* h);

/* labelParticleProcessor(RUN_ON_DEVICE) (Vector/vector_dist_comm.hpp:1398-1456):
 * the periodic wrap in place (apply_bc_each_part), then per particle the
 * processor of the first sub-domain of its geometric cell that contains it
 * (process_id_proc_each_part, Vector/cuda/vector_dist_cuda_funcs.cuh:55-75 ->
 * processorID_impl, CartDecomposition_gpu.cuh:13-56, with
 * isInsideNP_with_border: half-open, the upper face closed where a sub-domain
 * ends on the domain's upper face in a non-periodic dimension), and the
 * particles ordered by destination (the scan + reorder_lbl of :1440-1448;
 * stable here: processor ascending, own rank included, then particle index).
 * A particle inside no sub-domain is dropped and counted in *n_lost.
 * Synchronous: returns the staying count, the number of other processors with
 * at least one particle, and the lost count. */
int ofp_map_label(ofp_map* h, void* pos, size_t stride, size_t n, size_t* n_stay, int* n_dest,
                  size_t* n_lost, void* stream);

/* The last label's destination processors other than this rank (ascending) and
 * their particle counts (prc_sz); arrays of *n_dest entries. */
int ofp_map_procs(const ofp_map* h, uint32_t* procs, uint64_t* counts);

/* process_map_particles (Vector/cuda/vector_dist_cuda_funcs.cuh:76-88) and the
 * new local vector of fill_send_map_buf (Vector/vector_dist_comm.hpp:1240-1310):
 * the staying particles, in index order, to stay_pos (stride >= n_stay) and
 * the stay planes; the leaving ones, destination-major, to send_pos (stride >=
 * n - n_stay - n_lost) and the send planes.  planes_in gives the source of
 * both. */
int ofp_map_pack(const ofp_map* h, const void* pos, size_t stride, void* stay_pos, size_t stay_stride,
                 void* send_pos, size_t send_stride, const ofp_plane* planes_in,
                 const ofp_plane* stay_planes, const ofp_plane* send_planes, int n_planes,
                 void* stream);

/* ------------------------------------------------------------- cell list */

/* CellList_gpu(pbox, div) (vector_dist.hpp:1248): box and div as produced by
 * ofp_cl_param, one padding layer. */
int ofp_cl_create(ofp_cl** h, int dtype, const void* box_lo, const void* box_hi,
                  const uint32_t* div, uint32_t opt);
int ofp_cl_destroy(ofp_cl* h);

/* CellList_gpu::construct as called by updateCellListGPU (vector_dist.hpp:
 * 1334-1402): cell ids, counting sort with canonical (ascending index) order
 * inside each cell, permutations, domain sort ids; if pos_sorted != NULL the
 * reordered positions (CL_GPU_REORDER_POSITION); planes_in -> planes_out the
 * reordered properties (CL_GPU_REORDER_PROPERTY).  n_tot = domain + ghost. */
int ofp_cl_construct(ofp_cl* h, const void* pos, size_t stride, size_t n_tot,
                     size_t ghost_marker, void* pos_sorted, size_t sorted_stride,
                     const ofp_plane* planes_in, const ofp_plane* planes_out, int n_planes,
                     void* stream);

int ofp_cl_get_view(const ofp_cl* h, ofp_cl_view* out);

/* CellList_gpu<..., is_sparse = true> (CELLLIST_GPU_SPARSE, vector_dist.hpp:49;
 * tests vector_dist_gpu_unit_tests.cu:570-597, 1370-1378): the sparse form of
 * the last construct, built from it (non-empty cells compacted by a scan, the
 * neighbour table through a dense inverse map).  Each neighbour set it gives
 * is the dense list's; valid until the next construct.  Synchronous (the
 * non-empty count is read back).  Not for ensemble cell lists. */
int ofp_cl_build_sparse(ofp_cl* h, void* stream);
int ofp_cl_get_sparse_view(const ofp_cl* h, ofp_cl_sparse_view* out);

/* CellList_gpu::setBoxNN (vector_dist.hpp:1252): the neighbour stencil spans
 * box_nn cells on each side ((2 box_nn + 1)^3 cells; 1 = the 27-cell box).
 * Cells of size r/box_nn with this stencil cover a cut-off r.  1 <= box_nn <= 3. */
int ofp_cl_set_box_nn(ofp_cl* h, uint32_t box_nn);

/* restoreOrder (vector_dist.hpp:3042-3049): out[s2n[j]] = in[j] for every plane. */
int ofp_cl_restore(const ofp_cl* h, const ofp_plane* sorted_in, const ofp_plane* orig_out,
                   int n_planes, void* stream);

/* Reorders planes into sorted order without rebuilding: out[j] = in[s2n[j]]
 * (the CL_GPU_REORDER_PROPERTY step on its own). */
int ofp_cl_reorder(const ofp_cl* h, const ofp_plane* orig_in, const ofp_plane* sorted_out,
                   int n_planes, void* stream);

/* ---------------------------------------------------------------- Verlet */

int ofp_vl_create(ofp_vl** h);
int ofp_vl_destroy(ofp_vl* h);

/* getVerlet (vector_dist.hpp:1514-1545) on the GPU: for each domain particle,
 * every particle (self included) with distance2 < r_cut^2, walking the 27 cells
 * of `cl`, whose cell size must be >= r_cut.  pos_sorted are the positions in the
 * cell list's sorted order.  Two passes (count, fill) with one scan; the entry
 * total is read back (synchronous). */
int ofp_vl_build(ofp_vl* h, const ofp_cl* cl, const void* pos_sorted, size_t stride,
                 double r_cut, void* stream);
int ofp_vl_get_view(const ofp_vl* h, ofp_vl_view* out);

/* getVerletSym (vector_dist.hpp:1412-1416; VL_SYMMETRIC): from a full list `full`
 * (ofp_vl_build) on a cell list holding the ghosts of `g`, the half list in which
 * every interaction appears once.  Row r (domain particle of sorted index i)
 * keeps a domain neighbour j iff j > i, and a ghost neighbour (image of source
 * particle q, shift id s) iff sorted(q) > i, or q is the row's own particle and
 * s > 13 (the shift ids of s and -s are id and 26 - id).  A pass over it adds
 * each pair to both ends (ghost ends included); ofp_ghost_put_sorted then brings
 * the ghosts' sums to their sources, which gives the full list's sums (the
 * property vector_dist_symmetric_cell_list checks, Vector/tests/
 * vector_dist_cell_list_tests.cpp:443-649).  The reference's half is the
 * geometric one of CellListSym (openfpm_data); this rule is exact in integers.
 * Synchronous (the entry total is read back). */
int ofp_vl_build_sym(ofp_vl* h, const ofp_vl* full, const ofp_ghost* g, void* stream);

/* ---------------------------------------------------- pair interactions
 * The reference has no LJ/SPH code; these are the pair loops user kernels write
 * against getNNIteratorBox (template: Vector/cuda/vector_dist_gpu_unit_tests.cu:
 * 83-119, calc_forces Vector/performance/vector_dist_performance_common.hpp:19-84):
 * one result per domain particle, skip q == p, strict d^2 < r_c^2.
 * Outputs are indexed by sorted index (as the reordered vector); out planes are
 * 3 arrays of double (OFP_PREC_F64) or float (OFP_PREC_F32C) of n_part entries.
 * vl == NULL walks the 27 cells of cl (cell size >= cut-off); otherwise the
 * Verlet rows (the cut-off test is still applied, so a list built with a skin
 * works). */
int ofp_pair_lj(const ofp_cl* cl, const ofp_vl* vl, const void* pos_sorted, size_t stride,
                int precision, const ofp_lj_params* prm, void* force_x, void* force_y,
                void* force_z, void* energy, uint32_t* n_neigh, void* stream);

/* LJ forces over a symmetric list (ofp_vl_build_sym), Newton's third law: each
 * pair within r_cut adds its force to both particles and counts one neighbour
 * for each.  Outputs hold every particle of cl, ghosts included (sorted index),
 * FP64 forces; they are cleared first.  n_neigh may be NULL (no counts).
 * Follow with ofp_ghost_put_sorted of the planes to get each domain particle's
 * full sums. */
int ofp_pair_lj_sym(const ofp_cl* cl, const ofp_vl* vl, const void* pos_sorted, size_t stride,
                    const ofp_lj_params* prm, double* force_x, double* force_y, double* force_z,
                    uint32_t* n_neigh, void* stream);

/* rho_i = m W(0) + sum_j m W(d_ij), written as double for both precisions
 * (OFP_PREC_F32C: the pairs evaluated and summed in FP64 from the FP32 records). */
int ofp_sph_density(const ofp_cl* cl, const ofp_vl* vl, const void* pos_sorted,
                    size_t stride, int precision, const ofp_sph_params* prm, void* rho,
                    void* stream);

/* Tait pressure for n particles from double rho: prs = B((rho/rho0)^7 - 1)
 * (double, may be NULL) and the force term pterm = prs / rho^2 (double for
 * OFP_PREC_F64, float for OFP_PREC_F32C), evaluated in FP64. */
int ofp_sph_eos(int precision, const void* rho, size_t n, const ofp_sph_params* prm,
                void* prs, void* pterm, void* stream);

/* a_i = - sum_j m (P_i/rho_i^2 + P_j/rho_j^2) W'(d) (x_i - x_j)/d with
 * pterm = P/rho^2 in sorted order over domain + ghost particles. */
int ofp_sph_force(const ofp_cl* cl, const ofp_vl* vl, const void* pos_sorted, size_t stride,
                  int precision, const ofp_sph_params* prm, const void* pterm,
                  void* force_x, void* force_y, void* force_z, void* stream);

/* ------------------------------------------------------------- reductions */

/* reduce_local<prp, _add_/_max_> (Vector/cuda/vector_dist_cuda_funcs.cuh:
 * 247-270), plus _min_, over n values of type `dtype`, accumulated in FP64.
 * With `comp` != NULL the value of element i is x[i] + comp[i] (a compensated
 * pair).  Synchronous: the scalar is read back into *result. */
int ofp_reduce(int dtype, const void* x, const void* comp, size_t n, int op, double* result,
               void* stream);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, step loops that forbid hipMalloc):
 * ofp_reduce with the same kernels, grid and result, but allocating nothing
 * and reading nothing back.  The result is written to *result_dev (device
 * memory), stream-ordered; the partials live in a fixed per-GPU array of 1025
 * doubles inside the library (the most the partial kernel's grid writes),
 * looked up on the first call on each device.  Every call on one GPU shares
 * that array, so calls that overlap on one GPU (two streams, two host threads)
 * race; calls ordered on one stream do not.  n above
 * OFP_REDUCE_NOALLOC_MAX_N is rejected with OFP_E_ARG. */
#define OFP_REDUCE_NOALLOC_MAX_N (1ull << 40)
int ofp_reduce_noalloc(int dtype, const void* x, const void* comp, size_t n, int op, double* result_dev,
                       void* stream);

/* -------------------------------------------------------------- ensembles
 * Many independent fully periodic systems of uneven sizes driven by one set of
 * launches.  No counterpart in the reference, where each system is its own
 * vector_dist; per system the results are those of the one-system path above
 * (same ghosts, cell order, Verlet rows and pair arithmetic).
 *
 * Layout: the domain particles of system s occupy [first[s], first[s+1]) of one
 * SoA position array (x plane, then y, then z, `cap` apart); ghosts of all
 * systems follow at [N, N + G), ordered by source particle.  G, the Verlet
 * entry count and the rebuild decision stay in device memory, so a rebuild
 * and a step can be captured as hipGraphs.  Storage order is internal: after
 * a rebuild the particles are stored in cell order (ofp_ens_get returns them in
 * the order given to ofp_ens_set).
 *
 * LJ (truncated at r_cut, unshifted), velocity Verlet, Verlet lists at
 * r_cut (1 + skin) rebuilt when a drift has carried a particle beyond
 * skin r_cut / 2 of its position at the last build (checked on the drifted
 * positions, before their forces are computed).  Positions, velocities and forces are of the ensemble dtype;
 * the pair arithmetic is OFP_PREC_F64 or OFP_PREC_F32C. */
typedef struct ofp_ens ofp_ens;

typedef struct ofp_sys_desc {
  double box_lo[3], box_hi[3];
  double r_cut, sigma, eps, mass;
  uint64_t n;
} ofp_sys_desc;

/* Launch/geometry choices (the autotuner's knobs).  Zero fields take defaults:
 * skin 0.1, sub 2, lanes 16, waves_per_cu 32, headroom 1.25, pair_walk 0,
 * split_kick 0. */
typedef struct ofp_ens_opts {
  double skin;           /* Verlet radius r_cut (1 + skin) */
  uint32_t sub;          /* Verlet cell list: cells of r_verlet / sub, stencil sub (1..3) */
  uint32_t lanes;        /* lanes per Verlet row in the pair pass: 1, 2, 4, 8, 16 or 32 */
  uint32_t waves_per_cu; /* persistent grid of the Verlet build, waves per CU */
  uint32_t verlet_rows;  /* Verlet build kernels: 0 chosen by the mean candidates per target,
                            1 targets in the lanes (dense cells), 2 one thread per row (dilute) */
  double headroom;       /* ghost and Verlet capacities over the counts of a sizing rebuild */
  uint32_t pair_walk;    /* pair pass: 0 Verlet rows; 1 cell tiles over the cells of the last
                            rebuild (no Verlet list: their stencil covers r_cut (1 + skin), so
                            every pair within r_cut is found until the rebuild trigger) */
  uint32_t split_kick;   /* steps of a run over Verlet rows: 0 the pair kernel kicks and drifts
                            each row's particle as its force is summed; 1 the forces are stored
                            and a separate kernel kicks (same arithmetic, same results) */
} ofp_ens_opts;

typedef struct ofp_ens_state {
  uint64_t n_dom, n_ghost, n_entries;   /* of the last rebuild */
  uint64_t ghost_cap, entry_cap;
  uint64_t rebuilds, steps, recaptures; /* since creation */
  uint32_t rebuild_pending, overflow;
} ofp_ens_state;

/* Device arrays of an ensemble (storage order, `cap` = plane stride). */
typedef struct ofp_ens_arrays {
  void* pos;           /* 3 planes, domain then ghosts */
  void* vel;           /* 3 planes, domain */
  void* force;         /* 3 planes, domain, of the last force pass */
  const uint32_t* gid; /* index given to ofp_ens_set of each stored domain particle */
  const uint32_t* sid; /* system of each stored particle */
  /* The last pair pass in sorted order (slots of the cell list): forces as
   * computed (double for OFP_PREC_F64, float for OFP_PREC_F32C; 3 planes of
   * cap) and the number of neighbours within r_cut of each domain slot. */
  const void* force_sorted;
  const uint32_t* n_neigh;
  uint64_t cap;
  uint64_t n_dom;
} ofp_ens_arrays;

int ofp_ens_create(ofp_ens** out, int dtype, uint32_t n_sys, const ofp_sys_desc* sys,
                   const ofp_ens_opts* opts);
int ofp_ens_destroy(ofp_ens* e);
/* first[0..n_sys]: first domain particle of each system (host array). */
int ofp_ens_first(const ofp_ens* e, uint64_t* first);
int ofp_ens_arrays_get(const ofp_ens* e, ofp_ens_arrays* a);
/* The cell list (cells of all systems, global ids) and Verlet list of the last
 * rebuild: non-owning handles for the view functions above.  The Verlet view's
 * n_entries is 0 (the count is on the device at offsets[n_rows]).  With
 * pair_walk = 1 there is no Verlet list and *vl is NULL. */
int ofp_ens_lists(const ofp_ens* e, const ofp_cl** cl, const ofp_vl** vl);
/* cells per system: base[0..n_sys] global cell ids, divc: 3 per system. */
int ofp_ens_cells(const ofp_ens* e, uint32_t* base, uint32_t* divc, double* lo, double* spacing);
/* Domain positions (and velocities, may be NULL = 0) in user order, device SoA
 * with plane stride `stride`; the next step starts with a sizing rebuild. */
int ofp_ens_set(ofp_ens* e, const void* pos, const void* vel, size_t stride, void* stream);
/* Positions, velocities, forces (each may be NULL) back in user order. */
int ofp_ens_get(const ofp_ens* e, void* pos, void* vel, void* force, size_t stride, void* stream);
/* Rebuild now: wrap into the boxes, ghosts, cell list, Verlet list.  sizing != 0
 * reads the counts back and grows the capacities; 0 keeps them and reads back
 * the overflow flag only, redoing the rebuild as a sizing one if a count
 * outgrew its capacity.  Both are synchronous. */
int ofp_ens_rebuild(ofp_ens* e, int sizing, void* stream);
/* LJ forces on the current positions (ghost refresh, reorder, pair pass, back to
 * storage order); energy (per stored domain particle, may be NULL) of the ensemble dtype. */
int ofp_ens_forces(ofp_ens* e, int precision, void* energy, void* stream);
/* The LJ kernel alone, over the sorted positions of the last force pass (for
 * timing the pair loop; forces in force_sorted).  with_energy != 0 also writes the
 * per-particle energy and n_neigh (the force pass does); 0 is the forces-only
 * kernel that a step runs. */
int ofp_ens_pair_pass(ofp_ens* e, int precision, int with_energy, void* stream);
/* n_steps velocity-Verlet steps of size dt.  The run starts with the first
 * half kick and drift; each step is then one hipGraph replay (forces and the
 * second half kick, fused except after the last step with the next step's
 * first half kick and drift), preceded by a rebuild replay when the drift
 * flagged one; the host reads the flag between steps.  On return positions,
 * velocities and forces are those of the end of the last step.  use_graph = 0
 * launches the same kernels directly. */
int ofp_ens_run(ofp_ens* e, int precision, double dt, int n_steps, int use_graph, void* stream);
int ofp_ens_status(ofp_ens* e, ofp_ens_state* st, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* OFP_NN_H */
