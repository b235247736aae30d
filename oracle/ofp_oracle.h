/*
 * ofp_oracle.h -- CPU oracle for the vector_dist neighbour-search path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.  The
 * product path (libofp_nn.so, the C++ facade, the Python host layer) never
 * links or calls anything here.
 *
 * It is a plain-C restatement of the reference's CPU algorithms for the path,
 * instantiated for float (_f32) and double (_f64) positions.  Every function
 * cites the reference file:line it follows.  Where the algorithm lives in
 * openfpm_data (not mounted in the reference tree) the restatement follows the
 * reference's call sites, and the header of each function says which checks
 * pin it (golden vectors from the reference's own tests, closed forms, brute
 * force) or says "parity unpinned".
 *
 * Layout conventions (shared with the HIP library): positions are SoA planes,
 * coordinate k of particle i at pos[k*stride + i].  Cell grids are (div+2)^3
 * with one padding layer, linearised x-fastest.  Floating point: the library
 * is built with -ffp-contract=off so every expression rounds exactly as
 * written.
 */
#ifndef OFP_ORACLE_H
#define OFP_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OFO_DECL(T, S)                                                                       \
  void ofo_shift_vectors##S(const T lo[3], const T hi[3], T out[27 * 3]);                    \
  int ofo_local_ghost_boxes##S(const T lo[3], const T hi[3], const T g_lo[3],                \
                               const T g_hi[3], const int bc[3], T blo[26 * 3],               \
                               T bhi[26 * 3], uint32_t sv[26]);                               \
  void ofo_magnify##S(const T g_lo[3], const T g_hi[3], T factor, T out_lo[3],               \
                      T out_hi[3]);                                                          \
  void ofo_cl_param##S(const T lo[3], const T hi[3], T r_cut, const T genl_lo[3],            \
                       const T genl_hi[3], T out_lo[3], T out_hi[3], uint32_t div[3]);       \
  void ofo_cl_spacing##S(const T lo[3], const T hi[3], const uint32_t div[3], T sp[3]);      \
  size_t ofo_ghost_count##S(const T* pos, size_t stride, size_t n_dom, const T* blo,         \
                            const T* bhi, const uint32_t* sv, int nb, uint32_t* cnt);        \
  void ofo_ghost_emit##S(const T* pos, size_t stride, size_t n_dom, const T* blo,            \
                         const T* bhi, const uint32_t* sv, int nb, const T* shifts,          \
                         T* gpos, size_t gstride, uint32_t* op_p, uint32_t* op_s);           \
  void ofo_apply_bc##S(T* pos, size_t stride, size_t n, const T lo[3], const T hi[3],        \
                       const int bc[3]);                                                     \
  void ofo_cell_ids##S(const T* pos, size_t stride, size_t n, const T lo[3],                \
                       const T sp[3], const uint32_t divc[3], uint32_t* cid);                \
  void ofo_cl_fill##S(const T* pos, size_t stride, size_t n, const T lo[3], const T sp[3],   \
                      const uint32_t divc[3], uint32_t* cell_start, uint32_t* cell_ids);     \
  size_t ofo_verlet##S(const T* pos, size_t stride, size_t n_tot, size_t n_dom,              \
                       const T lo[3], const T sp[3], const uint32_t divc[3], T r_cut,        \
                       uint32_t* offsets, uint32_t* nbr, size_t cap);                        \
  size_t ofo_brute_nn##S(const T* pos, size_t stride, size_t n_tot, size_t n_dom, T r_cut,   \
                         uint32_t* offsets, uint32_t* nbr, size_t cap);                      \
  void ofo_lj##S(const T* pos, size_t stride, size_t n_tot, size_t n_dom, const T lo[3],     \
                 const T sp[3], const uint32_t divc[3], double r_cut, double sigma,          \
                 double eps, double* force, double* energy, uint32_t* nn);                   \
  void ofo_lj_walk##S(const T* pos, size_t stride, const uint32_t* targets, size_t n_targets, \
                      const T lo[3], const T sp[3], const uint32_t divc[3], const uint32_t* cs, \
                      const uint32_t* ci, double r_cut, double sigma, double eps,           \
                      double* force, double* energy, uint32_t* nn);                         \
  void ofo_sph_density##S(const T* pos, size_t stride, size_t n_tot, size_t n_dom,           \
                          const T lo[3], const T sp[3], const uint32_t divc[3], double h,    \
                          double mass, double* rho);                                         \
  void ofo_sph_force##S(const T* pos, size_t stride, size_t n_tot, size_t n_dom,             \
                        const T lo[3], const T sp[3], const uint32_t divc[3], double h,      \
                        double mass, const double* rho, const double* prs, double* force);

OFO_DECL(float, _f32)
OFO_DECL(double, _f64)

/* Tait equation of state P = B((rho/rho0)^7 - 1), B = c0^2 rho0 / 7 (textbook; the
 * reference has no SPH code, see the function header in the body). */
void ofo_tait(const double* rho, size_t n, double rho0, double c0, double* prs);

/* Remote ghosts (the send side of ghost_get on one rank), ofp_oracle_remote.c:
 * the geometric cell list of the internal ghost boxes (returns the number of
 * list entries; cell_start / cell_boxes may be NULL), the stable order by
 * processor, and the label (returns the number of outputs; out arrays may be
 * NULL to count), ordered by processor. */
size_t ofo_remote_geo(int n, const double* lo, const double* hi, double glo[3], double cs[3], uint32_t div[3],
                      uint32_t* cell_start, uint32_t* cell_boxes);
void ofo_order_by_proc(size_t n, const uint32_t* p_in, const uint64_t* id_in, uint32_t* p_out, uint64_t* id_out);
size_t ofo_remote_label_f32(const float* pos, size_t stride, size_t n, const float* blo, const float* bhi,
                            const uint32_t* proc, const uint32_t* shift, const double glo[3], const double cs[3],
                            const uint32_t div[3], const uint32_t* cell_start, const uint32_t* cell_boxes,
                            uint32_t* out_proc, uint64_t* out_id);
size_t ofo_remote_label_f64(const double* pos, size_t stride, size_t n, const double* blo, const double* bhi,
                            const uint32_t* proc, const uint32_t* shift, const double glo[3], const double cs[3],
                            const uint32_t div[3], const uint32_t* cell_start, const uint32_t* cell_boxes,
                            uint32_t* out_proc, uint64_t* out_id);
/* map: per particle the processor of the first containing sub-domain
 * (0xFFFFFFFF: none). */
void ofo_map_proc_f32(const float* pos, size_t stride, size_t n, const float* blo, const float* bhi,
                      const uint32_t* proc, const float dom_hi[3], const int bc[3], const double glo[3],
                      const double cs[3], const uint32_t div[3], const uint32_t* cell_start,
                      const uint32_t* cell_boxes, uint32_t* out_proc);
void ofo_map_proc_f64(const double* pos, size_t stride, size_t n, const double* blo, const double* bhi,
                      const uint32_t* proc, const double dom_hi[3], const int bc[3], const double glo[3],
                      const double cs[3], const uint32_t div[3], const uint32_t* cell_start,
                      const uint32_t* cell_boxes, uint32_t* out_proc);

/* Monaghan cubic spline in 3-D and its radial derivative (textbook, pinned by closed
 * forms: integral of W over R^3 is 1, W(0)=1/(pi h^3), W(2h)=0). */
double ofo_w3(double r, double h);
double ofo_dw3(double r, double h);

#ifdef __cplusplus
}
#endif
#endif
