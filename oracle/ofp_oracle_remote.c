/*
 * ofp_oracle_remote.c -- CPU oracle for the remote part of ghost_get (the send
 * side of one rank): label, order by destination processor; and map's
 * processor of each particle (ofo_map_proc, below).
 *
 * TEST INFRASTRUCTURE ONLY (the checker, never the product).
 *
 * ofo_remote_geo: the geometric cell list over the internal ghost boxes, by the
 *   rule the product documents for ofp_ghost_remote_geo (include/ofp_nn.h): a
 *   grid over the boxes' bounding box with div = clamp(floor(extent / narrowest
 *   positive box width), 1, 64) per dimension, and per cell, ascending, the
 *   boxes whose closed extent meets the cell widened by 1e-9 of the extent.
 *   The reference's geo_cell (ie_ghost's CellList over the internal ghost boxes)
 *   lives in openfpm_data's CellList, absent from the reference tree, so its
 *   grid is a choice of this build: parity unpinned against the reference; the
 *   tests pin it by the property that matters (a box that contains a point is
 *   in the list of the point's cell, checked against brute force over all
 *   boxes) and by equality with the product's host builder.
 * ofo_remote_label_*: per particle the walk of ghost_processorID_general_impl
 *   (Decomposition/cuda/ie_ghost_gpu.cuh:40-96) over its cell's boxes, with the
 *   outputs of num_proc_ghost_each_part / proc_label_id_ghost
 *   (Vector/cuda/vector_dist_cuda_funcs.cuh:19-43) in particle-major order;
 *   then ordered by processor, stably, as labelParticlesGhost_impl's sort
 *   (Vector/cuda/vector_dist_comm_util_funcs.cuh:110-118) leaves them when the
 *   sort keeps equal keys in input order.  The walk's same-pair rule is pinned
 *   by fixtures of overlapping boxes whose outputs follow from the rule by hand
 *   (tests/test_remote.py).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "ofp_oracle.h"

size_t ofo_remote_geo(int n, const double* lo, const double* hi, double glo[3], double cs[3], uint32_t div[3],
                      uint32_t* cell_start, uint32_t* cell_boxes) {
  double ghi[3], wmin[3], eps[3];
  for (int k = 0; k < 3; ++k) {
    glo[k] = INFINITY;
    ghi[k] = -INFINITY;
    wmin[k] = INFINITY;
  }
  for (int b = 0; b < n; ++b)
    for (int k = 0; k < 3; ++k) {
      if (lo[3 * b + k] < glo[k]) glo[k] = lo[3 * b + k];
      if (hi[3 * b + k] > ghi[k]) ghi[k] = hi[3 * b + k];
      double w = hi[3 * b + k] - lo[3 * b + k];
      if (w > 0.0 && w < wmin[k]) wmin[k] = w;
    }
  for (int k = 0; k < 3; ++k) {
    if (n == 0) {
      glo[k] = 0.0;
      ghi[k] = 1.0;
    }
    double ext = ghi[k] > glo[k] ? ghi[k] - glo[k] : 1.0;
    double d = isfinite(wmin[k]) ? floor(ext / wmin[k]) : 1.0;
    if (d > 64.0) d = 64.0;
    if (d < 1.0) d = 1.0;
    div[k] = (uint32_t)d;
    cs[k] = ext / d;
    eps[k] = 1e-9 * ext;
  }
  size_t m = 0;
  for (uint32_t cz = 0; cz < div[2]; ++cz)
    for (uint32_t cy = 0; cy < div[1]; ++cy)
      for (uint32_t cx = 0; cx < div[0]; ++cx) {
        size_t c = cx + (size_t)div[0] * (cy + (size_t)div[1] * cz);
        uint32_t ci[3] = {cx, cy, cz};
        if (cell_start) cell_start[c] = (uint32_t)m;
        for (int b = 0; b < n; ++b) {
          int meet = 1;
          for (int k = 0; k < 3; ++k) {
            double clo = glo[k] + (double)ci[k] * cs[k] - eps[k];
            double chi = glo[k] + (double)(ci[k] + 1) * cs[k] + eps[k];
            meet = meet && lo[3 * b + k] <= chi && hi[3 * b + k] >= clo;
          }
          if (meet) {
            if (cell_boxes) cell_boxes[m] = (uint32_t)b;
            ++m;
          }
        }
      }
  if (cell_start) cell_start[(size_t)div[0] * div[1] * div[2]] = (uint32_t)m;
  return m;
}

/* Stable order of the n outputs by processor: counting sort over the distinct
 * processors (ascending). */
static void order_by_proc(size_t n, const uint32_t* p_in, const uint64_t* id_in, uint32_t* p_out, uint64_t* id_out) {
  if (n == 0) return;
  uint32_t pmax = 0;
  for (size_t i = 0; i < n; ++i)
    if (p_in[i] > pmax) pmax = p_in[i];
  size_t* start = calloc((size_t)pmax + 2, sizeof(size_t));
  for (size_t i = 0; i < n; ++i) start[p_in[i] + 1]++;
  for (size_t q = 0; q <= pmax; ++q) start[q + 1] += start[q];
  for (size_t i = 0; i < n; ++i) {
    size_t d = start[p_in[i]]++;
    p_out[d] = p_in[i];
    id_out[d] = id_in[i];
  }
  free(start);
}

void ofo_order_by_proc(size_t n, const uint32_t* p_in, const uint64_t* id_in, uint32_t* p_out, uint64_t* id_out) {
  order_by_proc(n, p_in, id_in, p_out, id_out);
}

#define REMOTE_LABEL(T, S)                                                                                   \
  size_t ofo_remote_label##S(const T* pos, size_t stride, size_t n, const T* blo, const T* bhi,              \
                             const uint32_t* proc, const uint32_t* shift, const double glo[3],               \
                             const double cs[3], const uint32_t div[3], const uint32_t* cell_start,          \
                             const uint32_t* cell_boxes, uint32_t* out_proc, uint64_t* out_id) {             \
    size_t m = 0;                                                                                            \
    uint32_t* lp = NULL;                                                                                     \
    uint64_t* li = NULL;                                                                                     \
    for (int pass = 0; pass < 2; ++pass) {                                                                   \
      if (pass == 1) {                                                                                       \
        if (!out_proc) return m;                                                                             \
        lp = malloc((m ? m : 1) * sizeof(uint32_t));                                                         \
        li = malloc((m ? m : 1) * sizeof(uint64_t));                                                         \
        m = 0;                                                                                               \
      }                                                                                                      \
      for (size_t p = 0; p < n; ++p) {                                                                       \
        T x[3] = {pos[p], pos[stride + p], pos[2 * stride + p]};                                             \
        uint32_t c[3];                                                                                       \
        for (int k = 0; k < 3; ++k) {                                                                        \
          double f = floor(((double)x[k] - glo[k]) / cs[k]);                                                 \
          c[k] = f < 0.0 ? 0u : (f >= (double)div[k] ? div[k] - 1u : (uint32_t)f);                           \
        }                                                                                                    \
        size_t cell = c[0] + (size_t)div[0] * (c[1] + (size_t)div[1] * c[2]);                                \
        uint32_t b0 = cell_start[cell], b1 = cell_start[cell + 1];                                           \
        /* ie_ghost_gpu.cuh:49-95: switch_prc stays set while the pair repeats */                            \
        int switch_prc = 0;                                                                                  \
        uint32_t proc_prev = 0, shift_prev = 0;                                                              \
        for (uint32_t i = b0; i < b1; ++i) {                                                                 \
          uint32_t b = cell_boxes[i];                                                                        \
          uint32_t pa = proc[b], sa = shift[b];                                                              \
          if (i > b0) switch_prc = (pa == proc_prev && sa == shift_prev) && switch_prc;                      \
          int inside = 1;                                                                                    \
          for (int k = 0; k < 3; ++k) inside = inside && !(x[k] < blo[3 * b + k] || x[k] >= bhi[3 * b + k]); \
          if (inside && !switch_prc) {                                                                       \
            if (pass == 1) {                                                                                 \
              lp[m] = pa;                                                                                    \
              li[m] = (uint64_t)p | ((uint64_t)sa << 32);                                                    \
            }                                                                                                \
            switch_prc = 1;                                                                                  \
            ++m;                                                                                             \
          }                                                                                                  \
          proc_prev = pa;                                                                                    \
          shift_prev = sa;                                                                                   \
        }                                                                                                    \
      }                                                                                                      \
    }                                                                                                        \
    order_by_proc(m, lp, li, out_proc, out_id);                                                              \
    free(lp);                                                                                                \
    free(li);                                                                                                \
    return m;                                                                                                \
  }

REMOTE_LABEL(float, _f32)
REMOTE_LABEL(double, _f64)

/* map: processor of every particle (already wrapped by ofo_apply_bc), as
 * process_id_proc_each_part (Vector/cuda/vector_dist_cuda_funcs.cuh:55-75) ->
 * processorID_impl (Decomposition/cuda/CartDecomposition_gpu.cuh:13-56): the
 * first sub-domain of the particle's geometric cell, in list order, that
 * contains it by Box::isInsideNP_with_border (openfpm_data, absent from the
 * reference tree; restated from its published rule: per dimension
 * lo <= x < hi, except x <= hi where the dimension is non-periodic and the box's
 * upper face is the domain's).  0xFFFFFFFF where no sub-domain contains it (the
 * reference returns an unspecified processor there; the product drops it). */
#define MAP_PROC(T, S)                                                                                       \
  void ofo_map_proc##S(const T* pos, size_t stride, size_t n, const T* blo, const T* bhi,                    \
                       const uint32_t* proc, const T dom_hi[3], const int bc[3], const double glo[3],        \
                       const double cs[3], const uint32_t div[3], const uint32_t* cell_start,                \
                       const uint32_t* cell_boxes, uint32_t* out_proc) {                                     \
    for (size_t p = 0; p < n; ++p) {                                                                         \
      T x[3] = {pos[p], pos[stride + p], pos[2 * stride + p]};                                               \
      uint32_t c[3];                                                                                         \
      for (int k = 0; k < 3; ++k) {                                                                          \
        double f = floor(((double)x[k] - glo[k]) / cs[k]);                                                   \
        c[k] = f < 0.0 ? 0u : (f >= (double)div[k] ? div[k] - 1u : (uint32_t)f);                             \
      }                                                                                                      \
      size_t cell = c[0] + (size_t)div[0] * (c[1] + (size_t)div[1] * c[2]);                                  \
      uint32_t r = 0xFFFFFFFFu;                                                                              \
      for (uint32_t i = cell_start[cell]; i < cell_start[cell + 1] && r == 0xFFFFFFFFu; ++i) {               \
        uint32_t b = cell_boxes[i];                                                                          \
        int inside = 1;                                                                                      \
        for (int k = 0; k < 3; ++k) {                                                                        \
          T lo = blo[3 * b + k], hi = bhi[3 * b + k];                                                        \
          int closed = !bc[k] && hi == dom_hi[k];                                                            \
          if (x[k] < lo || (closed ? x[k] > hi : x[k] >= hi)) inside = 0;                                    \
        }                                                                                                    \
        if (inside) r = proc[b];                                                                             \
      }                                                                                                      \
      out_proc[p] = r;                                                                                       \
    }                                                                                                        \
  }

MAP_PROC(float, _f32)
MAP_PROC(double, _f64)
