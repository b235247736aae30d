// ofp_geom.hip -- host-side geometry of the one-rank path: periodic shift
// vectors, local ghost boxes, cell-list parameters.  No device code.
#include "ofp_internal.hpp"

namespace ofp {
namespace {

// comb<3>::lin(): sum_k (c_k + 1) 3^k (the id shift_vect_converter::linId_ld
// returns, Decomposition/shift_vect_converter.hpp:173-176).
inline int comb_lin(const int c[3]) { return (c[0] + 1) + 3 * (c[1] + 1) + 9 * (c[2] + 1); }
inline void comb_of(int l, int c[3]) {
  c[0] = l % 3 - 1;
  c[1] = (l / 3) % 3 - 1;
  c[2] = l / 9 - 1;
}

template <class T>
int shift_vectors(const T* lo, const T* hi, T* out) {
  for (int l = 0; l < 27; ++l) {
    int c[3];
    comb_of(l, c);
    for (int k = 0; k < 3; ++k) {
      T L = T(hi[k] - lo[k]);
      out[comb_lin(c) * 3 + k] = c[k] == 1 ? T(-L) : (c[k] == 0 ? T(0) : L);
    }
  }
  return OFP_OK;
}

// Single sub-domain == domain.  For every valid non-zero comb c (non-zero only
// along periodic dims): the border box bp of ie_loc_ghost::applyBC
// (ie_loc_ghost.hpp:150-213) selects the sub-domain, its image (sub + shift) is
// enlarged by the ghost and cut by the sub-domain (create_loc_ghost_ibox,
// :102-141).  Boxes come out in ascending comb id, the order createShiftBox sorts
// them into (vector_dist_comm.hpp:648-656).
template <class T>
int local_ghost_boxes(const T* lo, const T* hi, const T* glo, const T* ghi, const int* bc,
                      T* blo, T* bhi, uint32_t* sv, int* nb_out) {
  int nb = 0;
  for (int l = 0; l < 27; ++l) {
    int c[3];
    comb_of(l, c);
    if (c[0] == 0 && c[1] == 0 && c[2] == 0) continue;
    bool valid = true;
    for (int k = 0; k < 3; ++k)
      if (c[k] != 0 && !bc[k]) valid = false;
    if (!valid) continue;
    T bpl[3], bph[3], sh[3];
    for (int k = 0; k < 3; ++k) {
      if (c[k] == 1) {
        bpl[k] = T(hi[k] + glo[k]);
        bph[k] = hi[k];
        sh[k] = T(-hi[k] + lo[k]);
      } else if (c[k] == 0) {
        bpl[k] = lo[k];
        bph[k] = hi[k];
        sh[k] = T(0);
      } else {
        bpl[k] = lo[k];
        bph[k] = ghi[k];  // the reference takes ghost.getHigh(k) here
        sh[k] = T(hi[k] - lo[k]);
      }
    }
    bool inter = true;
    for (int k = 0; k < 3; ++k) {
      T a = lo[k] > bpl[k] ? lo[k] : bpl[k];
      T b = hi[k] < bph[k] ? hi[k] : bph[k];
      if (!(a < b)) inter = false;
    }
    if (!inter) continue;
    bool ok = true;
    T ol[3], oh[3];
    for (int k = 0; k < 3; ++k) {
      T il = T(lo[k] + sh[k]);
      T ih = T(hi[k] + sh[k]);
      il = T(il + glo[k]);
      ih = T(ih + ghi[k]);
      ol[k] = il > lo[k] ? il : lo[k];
      oh[k] = ih < hi[k] ? ih : hi[k];
      if (!(ol[k] < oh[k])) ok = false;
    }
    if (!ok) continue;
    for (int k = 0; k < 3; ++k) {
      blo[nb * 3 + k] = ol[k];
      bhi[nb * 3 + k] = oh[k];
    }
    sv[nb] = uint32_t(comb_lin(c));
    ++nb;
  }
  *nb_out = nb;
  return OFP_OK;
}

// getCellListGPU (vector_dist.hpp:1234-1248): ghostEnlarge = ghost.magnify(factor)
// with the factor converted to the dtype, then cl_param_calculate: enlarge the
// processor box by it, div = floor(extent / r_cut) + 1, high = low + div * r_cut.
template <class T>
int cl_param(const T* plo, const T* phi, double r_cut_d, const T* glo, const T* ghi,
             double factor_d, T* olo, T* ohi, uint32_t* div) {
  const T r_cut = T(r_cut_d);
  const T f = T(factor_d);
  OFP_CHECK(r_cut > T(0), OFP_E_ARG, "r_cut must be positive");
  for (int k = 0; k < 3; ++k) {
    T gl = T(glo[k] * f);
    T gh = T(ghi[k] * f);
    T l = T(plo[k] + gl);
    T h = T(phi[k] + gh);
    OFP_CHECK(h > l, OFP_E_ARG, "empty processor box along dim %d", k);
    double dd = double(T(T(h - l) / r_cut));
    OFP_CHECK(dd < 4.0e6, OFP_E_LIMIT, "too many cells along dim %d (%g)", k, dd);
    size_t d = size_t(T(T(h - l) / r_cut));
    d++;
    div[k] = uint32_t(d);
    olo[k] = l;
    ohi[k] = T(l + T(T(d) * r_cut));
  }
  return OFP_OK;
}

}  // namespace
}  // namespace ofp

using namespace ofp;

extern "C" {

int ofp_shift_vectors(int dtype, const void* lo, const void* hi, void* out) {
  OFP_CHECK(lo && hi && out, OFP_E_ARG, "ofp_shift_vectors: null argument");
  if (dtype == OFP_F32)
    return shift_vectors<float>((const float*)lo, (const float*)hi, (float*)out);
  if (dtype == OFP_F64)
    return shift_vectors<double>((const double*)lo, (const double*)hi, (double*)out);
  return set_error(OFP_E_ARG, "bad dtype %d", dtype);
}

int ofp_local_ghost_boxes(int dtype, const void* lo, const void* hi, const void* glo,
                          const void* ghi, const int* bc, void* blo, void* bhi,
                          uint32_t* sv, int* nb) {
  OFP_CHECK(lo && hi && glo && ghi && bc && blo && bhi && sv && nb, OFP_E_ARG,
            "ofp_local_ghost_boxes: null argument");
  if (dtype == OFP_F32)
    return local_ghost_boxes<float>((const float*)lo, (const float*)hi, (const float*)glo,
                                    (const float*)ghi, bc, (float*)blo, (float*)bhi, sv, nb);
  if (dtype == OFP_F64)
    return local_ghost_boxes<double>((const double*)lo, (const double*)hi,
                                     (const double*)glo, (const double*)ghi, bc,
                                     (double*)blo, (double*)bhi, sv, nb);
  return set_error(OFP_E_ARG, "bad dtype %d", dtype);
}

int ofp_cl_param(int dtype, const void* plo, const void* phi, double r_cut, const void* glo,
                 const void* ghi, double factor, void* olo, void* ohi, uint32_t* div) {
  OFP_CHECK(plo && phi && glo && ghi && olo && ohi && div, OFP_E_ARG,
            "ofp_cl_param: null argument");
  if (dtype == OFP_F32)
    return cl_param<float>((const float*)plo, (const float*)phi, r_cut, (const float*)glo,
                           (const float*)ghi, factor, (float*)olo, (float*)ohi, div);
  if (dtype == OFP_F64)
    return cl_param<double>((const double*)plo, (const double*)phi, r_cut,
                            (const double*)glo, (const double*)ghi, factor, (double*)olo,
                            (double*)ohi, div);
  return set_error(OFP_E_ARG, "bad dtype %d", dtype);
}

}  // extern "C"
