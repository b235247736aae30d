// User-code test of the C++ facade: the reference's GPU test kernels and call
// sequence (Vector/cuda/vector_dist_gpu_unit_tests.cu:46-119, :181-290,
// :425-568) compiled unchanged against include/openfpm_amd/vector_dist.hpp.
// Prints one line per check and exits non-zero on the first failure.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <random>

#include "openfpm_amd/vector_dist.hpp"

static int failures = 0;
#define CHECK(cond, ...)                      \
  do {                                        \
    if (!(cond)) {                            \
      std::printf("FAIL %s:%d ", __FILE__, __LINE__); \
      std::printf(__VA_ARGS__);               \
      std::printf("\n");                      \
      ++failures;                             \
    }                                         \
  } while (0)

// ---- the reference test kernels, verbatim in structure ----
template <typename T, typename CellList_type>
__global__ void calculate_force(vector_dist_ker<3, T, aggregate<T, T[3], T[3]>> vecDist, CellList_type cellList,
                                int rank) {
  size_t p = GET_PARTICLE(vecDist);

  Point<3, T> xp = vecDist.getPos(p);

  auto it = cellList.getNNIteratorBox(cellList.getCell(xp));

  Point<3, T> force({0.0, 0.0, 0.0});

  while (it.isNext()) {
    auto q = it.get();

    if (q == p) {
      ++it;
      continue;
    }

    Point<3, T> xq = vecDist.getPos(q);
    Point<3, T> r = xq - xp;

    if (r.norm() > 1e-6) {
      r /= r.norm();
      force += vecDist.template getProp<0>(q) * r;
    }

    ++it;
  }

  vecDist.template getProp<1>(p)[0] = force.get(0);
  vecDist.template getProp<1>(p)[1] = force.get(1);
  vecDist.template getProp<1>(p)[2] = force.get(2);
}

template <typename T, typename CellList_type>
__global__ void calculate_force_sort(vector_dist_ker<3, T, aggregate<T, T[3], T[3]>> vecDistSort,
                                     CellList_type cellList, int rank) {
  size_t p;
  GET_PARTICLE_SORT(p, cellList);

  Point<3, T> xp = vecDistSort.getPos(p);
  Point<3, T> force({0.0, 0.0, 0.0});

  auto it = cellList.getNNIteratorBox(cellList.getCell(xp));

  while (it.isNext()) {
    auto q = it.get_sort();

    if (q == p) {
      ++it;
      continue;
    }

    Point<3, T> xq = vecDistSort.getPos(q);
    Point<3, T> r = xq - xp;

    if (r.norm() > 1e-6) {
      r /= r.norm();
      force += vecDistSort.template getProp<0>(q) * r;
    }

    ++it;
  }

  vecDistSort.template getProp<1>(p)[0] = force.get(0);
  vecDistSort.template getProp<1>(p)[1] = force.get(1);
  vecDistSort.template getProp<1>(p)[2] = force.get(2);
}

// host check: same formula over the particles of the 27 cells around x_p
template <typename T, typename VD, typename CL>
static void check_force(VD& vd, const CL& NN, const char* what) {
  const size_t n = vd.size_local_with_ghost();
  auto cell = [&](size_t i, int* c) {
    for (int d = 0; d < 3; ++d) {
      const T sp = (NN.hi[d] - NN.lo[d]) / T(NN.div[d]);
      c[d] = int(std::floor((vd.getPos(i)[d] - NN.lo[d]) / sp)) + 1;
    }
  };
  std::vector<int> cells(3 * n);
  for (size_t i = 0; i < n; ++i) cell(i, &cells[3 * i]);
  double maxerr = 0;
  for (size_t p = 0; p < vd.size_local(); ++p) {
    T f[3] = {0, 0, 0};
    for (size_t q = 0; q < n; ++q) {
      if (q == p) continue;
      bool nb = true;
      for (int d = 0; d < 3; ++d) nb &= std::abs(cells[3 * q + d] - cells[3 * p + d]) <= 1;
      if (!nb) continue;
      T r[3], nr = 0;
      for (int d = 0; d < 3; ++d) {
        r[d] = vd.getPos(q)[d] - vd.getPos(p)[d];
        nr += r[d] * r[d];
      }
      nr = std::sqrt(nr);
      if (nr > 1e-6)
        for (int d = 0; d < 3; ++d) f[d] += vd.template getProp<0>(q) * (r[d] / nr);
    }
    for (int d = 0; d < 3; ++d)
      maxerr = std::max(maxerr, double(std::abs(vd.template getProp<1>(p)[d] - f[d])));
  }
  // the reference's tolerance (vector_dist_gpu_unit_tests.cu:160-162)
  CHECK(maxerr < 3e-4, "%s: max |F_gpu - F_host| = %g", what, maxerr);
  std::printf("%s: max abs force difference %.3g over %zu particles\n", what, maxerr, vd.size_local());
}

template <typename T>
static void test_ghost_and_forces(size_t n, bool sorted) {
  Box<3, T> domain({0.0, 0.0, 0.0}, {1.0, 1.0, 1.0});
  Ghost<3, T> g(0.1);
  size_t bc[3] = {PERIODIC, PERIODIC, PERIODIC};
  vector_dist_gpu<3, T, aggregate<T, T[3], T[3]>> vd(n, domain, bc, g);

  std::mt19937 rng(42);
  std::uniform_real_distribution<T> U(0, 1);
  auto it = vd.getDomainIterator();
  while (it.isNext()) {
    auto p = it.get();
    vd.getPos(p)[0] = U(rng);
    vd.getPos(p)[1] = U(rng);
    vd.getPos(p)[2] = U(rng);
    vd.template getProp<0>(p) = vd.getPos(p)[0] + vd.getPos(p)[1] + vd.getPos(p)[2];
    vd.template getProp<2>(p)[0] = vd.getPos(p)[0] + T(3.0) * vd.getPos(p)[1];
    ++it;
  }
  vd.hostToDevicePos();
  vd.template hostToDeviceProp<0, 2>();
  vd.map(RUN_ON_DEVICE);
  vd.template ghost_get<0, 2>(RUN_ON_DEVICE);
  vd.deviceToHostPos();
  vd.template deviceToHostProp<0, 2>();

  // ghosts: properties are the source's, position = source - shift, outside the domain
  const size_t ng = vd.size_local_with_ghost() - vd.size_local();
  CHECK(ng > 0, "no ghosts");
  size_t bad = 0;
  for (size_t q = vd.size_local(); q < vd.size_local_with_ghost(); ++q) {
    bool inside = true;
    for (int d = 0; d < 3; ++d) inside &= vd.getPos(q)[d] >= T(0) && vd.getPos(q)[d] < T(1);
    bad += inside;
    bad += vd.getPos(q)[0] < T(-0.1) - T(1e-6) || vd.getPos(q)[0] > T(1.1) + T(1e-6);
  }
  CHECK(bad == 0, "%zu ghosts out of place", bad);
  // domain + ghosts exceed the domain count (the reference test's assertion)
  CHECK(vd.size_local_with_ghost() > n, "size_local_with_ghost %zu", vd.size_local_with_ghost());

  auto NN = vd.getCellListGPU(T(0.1), sorted ? (CL_NON_SYMMETRIC | CL_GPU_REORDER) : CL_NON_SYMMETRIC);
  vd.template updateCellListGPU<0, 1, 2>(NN);
  auto ite = vd.getDomainIteratorGPU(32);
  if (sorted) {
    CUDA_LAUNCH((calculate_force_sort<T, decltype(NN.toKernel())>), ite, vd.toKernel(), NN.toKernel(), 0);
    vd.template restoreOrder<0, 1, 2>(NN);
  } else {
    CUDA_LAUNCH((calculate_force<T, decltype(NN.toKernel())>), ite, vd.toKernel(), NN.toKernel(), 0);
  }
  openfpm_amd::hip_check(hipDeviceSynchronize(), "sync");
  vd.template deviceToHostProp<0, 1, 2>();
  vd.deviceToHostPos();
  check_force<T>(vd, NN, sorted ? "calculate_force_sort" : "calculate_force");
}

template <typename T>
static void test_verlet(size_t n) {
  Box<3, T> domain({0.0, 0.0, 0.0}, {1.0, 1.0, 1.0});
  const T r = T(0.08);
  Ghost<3, T> g(r);
  size_t bc[3] = {PERIODIC, PERIODIC, PERIODIC};
  vector_dist_gpu<3, T, aggregate<T>> vd(n, domain, bc, g);
  std::mt19937 rng(7);
  std::uniform_real_distribution<T> U(0, 1);
  for (size_t p = 0; p < n; ++p)
    for (int d = 0; d < 3; ++d) vd.getPos(p)[d] = U(rng);
  vd.hostToDevicePos();
  vd.template ghost_get<>(RUN_ON_DEVICE);
  vd.deviceToHostPos();
  const size_t nt = vd.size_local_with_ghost();
  std::vector<T> P(3 * nt);
  for (size_t i = 0; i < nt; ++i)
    for (int d = 0; d < 3; ++d) P[3 * i + d] = vd.getPos(i)[d];
  for (int sub = 1; sub <= 2; ++sub) {
    auto vl = vd.getVerletGPU(r, sub);
    // rows: count per domain particle; compare as multiset of counts
    ofp_vl_view v;
    openfpm_amd::check(ofp_vl_get_view(vl.h, &v), "view");
    std::vector<uint64_t> off(v.n_rows + 1);
    openfpm_amd::hip_check(hipMemcpy(off.data(), v.offsets, off.size() * 8, hipMemcpyDeviceToHost), "off");
    std::vector<uint32_t> cnt_gpu(v.n_rows);
    for (size_t k = 0; k < v.n_rows; ++k) cnt_gpu[k] = uint32_t(off[k + 1] - off[k]);
    std::vector<uint32_t> cnt_host(n);
    for (size_t p = 0; p < n; ++p) {
      uint32_t c = 0;
      for (size_t q = 0; q < nt; ++q) {
        T d2 = 0;
        for (int d = 0; d < 3; ++d) {
          T dd = P[3 * q + d] - P[3 * p + d];
          d2 += dd * dd;
        }
        c += d2 < r * r;
      }
      cnt_host[p] = c;
    }
    std::sort(cnt_gpu.begin(), cnt_gpu.end());
    std::sort(cnt_host.begin(), cnt_host.end());
    size_t mism = 0;
    for (size_t i = 0; i < n; ++i) mism += cnt_gpu[i] != cnt_host[i];
    CHECK(v.n_rows == n && mism == 0, "Verlet sub=%d: %zu count mismatches", sub, mism);
    std::printf("getVerletGPU(sub=%d): %llu rows, %llu entries\n", sub, (unsigned long long)v.n_rows,
                (unsigned long long)v.n_entries);
    vd.template restoreOrder<>(vl.cl);
  }
}

int main() {
  try {
    test_ghost_and_forces<float>(1000, false);
    test_ghost_and_forces<float>(1000, true);
    test_ghost_and_forces<double>(2000, true);
    test_verlet<float>(3000);
    test_verlet<double>(3000);
  } catch (const std::exception& e) {
    std::printf("FAIL exception: %s\n", e.what());
    return 2;
  }
  std::printf(failures ? "FAILED (%d)\n" : "OK\n", failures);
  return failures ? 1 : 0;
}
