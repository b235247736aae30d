"""Writes the golden fixtures of tests/golden/ from the reference tests' own
known answers (the rules those tests assert), never from the oracle.

  shift_vectors.json  Decomposition/tests/shift_vect_converter_tests.cpp:17-51
  ghost_kat.json      Vector/cuda/vector_dist_cuda_func_test.cu:19-389
  lattice_shells.json Vector/tests/vector_dist_unit_test.cpp:1199-1322

Run: python tests/golden/make_golden.py
"""
import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def shift_vectors():
    # comb<3> c({a,b,c}) fills c[2]=a, c[1]=b, c[0]=c: the test's first case
    # comb {-1,-1,1} must map to shift (-1, 1, 1) for the unit box, which holds
    # only with that reversed initializer.  Entries list c[] in storage order.
    cases = [
        {"init": [-1, -1, 1], "c": [1, -1, -1], "shift": [-1.0, 1.0, 1.0]},
        {"init": [-1, 0, 1], "c": [1, 0, -1], "shift": [-1.0, 0.0, 1.0]},
        {"init": [0, 0, 1], "c": [1, 0, 0], "shift": [-1.0, 0.0, 0.0]},
    ]
    return {"domain_lo": [0.0, 0.0, 0.0], "domain_hi": [1.0, 1.0, 1.0], "n_shifts": 27, "cases": cases}


def ghost_kat(n=512, seed=1234):
    # Four boxes [0, t) x [0,1) x [0,1), t = 0.5, 0.3, 0.2, 0.1 with shift ids 0..3;
    # shift s = (10 (s+1), 0, 0).  The test asserts: count 0 if x >= 0.5, 1 if
    # x in [0.3,0.5), 2 if [0.2,0.3), 3 if [0.1,0.2), 4 below 0.1; ghost j of a
    # particle is at x - 10 (j+1) with o_part_loc (i, j), properties copied.
    rng = np.random.Generator(np.random.PCG64(seed))
    pos = rng.random((3, n)).astype(np.float32)
    x = pos[0]
    cnt = np.where(x >= 0.5, 0, np.where(x >= 0.3, 1, np.where(x >= 0.2, 2, np.where(x >= 0.1, 3, 4))))
    gx, gy, gz, op, os_ = [], [], [], [], []
    for i in range(n):
        for j in range(int(cnt[i])):
            gx.append(float(np.float32(x[i]) - np.float32(10.0 * (j + 1))))
            gy.append(float(pos[1, i]))
            gz.append(float(pos[2, i]))
            op.append(i)
            os_.append(j)
    return {
        "box_lo": [[0.0, 0.0, 0.0]] * 4,
        "box_hi": [[0.5, 1.0, 1.0], [0.3, 1.0, 1.0], [0.2, 1.0, 1.0], [0.1, 1.0, 1.0]],
        "box_sv": [0, 1, 2, 3],
        "shifts": [[10.0 * (s + 1), 0.0, 0.0] for s in range(4)],
        "pos": pos.tolist(),
        "expected_count": cnt.tolist(),
        "expected_ghost_pos": [gx, gy, gz],
        "expected_opart": [op, os_],
    }


def lattice_shells():
    # Verlet of radius third_dist*1.05 on an Ng^3 lattice of spacing 1/Ng,
    # ghost third_dist*1.1 (third_dist = sqrt(3) spacing): every particle has
    # 7 (6 + self) neighbours within first_dist*1.05, 12 more within
    # second_dist*1.05 and 8 more beyond.
    return {"cases": [{"ng": ng, "first": 7, "second": 12, "third": 8} for ng in (12, 17, 24)]}


def main():
    for name, obj in (("shift_vectors", shift_vectors()), ("ghost_kat", ghost_kat()),
                      ("lattice_shells", lattice_shells())):
        with open(os.path.join(HERE, name + ".json"), "w") as f:
            json.dump(obj, f)
        print("wrote", name)


if __name__ == "__main__":
    main()
