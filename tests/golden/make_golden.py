"""Writes tests/golden/ref_vectors.npz: input/output vectors produced by the
reference's own FFTW/Eigen-free translation units (compiled from the
reference sources into oracle/_ref/libbinder_ref.so by `make -C oracle ref`):

  - compute_jacobian_determinant_3D / 2D (BINDER/src/cpu/utils.cpp:10-345)
  - Spline::spline / log_spline, orders 1-3 (BINDER/src/cpu/Spline.cpp:44-145)
  - MILikelihood lookups, accumulate_info, update_parameters and compute_cost
    (BINDER/src/cpu/MILikelihood.cpp:8-206)

Run in a container where the reference sources exist:
    make -C oracle ref && python tests/golden/make_golden.py
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF = os.path.join(ROOT, "oracle", "_ref", "libbinder_ref.so")


def main():
    lib = ctypes.CDLL(REF)
    vp = ctypes.c_void_p
    d = ctypes.c_double

    def p(a):
        return a.ctypes.data_as(vp)

    out = {}
    rng = np.random.default_rng(20250101)
    # Jacobians: folded random fields, anisotropic voxels, degenerate sizes
    cases3 = [((6, 5, 4), (1.0, 1.0, 1.0), 0.7), ((9, 7, 8), (1.1, 0.9, 1.3), 1.2),
              ((2, 2, 2), (1.0, 2.0, 0.5), 0.5), ((1, 4, 3), (1.0, 1.0, 1.0), 0.3)]
    for i, (shape, vs, amp) in enumerate(cases3):
        pos = np.stack(np.indices(shape, dtype=np.float64), -1) * np.array(vs)
        f = np.ascontiguousarray(pos + rng.normal(0, amp, pos.shape))
        J = np.empty(shape)
        lib.ref_jacobian_3d(p(f), *shape, d(vs[0]), d(vs[1]), d(vs[2]), p(J))
        out[f"jac3_{i}_field"] = f
        out[f"jac3_{i}_vs"] = np.array(vs)
        out[f"jac3_{i}_J"] = J
    cases2 = [((7, 6), (1.0, 1.0), 0.6), ((5, 9), (1.3, 0.7), 1.0), ((1, 3), (1.0, 1.0), 0.2)]
    for i, (shape, vs, amp) in enumerate(cases2):
        pos = np.stack(np.indices(shape, dtype=np.float64), -1)
        f = np.ascontiguousarray(pos + rng.normal(0, amp, pos.shape))
        J = np.empty(shape)
        lib.ref_jacobian_2d(p(f), *shape, d(vs[0]), d(vs[1]), p(J))
        out[f"jac2_{i}_field"] = f
        out[f"jac2_{i}_vs"] = np.array(vs)
        out[f"jac2_{i}_J"] = J
    # Spline weights
    ds = np.array([0.0, 0.1, 0.2, 0.3, 0.4, 0.5, 0.77, 0.999, 1.0, -0.3])
    sp = np.empty((3, len(ds), 4))
    lsp = np.empty((3, len(ds), 4))
    for oi, order in enumerate((1, 2, 3)):
        for j, dv in enumerate(ds):
            a = np.zeros(4)
            b = np.zeros(4)
            lib.ref_spline(order, 0, d(dv), p(a), p(b))
            sp[oi, j] = a
            lsp[oi, j] = b
    out["spline_d"] = ds
    out["spline_w"] = sp
    out["spline_logw"] = lsp
    # MILikelihood: random lookups incl. out-of-range nodes
    K, L = 6, 5
    nshape, vshape = (5, 4, 3), (4, 4, 3)
    bn = rng.integers(0, K, nshape).astype(np.int32)
    bv = rng.integers(0, L, vshape).astype(np.int32)
    theta = rng.random((K, L)) + 0.1
    theta /= theta.sum(1, keepdims=True)
    nq = 200
    qv = rng.integers(0, np.prod(vshape), nq).astype(np.int32)
    qn = np.stack([rng.integers(-2, nshape[a] + 2, nq) for a in range(3)], -1).astype(np.int32)
    qp = rng.random(nq)
    for upd in (0, 1):
        th = theta.copy().ravel()
        lik = np.empty(nq)
        c = ctypes.CDLL(REF).ref_mi_step
        c.restype = d
        cost = c(d(2.0), p(th), K, L, p(bn), p(bv), *nshape, *vshape, nq, p(qv),
                 p(np.ascontiguousarray(qn)), p(qp), p(lik), upd)
        out[f"mi_theta_out_{upd}"] = th
        out[f"mi_lik_{upd}"] = lik
        out[f"mi_cost_{upd}"] = np.array(cost)
    out["mi_K"] = np.array(K)
    out["mi_L"] = np.array(L)
    out["mi_bn"] = bn
    out["mi_bv"] = bv
    out["mi_theta_in"] = theta
    out["mi_qv"] = qv
    out["mi_qn"] = qn
    out["mi_qp"] = qp
    out["mi_alpha"] = np.array(2.0)
    np.savez_compressed(os.path.join(HERE, "ref_vectors.npz"), **out)
    print("wrote", os.path.join(HERE, "ref_vectors.npz"))


if __name__ == "__main__":
    main()
