"""Writes the golden fixtures used by tests/test_oracle.py (run in the build container only).

1. embedding_q2_2d/3d, embedding_q1_2d/3d (.npy): the child embedding tables
   held as numbers in the reference source src/MGGeomEl.C:182-276 (Quad9,
   Quad4) and :322-690 (Hex27, Hex8), parsed from the text of that file.
2. laspack_*.npz: inputs/outputs of the reference's vendored LASPACK library
   (contrib/laspack, compiled from its sources into oracle/_ref by
   `make -C oracle ref`): Mul_QV, ILUPrecond and GMRESIter(+ILUPrecond) on a
   small assembled heat matrix (heat2d) and a small Navier-Stokes Oseen
   saddle-point matrix with a 1e-3 pressure diagonal (ns2d), both assembled by
   the oracle.
Usage: python tests/golden/make_fixtures.py
"""
import ctypes as C
import os
import re
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
REF = "/root/reference/src/MGGeomEl.C"


def parse_tables():
    txt = open(REF).read()
    out = {}
    for key, decl, dim, ns in (("embedding_q2_2d", "_embedding_matrix_q[4][9][9]", 2, 9),
                               ("embedding_q1_2d", "_embedding_matrix_l[4][4][4]", 2, 4),
                               ("embedding_q2_3d", "_embedding_matrix_q[8][27][27]", 3, 27),
                               ("embedding_q1_3d", "_embedding_matrix_l[8][8][8]", 3, 8)):
        j = txt.index("MGGeomEl::" + decl)
        j = txt.index("=", j)
        body = txt[j:]
        body = body[: body.index("};")]
        body = re.sub(r"//[^\n]*", "", body)
        nums = [float(v) for v in re.findall(r"[-+]?\d*\.\d+(?:[eE][-+]?\d+)?|[-+]?\d+\.?", body)]
        nch = 2 ** dim
        assert len(nums) == nch * ns * ns, (key, len(nums))
        out[key] = np.array(nums, dtype=np.float64).reshape(nch, ns, ns)
    return out


def laspack_vectors():
    from oracle import oracle as O
    from exfemus_amd import configs
    R = O.ref_lib()
    if R is None:
        raise SystemExit("oracle/_ref/liblaspack_ref.so missing: make -C oracle ref")
    res = {}
    # heat 2D, 2 levels, top matrix (k log-normal), non-symmetric because of Dirichlet rows
    m = O.Mesh(2, [2, 2], [0, 0], [1, 1], 2)
    dof, n = O.init_dof(m, 1, 1, 0)
    b0, b1 = O.genbc(m, 1, 0, dof, configs.heat_dirichlet_all)
    rp, ci = O.system_pattern(m, 1, 1, 0, dof)
    k = configs.lognormal_k(m.level(1)[0])
    p = O.HeatParams(1.0, 1.0, 0.1, 1.0, 1.0, 1.0, (0.3, -0.2, 0.0))
    v, rhs = O.heat_genmatrhs(m, 1, 1, p, dof, dof, b0, b1, np.linspace(0, 1, n), rp, ci, k_elem=k)
    cases = {"heat2d": (rp, ci, v, rhs)}
    # NS 2D cavity, top of 2 levels, Oseen matrix at a nonzero velocity (saddle point, non-symmetric); the
    # pressure rows have no diagonal in the reference matrix and LASPACK's ILU(0) divides by it, so 1e-3 is
    # added on the p-p diagonal (the pattern holds it: the element couples every DoF with every other)
    dn, nn = O.init_dof(m, 1, 2, 1)
    c0, c1 = O.genbc(m, 2, 1, dn, configs.cavity_bc)
    rpn, cin = O.system_pattern(m, 1, 2, 1, dn)
    xo = np.zeros(nn)
    xyz = m.coords()
    N = m.N
    for k in range(m.level(1)[1]):
        xo[dn[k]] = np.sin(np.pi * xyz[0, k]) * xyz[1, k]
        xo[dn[k + N]] = -0.5 * xyz[0, k] * (1 - xyz[1, k])
    pn = O.NSParams(0.01, 0.1, 1.0, 1.0, 0.0, (0.0, 0.0, 0.0), (0.0, 0.0, 0.0))
    vn, bn = O.ns_genmatrhs(m, 1, 1, pn, dn, dn, c0, c1, xo, rpn, cin)
    npr = m.level(1)[2]
    for k in range(npr):
        r = dn[k + 2 * N]
        q = rpn[r] + int(np.flatnonzero(cin[rpn[r]:rpn[r + 1]] == r)[0])
        vn[q] += 1e-3
    cases["ns2d"] = (rpn, cin, vn, bn)
    rng = np.random.default_rng(3)
    for name, (rp_, ci_, v_, b_) in cases.items():
        nn = len(rp_) - 1
        x = rng.standard_normal(nn)
        y = np.zeros(nn)
        R.lr_spmv(nn, rp_, ci_, v_, x, y)
        yi = np.zeros(nn)
        R.lr_ilu_apply(nn, rp_, ci_, v_, x, yi)
        xs = np.zeros(nn)
        acc = C.c_double(); its = C.c_int()
        R.lr_gmres(nn, rp_, ci_, v_, b_, xs, 8, 30, 1, 1e-20, C.byref(acc), C.byref(its))
        xs2 = np.zeros(nn)
        acc2 = C.c_double(); its2 = C.c_int()
        R.lr_gmres(nn, rp_, ci_, v_, b_, xs2, 45, 30, 1, 1e-20, C.byref(acc2), C.byref(its2))
        res[f"laspack_{name}"] = dict(rp=rp_, ci=ci_, v=v_, b=b_, x=x, spmv=y, ilu=yi, gmres8=xs,
                                      gmres8_acc=acc.value, gmres8_its=its.value, gmres45=xs2,
                                      gmres45_acc=acc2.value, gmres45_its=its2.value)
    return res


if __name__ == "__main__":
    for k, v in parse_tables().items():
        np.save(os.path.join(HERE, k + ".npy"), v)
        print(k, v.shape)
    for k, d in laspack_vectors().items():
        np.savez_compressed(os.path.join(HERE, k + ".npz"), **d)
        print(k, {kk: np.shape(vv) for kk, vv in d.items()})
