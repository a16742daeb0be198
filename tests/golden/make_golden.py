"""Writes the golden fixtures of this folder from the reference library.

Runs in the build container only: it loads oracle/_ref/libh2ref_*.so, the
reference (H2Lib) compiled from its own sources by oracle/Makefile, and
stores inputs and outputs as .npz files (plain arrays, no pickles).

    python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

from oracle_ffi import (Reference, ReferenceBem, ReferenceH2,  # noqa: E402
                        have_reference)


def save(name, **arrays):
    path = os.path.join(HERE, name)
    np.savez_compressed(path, **arrays)
    print("wrote", name, os.path.getsize(path), "bytes")


def meshes():
    for kind, name, split in [(0, "sphere", 4), (0, "sphere", 8),
                              (0, "sphere", 16), (1, "cube", 4),
                              (2, "cylinder", 2)]:
        x, t, n, g = Reference.mesh(kind, split)
        save("mesh_%s_s%d.npz" % (name, split), x=x, t=t, n=n, g=g,
             split=np.array(split), kind=np.array(kind))


def rules():
    out = {}
    for q, q2 in [(2, 4), (3, 5)]:
        for c, cname in enumerate(["dist", "vert", "edge", "id", "single"]):
            x, y, w, base = Reference.rule(q, q2, c)
            out["%s_%d_%d_x" % (cname, q, q2)] = x
            out["%s_%d_%d_y" % (cname, q, q2)] = y
            out["%s_%d_%d_w" % (cname, q, q2)] = w
    save("singquad.npz", **out)


def dense():
    """Dense matrices through nearfield(NULL, NULL, ...) on the s=4 sphere
    (128 triangles, 66 vertices), as Tests/test_laplacebem3d.c:202-205 does
    at larger size."""
    x, t, n, g = Reference.mesh(0, 4)
    out = {}
    cases = [("lap_slp_cc", 0, 0, 0, 0, 0.0, 0.0),
             ("lap_dlp_cc", 0, 1, 0, 0, 0.5, 0.0),
             ("lap_slp_ll", 0, 0, 1, 1, 0.0, 0.0),
             ("lap_dlp_ll", 0, 1, 1, 1, -0.5, 0.0),
             ("lap_slp_cl", 0, 0, 0, 1, 0.0, 0.0),
             ("lap_dlp_lc", 0, 1, 1, 0, 0.5, 0.0),
             ("helm_slp_cc", 1, 0, 0, 0, 0.0, 2.0),
             ("helm_dlp_cc", 1, 1, 0, 0, 0.5, 2.0),
             ("helm_slp_ll", 1, 0, 1, 1, 0.0, 2.0),
             ("helm_dlp_ll", 1, 1, 1, 1, 0.5, 2.0)]
    for name, kern, op, br, bc, alpha, k in cases:
        b = ReferenceBem(kern, op, br, bc, x, t, q=2, q2=4, alpha=alpha, k=k)
        out[name] = b.nearfield()
        out[name + "_far"] = b.nearfield(far=True)
        del b
    save("dense_sphere_s4.npz", **out)


def sampled():
    """64 entries per rule class of the cc Laplace SLP/DLP and Helmholtz SLP
    on the s=16 sphere (2048 triangles, the dense workload's mesh)."""
    x, t, n, g = Reference.mesh(0, 16)
    T = len(t)
    rng = np.random.default_rng(2024)
    tt = t.astype(np.int64)
    pairs = {0: [], 1: [], 2: [], 3: []}
    while min(len(v) for v in pairs.values()) < 64:
        i = int(rng.integers(T))
        if len(pairs[3]) < 64:
            pairs[3].append((i, i))
        j = int(rng.integers(T))
        share = len(set(tt[i]) & set(tt[j]))
        if len(pairs[share]) < 64:
            pairs[share].append((i, j))
        # neighbours of i for edge / vertex classes
        for j in np.nonzero(np.isin(tt, tt[i]).any(axis=1))[0]:
            share = len(set(tt[i]) & set(tt[int(j)]))
            if share in (1, 2) and len(pairs[share]) < 64:
                pairs[share].append((i, int(j)))
    rows, cols = [], []
    for c in range(4):
        for i, j in pairs[c][:64]:
            rows.append(i)
            cols.append(j)
    rows = np.array(rows, dtype=np.uint32)
    cols = np.array(cols, dtype=np.uint32)
    out = {"rows": rows, "cols": cols}
    for name, kern, op, alpha, k in [("lap_slp", 0, 0, 0.0, 0.0),
                                     ("lap_dlp", 0, 1, 0.5, 0.0),
                                     ("helm_slp", 1, 0, 0.0, 4.0)]:
        b = ReferenceBem(kern, op, 0, 0, x, t, q=2, q2=4, alpha=alpha, k=k)
        vals = []
        for i, j in zip(rows, cols):
            vals.append(b.nearfield(np.array([i]), np.array([j]))[0, 0])
        out[name] = np.array(vals)
        del b
    save("sampled_sphere_s16.npz", **out)


def h2():
    """Cluster tree, near-field blocks and H2 products of the Laplace SLP
    (linear basis) on the s=8 sphere, with clf=16, eta=2, m=3."""
    x, t, n, g = Reference.mesh(0, 8)
    b = ReferenceBem(0, 0, 1, 1, x, t, q=2, q2=4)
    h = ReferenceH2(b, 16, 2.0, 3, 0)
    nb = h.near_blocks()
    rng = np.random.default_rng(7)
    xv = rng.uniform(-1.0, 1.0, h.n)
    y = h.addeval(xv, np.zeros(h.n))
    yt = h.addeval(xv, np.zeros(h.n), trans=True)
    near0 = h.near_entries(0, int(nb[0, 1]), int(nb[0, 3]))
    save("h2_sphere_s8.npz", idx=h.idx(), near_blocks=nb, x=xv, y=y, yt=yt,
         near0=near0, sizes=h.sizes())
    del h, b


def dirichlet():
    """The interior Dirichlet problem of the Laplace test at full size: unit
    sphere split 64 (32768 triangles), linear basis, H2 by interpolation m=4,
    clf=64, eta=2, GMRES(30); the solution solved to residual 1e-12 |b|
    (the accuracy yardstick) and to 1e-8 |b| (the reference's own setting)."""
    from oracle_ffi import reference_dirichlet
    x, t, n, g = Reference.mesh(0, 64)
    f, b, xs, it, err = reference_dirichlet(x, t, 64, 2.0, 4, 1e-12, 5000, 30)
    _, _, xs8, it8, err8 = reference_dirichlet(x, t, 64, 2.0, 4, 1e-8, 5000,
                                               30)
    save("dirichlet_sphere_s64.npz", f=f, b=b, x=xs, steps=np.array(it),
         neumann_err=np.array(err), x_eps8=xs8, steps_eps8=np.array(it8),
         neumann_err_eps8=np.array(err8))


if __name__ == "__main__":
    if not have_reference(False) or not have_reference(True):
        sys.exit("build oracle/_ref first: make -C oracle")
    which = sys.argv[1:] or ["meshes", "rules", "dense", "sampled", "h2",
                             "dirichlet"]
    for name in which:
        globals()[name]()
