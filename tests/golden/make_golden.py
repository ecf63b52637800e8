"""Writes the committed parity fixtures from the oracle.

Run in the build container (it needs oracle/_ref, i.e. the reference's GENMMD
compiled from /root/reference by `make -C oracle`):

    python tests/golden/make_golden.py

Fixtures (all produced by the oracle: GENMMD itself for the ordering, the C
restatement of taucs_ccs_symbolic_elimination for the supernodes, the LAPACK
restatement of the multifrontal factor for the C1 solution):
  c1_mesh2d_100.npz     perm, invperm, etree parent, supernode sizes / tree /
                        row structure, fundamental and relaxed flop counts,
                        x for b = uniform[-1,1] (seed 1), its relative residual
  sweep_mesh2d.npz      perm and supernode digests for mesh2d sides 16..128
  sweep_mesh3d.npz      the same for mesh3d sides 3..40
  big.json              digests and statistics of mesh3d 60^3 and 80^3
  c2_mesh3d_60.npz      the FP64 oracle solution of mesh3d 60^3 for the
                        right-hand sides uniform[-1,1] of seeds 1..8 (in the
                        permuted space): x of seed 1 in full; for all eight a
                        sketch S x (S: 32 x n random signs), x at 4096 sampled
                        rows, |x|_2, and the oracle's relative residual and
                        backward error
  c3_varcoef_80.npz     the same for the 80^3 variable-coefficient operator
                        (seed 2024; conductivities 10^U[-2,2], shift 1e-6)
                        and right-hand sides of seeds 100..103
  c4_batch.npz          x in full for 8 of the 512 shifted mesh2d systems
                        (sides from numpy default_rng(7), matrix seed 10000+q,
                        right-hand side seed 500+q)

    python tests/golden/make_golden.py big-solutions   # only the last three
"""
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
from oracle import oracle as O  # noqa: E402

KEYS = ["sn_size", "sn_up_size", "first_child", "next_child", "struct_ptr", "struct_idx"]


def splitmix_uniform(n, seed, lo=-1.0, hi=1.0):
    """The seeded uniform stream of taucs_amd_random_uniform (splitmix64,
    53-bit mantissa), restated so fixtures do not depend on the product."""
    M = (1 << 64) - 1
    st = seed & M
    out = np.empty(n)
    for i in range(n):
        st = (st + 0x9E3779B97F4A7C15) & M
        z = st
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M
        z ^= z >> 31
        out[i] = lo + (hi - lo) * ((z >> 11) * (1.0 / 9007199254740992.0))
    return out


def digest(*arrays):
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a, dtype=np.int32).tobytes())
    return h.hexdigest()


def analyze(n, cp, ri, v):
    perm, inv = O.genmmd(n, cp, ri)
    pcp, pri, pv = O.permute(n, cp, ri, v, inv)
    parent = O.etree(n, pcp, pri)
    S = O.symbolic(n, pcp, pri)
    return perm, inv, (pcp, pri, pv), parent, S


SKETCH_ROWS, SKETCH_SEED, SAMPLE_ROWS = 32, 777, 4096


def sketch_matrix_apply(X, seed=SKETCH_SEED, rows=SKETCH_ROWS):
    """S X for the random-sign matrix S (rows x n) drawn from the splitmix
    stream: S[i, j] = +1 if u[i*n + j] < 1/2 else -1.  |S v|_2 / sqrt(rows)
    estimates |v|_2 (a dense projection: an error in any entry of x moves it)."""
    X = np.asarray(X).reshape(X.shape[0], -1)
    n = X.shape[0]
    out = np.empty((rows, X.shape[1]))
    for i in range(rows):
        s = np.where(O.splitmix_u01(n, seed + 7919 * i) < 0.5, 1.0, -1.0)
        out[i] = s @ X
    return out


def sample_rows(n, count=SAMPLE_ROWS, seed=SKETCH_SEED):
    return np.unique((O.splitmix_u01(count, seed) * n).astype(np.int64))


def oracle_solution(n, cp, ri, v, seeds):
    """Order with GENMMD, permute, factor and solve with the oracle; the
    right-hand sides are uniform[-1,1] of the given seeds in the permuted space."""
    perm, inv = O.genmmd(n, cp, ri)
    pcp, pri, pv = O.permute(n, cp, ri, v, inv)
    S = O.symbolic(n, pcp, pri)
    panels = O.factor(S, pcp, pri, pv)
    B = np.stack([O.uniform(n, sd) for sd in seeds], 1)
    X = O.solve(S, panels, B)
    del panels
    R = O.residual(n, pcp, pri, pv, X, B)
    import scipy.sparse as sp
    Lw = sp.csc_matrix((pv, pri, pcp), shape=(n, n))
    M = Lw + Lw.T - sp.diags(Lw.diagonal())
    anorm = abs(M).sum(1).max()
    relres = np.linalg.norm(R, axis=0) / np.linalg.norm(B, axis=0)
    bwd = abs(R).max(0) / (anorm * abs(X).max(0) + abs(B).max(0))
    return X, relres, bwd


def big_solutions():
    for name, gen, seeds, full in (
            ("c2_mesh3d_60.npz", lambda: O.mesh3d(60, 60, 60), list(range(1, 9)), True),
            ("c3_varcoef_80.npz", lambda: O.mesh3d_varcoef(80, 80, 80, 2024), list(range(100, 104)),
             False)):
        n, cp, ri, v = gen()
        X, relres, bwd = oracle_solution(n, cp, ri, v, seeds)
        idx = sample_rows(n)
        extra = {"x_first": X[:, 0]} if full else {}
        np.savez_compressed(os.path.join(HERE, name), seeds=np.array(seeds), n=n,
                            sketch=sketch_matrix_apply(X), sample_idx=idx, sample_x=X[idx],
                            xnorm=np.linalg.norm(X, axis=0), relres=relres, bwd=bwd, **extra)
        print(name, "relres", relres, "bwd", bwd, flush=True)
    sides = np.random.default_rng(7).integers(16, 129, 512)
    pick = sorted(set(np.argsort(sides, kind="stable")[-3:].tolist()) | {0, 1, 2, 255, 511})
    xs, ns = [], []
    for q in pick:
        n, cp, ri, v = O.mesh2d_shifted(int(sides[q]), 10_000 + q)
        perm, inv = O.genmmd(n, cp, ri)
        pcp, pri, pv = O.permute(n, cp, ri, v, inv)
        S = O.symbolic(n, pcp, pri)
        x = O.solve(S, O.factor(S, pcp, pri, pv), O.uniform(n, 500 + q))
        xs.append(x)
        ns.append(n)
    np.savez_compressed(os.path.join(HERE, "c4_batch.npz"), systems=np.array(pick),
                        sides=sides[pick], n=np.array(ns), x=np.concatenate(xs))
    print("c4 systems", pick, flush=True)


def main():
    if sys.argv[1:] == ["big-solutions"]:
        big_solutions()
        return
    # C1
    n, cp, ri, v = O.mesh2d(100)
    perm, inv, (pcp, pri, pv), parent, S = analyze(n, cp, ri, v)
    b = splitmix_uniform(n, 1)  # seed 1, stored with the fixture
    panels = O.factor(S, pcp, pri, pv)
    x = O.solve(S, panels, b)
    rel = np.linalg.norm(O.residual(n, pcp, pri, pv, x, b)) / np.linalg.norm(b)
    np.savez_compressed(os.path.join(HERE, "c1_mesh2d_100.npz"), perm=perm, invperm=inv,
                        parent=parent, b=b, x=x, relres=rel, n_sn=S["n_sn"],
                        fund_flops=S["fund_flops"], relax_flops=S["relax_flops"],
                        fund_nnz=S["fund_nnz"], relax_nnz=S["relax_nnz"],
                        **{k: S[k] for k in KEYS})
    for kind, sizes in (("mesh2d", range(16, 129)), ("mesh3d", range(3, 41))):
        rows = []
        for s in sizes:
            n, cp, ri, v = O.mesh2d(s) if kind == "mesh2d" else O.mesh3d(s, s, s)
            perm, inv, _, parent, S = analyze(n, cp, ri, v)
            rows.append((s, digest(perm), digest(parent), digest(*[S[k] for k in KEYS]), S["n_sn"],
                         S["fund_flops"], S["relax_flops"]))
        np.savez_compressed(os.path.join(HERE, f"sweep_{kind}.npz"),
                            side=np.array([r[0] for r in rows]),
                            perm_sha=np.array([r[1] for r in rows]),
                            parent_sha=np.array([r[2] for r in rows]),
                            sn_sha=np.array([r[3] for r in rows]),
                            n_sn=np.array([r[4] for r in rows]),
                            fund_flops=np.array([r[5] for r in rows]),
                            relax_flops=np.array([r[6] for r in rows]))
        print(kind, "done", flush=True)
    big = {}
    for s in (60, 80):
        n, cp, ri, v = O.mesh3d(s, s, s)
        perm, inv, _, parent, S = analyze(n, cp, ri, v)
        big[f"mesh3d_{s}"] = dict(perm_sha=digest(perm), parent_sha=digest(parent),
                                  sn_sha=digest(*[S[k] for k in KEYS]), n_sn=int(S["n_sn"]),
                                  depth=int(S["depth"]), fund_flops=S["fund_flops"],
                                  relax_flops=S["relax_flops"], fund_nnz=S["fund_nnz"],
                                  relax_nnz=S["relax_nnz"])
        print(s, big[f"mesh3d_{s}"], flush=True)
    with open(os.path.join(HERE, "big.json"), "w") as f:
        json.dump(big, f, indent=1)


if __name__ == "__main__":
    main()
