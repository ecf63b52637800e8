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


def main():
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
