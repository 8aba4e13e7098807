"""Writes tests/golden/golden.npz: input/output vectors of the bit-exact parts
of the path, produced by the CPU oracle (oracle/, FP64, no -ffast-math).

The reference commits no fixtures of its own (its .gitignore drops every
output; its tests recompute their expectations), so these vectors freeze the
oracle's answers once it has been pinned by the checks in
tests/test_oracle_*.py.  tests/test_golden.py then checks that the oracle
still reproduces them, and tests/test_gpu_golden.py checks the HIP path
against them directly.

Contents (key prefix: what it pins, reference lines):
  seeds_*     per-replica splitmix64 seeds            src/mc.cu:418-423
  stepu_*     Philox step uniforms (seed, sweep, step) this build's counter layout
  omega_*     FMC frequencies, uniform / comb          src/mc.cu:449-483
  mask_*      g4 FMC masks (colex order)              src/disorder.cu:130-189
  g2keep_*    g2 survivor pattern                      src/disorder.cu:199-216
  sparse_*    sparse quartet selection (i,j,k,l,ch)    src/mc_sparse.cu:340-349,477-549
  exch_*      PT exchange: energies in, permutations,
              counters and the next drand48 draw out   src/parallel_tempering.cu:316-352
  colex_*     quartet encode of sampled (i<j<k<l)      src/disorder.cu:159-171

Run from the repository root: python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden.npz")

# (N, fmc_mode, gamma, seed) of the mask / omega / g2 cases
MASK_CASES = [(12, 2, 0.2, 42), (16, 2, 0.05, 77777), (12, 1, 1.0, 42), (16, 1, 0.0, 42), (32, 2, 0.05, 42),
              (64, 2, 0.05, 42)]
# (N, fmc_mode, gamma, seed, J4, J4_0) of the sparse selection cases
SPARSE_CASES = [(16, 2, 0.05, 42, 1.0, 0.0), (24, 2, 0.05, 42, 1.0, 0.0), (12, 0, 0.0, 42, 1.0, 0.0),
                (20, 1, 1.0, 42, 1.0, 0.0)]
# (NT, nrep, rounds, Tmin, Tmax, seed48) of the exchange traces
EXCH_CASES = [(6, 1, 20, 0.2, 1.6, 5042), (8, 3, 20, 0.2, 1.6, 47), (32, 8, 8, 0.2, 1.6, 5042)]


def case_key(prefix, *parts):
    return prefix + "_" + "_".join(str(p).replace(".", "p") for p in parts)


def generate():
    import oracle as O

    O.build()
    out = {}
    for seed, n in ((42, 8), (77777, 5), (2024, 3)):
        out[case_key("seeds", seed, n)] = O.replica_seeds(seed, n)
    su = []
    for rs, sw, k in ((0x0123456789ABCDEF, 0, 0), (0x0123456789ABCDEF, 7, 31), (42, 1 << 33, 5)):
        su.append(np.concatenate([[float(rs), float(sw), float(k)], O.step_uniforms(rs, sw, k)]))
    out["stepu_rows"] = np.array(su)
    for N, fmc, gamma, seed in MASK_CASES:
        w = O.omega(N, fmc, seed)
        out[case_key("omega", N, fmc, seed)] = w
        out[case_key("mask", N, fmc, gamma, seed)] = O.g4_mask(N, fmc, w, gamma)
        if fmc == 2:
            out[case_key("g2keep", N, gamma, seed)] = O.g2_keep(N, w, gamma)
    for N, fmc, gamma, seed, J4, J40 in SPARSE_CASES:
        w = O.omega(N, fmc, seed) if fmc else np.zeros(N)
        mask = O.g4_mask(N, fmc, w, gamma)
        g_unit = np.array([O.normal(seed + 2000 + 12345, 4, q) for q in range(len(mask))])
        q, g, nnz = O.sparse_select(N, mask, g_unit, N, seed, J4, J40)
        out[case_key("sparse_q", N, fmc, seed)] = q
        out[case_key("sparse_g", N, fmc, seed)] = g
        out[case_key("sparse_nnz", N, fmc, seed)] = np.array([nnz], dtype=np.int64)
    for NT, nrep, rounds, Tmin, Tmax, seed48 in EXCH_CASES:
        T, beta = O.temperatures(NT, Tmin, Tmax, 1)
        rng = np.random.default_rng(1000 * NT + nrep)
        E = rng.normal(size=(rounds, NT * nrep)) * 3.0 + np.linspace(0, -5, NT * nrep)
        O.srand48(seed48)
        perm = np.arange(NT * nrep, dtype=np.int32)
        acc = np.zeros(NT - 1, dtype=np.int64)
        prop = np.zeros(NT - 1, dtype=np.int64)
        perms, betas = [], []
        for r in range(rounds):
            betas.append(O.pt_exchange(NT, nrep, beta, E[r], perm, acc, prop))
            perms.append(perm.copy())
        k = case_key("exch", NT, nrep, seed48)
        out[k + "_beta"] = beta
        out[k + "_E"] = E
        out[k + "_perm"] = np.array(perms)
        out[k + "_betas"] = np.array(betas)
        out[k + "_acc"] = acc
        out[k + "_prop"] = prop
        out[k + "_next"] = np.array([O.drand48()])
    rng = np.random.default_rng(7)
    quads = np.sort(np.array([rng.choice(128, 4, replace=False) for _ in range(64)]), axis=1)
    out["colex_quads"] = quads.astype(np.int32)
    out["colex_q"] = np.array([O.encode(*map(int, x)) for x in quads], dtype=np.int64)
    return out


if __name__ == "__main__":
    data = generate()
    np.savez_compressed(OUT, **data)
    print(f"wrote {OUT}: {len(data)} arrays, {os.path.getsize(OUT)} bytes")
