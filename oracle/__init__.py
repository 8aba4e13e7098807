"""CPU FP64 oracle for the 2+4 p-spin Monte Carlo path -- TEST INFRASTRUCTURE.

This package is the checker of the HIP product path.  Only ``tests/``,
``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline`` leg may import
it.  The product package ``cuda_pspin_amd`` never imports it.

It wraps ``oracle/_build/liboracle.so`` (built from ``pspin_oracle.cpp`` by
``oracle/Makefile``; each C function there cites the reference file:line it
restates) and, when built, ``oracle/_ref/libref_enum.so`` compiled from the
reference's own ``src/test_enum.cpp``.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")
_REF_PATH = os.path.join(_HERE, "_ref", "libref_enum.so")

_lib = None
_ref = None

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_u8p = np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_i64p = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
_u32p = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")
_u64p = np.ctypeslib.ndpointer(dtype=np.uint64, flags="C_CONTIGUOUS")
_f32p = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")


def build(force: bool = False) -> None:
    """Compile liboracle.so (and the reference enum shim when the reference is mounted)."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.check_call(["make", "-s", "-C", _HERE, "all"])
    if os.path.isdir("/root/reference/src") and (force or not os.path.exists(_REF_PATH)):
        subprocess.check_call(["make", "-s", "-C", _HERE, "ref"])


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        sig = {
            "or_splitmix64": (C.c_uint64, [C.POINTER(C.c_uint64)]),
            "or_replica_seeds": (None, [C.c_uint64, C.c_int, _u64p]),
            "or_n_quartets": (C.c_longlong, [C.c_int]),
            "or_n_pairs": (C.c_longlong, [C.c_int]),
            "or_encode": (C.c_longlong, [C.c_int] * 4),
            "or_decode": (None, [C.c_longlong, _i32p]),
            "or_omega": (None, [C.c_int, C.c_int, C.c_uint64, _dp]),
            "or_g2_keep": (None, [C.c_int, _dp, C.c_double, _u8p]),
            "or_g4_mask": (None, [C.c_int, C.c_int, _dp, C.c_double, _u8p]),
            "or_popcount_mask": (C.c_longlong, [_u8p, C.c_longlong]),
            "or_three_type": (C.c_longlong, [C.c_int, C.c_int, C.c_int, _i32p]),
            "or_energy": (None, [C.c_int, _dp, _dp, C.c_int, _dp, _u8p, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
            "or_energy_h4_sparse": (C.c_double, [_dp, C.c_int, _i32p, _dp]),
            "or_energy_h2": (C.c_double, [C.c_int, _dp, _dp]),
            "or_delta_e": (None, [C.c_int, _dp, _dp, C.c_int, _dp, _u8p, C.c_int, C.c_int, _dp, _dp,
                                  C.POINTER(C.c_double), C.POINTER(C.c_double)]),
            "or_delta_e_sparse": (None, [C.c_int, _dp, _dp, C.c_int, C.c_int, _i32p, _dp, C.c_int, C.c_int,
                                         _dp, _dp, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
            "or_pair_pick": (None, [C.c_int, C.c_double, C.c_double, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
            "or_propose_sphere": (None, [_dp, _dp, C.c_double, C.c_double, C.c_double, _dp, _dp]),
            "or_propose_cube": (None, [_dp, _dp, C.c_double, C.c_double, C.c_double, _dp, _dp]),
            "or_metropolis": (C.c_int, [C.c_double, C.c_double, C.c_double]),
            "or_temperatures": (None, [C.c_int, C.c_double, C.c_double, C.c_int, _dp, _dp]),
            "or_srand48": (None, [C.c_long]),
            "or_drand48": (C.c_double, []),
            "or_pt_exchange": (None, [C.c_int, C.c_int, _dp, _dp, _i32p, _i64p, _i64p, _dp]),
            "or_partial_shuffle": (None, [_i64p, C.c_longlong, C.c_int, C.c_uint64]),
            "or_sparse_select": (C.c_int, [C.c_int, _u8p, _dp, C.c_int, C.c_uint64, C.c_double, C.c_double,
                                           _i32p, _dp, C.POINTER(C.c_longlong)]),
            "or_philox4x32_10": (None, [_u32p, _u32p, _u32p]),
            "or_u32_to_unit": (C.c_double, [C.c_uint32]),
            "or_u64_to_unit": (C.c_double, [C.c_uint32, C.c_uint32]),
            "or_normal": (C.c_double, [C.c_uint64, C.c_uint32, C.c_uint64]),
            "or_step_uniforms": (None, [C.c_uint64, C.c_uint64, C.c_int, _dp]),
            "or_step_uniforms_grouped": (None, [C.c_uint64, C.c_uint64, C.c_uint64, C.c_int, _dp]),
            "or_pair_group_key": (C.c_uint64, [C.c_uint64, C.c_int]),
            "or_cpu_baseline_sweeps": (C.c_double, [C.c_int, C.c_int, _dp, _dp, C.c_int, _dp, _u8p, _dp, _u64p,
                                                    C.c_uint64, C.c_int, _dp, _i64p, C.c_int]),
            "or_max_threads": (C.c_int, []),
            "or_cpu_pt": (None, [C.c_int, C.c_int, C.c_int, _dp, _dp, C.c_int, _dp, _u8p, _dp, _u64p, C.c_int, C.c_long,
                                 C.c_int, _dp, _dp, _dp, _i64p, _i64p, _i64p, _i64p, C.POINTER(C.c_int), C.c_int,
                                 _dp]),
            "or_cpu_pt_lists": (None, [C.c_int, C.c_int, C.c_int, _dp, _dp, C.c_int, _dp, _u8p, _dp, _u64p, C.c_int,
                                       C.c_long, C.c_int, _dp, _dp, _dp, _i64p, _i64p, _i64p, _i64p,
                                       C.POINTER(C.c_int), C.c_int, _dp]),
            "or_cpu_pt_sparse": (None, [C.c_int, C.c_int, C.c_int, _dp, _dp, C.c_int, C.c_int, _i32p, _dp, _dp,
                                        _u64p, C.c_int, C.c_long, C.c_int, _dp, _dp, _dp, _i64p, _i64p, _i64p,
                                        _i64p, C.POINTER(C.c_int), C.c_int, _dp]),
            "or_log_jn": (C.c_double, [C.c_int, C.c_double]),
            "or_lp_log_density": (C.c_double, [C.c_int, _dp, _dp, _dp]),
            "or_lp_gauge": (C.c_int, [C.c_int, _dp, _dp, _dp]),
            "or_lp_draws": (None, [C.c_uint64, C.c_uint64, C.c_int, _dp, C.POINTER(C.c_double)]),
            "or_lp_propose": (C.c_double, [C.c_int, _dp, _f32p, C.c_double, C.c_double, C.c_double, C.c_double,
                                           C.c_double, C.c_uint64, C.c_uint64, _dp]),
            "or_lp_accept": (C.c_int, [C.c_int, C.c_double, _dp, _dp, C.c_double, C.c_double, _f32p, C.c_double,
                                       C.c_double, C.c_double, C.c_double, C.c_double, C.c_double, C.c_uint64,
                                       C.c_uint64, _dp]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def ref_lib():
    """The reference's test_enum.cpp compiled in oracle/_ref (None when unavailable)."""
    global _ref
    if _ref is None and os.path.exists(_REF_PATH):
        R = C.CDLL(_REF_PATH)
        R.ref_three_type.restype = C.c_longlong
        R.ref_three_type.argtypes = [C.c_int, C.c_int, C.c_int, _i32p]
        R.ref_test_enum_run.restype = C.c_int
        R.ref_test_enum_run.argtypes = []
        _ref = R
    return _ref


# ----------------------------------------------------------------------------
# numpy helpers
# ----------------------------------------------------------------------------
def n_quartets(N: int) -> int:
    return N * (N - 1) * (N - 2) * (N - 3) // 24


def n_inv(N: int) -> int:
    """Quartets touched by one pair move: 2*C(N-2,3) + C(N-2,2) (src/mc.cu:115-117)."""
    return 2 * (N - 2) * (N - 3) * (N - 4) // 6 + (N - 2) * (N - 3) // 2


def splitmix64_seq(seed: int, n: int) -> np.ndarray:
    st = C.c_uint64(seed)
    return np.array([lib().or_splitmix64(C.byref(st)) for _ in range(n)], dtype=np.uint64)


def replica_seeds(seed: int, nrep: int) -> np.ndarray:
    out = np.zeros(nrep, dtype=np.uint64)
    lib().or_replica_seeds(seed, nrep, out)
    return out


def omega(N: int, fmc_mode: int, seed: int) -> np.ndarray:
    w = np.zeros(N, dtype=np.float64)
    lib().or_omega(N, fmc_mode, seed, w)
    return w


def g2_keep(N: int, w: np.ndarray, gamma: float) -> np.ndarray:
    k = np.zeros(N * N, dtype=np.uint8)
    lib().or_g2_keep(N, np.ascontiguousarray(w, dtype=np.float64), gamma, k)
    return k.reshape(N, N)


def g4_mask(N: int, fmc_mode: int, w, gamma: float) -> np.ndarray:
    m = np.zeros(n_quartets(N), dtype=np.uint8)
    w = np.zeros(N) if w is None else np.ascontiguousarray(w, dtype=np.float64)
    lib().or_g4_mask(N, fmc_mode, w, gamma, m)
    return m


def decode(q: int):
    out = np.zeros(4, dtype=np.int32)
    lib().or_decode(q, out)
    return tuple(int(x) for x in out)


def encode(i, j, k, l) -> int:
    return int(lib().or_encode(i, j, k, l))


def three_type(N: int, i0: int, j0: int) -> np.ndarray:
    out = np.zeros(4 * (n_inv(N) + 1), dtype=np.int32)
    n = lib().or_three_type(N, i0, j0, out)
    return out[: 4 * n].reshape(n, 4)


def ref_three_type(N: int, i0: int, j0: int):
    R = ref_lib()
    if R is None:
        return None
    out = np.zeros(4 * (n_inv(N) + 1), dtype=np.int32)
    n = R.ref_three_type(N, i0, j0, out)
    return out[: 4 * n].reshape(n, 4)


def _spins(s) -> np.ndarray:
    s = np.asarray(s)
    if np.iscomplexobj(s):
        s = np.stack([s.real, s.imag], axis=-1)
    return np.ascontiguousarray(s, dtype=np.float64).reshape(-1)


def energy(N, spins, g2, h2_active, g4, mask):
    e2, e4 = C.c_double(), C.c_double()
    lib().or_energy(N, _spins(spins), np.ascontiguousarray(g2, dtype=np.float64).reshape(-1), int(h2_active),
                    np.ascontiguousarray(g4, dtype=np.float64), np.ascontiguousarray(mask, dtype=np.uint8),
                    C.byref(e2), C.byref(e4))
    return e2.value, e4.value


def energy_sparse(N, spins, g2, h2_active, quartets, gq):
    s = _spins(spins)
    e2 = lib().or_energy_h2(N, s, np.ascontiguousarray(g2, dtype=np.float64).reshape(-1)) if h2_active else 0.0
    q = np.ascontiguousarray(quartets, dtype=np.int32).reshape(-1)
    e4 = lib().or_energy_h4_sparse(s, len(gq), q, np.ascontiguousarray(gq, dtype=np.float64))
    return e2, e4


def delta_e(N, spins, g2, h2_active, g4, mask, i0, j0, ai_new, aj_new):
    dE, sc = C.c_double(), C.c_double()
    lib().or_delta_e(N, _spins(spins), np.ascontiguousarray(g2, dtype=np.float64).reshape(-1), int(h2_active),
                     np.ascontiguousarray(g4, dtype=np.float64), np.ascontiguousarray(mask, dtype=np.uint8),
                     int(i0), int(j0), _spins(ai_new), _spins(aj_new), C.byref(dE), C.byref(sc))
    return dE.value, sc.value


def delta_e_sparse(N, spins, g2, h2_active, quartets, gq, i0, j0, ai_new, aj_new):
    dE, sc = C.c_double(), C.c_double()
    q = np.ascontiguousarray(quartets, dtype=np.int32).reshape(-1)
    lib().or_delta_e_sparse(N, _spins(spins), np.ascontiguousarray(g2, dtype=np.float64).reshape(-1),
                            int(h2_active), len(gq), q, np.ascontiguousarray(gq, dtype=np.float64), int(i0), int(j0),
                            _spins(ai_new), _spins(aj_new), C.byref(dE), C.byref(sc))
    return dE.value, sc.value


def pair_pick(N, u1, u2):
    i0, j0 = C.c_int(), C.c_int()
    lib().or_pair_pick(N, u1, u2, C.byref(i0), C.byref(j0))
    return i0.value, j0.value


def propose_sphere(ai, aj, ua, u1, u2):
    oi, oj = np.zeros(2), np.zeros(2)
    lib().or_propose_sphere(_spins(ai), _spins(aj), ua, u1, u2, oi, oj)
    return oi, oj


def propose_cube(ai, aj, u, u1, u2):
    oi, oj = np.zeros(2), np.zeros(2)
    lib().or_propose_cube(_spins(ai), _spins(aj), u, u1, u2, oi, oj)
    return oi, oj


def temperatures(NT, Tmin, Tmax, log_temp):
    T, b = np.zeros(NT), np.zeros(NT)
    lib().or_temperatures(NT, Tmin, Tmax, int(log_temp), T, b)
    return T, b


def pt_exchange(NT, nrep, beta_sched, E, perm, acc, prop):
    """Sequential hot->cold exchange using glibc drand48 (seed it with srand48 first)."""
    betas = np.zeros(NT * nrep)
    lib().or_pt_exchange(NT, nrep, np.ascontiguousarray(beta_sched, dtype=np.float64),
                         np.ascontiguousarray(E, dtype=np.float64), perm, acc, prop, betas)
    return betas


def srand48(s: int) -> None:
    lib().or_srand48(s)


def drand48() -> float:
    return lib().or_drand48()


def sparse_select(N, mask, g4_unit, n_req, seed, J4, J4_0):
    q = np.zeros(5 * max(n_req, 1), dtype=np.int32)
    g = np.zeros(max(n_req, 1))
    nnz = C.c_longlong()
    n = lib().or_sparse_select(N, np.ascontiguousarray(mask, dtype=np.uint8),
                               np.ascontiguousarray(g4_unit, dtype=np.float64), n_req, seed, J4, J4_0, q, g,
                               C.byref(nnz))
    return q[: 5 * n].reshape(n, 5), g[:n], nnz.value


def philox(ctr, key) -> np.ndarray:
    out = np.zeros(4, dtype=np.uint32)
    lib().or_philox4x32_10(np.asarray(ctr, dtype=np.uint32), np.asarray(key, dtype=np.uint32), out)
    return out


def normal(seed: int, tag: int, k: int) -> float:
    return lib().or_normal(seed, tag, k)


def step_uniforms_grouped(rseed: int, gkey: int, sweep: int, step: int) -> np.ndarray:
    u = np.zeros(6)
    lib().or_step_uniforms_grouped(rseed, gkey, sweep, step, u)
    return u


def pair_group_key(seed: int, g: int) -> int:
    return int(lib().or_pair_group_key(seed, g))


def step_uniforms(rseed: int, sweep: int, step: int) -> np.ndarray:
    u = np.zeros(6)
    lib().or_step_uniforms(rseed, sweep, step, u)
    return u


def cpu_pt(N, NT, nr, spins, g2, h2_active, g4, mask, beta_sched, rseeds, n_sweeps, seed48, measure_from, qbins=0,
           lists=False):
    """The reference's PT loop on the CPU (FP64 sweeps, host exchange); returns per-slot sums
    (e, e2 of E/N over the slot's replicas; q2 and, with qbins > 0, a histogram of |q| over
    [0, 1] for the slot's first two replicas).  lists=True visits each site's active quartets
    from a table built once instead of the 3-type enumeration (same terms and per-term
    arithmetic, another summation order)."""
    sp = _spins(spins).copy()
    out = dict(e=np.zeros(NT), e2=np.zeros(NT), q2=np.zeros(NT), acc=np.zeros(NT, np.int64),
               prop=np.zeros(NT, np.int64), ex_acc=np.zeros(NT - 1, np.int64), ex_prop=np.zeros(NT - 1, np.int64),
               qh=np.zeros((NT, max(qbins, 1))))
    n = C.c_int()
    fn = lib().or_cpu_pt_lists if lists else lib().or_cpu_pt
    fn(N, NT, nr, sp, np.ascontiguousarray(g2, dtype=np.float64).reshape(-1), int(h2_active),
                    np.ascontiguousarray(g4, dtype=np.float64), np.ascontiguousarray(mask, dtype=np.uint8),
                    np.ascontiguousarray(beta_sched, dtype=np.float64), np.ascontiguousarray(rseeds, dtype=np.uint64),
                    n_sweeps, seed48, measure_from, out["e"], out["e2"], out["q2"], out["acc"], out["prop"],
                    out["ex_acc"], out["ex_prop"], C.byref(n), int(qbins), out["qh"])
    out["n"] = n.value
    out["spins"] = sp
    return out


def cpu_pt_sparse(N, NT, nr, spins, g2, h2_active, quartets, gq, beta_sched, rseeds, n_sweeps, seed48, measure_from,
                  qbins=0):
    """cpu_pt for a sparse (smoothed-cube) system: quartets [nq][5] (i, j, k, l, ch), couplings gq [nq]
    (src/mc_sparse.cu:63-253 chains in the reference's PT loop)."""
    sp = _spins(spins).copy()
    q = np.ascontiguousarray(quartets, dtype=np.int32).reshape(-1, 5)
    out = dict(e=np.zeros(NT), e2=np.zeros(NT), q2=np.zeros(NT), acc=np.zeros(NT, np.int64),
               prop=np.zeros(NT, np.int64), ex_acc=np.zeros(NT - 1, np.int64), ex_prop=np.zeros(NT - 1, np.int64),
               qh=np.zeros((NT, max(qbins, 1))))
    n = C.c_int()
    lib().or_cpu_pt_sparse(N, NT, nr, sp, np.ascontiguousarray(g2, dtype=np.float64).reshape(-1), int(h2_active),
                           len(q), q.reshape(-1), np.ascontiguousarray(gq, dtype=np.float64),
                           np.ascontiguousarray(beta_sched, dtype=np.float64),
                           np.ascontiguousarray(rseeds, dtype=np.uint64), n_sweeps, seed48, measure_from, out["e"],
                           out["e2"], out["q2"], out["acc"], out["prop"], out["ex_acc"], out["ex_prop"], C.byref(n),
                           int(qbins), out["qh"])
    out["n"] = n.value
    out["spins"] = sp
    return out


# ----------------------------------------------------------------------------
# Learned global move (projected-Gaussian proposal, exact MH)
# ----------------------------------------------------------------------------
def log_jn(n: int, alpha: float) -> float:
    """log of int_0^inf s^n exp(-s^2/2 + alpha s) ds."""
    return lib().or_log_jn(int(n), float(alpha))


def lp_log_density(y, mu, ls) -> float:
    """log density of unit vector y under the radially projected N(mu, diag exp(2 ls))."""
    y = np.ascontiguousarray(y, dtype=np.float64)
    return lib().or_lp_log_density(len(y), y, np.ascontiguousarray(mu, dtype=np.float64),
                                   np.ascontiguousarray(ls, dtype=np.float64))


def lp_gauge(N: int, a):
    """(k*, c, u): gauge phase c = conj(a_k*)/|a_k*| and u = a c as [N][2]."""
    c = np.zeros(2)
    u = np.zeros(2 * N)
    k = lib().or_lp_gauge(N, _spins(a), c, u)
    return k, c, u


def lp_draws(seed: int, move: int, N: int):
    xi = np.zeros(2 * N)
    u = C.c_double()
    lib().or_lp_draws(seed, move, N, xi, C.byref(u))
    return xi, u.value


def lp_propose(N, a, O_f, ls0, ls_min, ls_max, log_scale, seed, move, mean_scale=1.0):
    prop = np.zeros(2 * N)
    lq = lib().or_lp_propose(N, _spins(a), np.ascontiguousarray(O_f, dtype=np.float32).reshape(-1), ls0, ls_min,
                             ls_max, log_scale, mean_scale, seed, move, prop)
    return prop, lq


def lp_accept(N, beta, a, prop, E, E_new, O_r, ls0, ls_min, ls_max, log_scale, lq_f, seed, move, mean_scale=1.0):
    out = np.zeros(8)
    lib().or_lp_accept(N, beta, _spins(a), _spins(prop), E, E_new,
                       np.ascontiguousarray(O_r, dtype=np.float32).reshape(-1), ls0, ls_min, ls_max, log_scale,
                       mean_scale, lq_f, seed, move, out)
    return out
