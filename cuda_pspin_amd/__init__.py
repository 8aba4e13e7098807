"""cuda_pspin_amd -- MI355X-native 2+4 p-spin Monte Carlo (gfx950 HIP kernels).

Python view of the C ABI in ``include/pspin.h`` (``libpspin.so`` in this
directory).  It mirrors the reference's ``MCState`` / ``MCStateSparse`` API
(include/mc.h:33-56, include/mc_sparse.h:42-61) plus the parallel-tempering
loop that the reference keeps in its driver (src/parallel_tempering.cu:273-352).

The shared library is required: importing this package on a machine where
``libpspin.so`` was not built raises ImportError.  There is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

__all__ = ["SimConfig", "PSpin", "PSpinBatch", "load_library", "build", "call", "DENSE", "SPARSE", "STEP_DTYPE",
           "PSpinError"]

_HERE = os.path.dirname(os.path.abspath(__file__))
# CUDA_PSPIN_AMD_LIB selects another build of the same library (e.g. a
# differently tuned one); the default is the in-tree build.
LIB_PATH = os.environ.get("CUDA_PSPIN_AMD_LIB", os.path.join(_HERE, "libpspin.so"))

DENSE = 0
SPARSE = 1
LAUNCH_AUTO, LAUNCH_GRAPHS, LAUNCH_DIRECT = 0, 1, 2
KERNEL_AUTO, KERNEL_LISTS, KERNEL_PAIRS, KERNEL_GROUPS = 0, 1, 2, 3

STEP_DTYPE = np.dtype([("i0", "<i4"), ("j0", "<i4"), ("acc", "<i4"), ("dE", "<f4"),
                       ("ai", "<f8", (2,)), ("aj", "<f8", (2,))])
assert STEP_DTYPE.itemsize == 48


class PSpinError(RuntimeError):
    pass


class SimConfig(C.Structure):
    """Same fields as the reference's SimConfig (include/config.h:9-25)."""

    _fields_ = [("save_freq", C.c_int), ("verbose", C.c_int), ("N", C.c_int), ("nrep", C.c_int),
                ("T", C.c_double), ("J", C.c_double), ("J0", C.c_double), ("alpha", C.c_double),
                ("alpha0", C.c_double), ("mc_iterations", C.c_int), ("seed", C.c_uint64), ("label", C.c_int),
                ("dev", C.c_int), ("fmc_mode", C.c_int), ("gamma", C.c_double)]

    @classmethod
    def default(cls, **kw) -> "SimConfig":
        c = cls()
        load_library().pspin_default_config(C.byref(c))
        for k, v in kw.items():
            if not hasattr(c, k):
                raise AttributeError(k)
            setattr(c, k, v)
        return c

    @classmethod
    def parse_args(cls, argv) -> "SimConfig":
        """parse_args (src/config.cu:6-70); raises PSpinError where the reference exits 1."""
        c = cls()
        args = [b"prog"] + [a.encode() for a in argv]
        arr = (C.c_char_p * len(args))(*args)
        rc = load_library().pspin_parse_args(len(args), arr, C.byref(c))
        _check(rc)
        return c


class Info(C.Structure):
    _fields_ = [("N", C.c_int), ("nrep", C.c_int), ("variant", C.c_int), ("h2_active", C.c_int),
                ("n_pairs_active", C.c_longlong), ("n_quart_active", C.c_longlong), ("n_quartets", C.c_int),
                ("n_site_entries", C.c_longlong), ("coupling_format", C.c_int), ("kernel", C.c_int)]


_lib = None
_EXPORTS = None


def build(verbose: bool = False) -> None:
    """Compile libpspin.so for gfx950 with hipcc (make in csrc/)."""
    cmd = ["make", "-C", os.path.join(_HERE, "csrc"), "-j8"]
    if not verbose:
        cmd.insert(1, "-s")
    subprocess.check_call(cmd)


def load_library() -> C.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"libpspin.so not built at {LIB_PATH}; run __graft_entry__.build() "
                          "(cuda_pspin_amd has no CPU fallback)")
    L = C.CDLL(LIB_PATH)
    P = C.c_void_p
    dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
    i64p = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
    i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
    u8p = np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS")
    u64p = np.ctypeslib.ndpointer(dtype=np.uint64, flags="C_CONTIGUOUS")
    f32p = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
    sigs = {
        "pspin_last_error": (C.c_char_p, []),
        "pspin_default_config": (None, [P]),
        "pspin_parse_args": (C.c_int, [C.c_int, C.POINTER(C.c_char_p), P]),
        "pspin_create": (C.c_int, [P, C.c_int, C.POINTER(P)]),
        "pspin_destroy": (C.c_int, [P]),
        "pspin_get_info": (C.c_int, [P, P]),
        "pspin_set_betas": (C.c_int, [P, dp]),
        "pspin_set_uniform_beta": (C.c_int, [P, C.c_double]),
        "pspin_sweep": (C.c_int, [P, C.c_int]),
        "pspin_get_results": (C.c_int, [P, dp, i64p, i64p]),
        "pspin_reset_counters": (C.c_int, [P]),
        "pspin_get_spins": (C.c_int, [P, dp]),
        "pspin_set_spins": (C.c_int, [P, dp]),
        "pspin_split_energies": (C.c_int, [P, dp, dp]),
        "pspin_resync_energy": (C.c_int, [P]),
        "pspin_set_energies": (C.c_int, [P, dp]),
        "pspin_get_omega": (C.c_int, [P, dp]),
        "pspin_get_g2": (C.c_int, [P, dp]),
        "pspin_get_g4": (C.c_int, [P, dp, u8p]),
        "pspin_get_sparse_quartets": (C.c_int, [P, i32p, dp]),
        # free functions on device buffers (pointers as integers)
        "pspin_generate_g2": (C.c_int, [P, C.c_int, C.c_double, C.c_ulonglong]),
        "pspin_generate_g4": (C.c_int, [P, C.c_int, C.c_double, C.c_ulonglong]),
        "pspin_rescale_g2": (C.c_int, [P, C.c_int, C.c_double, C.c_longlong]),
        "pspin_rescale_g4": (C.c_int, [P, C.c_int, C.c_double, C.c_longlong]),
        "pspin_shift_mean_couplings": (C.c_int, [P, C.c_longlong, C.c_double]),
        "pspin_apply_fmc_g2": (C.c_int, [P, C.c_int, P, C.c_double]),
        "pspin_init_g4_mask": (C.c_int, [P, C.c_int]),
        "pspin_apply_fmc_g4": (C.c_int, [P, C.c_int, P, C.c_double]),
        "pspin_count_active_g4": (C.c_int, [P, C.c_int, C.POINTER(C.c_longlong)]),
        "pspin_compute_energy": (C.c_int, [P, P, P, P, C.c_int, C.POINTER(C.c_double)]),
        "pspin_compute_energy_noalloc": (C.c_int, [P, P, P, P, C.c_int, C.POINTER(C.c_double)]),
        "pspin_init_spins_uniform": (C.c_int, [P, C.c_int, C.c_ulonglong]),
        "pspin_init_spins_cube": (C.c_int, [P, C.c_int, C.c_ulonglong]),
        "pspin_decode_quartets": (C.c_int, [P, P, C.c_longlong]),
        "pspin_pt_init": (C.c_int, [P, C.c_int, C.c_int, dp, C.c_long]),
        "pspin_pt_run": (C.c_int, [P, C.c_int, C.c_int]),
        "pspin_pt_exchange": (C.c_int, [P]),
        "pspin_pt_get_perm": (C.c_int, [P, i32p]),
        "pspin_pt_get_exchange_stats": (C.c_int, [P, i64p, i64p, i64p, i64p]),
        "pspin_pt_reset_window": (C.c_int, [P]),
        "pspin_pt_get_rng48": (C.c_int, [P, C.POINTER(C.c_uint64)]),
        "pspin_pt_enable_series": (C.c_int, [P, C.c_int]),
        "pspin_pt_get_series": (C.c_int, [P, P, C.POINTER(C.c_int)]),
        "pspin_enable_step_log": (C.c_int, [P, C.c_int]),
        "pspin_get_step_log": (C.c_int, [P, P, C.POINTER(C.c_int)]),
        "pspin_get_rng_info": (C.c_int, [P, u64p, C.POINTER(C.c_uint64)]),
        "pspin_set_profiling": (C.c_int, [P, C.c_int]),
        "pspin_set_launch_mode": (C.c_int, [P, C.c_int]),
        "pspin_set_wait_limit": (C.c_int, [P, C.c_uint]),
        "pspin_get_profile_launches": (C.c_int, [P, C.POINTER(C.c_longlong)]),
        "pspin_get_profile": (C.c_int, [P, C.POINTER(C.c_double), C.POINTER(C.c_longlong),
                                        C.POINTER(C.c_double), C.POINTER(C.c_longlong)]),
        "pspin_set_kernel": (C.c_int, [P, C.c_int, C.c_int]),
        "pspin_set_pair_schedule": (C.c_int, [P, C.c_int]),
        "pspin_set_pair_tiles": (C.c_int, [P, C.c_int, C.c_int]),
        "pspin_set_list_batch": (C.c_int, [P, C.c_int]),
        "pspin_get_pair_keys": (C.c_int, [P, u64p]),
        "pspin_autotune": (C.c_int, [P, C.c_int, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                     C.POINTER(C.c_double)]),
        "pspin_set_resync_interval": (C.c_int, [P, C.c_int]),
        "pspin_set_coupling_format": (C.c_int, [P, C.c_int]),
        "pspin_batch_create": (C.c_int, [P, C.c_int, i32p, u64p, C.c_int, C.c_int, dp, C.POINTER(P)]),
        "pspin_batch_destroy": (C.c_int, [P]),
        "pspin_batch_set_resync_interval": (C.c_int, [P, C.c_int]),
        "pspin_batch_run": (C.c_int, [P, C.c_int, C.c_int]),
        "pspin_batch_info": (C.c_int, [P, C.POINTER(C.c_int), C.POINTER(C.c_longlong), C.POINTER(C.c_longlong)]),
        "pspin_batch_get_results": (C.c_int, [P, C.c_int, dp, i64p, i64p]),
        "pspin_batch_get_spins": (C.c_int, [P, C.c_int, dp]),
        "pspin_batch_get_perm": (C.c_int, [P, C.c_int, i32p]),
        "pspin_batch_get_exchange_stats": (C.c_int, [P, C.c_int, i64p, i64p]),
        "pspin_batch_set_chunk": (C.c_int, [P, C.c_int]),
        "pspin_batch_autotune": (C.c_int, [P, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_int),
                                           C.POINTER(C.c_double)]),
        "pspin_lp_default_config": (C.c_int, [P]),
        "pspin_lp_create": (C.c_int, [P, P, C.POINTER(P)]),
        "pspin_lp_destroy": (C.c_int, [P]),
        "pspin_lp_n_params": (C.c_longlong, [P]),
        "pspin_lp_dims": (C.c_int, [P, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int),
                                    C.POINTER(C.c_int)]),
        "pspin_lp_get_params": (C.c_int, [P, f32p]),
        "pspin_lp_set_params": (C.c_int, [P, f32p]),
        "pspin_lp_set_scale": (C.c_int, [P, C.c_double]),
        "pspin_lp_set_mean_scale": (C.c_int, [P, C.c_double]),
        "pspin_lp_forward": (C.c_int, [P, f32p, C.c_int, f32p]),
        "pspin_lp_loss_grad": (C.c_int, [P, f32p, f32p, f32p, C.c_int, C.POINTER(C.c_double), f32p]),
        "pspin_lp_record": (C.c_int, [P]),
        "pspin_lp_collect": (C.c_int, [P, C.c_longlong, C.c_int, C.c_int]),
        "pspin_lp_data_size": (C.c_longlong, [P]),
        "pspin_lp_train": (C.c_int, [P, C.c_longlong, C.c_int, C.c_int, C.c_double, P]),
        "pspin_lp_move": (C.c_int, [P]),
        "pspin_lp_run": (C.c_int, [P, C.c_longlong, C.c_int, C.c_int]),
        "pspin_lp_get_stats": (C.c_int, [P, i64p, i64p]),
        "pspin_lp_reset_stats": (C.c_int, [P]),
        "pspin_lp_get_last": (C.c_int, [P, P, P, P, P]),
        "pspin_lp_get_move_counters": (C.c_int, [P, u64p]),
    }
    for name, (res, args) in sigs.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def _check(rc: int) -> None:
    if rc != 0:
        msg = load_library().pspin_last_error()
        raise PSpinError(f"pspin error {rc}: {msg.decode() if msg else ''}")


def call(name: str, *args) -> None:
    """Call C-ABI entry point `name` (device buffers as integer addresses)
    and raise PSpinError when it fails."""
    _check(getattr(load_library(), name)(*args))


class PSpin:
    """One MCState (variant=DENSE) or MCStateSparse (variant=SPARSE) on the GPU."""

    def __init__(self, cfg: SimConfig, variant: int = DENSE):
        self._L = load_library()
        self._h = C.c_void_p()
        _check(self._L.pspin_create(C.byref(cfg), variant, C.byref(self._h)))
        self.cfg = cfg
        info = Info()
        _check(self._L.pspin_get_info(self._h, C.byref(info)))
        self.info = info
        self.N = info.N
        self.nrep = info.nrep
        self.variant = variant
        self.NT = None
        self.nrep_per_temp = None

    # -- lifetime
    def close(self) -> None:
        if self._h:
            self._L.pspin_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # -- MCState API
    def set_betas(self, betas) -> None:
        _check(self._L.pspin_set_betas(self._h, np.ascontiguousarray(betas, dtype=np.float64)))

    def set_uniform_beta(self, beta: float) -> None:
        _check(self._L.pspin_set_uniform_beta(self._h, float(beta)))

    def sweep(self, n: int = 1) -> None:
        _check(self._L.pspin_sweep(self._h, int(n)))

    def results(self):
        E = np.zeros(self.nrep)
        a = np.zeros(self.nrep, dtype=np.int64)
        p = np.zeros(self.nrep, dtype=np.int64)
        _check(self._L.pspin_get_results(self._h, E, a, p))
        return E, a, p

    def reset_counters(self) -> None:
        _check(self._L.pspin_reset_counters(self._h))

    def spins(self) -> np.ndarray:
        """[nrep, N] complex128."""
        s = np.zeros(2 * self.N * self.nrep)
        _check(self._L.pspin_get_spins(self._h, s))
        s = s.reshape(self.nrep, self.N, 2)
        return s[..., 0] + 1j * s[..., 1]

    def set_spins(self, z) -> None:
        z = np.asarray(z).reshape(self.nrep, self.N)
        s = np.ascontiguousarray(np.stack([z.real, z.imag], axis=-1), dtype=np.float64).reshape(-1)
        _check(self._L.pspin_set_spins(self._h, s))

    def split_energies(self):
        e2, e4 = np.zeros(self.nrep), np.zeros(self.nrep)
        _check(self._L.pspin_split_energies(self._h, e2, e4))
        return e2, e4

    def resync_energy(self) -> None:
        _check(self._L.pspin_resync_energy(self._h))

    def set_energies(self, E) -> None:
        _check(self._L.pspin_set_energies(self._h, np.ascontiguousarray(E, dtype=np.float64)))

    # -- disorder
    def omega(self) -> np.ndarray:
        w = np.zeros(self.N)
        _check(self._L.pspin_get_omega(self._h, w))
        return w

    def g2(self) -> np.ndarray:
        g = np.zeros(self.N * self.N)
        _check(self._L.pspin_get_g2(self._h, g))
        return g.reshape(self.N, self.N)

    def g4(self):
        nq = self.N * (self.N - 1) * (self.N - 2) * (self.N - 3) // 24
        g = np.zeros(nq)
        m = np.zeros(nq, dtype=np.uint8)
        _check(self._L.pspin_get_g4(self._h, g, m))
        return g, m

    def sparse_quartets(self):
        n = self.info.n_quartets
        q = np.zeros(5 * max(n, 1), dtype=np.int32)
        g = np.zeros(max(n, 1))
        _check(self._L.pspin_get_sparse_quartets(self._h, q, g))
        return q[: 5 * n].reshape(n, 5), g[:n]

    # -- parallel tempering
    def pt_init(self, NT: int, nrep_per_temp: int, beta_sched, seed48: int) -> None:
        _check(self._L.pspin_pt_init(self._h, NT, nrep_per_temp,
                                     np.ascontiguousarray(beta_sched, dtype=np.float64), int(seed48)))
        self.NT, self.nrep_per_temp = NT, nrep_per_temp

    def pt_run(self, n_sweeps: int, pt_freq: int = 1) -> None:
        _check(self._L.pspin_pt_run(self._h, int(n_sweeps), int(pt_freq)))

    def pt_exchange(self) -> None:
        _check(self._L.pspin_pt_exchange(self._h))

    def pt_perm(self) -> np.ndarray:
        p = np.zeros(self.NT * self.nrep_per_temp, dtype=np.int32)
        _check(self._L.pspin_pt_get_perm(self._h, p))
        return p

    def pt_exchange_stats(self):
        n = self.NT - 1
        arrs = [np.zeros(n, dtype=np.int64) for _ in range(4)]
        _check(self._L.pspin_pt_get_exchange_stats(self._h, *arrs))
        return tuple(arrs)

    def pt_reset_window(self) -> None:
        _check(self._L.pspin_pt_reset_window(self._h))

    def pt_rng48(self) -> int:
        x = C.c_uint64()
        _check(self._L.pspin_pt_get_rng48(self._h, C.byref(x)))
        return x.value

    def pt_enable_series(self, capacity: int) -> None:
        _check(self._L.pspin_pt_enable_series(self._h, int(capacity)))
        self._series_cap = capacity

    def pt_series(self) -> np.ndarray:
        out = np.zeros((max(self._series_cap, 1), self.nrep_per_temp))
        n = C.c_int()
        _check(self._L.pspin_pt_get_series(self._h, out.ctypes.data_as(C.c_void_p), C.byref(n)))
        return out[: n.value]

    # -- hooks
    def enable_step_log(self, capacity_sweeps: int) -> None:
        _check(self._L.pspin_enable_step_log(self._h, int(capacity_sweeps)))
        self._log_cap = capacity_sweeps

    def step_log(self):
        steps = self.N // 2
        out = np.zeros(self.nrep * self._log_cap * steps, dtype=STEP_DTYPE)
        n = C.c_int()
        _check(self._L.pspin_get_step_log(self._h, out.ctypes.data_as(C.c_void_p), C.byref(n)))
        return out.reshape(self.nrep, self._log_cap, steps)[:, : n.value], n.value

    def rng_info(self):
        s = np.zeros(self.nrep, dtype=np.uint64)
        c = C.c_uint64()
        _check(self._L.pspin_get_rng_info(self._h, s, C.byref(c)))
        return s, c.value

    def set_launch_mode(self, mode: int) -> None:
        """LAUNCH_AUTO (cooperative multi-round launches when they fit), LAUNCH_GRAPHS or LAUNCH_DIRECT."""
        _check(self._L.pspin_set_launch_mode(self._h, int(mode)))

    def set_wait_limit(self, polls: int) -> None:
        """Bound on the polls of a cross-workgroup wait in cooperative launches (0: the default)."""
        _check(self._L.pspin_set_wait_limit(self._h, int(polls)))

    def set_profiling(self, on: bool) -> None:
        _check(self._L.pspin_set_profiling(self._h, int(bool(on))))

    def profile(self):
        a, c = C.c_double(), C.c_double()
        b, d = C.c_longlong(), C.c_longlong()
        _check(self._L.pspin_get_profile(self._h, C.byref(a), C.byref(b), C.byref(c), C.byref(d)))
        e = C.c_longlong()
        _check(self._L.pspin_get_profile_launches(self._h, C.byref(e)))
        return {"sweep_ms": a.value, "n_sweep": b.value, "exch_ms": c.value, "n_exch": d.value,
                "n_launch": e.value}

    def set_coupling_format(self, fmt: int) -> None:
        """0 = FP32, 1 = BF16 hi/lo pairs (list kernel only)."""
        _check(self._L.pspin_set_coupling_format(self._h, int(fmt)))
        info = Info()
        _check(self._L.pspin_get_info(self._h, C.byref(info)))
        self.info = info

    def set_resync_interval(self, sweeps: int) -> None:
        _check(self._L.pspin_set_resync_interval(self._h, int(sweeps)))

    def set_pair_schedule(self, groups: int) -> None:
        """0: per-replica pairs; G > 0: replicas r = g (mod G) share a pair stream."""
        _check(self._L.pspin_set_pair_schedule(self._h, int(groups)))

    def set_pair_tiles(self, nwaves: int, rows: int) -> None:
        _check(self._L.pspin_set_pair_tiles(self._h, int(nwaves), int(rows)))

    def set_list_batch(self, batch: int) -> None:
        _check(self._L.pspin_set_list_batch(self._h, int(batch)))

    def pair_keys(self) -> np.ndarray:
        k = np.zeros(self.nrep, dtype=np.uint64)
        _check(self._L.pspin_get_pair_keys(self._h, k))
        return k

    def autotune(self, trial_sweeps: int = 8, verbose: bool = False):
        """Time the sweep kernel candidates on the live state; keep the fastest.
        Returns (kernel, block_size, ms_per_sweep)."""
        k, b, ms = C.c_int(), C.c_int(), C.c_double()
        _check(self._L.pspin_autotune(self._h, int(trial_sweeps), int(verbose), C.byref(k), C.byref(b), C.byref(ms)))
        info = Info()
        _check(self._L.pspin_get_info(self._h, C.byref(info)))
        self.info = info
        return k.value, b.value, ms.value

    def set_kernel(self, kernel: int = KERNEL_AUTO, block_size: int = 0) -> None:
        _check(self._L.pspin_set_kernel(self._h, int(kernel), int(block_size)))
        info = Info()
        _check(self._L.pspin_get_info(self._h, C.byref(info)))
        self.info = info


class PSpinBatch:
    """Many independent sparse (smoothed-cube) PT systems of uneven N in one launch.

    Sample s uses N = Ns[s] and seed = seeds[s]; every sample has NT x nrep_per_temp
    chains on beta_sched.  Mirrors running the reference's parallel_tempering_sparse
    once per sample (one process per label in scan_samples_pt).
    """

    def __init__(self, base: SimConfig, Ns, seeds, NT: int, nrep_per_temp: int, beta_sched):
        self._L = load_library()
        self._h = C.c_void_p()
        self.Ns = np.ascontiguousarray(Ns, dtype=np.int32)
        self.seeds = np.ascontiguousarray(seeds, dtype=np.uint64)
        self.NT, self.nrep_per_temp = NT, nrep_per_temp
        self.chains_per_sample = NT * nrep_per_temp
        _check(self._L.pspin_batch_create(C.byref(base), len(self.Ns), self.Ns, self.seeds, NT, nrep_per_temp,
                                          np.ascontiguousarray(beta_sched, dtype=np.float64), C.byref(self._h)))
        n, ch, pu = C.c_int(), C.c_longlong(), C.c_longlong()
        _check(self._L.pspin_batch_info(self._h, C.byref(n), C.byref(ch), C.byref(pu)))
        self.n_samples, self.chains, self.pair_updates_per_sweep = n.value, ch.value, pu.value

    def close(self):
        if self._h:
            self._L.pspin_batch_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_resync_interval(self, sweeps: int):
        _check(self._L.pspin_batch_set_resync_interval(self._h, int(sweeps)))

    def run(self, n_sweeps: int, pt_freq: int = 1):
        _check(self._L.pspin_batch_run(self._h, int(n_sweeps), int(pt_freq)))

    def set_chunk(self, sweeps_per_launch: int):
        _check(self._L.pspin_batch_set_chunk(self._h, int(sweeps_per_launch)))

    def autotune(self, trial_sweeps: int = 256, pt_freq: int = 1, verbose: bool = False):
        """Time sweeps-per-launch candidates; keep the fastest. Returns (chunk, ms_per_sweep)."""
        c, ms = C.c_int(), C.c_double()
        _check(self._L.pspin_batch_autotune(self._h, int(trial_sweeps), int(pt_freq), int(verbose), C.byref(c),
                                            C.byref(ms)))
        return c.value, ms.value

    def results(self, s: int):
        c = self.chains_per_sample
        E, a, p = np.zeros(c), np.zeros(c, dtype=np.int64), np.zeros(c, dtype=np.int64)
        _check(self._L.pspin_batch_get_results(self._h, s, E, a, p))
        return E, a, p

    def spins(self, s: int):
        c, N = self.chains_per_sample, int(self.Ns[s])
        out = np.zeros(2 * N * c)
        _check(self._L.pspin_batch_get_spins(self._h, s, out))
        out = out.reshape(c, N, 2)
        return out[..., 0] + 1j * out[..., 1]

    def perm(self, s: int):
        p = np.zeros(self.chains_per_sample, dtype=np.int32)
        _check(self._L.pspin_batch_get_perm(self._h, s, p))
        return p

    def exchange_stats(self, s: int):
        a, p = np.zeros(self.NT - 1, dtype=np.int64), np.zeros(self.NT - 1, dtype=np.int64)
        _check(self._L.pspin_batch_get_exchange_stats(self._h, s, a, p))
        return a, p


class LPConfig(C.Structure):
    """pspin_lp_config (include/pspin.h)."""
    _fields_ = [("hidden", C.c_int), ("batch_cap", C.c_int), ("data_cap", C.c_longlong),
                ("log_sigma0", C.c_double), ("log_sigma_min", C.c_double), ("log_sigma_max", C.c_double),
                ("scale", C.c_double), ("init_gain", C.c_double), ("seed", C.c_uint64),
                ("mean_scale", C.c_double)]

    @classmethod
    def default(cls, **kw) -> "LPConfig":
        c = cls()
        _check(load_library().pspin_lp_default_config(C.byref(c)))
        for k, v in kw.items():
            setattr(c, k, v)
        return c


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


class LearnedMove:
    """Learned global-move proposal attached to a dense PSpin state."""

    def __init__(self, st: PSpin, cfg: LPConfig = None):
        self._L = load_library()
        self.st = st
        self.cfg = cfg if cfg is not None else LPConfig.default()
        self._h = C.c_void_p()
        _check(self._L.pspin_lp_create(st._h, C.byref(self.cfg), C.byref(self._h)))
        a, b, c, d = C.c_int(), C.c_int(), C.c_int(), C.c_int()
        _check(self._L.pspin_lp_dims(self._h, C.byref(a), C.byref(b), C.byref(c), C.byref(d)))
        self.in_dim, self.in_pad, self.hidden, self.out_dim = a.value, b.value, c.value, d.value
        self.n_params = int(self._L.pspin_lp_n_params(self._h))

    def close(self):
        if self._h:
            self._L.pspin_lp_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def params(self) -> np.ndarray:
        p = np.zeros(self.n_params, dtype=np.float32)
        _check(self._L.pspin_lp_get_params(self._h, p))
        return p

    def set_params(self, p) -> None:
        _check(self._L.pspin_lp_set_params(self._h, np.ascontiguousarray(p, dtype=np.float32)))

    def set_scale(self, scale: float) -> None:
        _check(self._L.pspin_lp_set_scale(self._h, float(scale)))
        self.cfg.scale = float(scale)

    def set_mean_scale(self, m: float) -> None:
        _check(self._L.pspin_lp_set_mean_scale(self._h, float(m)))
        self.cfg.mean_scale = float(m)

    def forward(self, x) -> np.ndarray:
        x = np.ascontiguousarray(x, dtype=np.float32)
        out = np.zeros((x.shape[0], self.out_dim), dtype=np.float32)
        _check(self._L.pspin_lp_forward(self._h, x, x.shape[0], out))
        return out

    def loss_grad(self, x, u, y):
        x = np.ascontiguousarray(x, dtype=np.float32)
        loss = C.c_double()
        g = np.zeros(self.n_params, dtype=np.float32)
        _check(self._L.pspin_lp_loss_grad(self._h, x, np.ascontiguousarray(u, dtype=np.float32),
                                          np.ascontiguousarray(y, dtype=np.float32), x.shape[0], C.byref(loss), g))
        return loss.value, g

    def record(self) -> None:
        _check(self._L.pspin_lp_record(self._h))

    def collect(self, n_snapshots: int, every: int = 1, pt_freq: int = 1) -> None:
        _check(self._L.pspin_lp_collect(self._h, int(n_snapshots), int(every), int(pt_freq)))

    def data_size(self) -> int:
        return int(self._L.pspin_lp_data_size(self._h))

    def train(self, n_steps: int, batch: int = 1024, lag: int = 1, lr: float = 1e-3) -> np.ndarray:
        hist = np.zeros(max(1, int(n_steps)))
        _check(self._L.pspin_lp_train(self._h, int(n_steps), int(batch), int(lag), float(lr), _ptr(hist)))
        return hist[:n_steps]

    def move(self) -> None:
        _check(self._L.pspin_lp_move(self._h))

    def run(self, n_sweeps: int, pt_freq: int = 1, move_every: int = 8) -> None:
        _check(self._L.pspin_lp_run(self._h, int(n_sweeps), int(pt_freq), int(move_every)))

    def stats(self):
        n = self.st.NT if self.st.NT else 1
        a, p = np.zeros(n, dtype=np.int64), np.zeros(n, dtype=np.int64)
        _check(self._L.pspin_lp_get_stats(self._h, a, p))
        return a, p

    def reset_stats(self) -> None:
        _check(self._L.pspin_lp_reset_stats(self._h))

    def last(self):
        """(proposal [R][N] complex, dbg [R][8], out_fwd [R][4N], out_rev [R][4N])."""
        R, N = self.st.nrep, self.st.N
        prop = np.zeros((R, N, 2))
        dbg = np.zeros((R, 8))
        of = np.zeros((R, self.out_dim), dtype=np.float32)
        orv = np.zeros((R, self.out_dim), dtype=np.float32)
        _check(self._L.pspin_lp_get_last(self._h, _ptr(prop), _ptr(dbg), _ptr(of), _ptr(orv)))
        return prop[..., 0] + 1j * prop[..., 1], dbg, of, orv

    def move_counters(self) -> np.ndarray:
        c = np.zeros(self.st.nrep, dtype=np.uint64)
        _check(self._L.pspin_lp_get_move_counters(self._h, c))
        return c
