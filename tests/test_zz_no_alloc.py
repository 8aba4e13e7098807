"""pspin_compute_energy_noalloc against pspin_compute_energy through the C
ABI: the same kernels in the same order, so the energies must agree bit for
bit, on two calls in a row (the second at a smaller N, so the shared
workspace holds the first call's partials when the second starts), and the
entry point must refuse N above its workspace's maximum."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _dev(a):
    t = torch.from_numpy(np.ascontiguousarray(a)).to("cuda")
    torch.cuda.synchronize()
    return t


def _system(N, seed):
    rng = np.random.default_rng(seed)
    nq = N * (N - 1) * (N - 2) * (N - 3) // 24
    z = rng.normal(size=(N, 2))
    g2 = rng.normal(size=(N * N, 2))
    g4 = rng.normal(size=(nq, 2))
    mask = rng.choice(np.array([0, 1, 2, 4, 5], dtype=np.uint8), size=nq)
    return [_dev(z), _dev(g2), _dev(g4), _dev(mask)]


def _energy(P, entry, bufs, N):
    e = C.c_double()
    P.call(entry, *[b.data_ptr() for b in bufs], N, C.byref(e))
    return e.value


def test_noalloc_energy_matches_allocating_entry(pspin):
    for N, seed in ((128, 1), (40, 2)):  # 1024 then 358 partials in the shared workspace
        print(f"[no-alloc energy] N={N}", flush=True)
        bufs = _system(N, seed)
        a = _energy(pspin, "pspin_compute_energy", bufs, N)
        b = _energy(pspin, "pspin_compute_energy_noalloc", bufs, N)
        assert a == b and np.isfinite(a), (N, a, b)
    e = C.c_double()
    with pytest.raises(pspin.PSpinError, match="maximum"):
        pspin.call("pspin_compute_energy_noalloc", bufs[0].data_ptr(), 0, bufs[2].data_ptr(), bufs[3].data_ptr(),
                   257, C.byref(e))
