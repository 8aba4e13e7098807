"""Test infrastructure (the oracle): the reference PT drivers' output files,
restated from their printf formats so a test can rebuild every file byte for
byte from independently computed values (or, for values only a run knows,
from the numbers the run wrote) and compare it with what bin/parallel_tempering
and bin/parallel_tempering_sparse wrote.

Formats follow src/parallel_tempering.cu:120-176 (params.txt, file headers,
configs.bin header), :243-271 (frequencies.txt, temperatures.txt),
:355-416 (energy_accept.txt, exchanges.txt, energy_split.txt, configs.bin
snapshots), :495-505 (time.txt); src/parallel_tempering_sparse.cu:277-296
(quartets.txt) and :505-517 (its time.txt).  Python's printf-style
formatting follows C's for these conversions (%d, %.Nf, %.Ng, %.Ne).
Only tests may import this module.
"""
import struct

import numpy as np


def params_txt(N, J, J0, alpha, alpha0, NT, nrep, label, seed, nsweeps, Tmin, Tmax):
    """src/parallel_tempering.cu:120-136 (label -1 is written as 0, :111)."""
    label = label if label >= 0 else 0
    return (f"N {N}\n" + "J %.10g\n" % J + "J0 %.10g\n" % J0 + "alpha %.10g\n" % alpha + "alpha0 %.10g\n" % alpha0
            + f"NT {NT}\nnrep {nrep}\nlabel {label}\nseed {seed}\nnsweeps {nsweeps}\n"
            + "Tmin %.10g\n" % Tmin + "Tmax %.10g\n" % Tmax)


def temperatures_txt(T, beta):
    """src/parallel_tempering.cu:262-270."""
    return "# Tidx T beta\n" + "".join("%d\t%.12f\t%.12f\n" % (t, T[t], beta[t]) for t in range(len(T)))


def frequencies_txt(omega):
    """src/parallel_tempering.cu:249-256 (written when fmc > 0)."""
    return "# i omega_i\n" + "".join("%d\t%.12f\n" % (i, w) for i, w in enumerate(omega))


def energy_accept_header(nrep):
    """src/parallel_tempering.cu:163-166."""
    return "# sweep\tTidx\tT" + "".join(f"\tE{r}/N\tacc{r}" for r in range(nrep)) + "\n"


def energy_accept_line(sweep, t, T_t, e_over_n, acc):
    """src/parallel_tempering.cu:375-382: one slot t, its replicas r."""
    return ("%d\t%d\t%.8f" % (sweep, t, T_t) + "".join("\t%.8f\t%.5f" % (e, a) for e, a in zip(e_over_n, acc))
            + "\n")


def energy_accept_rows(sweep, T, e_over_n, acc):
    """src/parallel_tempering.cu:373-382: e_over_n[t][r], acc[t][r] by slot."""
    return "".join(energy_accept_line(sweep, t, T[t], e_over_n[t], acc[t]) for t in range(len(T)))


def exchanges_header():
    """src/parallel_tempering.cu:168."""
    return "# sweep\tTidx\tT_high\tT_low\tn_acc\tn_prop\trate\n"


def exchanges_rows(sweep, T, n_acc, n_prop):
    """src/parallel_tempering.cu:385-394: window counts per level pair."""
    out = []
    for t in range(len(T) - 1):
        rate = n_acc[t] / n_prop[t] if n_prop[t] > 0 else 0.0
        out.append("%d\t%d\t%.8f\t%.8f\t%d\t%d\t%.5f\n" % (sweep, t, T[t], T[t + 1], n_acc[t], n_prop[t], rate))
    return "".join(out)


def energy_split_header(nrep):
    """src/parallel_tempering.cu:170-173."""
    return "# sweep\tTidx\tT" + "".join(f"\tE2_{r}/N\tE4_{r}/N" for r in range(nrep)) + "\n"


def energy_split_rows(sweep, T, e2_over_n, e4_over_n):
    """src/parallel_tempering.cu:397-405."""
    out = []
    for t in range(len(T)):
        out.append("%d\t%d\t%.8f" % (sweep, t, T[t])
                   + "".join("\t%.8f\t%.8f" % (e2_over_n[t][r], e4_over_n[t][r]) for r in range(len(e2_over_n[t])))
                   + "\n")
    return "".join(out)


def configs_bin(N, NT, nrep, snapshots):
    """src/parallel_tempering.cu:154-159,412-428: int32[3] {N, NT, nrep}, then
    per snapshot int32 sweep and NT*nrep*N complex FP64 (re, im) in slot
    order t * nrep + r.  snapshots: [(sweep, array [NT*nrep, N] complex)]."""
    b = struct.pack("<3i", N, NT, nrep)
    for sweep, z in snapshots:
        z = np.asarray(z, dtype=np.complex128).reshape(NT * nrep, N)
        b += struct.pack("<i", sweep) + np.stack([z.real, z.imag], -1).astype("<f8").tobytes()
    return b


def read_configs_bin(data):
    """Inverse of configs_bin."""
    N, NT, nrep = struct.unpack_from("<3i", data, 0)
    off, snaps = 12, []
    size = NT * nrep * N * 16
    while off < len(data):
        (sweep,) = struct.unpack_from("<i", data, off)
        a = np.frombuffer(data, dtype="<f8", count=size // 8, offset=off + 4).reshape(NT * nrep, N, 2)
        snaps.append((sweep, a[..., 0] + 1j * a[..., 1]))
        off += 4 + size
    return (N, NT, nrep), snaps


def time_txt(N, NT, nrep, total, iters, pt_freq, save_freq, Tmax, Tmin, elapsed, n_sparse=None):
    """src/parallel_tempering.cu:495-505; the sparse driver adds n_sparse
    before the times (src/parallel_tempering_sparse.cu:505-517)."""
    if n_sparse is None:
        head = ("# N NT nrep_per_temp total_replicas mc_iterations pt_freq save_freq Tmax Tmin total_time_s "
                "ms_per_sweep\n")
        lead = "%d %d %d %d %d %d %d %.6f %.6f " % (N, NT, nrep, total, iters, pt_freq, save_freq, Tmax, Tmin)
    else:
        head = ("# N NT nrep_per_temp total_replicas mc_iterations pt_freq save_freq Tmax Tmin n_sparse total_time_s "
                "ms_per_sweep\n")
        lead = "%d %d %d %d %d %d %d %.6f %.6f %d " % (N, NT, nrep, total, iters, pt_freq, save_freq, Tmax, Tmin,
                                                       n_sparse)
    return head + lead + "%.6f %.6f\n" % (elapsed, elapsed / iters * 1e3)


def quartets_txt(quartets, g):
    """src/parallel_tempering_sparse.cu:284-294: quartets [n][5] (i,j,k,l,ch)."""
    return "# idx i j k l ch Re(g)\n" + "".join(
        "%d\t%d\t%d\t%d\t%d\t%d\t%.12e\n" % (s, *[int(x) for x in q], gv) for s, (q, gv) in enumerate(zip(quartets, g)))
