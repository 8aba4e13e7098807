"""The oracle's restatement of the PT drivers' output formats
(oracle/outputs.py) against the C library's own snprintf on the reference's
format strings, character for character (src/parallel_tempering.cu:120-136,
262-270,249-256,373-405,495-505; src/parallel_tempering_sparse.cu:284-294).
Values include halfway-rounding, negative, tiny and large cases."""
import ctypes as C
import re

import numpy as np

from oracle import outputs as F

libc = C.CDLL("libc.so.6")


def cfmt(fmt, *args):
    """C's snprintf(fmt, args...), each argument typed by its conversion."""
    specs = re.findall(r"%[-+ #0-9.]*(lld|lu|d|f|g|e)", fmt)
    assert len(specs) == len(args)
    ctype = {"lld": C.c_longlong, "lu": C.c_ulong, "d": C.c_int, "f": C.c_double, "g": C.c_double, "e": C.c_double}
    buf = C.create_string_buffer(4096)
    libc.snprintf(buf, C.c_size_t(4096), fmt.encode(), *[ctype[k](a) for k, a in zip(specs, args)])
    return buf.value.decode()


VALUES = [1.6, 0.4, 0.1, 2.0, 0.0, 1e-7, 123456789.123456789, 0.125, 2.5e-9, 1.0 / 3, 0.999999999995,
          -0.30000000000000004, 5e-324, 7.000000005, -12.3456785]


def test_params_txt_matches_printf():
    for J, J0, a, a0, Tmin, Tmax in zip(VALUES, VALUES[3:] + VALUES[:3], VALUES[5:] + VALUES[:5],
                                        VALUES[7:] + VALUES[:7], VALUES[2:] + VALUES[:2], VALUES[9:] + VALUES[:9]):
        seed = 18446744073709551615
        want = (cfmt("N %d\n", 12) + cfmt("J %.10g\n", J) + cfmt("J0 %.10g\n", J0) + cfmt("alpha %.10g\n", a)
                + cfmt("alpha0 %.10g\n", a0) + cfmt("NT %d\n", 4) + cfmt("nrep %d\n", 2) + cfmt("label %d\n", 3)
                + cfmt("seed %lu\n", seed) + cfmt("nsweeps %d\n", 40) + cfmt("Tmin %.10g\n", Tmin)
                + cfmt("Tmax %.10g\n", Tmax))
        assert F.params_txt(12, J, J0, a, a0, 4, 2, 3, seed, 40, Tmin, Tmax) == want


def test_table_lines_match_printf():
    rng = np.random.default_rng(3)
    vals = np.concatenate([VALUES, rng.normal(scale=3, size=40), np.round(rng.normal(size=20), 8) + 5e-9])
    T = [float(v) for v in np.abs(vals[:6]) + 0.1]
    beta = [1.0 / t for t in T]
    assert F.temperatures_txt(T, beta) == "# Tidx T beta\n" + "".join(
        cfmt("%d\t%.12f\t%.12f\n", t, T[t], beta[t]) for t in range(len(T)))
    assert F.frequencies_txt(vals[:9]) == "# i omega_i\n" + "".join(
        cfmt("%d\t%.12f\n", i, float(v)) for i, v in enumerate(vals[:9]))
    for k in range(0, len(vals) - 4, 4):
        e, a = [float(vals[k]), float(vals[k + 1])], [float(abs(vals[k + 2]) % 1), float(abs(vals[k + 3]) % 1)]
        want = cfmt("%d\t%d\t%.8f", 30, 2, T[2]) + cfmt("\t%.8f\t%.5f", e[0], a[0]) + cfmt("\t%.8f\t%.5f", e[1], a[1])
        assert F.energy_accept_line(30, 2, T[2], e, a) == want + "\n"
        want = cfmt("%d\t%d\t%.8f", 40, 0, T[0]) + cfmt("\t%.8f\t%.8f", e[0], a[0]) + "\n"
        assert F.energy_split_rows(40, T[:1], [[e[0]]], [[a[0]]]) == want
        q = [[1, 4, 7, 9, 2]]
        assert F.quartets_txt(q, [float(vals[k])]).splitlines(keepends=True)[1] == cfmt(
            "%d\t%d\t%d\t%d\t%d\t%d\t%.12e\n", 0, 1, 4, 7, 9, 2, float(vals[k]))
    n_acc, n_prop = [3, 0, 20], [20, 0, 20]
    want = "".join(cfmt("%d\t%d\t%.8f\t%.8f\t%lld\t%lld\t%.5f\n", 10, t, T[t], T[t + 1], n_acc[t], n_prop[t],
                        n_acc[t] / n_prop[t] if n_prop[t] else 0.0) for t in range(3))
    assert F.exchanges_rows(10, T[:4], n_acc, n_prop) == want
    line = F.time_txt(12, 4, 2, 8, 40, 1, 10, 1.6, 0.4, 0.1234565).splitlines(keepends=True)[1]
    assert line == cfmt("%d %d %d %d %d %d %d %.6f %.6f %.6f %.6f\n", 12, 4, 2, 8, 40, 1, 10, 1.6, 0.4, 0.1234565,
                        0.1234565 / 40 * 1e3)
