"""Byte-level check of every file the PT drivers write
(src/parallel_tempering.cu:120-176,243-271,355-416,495-505;
src/parallel_tempering_sparse.cu:277-296,505-517), against the oracle's
restatement of the reference's printf formats (oracle/outputs.py).

Files whose values are known without the run are rebuilt from independent
values: params.txt from the flags, temperatures.txt from the oracle's
schedule, frequencies.txt from the oracle's omega, quartets.txt from the
oracle's sparse selection.  Run-dependent files are rebuilt from the numbers
the run wrote (which checks every separator, width and precision) and their
values are checked against the physics: energy_split.txt against the FP64
oracle energies of the configs.bin snapshot of the same sweep and slot,
energy_accept.txt against the same to the sweep's FP32 accuracy, the
exchange windows' proposal counts against the schedule.
"""
import os
import subprocess

import numpy as np
import pytest

from oracle import outputs as F

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "bin")


def _run(args, cwd):
    p = subprocess.run(args, cwd=cwd, capture_output=True, text=True, timeout=120)
    assert p.returncode == 0, p.stderr + p.stdout
    return p


def _rows(path):
    return [line.rstrip("\n").split("\t") for line in open(path) if not line.startswith("#")]


def _same_text(got, want, tol):
    """Equal text, except numeric fields of two FP64 evaluations of the same
    value (another summation order) that differ in the last printed digit:
    those must agree within tol.  Every other byte must match."""
    if got == want:
        return
    gl, wl = got.splitlines(keepends=True), want.splitlines(keepends=True)
    assert len(gl) == len(wl)
    for a, b in zip(gl, wl):
        if a == b:
            continue
        fa, fb = a.split("\t"), b.split("\t")
        assert len(fa) == len(fb), (a, b)
        for x, y in zip(fa, fb):
            if x != y:
                assert len(x) == len(y) and abs(float(x) - float(y)) <= tol, (a, b)


def _time_line_ok(text, want_lead, iters):
    """time.txt: header and leading fields exact; the run's own time in
    %.6f, ms_per_sweep = time / iters * 1e3 in %.6f."""
    head, line = text.splitlines(keepends=True)
    assert head == want_lead.splitlines(keepends=True)[0]
    lead = want_lead.splitlines()[1].rsplit(" ", 2)[0] + " "
    assert line.startswith(lead) and line.endswith("\n")
    t, ms = line[len(lead):].split()
    assert t == "%.6f" % float(t) and ms == "%.6f" % float(ms)
    assert abs(float(ms) - float(t) / iters * 1e3) <= 1e-6 + 0.5e-6 * 1e3 / iters


def test_pt_driver_files_byte_exact(pspin, oracle, tmp_path):
    N, NT, nrep, iters, save, seed, label = 12, 4, 2, 40, 10, 42, 3
    Tmin, Tmax, gamma = 0.4, 1.6, 0.3
    _run([os.path.join(BIN, "parallel_tempering"), "-N", str(N), "-Tmax", str(Tmax), "-Tmin", str(Tmin), "-NT",
          str(NT), "-pt_freq", "1", "-iter", str(iters), "-nrep", str(nrep), "-save_freq", str(save), "-seed",
          str(seed), "-fmc", "2", "-gamma", str(gamma), "-log_temp", "-label", str(label)], tmp_path)
    d = tmp_path / "data" / f"PT_N{N}_a0.5_Rinf_a00.5_NT{NT}_NR{nrep}_S{label}"
    cfg = pspin.SimConfig.default()
    T, beta = oracle.temperatures(NT, Tmin, Tmax, 1)
    # files known before the run
    assert (d / "params.txt").read_text() == F.params_txt(N, cfg.J, cfg.J0, cfg.alpha, cfg.alpha0, NT, nrep, label,
                                                          seed, iters, Tmin, Tmax)
    assert (d / "temperatures.txt").read_text() == F.temperatures_txt(T, beta)
    assert (d / "frequencies.txt").read_text() == F.frequencies_txt(oracle.omega(N, 2, seed))
    # configs.bin: header, the snapshots of the second half (sweeps 30 and 40)
    raw = (d / "configs.bin").read_bytes()
    (hN, hT, hr), snaps = F.read_configs_bin(raw)
    assert (hN, hT, hr) == (N, NT, nrep) and [s for s, _ in snaps] == [30, 40]
    assert F.configs_bin(N, NT, nrep, snaps) == raw
    for _, z in snaps:
        np.testing.assert_allclose(np.sum(np.abs(z) ** 2, axis=1), N, rtol=1e-12)
    # the run's disorder (pspin_create with the same flags; its couplings are
    # checked against the oracle's in test_gpu_disorder.py)
    st = pspin.PSpin(pspin.SimConfig.default(N=N, nrep=NT * nrep, seed=seed, fmc_mode=2, gamma=gamma), pspin.DENSE)
    g2, (g4, mask), h2 = np.asarray(st.g2()).reshape(-1), st.g4(), st.info.h2_active
    st.close()
    e_snap = {}
    for sw, z in snaps:
        e_snap[sw] = np.array([oracle.energy(N, z[i], g2, h2, g4, mask) for i in range(NT * nrep)]) / N
    # energy_split.txt: the rows of the snapshot sweeps rebuilt from the FP64
    # oracle energies of the same slot's configuration
    es_txt = (d / "energy_split.txt").read_text()
    assert es_txt.startswith(F.energy_split_header(nrep))
    es = es_txt.splitlines(keepends=True)[1:]
    assert len(es) == NT * iters // save
    for sw in (30, 40):
        got = "".join(line for line in es if int(line.split("\t")[0]) == sw)
        e = e_snap[sw]
        _same_text(got, F.energy_split_rows(sw, T, e[:, 0].reshape(NT, nrep), e[:, 1].reshape(NT, nrep)), 1.5e-8)
    # energy_accept.txt: every line rebuilt from the values it holds; E/N at
    # the snapshot sweeps against FP64 (the running energy sums FP32 Delta E)
    ea_txt = (d / "energy_accept.txt").read_text()
    ea = _rows(d / "energy_accept.txt")
    assert len(ea) == NT * iters // save
    assert ea_txt == F.energy_accept_header(nrep) + "".join(
        F.energy_accept_line(int(r[0]), int(r[1]), T[int(r[1])], [float(x) for x in r[3::2]],
                             [float(x) for x in r[4::2]]) for r in ea)
    assert [int(r[0]) for r in ea] == [(w // NT + 1) * save for w in range(len(ea))]
    for sw in (30, 40):
        eN = np.array([[float(x) for x in r[3::2]] for r in ea if int(r[0]) == sw]).reshape(-1)
        np.testing.assert_allclose(eN, e_snap[sw].sum(axis=1), atol=2e-6)
    # exchanges.txt: every window proposes save * nrep swaps per level pair
    ex = _rows(d / "exchanges.txt")
    assert len(ex) == (NT - 1) * iters // save
    want = F.exchanges_header()
    for w in range(iters // save):
        rows = ex[w * (NT - 1):(w + 1) * (NT - 1)]
        want += F.exchanges_rows((w + 1) * save, T, [int(r[4]) for r in rows], [save * nrep] * (NT - 1))
    assert (d / "exchanges.txt").read_text() == want
    _time_line_ok((d / "time.txt").read_text(), F.time_txt(N, NT, nrep, NT * nrep, iters, 1, save, Tmax, Tmin, 0.0),
                  iters)


def test_sparse_driver_quartets_byte_exact(pspin, oracle, tmp_path):
    N, NT, nrep, iters, save, seed, gamma = 16, 3, 2, 20, 10, 11, 0.3
    _run([os.path.join(BIN, "parallel_tempering_sparse"), "-N", str(N), "-Tmax", "1.5", "-Tmin", "0.5", "-NT",
          str(NT), "-pt_freq", "2", "-iter", str(iters), "-nrep", str(nrep), "-save_freq", str(save), "-seed",
          str(seed), "-fmc", "2", "-gamma", str(gamma)], tmp_path)
    d = tmp_path / "data" / f"PTS_N{N}_a0.5_Rinf_a00.5_NT{NT}_NR{nrep}_S0"
    cfg = pspin.SimConfig.default()
    J4 = cfg.alpha * cfg.J
    w = oracle.omega(N, 2, seed)
    mask = oracle.g4_mask(N, 2, w, gamma)
    g_unit = np.array([oracle.normal(seed + 2000 + 12345, 4, q) for q in range(len(mask))])
    q, g, _ = oracle.sparse_select(N, mask, g_unit, N, seed, J4, cfg.alpha0 * cfg.J0)
    rows = _rows(d / "quartets.txt")
    assert [r[:6] for r in rows] == [[str(s)] + [str(int(x)) for x in qq] for s, qq in enumerate(q)]
    # g: the oracle's normals agree with the device's to ~1e-15 relative
    _same_text((d / "quartets.txt").read_text(), F.quartets_txt(q, g), 1e-12 * float(np.abs(g).max()))
    _time_line_ok((d / "time.txt").read_text(),
                  F.time_txt(N, NT, nrep, NT * nrep, iters, 2, save, 1.5, 0.5, 0.0, n_sparse=len(q)), iters)
    T, beta = oracle.temperatures(NT, 0.5, 1.5, 0)
    assert (d / "temperatures.txt").read_text() == F.temperatures_txt(T, beta)
    assert (d / "frequencies.txt").read_text() == F.frequencies_txt(w)
