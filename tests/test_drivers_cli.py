"""Driver binaries (bin/parallel_tempering, ..._sparse, montecarlo,
simulated_annealing): argument handling and exit codes that need no GPU
(src/parallel_tempering.cu:58-70, src/simulated_annealing.cu:53-66,
src/config.cu:45-59), plus a GPU run of each that checks the output layout
the reference's analysis reads (src/parallel_tempering.cu:120-176,355-505)."""
import os
import struct
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "bin")


def _run(args, cwd):
    return subprocess.run(args, cwd=cwd, capture_output=True, text=True, timeout=120)


def _need(name):
    path = os.path.join(BIN, name)
    if not os.path.exists(path):
        pytest.skip(f"{path} not built")
    return path


@pytest.mark.parametrize("prog,args,code", [
    ("parallel_tempering", ["-N", "8"], 1),                                         # missing PT flags
    ("parallel_tempering", ["-Tmax", "1", "-Tmin", "2", "-NT", "4", "-pt_freq", "1"], 1),  # Tmin > Tmax
    ("parallel_tempering", ["-Tmax", "2", "-Tmin", "1", "-NT", "1", "-pt_freq", "1"], 1),  # NT < 2
    ("parallel_tempering", ["-Tmax", "2", "-Tmin", "1", "-NT", "4", "-pt_freq", "1", "-bogus"], 1),
    ("parallel_tempering", ["-Tmax", "2", "-Tmin", "1", "-NT", "4", "-pt_freq", "1", "-N", "3"], 1),
    ("parallel_tempering_sparse", ["-Tmax", "2", "-NT", "4", "-pt_freq", "1"], 1),
    ("simulated_annealing", ["-Tmax", "2", "-Tmin", "1"], 1),                      # missing -ntemp
    ("simulated_annealing", ["-Tmax", "1", "-Tmin", "2", "-ntemp", "3"], 1),
    ("montecarlo", ["-what"], 1),
    ("montecarlo", ["-N", "2"], 1),
])
def test_cli_errors(tmp_path, prog, args, code):
    p = _run([_need(prog)] + args, tmp_path)
    assert p.returncode == code, p.stderr


@pytest.mark.gpu
def test_pt_driver_outputs(tmp_path):
    exe = _need("parallel_tempering")
    N, NT, nrep, iters, save = 10, 4, 2, 40, 10
    p = _run([exe, "-N", str(N), "-Tmax", "1.6", "-Tmin", "0.4", "-NT", str(NT), "-pt_freq", "1", "-iter",
              str(iters), "-nrep", str(nrep), "-save_freq", str(save), "-seed", "7", "-fmc", "2", "-gamma", "0.3",
              "-log_temp", "-label", "3", "-verbose", "2"], tmp_path)
    assert p.returncode == 0, p.stderr + p.stdout
    d = tmp_path / "data" / f"PT_N{N}_a0.5_Rinf_a00.5_NT{NT}_NR{nrep}_S3"
    assert d.is_dir(), list((tmp_path / "data").iterdir())
    names = {x.name for x in d.iterdir()}
    assert {"params.txt", "energy_accept.txt", "exchanges.txt", "energy_split.txt", "configs.bin", "time.txt",
            "temperatures.txt", "frequencies.txt"} <= names
    ea = np.loadtxt(d / "energy_accept.txt")
    assert ea.shape == (NT * iters // save, 3 + 2 * nrep)
    assert np.all((ea[:, 4::2] >= 0) & (ea[:, 4::2] <= 1))
    ex = np.loadtxt(d / "exchanges.txt")
    assert ex.shape == ((NT - 1) * iters // save, 7) and np.all(ex[:, 5] == save * nrep)
    es = np.loadtxt(d / "energy_split.txt")
    # E2/N + E4/N equals E/N of the same physical replica (the running energy
    # agrees with the FP64 recomputation to the sweep's FP32 accuracy)
    np.testing.assert_allclose(es[:, 3::2] + es[:, 4::2], ea[:, 3::2], atol=1e-4)
    raw = (d / "configs.bin").read_bytes()
    hdr = struct.unpack("3i", raw[:12])
    assert hdr == (N, NT, nrep)
    snap = 4 + NT * nrep * 2 * N * 8
    assert (len(raw) - 12) % snap == 0 and (len(raw) - 12) // snap == 2  # saves at 30 and 40
    T = np.loadtxt(d / "temperatures.txt")
    np.testing.assert_allclose(T[:, 1], 1.6 * (0.4 / 1.6) ** (np.arange(NT) / (NT - 1)), rtol=1e-11)
    params = dict(line.split() for line in open(d / "params.txt"))
    assert params["N"] == str(N) and params["nsweeps"] == str(iters) and params["seed"] == "7"


@pytest.mark.gpu
def test_sparse_mc_sa_drivers(tmp_path):
    p = _run([_need("parallel_tempering_sparse"), "-N", "12", "-Tmax", "1.5", "-Tmin", "0.5", "-NT", "3",
              "-pt_freq", "2", "-iter", "20", "-nrep", "2", "-save_freq", "10"], tmp_path)
    assert p.returncode == 0, p.stderr
    d = tmp_path / "data" / "PTS_N12_a0.5_Rinf_a00.5_NT3_NR2_S0"
    q = np.loadtxt(d / "quartets.txt")
    assert q.shape == (12, 7) and np.all(np.diff(q[:, 1:5], axis=1) > 0)
    p = _run([_need("montecarlo"), "-N", "10", "-nrep", "3", "-T", "0.7", "-iter", "20", "-save_freq", "5",
              "-label", "1"], tmp_path)
    assert p.returncode == 0, p.stderr
    ea = np.loadtxt(tmp_path / "data" / "N10_NR3_S1" / "energy_accept.txt")
    assert ea.shape == (4, 7)
    assert len(list((tmp_path / "data" / "N10_NR3_S1" / "configs").iterdir())) == 3 * 2
    p = _run([_need("simulated_annealing"), "-N", "10", "-nrep", "2", "-Tmax", "2", "-Tmin", "0.5", "-ntemp", "4",
              "-iter", "3", "-save_freq", "4"], tmp_path)
    assert p.returncode == 0, p.stderr
    sa = tmp_path / "data" / "SA_N10_NR2_S0"
    ea = np.loadtxt(sa / "energy_accept.txt")
    assert ea.shape == (4 * 2, 2 + 4)
    assert len(list((sa / "configs").iterdir())) == 2 * 2
