"""C-ABI checks that need no GPU: libpspin.so loads, exports every function
declared in include/pspin.h, and the host-only entry points behave like the
reference's parse_args / default_config (include/config.h:28-46,
src/config.cu:6-70)."""
import ctypes as C
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    names = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if fn.endswith(".h"):
            txt = open(os.path.join(ROOT, "include", fn)).read()
            names |= set(re.findall(r"^\s*(?:const\s+)?\w+\s*\**\s*(pspin_\w+)\s*\(", txt, re.M))
    return names


def test_exports_every_declared_symbol(pspin):
    names = _declared()
    assert len(names) >= 30
    L = C.CDLL(pspin.LIB_PATH)
    missing = [n for n in sorted(names) if not hasattr(L, n)]
    assert not missing, missing


def test_default_config(pspin):
    c = pspin.SimConfig.default()
    assert (c.N, c.nrep, c.T, c.J, c.J0, c.alpha, c.alpha0) == (64, 1, 1.0, 2.0, 0.0, 0.5, 0.5)
    assert (c.mc_iterations, c.seed, c.save_freq, c.verbose, c.label, c.dev, c.fmc_mode, c.gamma) == \
        (10000, 42, 100, 0, -1, -1, 0, 0.0)


def test_parse_args_flags_and_auto_gamma(pspin):
    c = pspin.SimConfig.parse_args(["-N", "32", "-nrep", "4", "-T", "0.5", "-J", "1.5", "-J0", "0.25",
                                    "-alpha", "0.3", "-alpha0", "0.7", "-iter", "99", "-seed", "18446744073709551615",
                                    "-save_freq", "7", "-label", "3", "-dev", "0", "-fmc", "2", "-verbose"])
    assert (c.N, c.nrep, c.T, c.J, c.J0, c.alpha, c.alpha0, c.mc_iterations) == (32, 4, 0.5, 1.5, 0.25, 0.3, 0.7, 99)
    assert c.seed == 2 ** 64 - 1 and c.save_freq == 7 and c.label == 3 and c.dev == 0
    assert c.verbose == 1
    assert abs(c.gamma - 1.0 / (2 * 31)) < 1e-15  # uniform auto-gamma (src/config.cu:65-66)
    c = pspin.SimConfig.parse_args(["-fmc", "1", "-verbose", "2"])
    assert c.gamma == 0.0 and c.verbose == 2
    c = pspin.SimConfig.parse_args(["-fmc", "2", "-gamma", "0.05"])
    assert c.gamma == 0.05


@pytest.mark.parametrize("argv", [["-bogus"], ["-N", "3"], ["-N"]])
def test_parse_args_errors(pspin, argv):
    with pytest.raises(pspin.PSpinError):
        pspin.SimConfig.parse_args(argv)
