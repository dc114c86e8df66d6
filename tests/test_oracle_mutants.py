"""The oracle's advection pins catch plausible mistakes.

Each case copies oracle/gemini_oracle.c, plants one mistake by a text substitution whose
target must occur exactly as often as stated (so the case notices when the oracle's code
moves), builds the copy with the oracle's own flags, and runs
tests/test_oracle_advection.py against it in a child process ($GEMINI_ORACLE_LIB).  The
pins must fail on every copy, and pass on the unmodified oracle.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "oracle", "gemini_oracle.c")
PINS = os.path.join(ROOT, "tests", "test_oracle_advection.py")
CFLAGS = ["-O1", "-fPIC", "-std=c11", "-ffp-contract=off", "-fno-fast-math", "-fopenmp",
          "-shared", "-w"]

# name -> (text in gemini_oracle.c, planted replacement, occurrences)
MUTANTS = {
    "plain minmod slope instead of MC": (
        "slope[i] = orc_minmod(cslope, orc_minmod(2.0 * lslope, 2.0 * rslope));",
        "slope[i] = orc_minmod(lslope, rslope);", 1),
    "MC without the factor 2 on the one-sided slopes": (
        "orc_minmod(2.0 * lslope, 2.0 * rslope)", "orc_minmod(lslope, rslope)", 1),
    "v*h1i instead of v/h1i in the flux": ("(v / H1I(i)) * dt", "(v * H1I(i)) * dt", 2),
    "upwind slope of the wrong cell (v > 0)": (
        "(v / H1I(i)) * dt)) * slope[i - 1];", "(v / H1I(i)) * dt)) * slope[i];", 1),
    "factor 1 for 0.5 in the v > 0 flux": (
        "phi[i] = F(i - 1) * v + ((0.5 * v)", "phi[i] = F(i - 1) * v + ((1.0 * v)", 1),
    "x1 face product pairs h1 with h3": (
        "ha2i[k - 1] = g->h2x1i[ci] * g->h3x1i[ci];",
        "ha2i[k - 1] = g->h1x1i[ci] * g->h3x1i[ci];", 1),
    "x1 sweep honours the tensor rank": (
        "if (dir == 1 || frank == 0) ha1[k + 1]", "if (frank == 0) ha1[k + 1]", 1),
    "x2 rank-1 face product without h1 squared": (
        "ha2i[k - 1] = frank == 0 ? p * g->h3x2i[ci] : (p * p) * g->h3x2i[ci];",
        "ha2i[k - 1] = frank == 0 ? p * g->h3x2i[ci] : p * g->h3x2i[ci];", 1),
    "x3 face product pairs h1 with h3": (
        "ha2i[k - 1] = frank == 0 ? p * g->h2x3i[ci] : (p * p) * g->h2x3i[ci];",
        "ha2i[k - 1] = frank == 0 ? p * g->h3x3i[ci] : (p * p) * g->h3x3i[ci];", 1),
    "rank-1 volume without h1 squared": (
        "else ha1[k + 1] = ((h1 * h1) * h2) * h3;", "else ha1[k + 1] = (h1 * h2) * h3;", 1),
    "x2 sweep scales v by h1 instead of h2": (
        "h1i[k - 1] = g->h2x2i[ci];", "h1i[k - 1] = g->h1x2i[ci];", 1),
    "x3 sweep scales v by h2 instead of h3": (
        "h1i[k - 1] = g->h3x3i[ci];", "h1i[k - 1] = g->h2x3i[ci];", 1),
    "composite step sweeps x2 momentum as a scalar": (
        "orc_sweep_allspec(g, 2, rhovs1, vs2i, dt, 1, 6);",
        "orc_sweep_allspec(g, 2, rhovs1, vs2i, dt, 0, 6);", 2),
    "composite step leaves the electron energy out of the x1 sweep": (
        "orc_sweep_allspec(g, 1, rhoes, vs1i, dt, 0, 7);",
        "orc_sweep_allspec(g, 1, rhoes, vs1i, dt, 0, 6);", 2),
    "T2rhoe with the atomic gamma for every species": (
        "    double gm1 = gammas_(isp) - 1.0;\n    for (long k = 0; k < d.N; ++k) rhoes",
        "    double gm1 = gammas_(1) - 1.0;\n    for (long k = 0; k < d.N; ++k) rhoes", 1),
    "rhoe2T without its 100 K floor": (
        "Ts[o + k] = fmax_f(t, 100.0);", "Ts[o + k] = t;", 1),
    "v12rhov1 with the O+ mass for every species": (
        "rhovs1[o + k] = (ns[o + k] * ms_(isp)) * vs1[o + k];",
        "rhovs1[o + k] = (ns[o + k] * ms_(1)) * vs1[o + k];", 1),
    "electron velocity without its sign": (
        "vs1[o7] = -(((1.0 / fmax_f(ns[o7], mindensdiv)) / qs_(LSP)) * chrgflux);",
        "vs1[o7] = (((1.0 / fmax_f(ns[o7], mindensdiv)) / qs_(LSP)) * chrgflux);", 1),
    "electron density without H+": (
        "/* sum(ns(:,:,:,1:lsp-1),4) */\n    for (int isp = 1; isp <= LSP - 1; ++isp)",
        "/* sum(ns(:,:,:,1:lsp-1),4) */\n    for (int isp = 1; isp <= LSP - 2; ++isp)", 1),
    "second x1 ghost velocity copies the first": (
        "A3(r, -1, i2, i3) = (A3(r, 0, i2, i3) + A3(r, 0, i2, i3)) - A3(u, 1, i2, i3);",
        "A3(r, -1, i2, i3) = A3(r, 0, i2, i3);", 1),
    "first x1 ghost velocity reflects the edge cell": (
        "A3(r, lx1 + 1, i2, i3) = 2.0 * V1I(lx1 + 1, i2, i3) - A3(u, lx1, i2, i3);",
        "A3(r, lx1 + 1, i2, i3) = A3(u, lx1, i2, i3);", 1),
}


def _build(tmp, name, text):
    base = os.path.join(tmp, name)
    with open(base + ".c", "w") as f:
        f.write(text)
    subprocess.run(["gcc", *CFLAGS, "-I", os.path.join(ROOT, "oracle"), "-o", base + ".so",
                    base + ".c", "-lm"], check=True)
    return base + ".so"


def _run_pins(libpath):
    env = dict(os.environ, GEMINI_ORACLE_LIB=libpath, OMP_NUM_THREADS="1")
    r = subprocess.run([sys.executable, "-m", "pytest", PINS, "-q", "-p", "no:cacheprovider", "-x"], cwd=ROOT, env=env, capture_output=True,
                       text=True, timeout=300)
    return r.returncode, r.stdout + r.stderr


@pytest.fixture(scope="module")
def mutant_libs(tmp_path_factory):
    tmp = str(tmp_path_factory.mktemp("mutants"))
    src = open(SRC).read()
    jobs = {"unmodified": src}
    for i, (name, (old, new, count)) in enumerate(MUTANTS.items()):
        assert src.count(old) == count, f"{name}: target found {src.count(old)} times"
        jobs[f"m{i:02d}"] = src.replace(old, new)
    with ThreadPoolExecutor(8) as ex:
        paths = dict(zip(jobs, ex.map(lambda k: _build(tmp, k, jobs[k]), jobs)))
    return paths


def test_pins_pass_on_the_oracle_as_built_here(mutant_libs):
    rc, out = _run_pins(mutant_libs["unmodified"])
    assert rc == 0, out[-2000:]


def test_pins_fail_on_every_planted_mistake(mutant_libs):
    names = list(MUTANTS)
    with ThreadPoolExecutor(4) as ex:
        results = list(ex.map(lambda i: _run_pins(mutant_libs[f"m{i:02d}"]), range(len(names))))
    survivors = []
    for name, (rc, out) in zip(names, results):
        # pytest exits 1 when tests fail; anything else (a crash, a collection error)
        # would not show that a pin caught the mistake
        if rc != 1 or " failed" not in out:
            survivors.append((name, rc, out[-600:]))
    assert not survivors, survivors
