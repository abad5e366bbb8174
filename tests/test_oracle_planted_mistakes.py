"""Each plausible mistake in the oracle's fluid passes fails a pin.

CPU only. For every mistake below, the oracle's sources are copied to a
temporary folder, the mistake is planted in the copy (one textual change),
the copy is built with the oracle's own flags, and the checks of
tests/fluid_pins.py run against it: at least one must fail. The unchanged
oracle passes every check (tests/test_oracle_fluid_pins.py), so the checks
pin what each planted line computes.
"""
import os
import shutil
import subprocess
import tempfile
from concurrent.futures import ThreadPoolExecutor

import pytest

import fluid_pins
import oracle as orc

ORACLE = os.path.dirname(os.path.abspath(orc.__file__))

# name: (file, text in the oracle, text with the mistake)
MISTAKES = {
    "lambda: w_i dropped from the denominator (lambda.ts:60-66)":
        ("orc_sim.h", "SFX(dot)(gs, gs) * s->w[i] + gss", "SFX(dot)(gs, gs) + gss"),
    "lambda: not unilateral (lambda.ts:59-63)":
        ("orc_sim.h", "if (!(C <= R(0.0))) {", "if (1) {"),
    "delta: lambda_i instead of lambda_i + lambda_j (positionDelta.ts:50)":
        ("orc_sim.h", "REAL sc = (li + lj) * SFX(nmass)(s, F, j);", "REAL sc = li * SFX(nmass)(s, F, j);"),
    "XSPH: v_i - v_j (velocity.ts:59-66)":
        ("orc_sim.h", "((vj[a] - vi[a]) * W) * V", "((vi[a] - vj[a]) * W) * V"),
    "vorticity: grad W x (v_i - v_j) (velocity.ts:54)":
        ("orc_sim.h", "SFX(cross)(dv, g, c);", "SFX(cross)(g, dv, c);"),
    "confinement: omega x eta_hat (velocity.ts:110-113)":
        ("orc_sim.h", "SFX(cross)(nrm, F->omega + 3 * (size_t)i, f);", "SFX(cross)(F->omega + 3 * (size_t)i, nrm, f);"),
    "eta: |omega_i| - |omega_j| (velocity.ts:86-88)":
        ("orc_sim.h", "REAL sc = (oj - oi) * F->pv;", "REAL sc = (oi - oj) * F->pv;"),
    "cohesion constant 32 / (pi h^6) (cohesion.ts:47)":
        ("pf_oracle.c", "F->coh_coef = (T)(32.0 / (M_PI * pow(h, 9)));", "F->coh_coef = (T)(32.0 / (M_PI * pow(h, 6)));"),
    "cohesion: density correction 1 / (rho_i + rho_j) (cohesion.ts:88)":
        ("orc_sim.h", "REAL corr = (F->rho0 * R(2.0)) / RFMAX(s->rho[i] + s->rho[j], R(1.0));",
         "REAL corr = R(2.0) / RFMAX(s->rho[i] + s->rho[j], R(1.0));"),
    "adhesion constant 0.007 / h^3 (adhesion.ts:17)":
        ("pf_oracle.c", "(0.007 / pow(h, 3.25))", "(0.007 / pow(h, 3.0))"),
    "neighbour filter <= h^2 (neighborList.ts:55-60)":
        ("orc_sim.h", "if (SFX(dot)(d, d) < F->h2) {", "if (SFX(dot)(d, d) <= F->h2) {"),
    "neighbour walk dx outermost (query.ts:37-39)":
        ("orc_sim.h",
         "uint32_t b = orc_morton((int32_t)((uint32_t)q[0] + (uint32_t)dx), (int32_t)((uint32_t)q[1] + (uint32_t)dy),\n"
         "                                  (int32_t)((uint32_t)q[2] + (uint32_t)dz))",
         "uint32_t b = orc_morton((int32_t)((uint32_t)q[0] + (uint32_t)dz), (int32_t)((uint32_t)q[1] + (uint32_t)dy),\n"
         "                                  (int32_t)((uint32_t)q[2] + (uint32_t)dx))"),
    "poly6 squared instead of cubed (kernels.ts:64-69)":
        ("orc_sim.h", "return ((d * d) * d) * coef;", "return (d * d) * coef;"),
    "spiky gradient sign (kernels.ts:89-107)":
        ("orc_sim.h", "REAL mag = -((t * t) * coef);", "REAL mag = ((t * t) * coef);"),
}


def _cflags():
    text = open(os.path.join(ORACLE, "Makefile")).read()
    line = next(l for l in text.splitlines() if l.startswith("CFLAGS"))
    return line.split("=", 1)[1].split()


def _build(name, tmp):
    fname, old, new = MISTAKES[name]
    d = os.path.join(tmp, str(sorted(MISTAKES).index(name)))
    os.makedirs(d)
    for f in ("pf_oracle.c", "orc_sim.h"):
        shutil.copy(os.path.join(ORACLE, f), d)
    path = os.path.join(d, fname)
    src = open(path).read()
    assert src.count(old) >= 1, f"{name}: the line to change is not in {fname}"
    open(path, "w").write(src.replace(old, new, 1))
    so = os.path.join(d, "liboracle.so")
    flags = [f for f in _cflags() if f != "-O2"] + ["-O1"]
    subprocess.run(["gcc", *flags, "-o", so, os.path.join(d, "pf_oracle.c"), "-lm"], check=True)
    return so


@pytest.fixture(scope="module")
def mutants():
    tmp = tempfile.mkdtemp(prefix="orc_mut_")
    try:
        with ThreadPoolExecutor(max_workers=6) as ex:
            built = dict(zip(MISTAKES, ex.map(lambda n: _build(n, tmp), MISTAKES)))
        yield built
    finally:
        shutil.rmtree(tmp, ignore_errors=True)


@pytest.mark.parametrize("name", list(MISTAKES))
def test_planted_mistake_fails_a_pin(mutants, name):
    library = orc.lib(mutants[name])
    failed = []
    for check in fluid_pins.CHECKS:
        try:
            check(library)
        except AssertionError:
            failed.append(check.__name__)
    assert failed, f"no pin notices the planted mistake: {name}"
