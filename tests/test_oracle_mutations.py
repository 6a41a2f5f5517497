"""The SI pins of tests/oracle_pins.py are sharp enough to catch a wrong constant: every constant of
the oracle's thermodynamic equations, Koehler factors, ventilation fit, Rogers-Yau / Rogers-GK
velocities and the physical constants they use is scaled by 1 +- 1e-6 in a copy of
oracle/cleo_oracle.cpp (on the CPU, in a temporary directory), the copy is compiled, and at least
one pin must fail on it.  The unmutated copy, built the same way, must pass every pin."""
import ctypes as C
import os
import re
import subprocess
import tempfile
from concurrent.futures import ThreadPoolExecutor

import pytest

from cleo_amd.abi import Backend
from tests import oracle_pins

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "oracle", "cleo_oracle.cpp")

# (start of the region in the oracle source, text inside it that holds the constant, the constant)
MUTATIONS = [
    ("static inline double saturation_pressure(", "A = 17.4146", "17.4146"),
    ("static inline double saturation_pressure(", "B = 33.639", "33.639"),
    ("static inline double saturation_pressure(", "TREF = 273.16", "273.16"),
    ("static inline double saturation_pressure(", "PREF = 611.655", "611.655"),
    ("static inline double diffusion_factor(", "A = 7.11756e-5", "7.11756e-5"),
    ("static inline double diffusion_factor(", "B = 4.38127686e-3", "4.38127686e-3"),
    ("static inline double diffusion_factor(", "D = 4.012182971e-5", "4.012182971e-5"),
    ("static inline double diffusion_factor(", "pow(TEMP, 1.94)", "1.94"),
    ("static inline double diffusion_factor(", "pow(TEMP, 2.0)", "2.0"),
    ("static inline double ventilation_factor(", "vent_factor_max = 20", "20"),
    ("static inline double ventilation_factor(", "C1 = 6.954e+7", "6.954e+7"),
    ("static inline double ventilation_factor(", "A = 1.963", "1.963"),
    ("static inline double ventilation_factor(", "C2 = 1.069e+3", "1.069e+3"),
    ("static inline double ventilation_factor(", "B = 0.702", "0.702"),
    ("static inline void kohler_factors(", "3.3e-7", "3.3e-7"),
    ("static inline void kohler_factors(", "4.3e-6", "4.3e-6"),
    ("static inline double tv_rogersyau(", "r1 = 3e-5", "3e-5"),
    ("static inline double tv_rogersyau(", "r2 = 6e-4", "6e-4"),
    ("static inline double tv_rogersyau(", "r3 = 2e-3", "2e-3"),
    ("static inline double tv_rogersyau(", "k1 = 1.19e8", "1.19e8"),
    ("static inline double tv_rogersyau(", "k2 = 8000", "8000"),
    ("static inline double tv_rogersyau(", "k3 = 201", "201"),
    ("static inline double tv_rogersyau(", "k4 = 9", "9"),
    ("static inline double tv_rogersgk(", "radius0 = 3.725", "3.725"),
    ("static inline double tv_rogersgk(", "2.0 * 4.0", "4.0"),
    ("static inline double tv_rogersgk(", "2.0 * 12.0", "12.0"),
    ("static inline double tv_rogersgk(", "acaps = 9.65", "9.65"),
    ("static inline double tv_rogersgk(", "bcaps = 10.43", "10.43"),
    ("static inline double tv_rogersgk(", "2.0 * 0.6", "0.6"),
    ("namespace DC {", "RGAS_UNIV = 8.314462618", "8.314462618"),
    ("namespace DC {", "MR_WATER = 0.01801528", "0.01801528"),
    ("namespace DC {", "MR_DRY = 0.028966216", "0.028966216"),
    ("namespace DC {", "LATENT_V = 2500930", "2500930"),
    ("namespace DC {", "CP_DRY = 1004.64", "1004.64"),
    ("namespace DC {", "CP_V = 1865.01", "1865.01"),
    ("namespace DC {", "C_L = 4192.664", "4192.664"),
    ("namespace DC {", "RHO_L = 998.203", "998.203"),
]


def mutate(src, start, text, const, factor):
    i = src.index(start)
    j = src.index("\n}", i)
    region = src[i:j]
    assert region.count(text) == 1, f"{text!r} is not unique after {start!r}"
    return src[:i] + region.replace(text, text.replace(const, f"({const}*{factor!r})")) + src[j:]


def build(src_text, tag, tmp):
    path = os.path.join(tmp, f"orc_{tag}.cpp")
    with open(path, "w") as f:
        f.write(src_text)
    so = os.path.join(tmp, f"liborc_{tag}.so")
    # the oracle includes "../include/cleo_hip.h" relative to oracle/
    subprocess.run(["g++", "-std=c++20", "-O1", "-fPIC", "-shared", "-ffp-contract=off",
                    "-iquote", os.path.join(ROOT, "oracle"), "-o", so, path], check=True, capture_output=True)
    return so


def failing_pins(so):
    b = Backend(C.CDLL(so), "orc_")
    lib = b.lib
    for n in ("orc_saturation_pressure", "orc_ventilation_factor"):
        getattr(lib, n).argtypes = [C.c_double]
        getattr(lib, n).restype = C.c_double
    lib.orc_diffusion_factor.argtypes = [C.c_double] * 3
    lib.orc_diffusion_factor.restype = C.c_double
    lib.orc_supersaturation_ratio.argtypes = [C.c_double] * 3
    lib.orc_supersaturation_ratio.restype = C.c_double
    lib.orc_terminalv.argtypes = [C.c_int, C.c_double]
    lib.orc_terminalv.restype = C.c_double
    bad = []
    for pin in oracle_pins.PINS:
        try:
            pin(b)
        except AssertionError:
            bad.append(pin.__name__)
    return bad


def test_every_constant_perturbation_fails_a_pin():
    with open(SRC) as f:
        src = f.read()
    with tempfile.TemporaryDirectory() as tmp:
        jobs = [("clean", src)]
        for k, (start, text, const) in enumerate(MUTATIONS):
            for sgn, fac in (("p", 1 + 1e-6), ("m", 1 - 1e-6)):
                jobs.append((f"{k}{sgn}", mutate(src, start, text, const, fac)))
        with ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 1)) as ex:
            sos = dict(zip([t for t, _ in jobs], ex.map(lambda j: build(j[1], j[0], tmp), jobs)))
        assert failing_pins(sos["clean"]) == []
        survivors = []
        for k, (start, text, const) in enumerate(MUTATIONS):
            for sgn in ("p", "m"):
                if not failing_pins(sos[f"{k}{sgn}"]):
                    survivors.append(f"{text} x (1 {'+' if sgn == 'p' else '-'} 1e-6)")
        assert not survivors, f"perturbations no pin detects: {survivors}"


@pytest.mark.parametrize("start,text,const", MUTATIONS[:1])
def test_mutation_helper_edits_one_constant(start, text, const):
    with open(SRC) as f:
        src = f.read()
    out = mutate(src, start, text, const, 2.0)
    assert out.count(f"({const}*2.0)") == 1 and len(re.findall(re.escape(text), out)) == len(re.findall(
        re.escape(text), src)) - 1
