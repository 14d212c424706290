"""Writes tests/golden/floquet_qutip.npz, the reference's own golden vectors
for the Floquet rates of the spectral path (integratorType 6).

The reference computes them with QuTiP (scripts/qutipFloquet.py: H = delta/2
sigma_x + A/2 cos(w t) sigma_z with delta = gamma*3e-6, A = gamma*3e-5,
w = 2 pi * 1 kHz, c_op = sigma_y, spectrum 1/((w T)^2 + 1), kmax = 5) and
holds them in two places, both read here as data:
* scripts/floquet_matrices/{Delta,Gamma,A,b}.txt;
* the arrays X_re_ref and X_im_ref of the test that checks
  floquet_master_equation_rates against them (tests/floquet_validation.cpp:
  105-117), which the .txt files carry only as |X|^2 (X.txt).
Nothing is computed: the numbers are parsed and stored.

Run from the repository root (needs the reference checkout at
/root/reference):  python tests/golden/make_floquet_golden.py
"""
import os
import re

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REF = "/root/reference"
NK = 11


def numbers(text):
    return [float(v) for v in re.findall(r"[-+]?\d+\.?\d*(?:[eE][-+]?\d+)?", text)]


def txt(name):
    with open(os.path.join(REF, "scripts", "floquet_matrices", name)) as f:
        return np.array(numbers(f.read()))


def cpp_array(src, name):
    m = re.search(name + r"\[2\]\[2\]\[NK\]\s*=\s*(\{.*?\});", src, re.S)
    return np.array(numbers(m.group(1))).reshape(2, 2, NK)


def main():
    with open(os.path.join(REF, "tests", "floquet_validation.cpp")) as f:
        src = f.read()
    out = dict(
        Delta=txt("Delta.txt").reshape(2, 2, NK),
        Xsq=txt("X.txt").reshape(2, 2, NK),
        Gamma=txt("Gamma.txt").reshape(2, 2, NK),
        A=txt("A.txt").reshape(2, 2),
        b=txt("b.txt"),
        X_re=cpp_array(src, "X_re_ref"),
        X_im=cpp_array(src, "X_im_ref"),
    )
    # the test's copies of the .txt arrays must be the same numbers
    assert np.array_equal(cpp_array(src, "Delta_ref"), out["Delta"])
    assert np.array_equal(cpp_array(src, "Gamma_ref"), out["Gamma"])
    assert np.array_equal(cpp_array(src, "X_ref"), out["Xsq"])
    np.savez(os.path.join(HERE, "floquet_qutip.npz"), **out)
    print({k: v.shape for k, v in out.items()})


if __name__ == "__main__":
    main()
