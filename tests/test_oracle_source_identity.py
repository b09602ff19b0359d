"""The oracle's GRP against the reference's source text, token for token.

The star-region time derivatives of the relaxation GRP carry terms in the
impedance slope c_x at u* != u_K whose coefficients (the 0.5 of k_l, k_r,
the (u - u_K) c_x terms of d_l, d_r, the 1.5 of rho_t, the 2 (e_K - e) c_x
term of e_t; src/basicFuns.cuh:151-154, 165, 168, 179, 182) are the
author's modelling choices: no mathematical property fixes them (the other
terms are pinned in test_oracle_pins.py), and the reference cannot be
compiled here (CUDA runtime, thrust and OpenCV).  What defines them is the
reference's source, so this test compares the oracle's `grp`
(oracle/fvm2d_oracle.cpp) with the body of the reference's GRP kernel
`rel_RP_posi_fix` (src/basicFuns.cuh:86-189) token for token, from the
slope preparation (e_lx) to the end of the star region.  The only
differences allowed are spelled out in normalise() (std::sqrt for sqrt, the
oracle's tmax for thrust's max, and the oracle's branch index).  It reads
the reference checkout, so it runs where one is present (FVM2D_REFERENCE_DIR,
default /root/reference) and is skipped elsewhere.
"""
import os
import re

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SRC = os.path.join(HERE, "..", "oracle", "fvm2d_oracle.cpp")
REF_DIR = os.environ.get("FVM2D_REFERENCE_DIR", "/root/reference")
REF_SRC = os.path.join(REF_DIR, "src", "basicFuns.cuh")

TOKEN = re.compile(r"[A-Za-z_][A-Za-z_0-9]*|\d+\.\d*(?:[eE][-+]?\d+)?|\d+(?:[eE][-+]?\d+)?|\.\d+"
                   r"|::|<=|>=|==|!=|&&|\|\||[-+*/=<>(){}\[\];,.!?:]")


def tokens(text):
    text = re.sub(r"/\*.*?\*/", " ", text, flags=re.S)
    text = re.sub(r"//[^\n]*", " ", text)
    return TOKEN.findall(text)


def region(text, start, end):
    i = text.index(start)
    return text[i:text.index(end, i)]


def normalise(toks):
    """The oracle's spelling of the reference's calls, and without its branch
    index (`int br;` and `br = N;`)."""
    out, i = [], 0
    while i < len(toks):
        t = toks[i:i + 4]
        if t[:3] == ["std", "::", "sqrt"]:
            out.append("sqrt")
            i += 3
        elif t[:1] == ["tmax"]:
            out.append("max")
            i += 1
        elif t[:3] == ["int", "br", ";"]:
            i += 3
        elif len(t) == 4 and t[0] == "br" and t[1] == "=" and t[3] == ";":
            i += 4
        else:
            out.append(toks[i])
            i += 1
    return out


def reference_grp_tokens():
    if not os.path.exists(REF_SRC):
        pytest.skip(f"reference checkout not present ({REF_SRC})")
    src = open(REF_SRC).read()
    body = region(src, "__global__ void rel_RP_posi_fix", "u_u_t_interface[2 * idx] =")
    return tokens(body[body.index("double e_lx ="):])


def oracle_grp_tokens(text=None):
    src = text if text is not None else open(ORACLE_SRC).read()
    body = region(src, "void grp(const V5 &fl", "W = V5{rho, pi, u, v, e};")
    return normalise(tokens(body[body.index("double e_lx ="):]))


def first_difference(a, b):
    for k, (x, y) in enumerate(zip(a, b)):
        if x != y:
            return k, " ".join(a[max(0, k - 8):k + 8]), " ".join(b[max(0, k - 8):k + 8])
    return min(len(a), len(b)), "<end>", "<end>"


def test_oracle_grp_is_the_reference_source_token_for_token():
    ref = reference_grp_tokens()
    orc = oracle_grp_tokens()
    assert len(ref) > 900, "reference region not found"
    if ref != orc:
        k, r, o = first_difference(ref, orc)
        pytest.fail(f"oracle grp differs from src/basicFuns.cuh at token {k}:\n"
                    f"  reference: {r}\n  oracle:    {o}")


def test_oracle_flux_is_the_reference_source_token_for_token():
    """The time-accurate flux (src/basicFuns.cuh:14-34; its return
    expression, braces of the aggregate as the oracle's constructor call)."""
    if not os.path.exists(REF_SRC):
        pytest.skip(f"reference checkout not present ({REF_SRC})")
    ref = region(open(REF_SRC).read(), "__device__ vec5d flux(", "// MUSCL reconstruction")
    orc = region(open(ORACLE_SRC).read(), "V5 flux(const V5 &U", "// basicFuns.cuh:49-68")

    def expr(text):
        t = tokens(text[text.index("return"):])
        t = t[:t.index(";")]
        return ["(" if x == "{" else ")" if x == "}" else x for x in t if x != "V5"]

    assert len(expr(ref)) > 150
    assert expr(orc) == expr(ref)


@pytest.mark.parametrize("old,new", [
    ("+ 1.5 * (u - u_r) * c_rx", "+ 1.6 * (u - u_r) * c_rx"),      # rho_t coefficient
    ("+ 0.5 * (u - u_l) * c_lx", "+ 0.5 * (u - u_r) * c_lx"),      # k_l side mix-up
    ("2.0 * (e_l - e) * c_l * c_l * c_lx", "2.0 * (e - e_l) * c_l * c_l * c_lx"),  # e_t sign
    ("c_l / rho_l * (u - u_l) * c_lx", "c_l / rho_l * (u - u_l) * c_rx"),          # d_l slope
])
def test_a_planted_mistake_is_caught(old, new):
    """The comparison has teeth: each single-token mistake in the
    c_x terms, planted in a copy of the oracle's text, fails it."""
    ref = reference_grp_tokens()
    src = open(ORACLE_SRC).read()
    assert src.count(old) == 1
    assert oracle_grp_tokens(src.replace(old, new)) != ref
