"""The oracle's pin checks (tests/test_oracle_pins.py, tests/test_oracle_br2.py)
must fail on an oracle with a plausible restatement mistake.  Each case
compiles the oracle sources with one planted mistake (a preprocessor switch
the Makefile never sets) into a temporary directory and runs the pin checks
against that build in a child process (CHARM_ORACLE_LIB), expecting the named
checks to fail and every other pin check to pass.  CPU only; gcc builds each
variant in about a second.
"""
import os
import subprocess
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE = os.path.join(os.path.dirname(HERE), "oracle")
SOURCES = ["or_numerics.c", "or_model.c", "or_ns.c", "or_amr.c"]
PIN_FILES = ["test_oracle_pins.py", "test_oracle_br2.py"]

# macro -> the checks it must fail
MUTANTS = {
    # sign of (S* - u_k) in the HLLC star-region flux (src/charm_fluxes.c:319-321)
    "OR_MUTANT_HLLC_SIGN": ["test_hllc_star_region_equals_toro_star_states"],
    # psi evaluated at the face centres instead of the vertices
    # (src/limiters/charm_limiter_bj.c:279-299)
    "OR_MUTANT_BJ_FACE_CENTRES": ["test_bj_vertex_psi_hand_value"],
    # BR2: BR1's auxiliary tau_h in place of the own-trace stress, and the
    # lifting's sign flipped
    "OR_MUTANT_BR2_TAU_H": ["test_br2_shear_layer_closed_form", "test_br2_hanging_shear_layer_closed_form"],
    "OR_MUTANT_BR2_LIFT_SIGN": ["test_br2_shear_layer_closed_form", "test_br2_hanging_shear_layer_closed_form"],
}


def _build(tmp_path, macro):
    out = os.path.join(str(tmp_path), "liboracle_%s.so" % macro.lower())
    cmd = ["gcc", "-O2", "-fPIC", "-std=c99", "-ffp-contract=off", "-shared", "-D" + macro, "-o", out]
    cmd += [os.path.join(ORACLE, s) for s in SOURCES] + ["-lm"]
    subprocess.run(cmd, check=True, capture_output=True)
    return out


def _run_pins(lib, select):
    env = dict(os.environ, CHARM_ORACLE_LIB=lib)
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-p", "no:cacheprovider"] +
                       [os.path.join(HERE, f) for f in PIN_FILES] + ["-k", select],
                       env=env, capture_output=True, text=True, cwd=HERE)
    return r.returncode, r.stdout + r.stderr


@pytest.mark.parametrize("macro", sorted(MUTANTS))
def test_pin_checks_catch_planted_mistake(tmp_path, macro):
    lib = _build(tmp_path, macro)
    targets = MUTANTS[macro]
    for t in targets:
        rc, out = _run_pins(lib, t)
        assert rc == 1 and "failed" in out, (t, out[-2000:])
    # the same build passes the checks the mistake does not touch
    rc, out = _run_pins(lib, " and ".join("not %s" % t for t in targets))
    assert rc == 0, out[-2000:]
