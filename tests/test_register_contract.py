"""The native bodies write only registers their templates give them: every destination of a
generated body lies in the clobber list of its template placeholder (read from the template
generator itself), so the template's own values (loop counter, word index, lane masks) survive
the body. Runs on the host: the bodies are generated and read, nothing is launched."""
import importlib.util
import os
import re

import pytest

import naivesat_amd as ns

_GEN = os.path.join(os.path.dirname(__file__), "..", "naivesat_amd", "csrc", "templates", "gen_templates.py")
_spec = importlib.util.spec_from_file_location("gen_templates", _GEN)
gen = importlib.util.module_from_spec(_spec)
_spec.loader.exec_module(gen)

VCC = {"vcc", "vcc_lo", "vcc_hi"}
NO_DEST = ("s_nop", "s_branch", "s_cbranch_vccz")


def destinations(body):
    """(SGPR names, VGPR numbers, AGPR numbers) written by the body's instructions"""
    sg, vg, ag = set(), set(), set()
    for ln in body.splitlines():
        t = ln.strip()
        if not t or t.startswith(";") or t.endswith(":"):
            continue
        op, rest = t.split(None, 1) if " " in t or "\t" in t else (t, "")
        if op in NO_DEST:
            continue
        d = rest.split(",")[0].strip()
        if re.fullmatch(r"s\d+", d) or d in VCC:
            sg.add(d)
        elif re.fullmatch(r"v\d+", d):
            vg.add(int(d[1:]))
        elif re.fullmatch(r"a\d+", d):
            ag.add(int(d[1:]))
        else:
            raise AssertionError(f"unexpected destination in {t!r}")
    return sg, vg, ag


def check(body, cls, sgprs, vgpr_lo, vgpr_hi, outputs, vcc):
    budget, hi = gen.CLASSES[f"r{cls}"]
    sg, vg, ag = destinations(body)
    allowed_s = {f"s{r}" for r in sgprs} | (VCC if vcc else set())
    assert sg <= allowed_s, sorted(sg - allowed_s)
    assert "s20" not in sg and "s21" not in sg  # the word index
    allowed_v = set(range(vgpr_lo, vgpr_hi + 1)) | set(outputs)
    assert vg <= allowed_v, sorted(vg - allowed_v)
    assert not ag or cls == 512, sorted(ag)  # AGPR spill slots only in the r512 class


def sat_pool_hi(cls):
    budget, hi = gen.CLASSES[f"r{cls}"]
    return hi if budget >= 256 else hi - gen.SAT_VCC_POOL_CUT


@pytest.mark.parametrize("kind", [0, 7])
def test_sat_bodies_of_the_36_input_search(kind):
    # the 36-input, 20,000-gate search: most SGPR pressure (grouped output, kept values)
    c, _ = ns.Circuit.gen_sat(2, 36, 20000, 40)
    body, cls, _ = c.native_body(kind, 0)
    check(body, cls, gen.SAT_SGPR_POOL, 8, sat_pool_hi(cls), [0], vcc=True)
    if kind == 0:  # without group exits the body also takes vcc_lo / vcc_hi and s0, s1
        sg, _, _ = destinations(body)
        assert sg & {"vcc_lo", "vcc_hi"} and sg & {"s0", "s1"}
    else:  # the exit tests write vcc, so no value lives there
        sg, _, _ = destinations(body)
        assert not sg & {"vcc_lo", "vcc_hi"}
    assert "s2" not in destinations(body)[0]  # the template's loop counter


def test_small_sat_body():
    c, _ = ns.Circuit.gen_sat(1, 24, 2000, 32)
    body, cls, _ = c.native_body(0, 0)
    assert cls in (64, 128)
    check(body, cls, gen.SAT_SGPR_POOL, 8, sat_pool_hi(cls), [0], vcc=True)


@pytest.mark.parametrize("kind", [1, 2])
def test_qsat_bodies(kind):
    c = ns.Circuit.gen_qsat(3, 32, 10000, 200)
    body, cls, _ = c.native_body(kind, 0)
    hi = gen.CLASSES[f"r{cls}"][1] - (gen.QSATB_POOL_CUT if kind == 2 else 0)
    check(body, cls, gen.SGPR_POOL, 8, hi, [0], vcc=False)


def test_batch_body():
    c, _ = ns.Circuit.gen_sat(5, 24, 4000, 24)
    body, cls, _ = c.native_body(4, 0)
    sg, _, _ = destinations(body)
    # the body's pool; s36..s39 hold the call target and the return address
    assert sg <= {f"s{r}" for r in range(48, 102)}, sorted(sg)
