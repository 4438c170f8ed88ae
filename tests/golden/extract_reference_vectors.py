"""Transcribes golden vectors from the reference's own unit tests into JSON fixtures.

Reads the Rust test tables under /root/reference (data only — quantifier strings, truth
tables, push orders and expected results) and writes:
  quant_reducer_cases.json  <- test_quant_reducer, src/qsat.rs:2035-2466 (30 cases)

Run once in a container that has the reference mounted:
    python tests/golden/extract_reference_vectors.py
"""
import json
import os
import re
import sys

REF = "/root/reference"
OUT = os.path.dirname(os.path.abspath(__file__))


def strings_in(text):
    return re.findall(r'"([^"]*)"', text)


def matching_paren(text, start):
    depth = 0
    for i in range(start, len(text)):
        if text[i] == "(":
            depth += 1
        elif text[i] == ")":
            depth -= 1
            if depth == 0:
                return i
    raise ValueError("unbalanced parentheses")


def parse_order(text):
    m = re.match(r"\s*vec!\[([^\]]*)\]", text)
    if m:
        return [int(x) for x in m.group(1).replace("\n", " ").split(",") if x.strip()]
    m = re.match(r"\s*\((\d+)\.\.(\d+)\)", text)
    if not m:
        raise ValueError("unknown ordering expression: " + text[:80])
    order = list(range(int(m.group(1)), int(m.group(2))))
    rest = text[m.end():]
    mc = re.match(r"\s*\.chain\(\((\d+)\.\.(\d+)\)\.map\(\|i\|\s*(.*?)\)\)\s*\.collect", rest, re.S)
    if mc:
        expr = mc.group(3)
        if not re.fullmatch(r"[0-9i<>&|+() ]+", expr):
            raise ValueError("unsupported map expression: " + expr)
        fn = eval("lambda i: " + expr)  # arithmetic on i only (checked above)
        order += [fn(i) for i in range(int(mc.group(1)), int(mc.group(2)))]
    return order


def quant_reducer_cases():
    src = open(os.path.join(REF, "src/qsat.rs")).read()
    a = src.index("fn test_quant_reducer()")
    b = src.index("fn test_get_aggr_output_cpu_code_defs()")
    body = src[a:b]
    starts = [m.start() for m in re.finditer(r"str_to_quants\(", body)]
    cases = []
    for k, s in enumerate(starts):
        chunk = body[s: starts[k + 1] if k + 1 < len(starts) else len(body)]
        quants = strings_in(chunk[: chunk.index(")")])[0]
        bi = chunk.index("str_to_bools(")
        be = matching_paren(chunk, bi + len("str_to_bools"))
        bools = "".join(ch for lit in strings_in(chunk[bi:be]) for ch in lit if ch in "01")
        rest = chunk[be + 1:].lstrip().lstrip(",")
        order = parse_order(rest)
        mo = re.search(r"(None|Some\((\d+)\)),\s*FinalResult", rest)
        opt = None if mo.group(1) == "None" else int(mo.group(2))
        mf = re.search(r"reversed:\s*(true|false),\s*solution_bits:\s*(\d+),\s*solution:\s*(None|Some\(([0-9a-fx_b]+)\))",
                       rest)
        sol = None if mf.group(3) == "None" else int(mf.group(4).replace("_", ""), 0)
        cases.append({
            "quants": quants,
            "items": bools,
            "order": order,
            "final_from": opt,
            "expected": {"reversed": mf.group(1) == "true", "solution_bits": int(mf.group(2)),
                         "solution": sol},
        })
    return cases


def main():
    cases = quant_reducer_cases()
    if len(cases) != 30:
        sys.exit(f"expected 30 QuantReducer cases, found {len(cases)}")
    with open(os.path.join(OUT, "quant_reducer_cases.json"), "w") as f:
        json.dump({"source": "src/qsat.rs:2035-2466 (test_quant_reducer)", "cases": cases}, f, indent=1)
    print("quant_reducer_cases.json:", len(cases), "cases")


if __name__ == "__main__":
    main()
