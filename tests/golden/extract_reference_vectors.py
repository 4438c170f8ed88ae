"""Transcribes golden vectors from the reference's own unit tests into JSON fixtures.

Reads the Rust test tables under /root/reference (data only — quantifier strings, truth
tables, push orders and expected results) and writes:
  quant_reducer_cases.json  <- test_quant_reducer, src/qsat.rs:2035-2466 (30 cases)
  mem_image_find_solution.json <- test_mem_image_find_solution,
                               src/greedysatostop.rs:1203-1268 (10 cases)

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


def int_expr(text):
    """value of an integer literal expression of the test tables (<<, *, + and parentheses)"""
    t = text.strip()
    if not re.fullmatch(r"[0-9_<>*+() ]+", t):
        raise ValueError("unsupported expression: " + t)
    return int(eval(t.replace("_", "")))  # digits and integer operators only (checked above)


def mem_image_cases():
    src = open(os.path.join(REF, "src/greedysatostop.rs")).read()
    a = src.index("fn test_mem_image_find_solution()")
    b = src.index(".into_iter()", a)
    body = src[src.index("[", a) + 1: b]
    body = body[: body.rindex("]")]
    cases = []
    i = 0
    while True:
        i = body.find("(", i)
        if i < 0:
            break
        j = matching_paren(body, i)
        tup = body[i + 1: j]
        i = j + 1
        head, _, sol = tup.partition("Some(")
        parts = [x for x in head.split(",")]
        vals = [int_expr(x) for x in parts[:6]]
        if sol:
            ms = re.search(r"start:\s*([^,]+),\s*end:\s*([^,}]+)", sol)
            exp = {"start": int_expr(ms.group(1)), "end": int_expr(ms.group(2))}
        else:
            if parts[6].strip() != "None":
                raise ValueError("unexpected solution field: " + parts[6])
            exp = None
        cases.append({"state_len": vals[0], "start": vals[1], "len": vals[2], "unknowns": vals[3],
                      "pos_sol": vals[4], "next": vals[5], "expected": exp})
    return cases


def main():
    mi = mem_image_cases()
    if len(mi) != 10:
        sys.exit(f"expected 10 mem-image cases, found {len(mi)}")
    with open(os.path.join(OUT, "mem_image_find_solution.json"), "w") as f:
        json.dump({"source": "src/greedysatostop.rs:1203-1268 (test_mem_image_find_solution)", "cases": mi}, f,
                  indent=1)
    print("mem_image_find_solution.json:", len(mi), "cases")
    cases = quant_reducer_cases()
    if len(cases) != 30:
        sys.exit(f"expected 30 QuantReducer cases, found {len(cases)}")
    with open(os.path.join(OUT, "quant_reducer_cases.json"), "w") as f:
        json.dump({"source": "src/qsat.rs:2035-2466 (test_quant_reducer)", "cases": cases}, f, indent=1)
    print("quant_reducer_cases.json:", len(cases), "cases")


if __name__ == "__main__":
    main()
