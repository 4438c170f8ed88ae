"""Transcribes the hash-map test data of the reference's naivesatostop into a JSON fixture.

Reads the data-building functions of src/satostop.rs under /root/reference (test data only:
state widths, output words, hash entries, unknown-fill counters and expected results) and
writes hashmap_cases.json:
  join_to_hashmap            <- join_to_hashmap_cpu_testcase_data_1..3 (src/satostop.rs:1917-2607)
  join_hashmap_itself        <- join_hashmap_itself_and_check_solution_data (2772-3443)
  add_to_hashmap             <- add_to_hashmap_and_check_solution_data_1..2 (3706-4146)
Those functions are written as straight-line Rust statements (lets, indexed stores,
`hashmap_insert` calls); this script evaluates that statement subset. `hashmap_insert` places
an entry at hash_function_64(state_len, current) >> (state_len - hbits), the hash of
src/satostop.rs:127-136, restated below. Only non-default entries are written out.

Run once in a container that has the reference mounted:
    python tests/golden/extract_hashmap_vectors.py
"""
import json
import os
import re
import sys

REF = "/root/reference"
OUT = os.path.dirname(os.path.abspath(__file__))
M64 = (1 << 64) - 1
STATES = {"HASH_STATE_UNUSED": 0, "HASH_STATE_USED": 1, "HASH_STATE_STOPPED": 2, "HASH_STATE_LOOPED": 3}


def hash_function_64(bits, value):
    """src/satostop.rs:127-136 (u64 wrapping arithmetic)"""
    mask = (1 << bits) - 1
    half = bits >> 1
    temp = (value * 9615409803190489167) & M64
    temp = ((temp << half) & M64) | (temp >> (bits - half))
    h = ((value * 6171710485021949031) & M64) ^ temp ^ 0xb89d2ecda078ca1f
    return h & mask


def function_body(src, name):
    a = src.index("fn " + name + "(")
    b = src.index("{", src.index(")", a))
    depth, i = 0, b
    while True:
        if src[i] == "{":
            depth += 1
        elif src[i] == "}":
            depth -= 1
            if depth == 0:
                return src[b + 1:i]
        i += 1


def strip_comments(text):
    return "\n".join(line.split("//")[0] for line in text.splitlines())


def statements(text):
    """top-level statements split at ';' (brace/paren depth 0)"""
    out, cur, depth = [], "", 0
    for ch in text:
        cur += ch
        if ch in "{([":
            depth += 1
        elif ch in "})]":
            depth -= 1
            if ch == "}" and depth == 0 and cur.strip().startswith("for "):
                out.append(cur.strip())  # a loop statement ends with its block
                cur = ""
        elif ch == ";" and depth == 0:
            out.append(cur.strip()[:-1].strip())
            cur = ""
    if cur.strip():
        out.append(cur.strip())
    return out


class Env:
    def __init__(self):
        self.scal = {}
        self.vecs = {}

    def ev(self, expr):
        e = re.sub(r"\bas\s+(u32|u64|usize)\b", "", expr)
        e = re.sub(r"\b(\d+)(u32|u64|usize)\b", r"\1", e)
        e = re.sub(r"(0x[0-9a-fA-F_]+|\b\d[\d_]*)", lambda m: m.group(1).replace("_", ""), e)
        for name in re.findall(r"[A-Za-z_][A-Za-z_0-9]*", e):
            if not (name.startswith("0x") or name in self.scal):
                if not re.fullmatch(r"x[0-9a-fA-F]+", name):
                    raise ValueError(f"unknown name {name} in {expr}")
        if not re.fullmatch(r"[0-9a-fA-Fx_A-Za-z\s()+\-*<>|&^]+", e):
            raise ValueError("unsupported expression: " + expr)
        e = " ".join(e.split())
        return int(eval(e, {"__builtins__": {}}, dict(self.scal)))  # integer arithmetic only (checked above)

    def entry(self, text):
        fields = dict(re.findall(r"(\w+)\s*:\s*([^,}]+)", text))
        st = fields["state"].strip()
        return {"current": self.ev(fields["current"]) & M64, "next": self.ev(fields["next"]) & M64,
                "steps": self.ev(fields["steps"]) & M64, "predecessors": self.ev(fields["predecessors"]),
                "state": STATES[st] if st in STATES else self.ev(st)}


def run(env, stmts):
    for s in stmts:
        s = s.strip()
        if not s:
            continue
        m = re.fullmatch(r"let\s+(mut\s+)?(\w+)\s*(:\s*\w+)?\s*=\s*([\s\S]+)", s)
        if m:
            name, rhs = m.group(2), m.group(4).strip()
            mv = re.fullmatch(r"vec!\[\s*0u32\s*;\s*([\s\S]+)\]", rhs)
            mh = re.fullmatch(r"vec!\[\s*(HashEntry::default\(\)|HashEntry\s*\{[\s\S]*?\})\s*;\s*([\s\S]+)\]", rhs)
            if mv:
                env.vecs[name] = {"kind": "u32", "len": env.ev(mv.group(1)), "data": {}}
            elif mh:
                env.vecs[name] = {"kind": "hash", "len": env.ev(mh.group(2)), "data": {}}
            elif rhs.startswith("{"):
                # block expression: its own scope, value = the last expression
                inner = statements(rhs[1:rhs.rindex("}")])
                saved = dict(env.vecs)
                run(env, inner[:-1])
                value = env.vecs[inner[-1].strip()]
                env.vecs = saved
                env.vecs[name] = value
            elif rhs.startswith("create_vec_of_atomic_u32("):
                env.vecs[name] = {"kind": "u32", "len": env.ev(rhs[len("create_vec_of_atomic_u32("):-1]), "data": {}}
            elif "filter(" in rhs and ".count()" in rhs:
                src = re.match(r"(?:Arc::new\(AtomicU64::new\()?\s*(\w+)\s*\.iter\(\)", rhs).group(1)
                ufb = env.scal["unknown_fill_bits"]
                env.scal[name] = sum(1 for v in env.vecs[src]["data"].values() if (v >> ufb) != 0)
            elif rhs.startswith("Mutex::new"):
                pass
            elif rhs.startswith("Some(Solution"):
                f = dict(re.findall(r"(\w+)\s*:\s*([^,}]+)", rhs))
                env.scal[name] = {"start": env.ev(f["start"]), "end": env.ev(f["end"]), "steps": env.ev(f["steps"])}
            elif rhs == "None":
                env.scal[name] = None
            else:
                env.scal[name] = env.ev(rhs)
            continue
        m = re.fullmatch(r"(\w+)\[([\s\S]+?)\]\s*=\s*([\s\S]+)", s)
        if m:
            v = env.vecs[m.group(1)]
            idx = env.ev(m.group(2))
            assert 0 <= idx < v["len"]
            v["data"][idx] = env.entry(m.group(3)) if v["kind"] == "hash" else env.ev(m.group(3)) & 0xFFFFFFFF
            continue
        m = re.fullmatch(r"hashmap_insert\(\s*state_len\s*,\s*hbits\s*,\s*&mut\s+(\w+)\s*,\s*(HashEntry\s*\{[\s\S]*\})\s*,?\s*\)", s)
        if m:
            v = env.vecs[m.group(1)]
            e = env.entry(m.group(2))
            sl, hb = env.scal["state_len"], env.scal["hbits"]
            assert e["current"] >> sl == 0 and e["next"] >> sl == 0
            idx = hash_function_64(sl, e["current"]) >> (sl - hb)
            assert idx not in v["data"], ("slot taken", idx)
            v["data"][idx] = e
            continue
        m = re.fullmatch(r"(\w+)\[([\s\S]+?)\]\.store\(([\s\S]+),\s*atomic::Ordering::SeqCst\s*,?\s*\)", s)
        if m:
            env.vecs[m.group(1)]["data"][env.ev(m.group(2))] = env.ev(m.group(3))
            continue
        m = re.fullmatch(r"for i in 0\.\.([\s\S]+?)\{\s*(\w+)\[i\]\.store\(\s*(\w+)\[i\]\.load[\s\S]*\}", s)
        if m:
            env.vecs[m.group(2)]["data"] = dict(env.vecs[m.group(3)]["data"])
            continue
        if re.match(r"\(|\w+Data\s*\{|JoinHashMap|AddToHashMap", s):
            continue  # the returned tuple / struct
        raise ValueError("unsupported statement: " + s[:120])


def dataset(src, name):
    env = Env()
    run(env, statements(strip_comments(function_body(src, name))))
    return env


def hash_vec(v):
    return {"len": v["len"], "entries": {str(k): e for k, e in sorted(v["data"].items())}}


def u32_vec(v):
    return {"len": v["len"], "values": {str(k): x for k, x in sorted(v["data"].items()) if x}}


def main():
    src = strip_comments(open(os.path.join(REF, "src/satostop.rs")).read())
    out = {"source": "src/satostop.rs (hash-map test data)", "join_to_hashmap": [], "add_to_hashmap": []}
    for k in (1, 2, 3):
        e = dataset(src, f"join_to_hashmap_cpu_testcase_data_{k}")
        out["join_to_hashmap"].append({
            "state_len": e.scal["state_len"], "arg_bit_place": e.scal["arg_bit_place"], "arg": e.scal["arg"],
            "outputs": u32_vec(e.vecs["outputs"]), "hashmap": hash_vec(e.vecs["hashmap"]),
            "expected_hashmap": hash_vec(e.vecs["expected_hashmap"])})
    e = dataset(src, "join_hashmap_itself_and_check_solution_data")
    out["join_hashmap_itself"] = {
        "state_len": e.scal["state_len"], "hbits": e.scal["hbits"], "unknown_bits": e.scal["unknown_bits"],
        "unknown_fill_bits": e.scal["unknown_fill_bits"], "hashmap": hash_vec(e.vecs["hashmap"]),
        "expected_hashmap": hash_vec(e.vecs["expected_hashmap"]),
        "unknown_fills": u32_vec(e.vecs["unknown_fills"]),
        "expected_unknown_fills": u32_vec(e.vecs["expected_unknown_fills"]),
        "resolved_unknowns": e.scal["resolved_unknowns"],
        "expected_resolved_unknowns": e.scal["expected_resolved_unknowns"],
        "expected_solution": e.scal["expected_solution"]}
    for k in (1, 2):
        e = dataset(src, f"add_to_hashmap_and_check_solution_data_{k}")
        out["add_to_hashmap"].append({
            "state_len": e.scal["state_len"], "arg_bit_place": e.scal["arg_bit_place"], "arg": e.scal["arg"],
            "hbits": e.scal["hbits"], "unknown_bits": e.scal["unknown_bits"],
            "unknown_fill_bits": e.scal["unknown_fill_bits"], "max_predecessors": 1,
            "outputs": u32_vec(e.vecs["outputs"]), "hashmap": hash_vec(e.vecs["hashmap"]),
            "expected_hashmap": hash_vec(e.vecs["expected_hashmap"]),
            "unknown_fills": u32_vec(e.vecs["unknown_fills"]),
            "expected_unknown_fills": u32_vec(e.vecs["expected_unknown_fills"]),
            "resolved_unknowns": e.scal["resolved_unknowns"],
            "expected_resolved_unknowns": e.scal["expected_resolved_unknowns"],
            "expected_solution": e.scal["expected_solution"]})
    with open(os.path.join(OUT, "hashmap_cases.json"), "w") as f:
        json.dump(out, f, indent=1)
    print("hashmap_cases.json:", {k: len(v) if isinstance(v, list) else 1 for k, v in out.items() if k != "source"})


if __name__ == "__main__":
    sys.exit(main())
