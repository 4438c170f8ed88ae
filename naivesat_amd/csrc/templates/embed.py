"""Embeds the compiled template assembly files into a C++ source (templates_data.cpp).

Usage: python embed.py OUT.cpp tpl_<kernel>.s ...
"""
import os
import sys

out = ["// generated by embed.py: assembly of the native-path kernel templates", "#include <cstring>",
       "#include <string>", "namespace ns {", "namespace {", "struct Tpl { const char* name; const char* text; };",
       "const Tpl kTemplates[] = {"]
for path in sys.argv[2:]:
    name = os.path.basename(path)[len("tpl_"):-len(".s")]
    text = open(path).read()
    # the batch templates call their bodies (s_swappc_b64 into the region appended at run
    # time); the others hold one body placeholder
    batch = name == "ns_nat_batch"  # 2 kernels x 4 register classes
    ok = text.count("s_swappc_b64 s[38:39], s[36:37]") == 8 if batch else text.count("; NS_BODY " + name) == 1
    if ")NSTPL" in text or not ok:
        sys.exit(f"{path}: placeholder missing or text unusable")
    # every kernel stays within its register class (the suffix _r64 .. _r512): the VGPRs
    # (with AGPRs) its metadata declares decide its occupancy
    kname = None
    for line in text.splitlines():
        t = line.strip()
        if t.startswith(".name:"):
            kname = t.split()[1]
        elif t.startswith(".vgpr_count:") and kname:
            cls = kname.rsplit("_r", 1)[-1]
            if cls.isdigit() and int(t.split()[1]) > int(cls):
                sys.exit(f"{path}: {kname} declares {t.split()[1]} VGPRs, over its class of {cls}")
            kname = None
    out.append(f'  {{"{name}", R"NSTPL({text})NSTPL"}},')
out += ["};", "}  // namespace", "const char* native_template(const std::string& name) {",
        "  for (const Tpl& t : kTemplates)", "    if (name == t.name) return t.text;", "  return nullptr;", "}",
        "}  // namespace ns", ""]
open(sys.argv[1], "w").write("\n".join(out))
