"""Writes programs_expected.json: the expected output lines of the blocks of the reference's
tests/programs/*.dd files that the plan-replay tests run (data only: each block's name, kind
and the lines after its `----` separator). Run in a checkout with the reference at
/root/reference; the GPU tests read only the JSON."""
import json
import os
import sys

REF = "/root/reference/tests/programs"
FILES = ["arrays.dd", "arith.dd", "zip_map.dd", "grouping.dd", "sum_rewrites.dd"]


def blocks(text):
    out, name, kind, lines, after = {}, None, None, [], False
    for line in text.splitlines() + ["# __end__"]:
        if line.startswith("# ") and (line[2:].strip().replace("_", "").isalnum() or line == "# __end__"):
            if name is not None and kind is not None:
                out[name] = {"kind": kind, "expected": [x for x in lines if x.strip()]}
            name, kind, lines, after = line[2:].strip(), None, [], False
            continue
        if name is None:
            continue
        if kind is None and line.strip() in ("run", "error"):
            kind = line.strip()
        elif line.strip() == "----":
            after = True
        elif after:
            lines.append(line)
    return out


def main():
    res = {}
    for f in FILES:
        text = open(os.path.join(REF, f)).read()
        for name, b in blocks(text).items():
            res[f"{f}:{name}"] = b
    dst = os.path.join(os.path.dirname(os.path.abspath(__file__)), "programs_expected.json")
    json.dump(res, open(dst, "w"), indent=1, sort_keys=True)
    print(f"{len(res)} blocks -> {dst}", file=sys.stderr)


if __name__ == "__main__":
    main()
