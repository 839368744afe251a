"""Writes the reference-side IForce3 adapter shown in INTEGRATION.md (the first
C++ block of its section 1) to a header, so that oracle/Makefile can compile it
against the reference's own TetrahedronElementSet3.h. Test infrastructure.

    python3 extract_adapter.py ../INTEGRATION.md _ref/TetrahedronElementSet3_MI355X.h
"""
import re
import sys


def adapter_block(markdown):
    sec = re.search(r"^## 1\. .*?(?=^## 2\. )", markdown, flags=re.S | re.M)
    if not sec:
        raise SystemExit("INTEGRATION.md has no section 1")
    code = re.search(r"```cpp\n(.*?)```", sec.group(0), flags=re.S)
    if not code:
        raise SystemExit("section 1 of INTEGRATION.md has no cpp block")
    return code.group(1)


if __name__ == "__main__":
    with open(sys.argv[1]) as f:
        text = adapter_block(f.read())
    with open(sys.argv[2], "w") as f:
        f.write("#pragma once\n" + text)
