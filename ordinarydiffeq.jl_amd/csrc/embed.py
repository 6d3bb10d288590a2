"""Embed text files as one C++ string literal (a sequence of raw-string chunks).

usage: embed.py OUT IN1 [IN2 ...]; an `#include "x.h"` line naming a file of
the same folder that is not itself one of the inputs is replaced by that
file's text (hipRTC sees one self-contained source)."""
import os
import re
import sys

dst, srcs = sys.argv[1], sys.argv[2:]
names = {os.path.basename(s) for s in srcs}


def inline(path):
    out = []
    for line in open(path).read().replace("#pragma once\n", "", 1).splitlines(keepends=True):
        m = re.match(r'\s*#include "([^"]+)"', line)
        if m and m.group(1) not in names:
            inc = os.path.join(os.path.dirname(path), m.group(1))
            if os.path.exists(inc):
                out.append(inline(inc))
                continue
        out.append(line)
    return "".join(out)


text = "#define ODEMI_RTC 1\n" + "\n".join(inline(s) for s in srcs)
chunks = [text[i:i + 8000] for i in range(0, len(text), 8000)]
with open(dst, "w") as f:
    for c in chunks:
        assert ")ODEMI\"" not in c
        f.write('R"ODEMI(' + c + ')ODEMI"\n')
