"""Embed text files as one C++ string literal (a sequence of raw-string chunks).

usage: embed.py OUT IN1 [IN2 ...]; `#include "odemi_wg.h"`-style lines that
reference the embedded files themselves are dropped (the text is inlined)."""
import sys

dst, srcs = sys.argv[1], sys.argv[2:]
text = "#define ODEMI_RTC 1\n" + "\n".join(open(s).read().replace("#pragma once\n", "", 1) for s in srcs)
chunks = [text[i:i + 8000] for i in range(0, len(text), 8000)]
with open(dst, "w") as f:
    for c in chunks:
        assert ")ODEMI\"" not in c
        f.write('R"ODEMI(' + c + ')ODEMI"\n')
