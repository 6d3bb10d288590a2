"""Embed a text file as a C++ string literal (a sequence of raw-string chunks)."""
import sys

src, dst = sys.argv[1], sys.argv[2]
text = open(src).read()
chunks = [text[i:i + 8000] for i in range(0, len(text), 8000)]
with open(dst, "w") as f:
    for c in chunks:
        assert ")ODEMI\"" not in c
        f.write('R"ODEMI(' + c + ')ODEMI"\n')
