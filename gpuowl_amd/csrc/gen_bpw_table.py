"""Transcribes the reference's measured max-bits-per-word table (src/fftbpw.h,
one row per FFT shape, 6 values for variants 000,101,202,010,111,212) into
bpw_table.inc, a plain data file in this project's own format:
    {WIDTH, MIDDLE, HEIGHT, {b0, b1, b2, b3, b4, b5}},
The numbers are kept as they are because FFT-size selection must match the
reference's. Run once in a tree that has the reference at /root/reference."""
import os, re
SRC = "/root/reference/src/fftbpw.h"
here = os.path.dirname(os.path.abspath(__file__))

def num(s):
    s = s.strip()
    mult = 1024 if s[-1] in "kK" else (1024 * 1024 if s[-1] in "mM" else 1)
    return int(float(s.rstrip("kKmM")) * mult)

rows = []
for line in open(SRC):
    m = re.match(r'\{\s*"([^"]+)",\s*\{([^}]*)\}\}', line.strip())
    if not m:
        continue
    w, mid, h = (num(x) for x in m.group(1).split(":"))
    vals = [v.strip() for v in m.group(2).split(",") if v.strip()]
    assert len(vals) == 6
    rows.append(f"{{{w}, {mid}, {h}, {{{', '.join(vals)}}}}},")
with open(os.path.join(here, "bpw_table.inc"), "w") as f:
    f.write("// Measured max bits/word per FFT shape (WIDTH, MIDDLE, HEIGHT) for variants\n")
    f.write("// 000,101,202,010,111,212. Data transcribed from the reference src/fftbpw.h:1-93.\n")
    f.write("\n".join(rows) + "\n")
print(len(rows), "rows")
