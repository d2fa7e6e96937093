"""Fits this engine's own max bits per word for the FFT shapes whose round-off, at the top of the
range the reference's table gives them, is above the reference's criterion (Gumbel Z = 28 over
per-iteration max round-off, src/tune.cpp:106-172, src/Gpu.h:62-81).

For every shape that the library's selection hands out for exponents in [1M, LIMIT] and that has
no row in engine_bpw.inc yet, it runs Gpu::measureROE (2,000 squarings after 50) at the largest
prime at or below the shape's maxExp (the larger of its variants 101 and 202). Where Z < 32 it
measures again 0.15 and 0.30 bits per word lower and interpolates linearly the bits per word at
Z = 30 (the reference's criterion plus a margin of 2 for the noise of one 2,000-squaring sample);
shapes whose fit lies below the reference's limit get a row {WIDTH, MIDDLE, HEIGHT, bpw},
which FFTConfig takes as a cap on the shape's limit. Capping a shape hands its top exponents to
other shapes, so the library is rebuilt with the new rows and the script run again until it adds
none. Needs a GPU; run from the repository root:
    python gpuowl_amd/csrc/gen_engine_bpw.py [LIMIT] > new.inc && mv new.inc gpuowl_amd/csrc/engine_bpw.inc"""
import os
import sys

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__)))))
import gpuowl_amd as owl  # noqa: E402
from gpuowl_amd import workloads  # noqa: E402

TARGET = 30.0


def parse_num(s):
    return int(float(s.rstrip("KM")) * (1024 if s.endswith("K") else 1024 * 1024 if s.endswith("M") else 1))


def prev_prime(n):
    while not workloads.is_prime(n):
        n -= 1
    return n


def z_at(E, spec):
    with owl.Gpu(E, spec) as g:
        return g.measure_roe(2000, 50)["z"]


HERE = os.path.dirname(os.path.abspath(__file__))


def main():
    limit = int(float(sys.argv[1])) if len(sys.argv) > 1 else 150_000_000
    rows, capped = [], set()
    for line in open(os.path.join(HERE, "engine_bpw.inc")):
        if line.startswith("{"):
            rows.append(line.strip())
            w, m, h = (int(x) for x in line.strip("{},\n ").split(",")[:3])
            capped.add((w, m, h))
    top = {}
    E = 1_000_003
    while E < limit:
        spec, N, mx = owl.best_fit(E)
        shape = ":".join(spec.split(":")[:3])
        if shape not in top or mx > top[shape][2]:
            top[shape] = (spec, N, mx)
        E = mx + 1
    for shape, (spec, N, mx) in top.items():
        if tuple(parse_num(x) for x in shape.split(":")) in capped:
            continue
        E = prev_prime(mx)
        pts = [(E / N, z_at(E, spec))]
        if pts[0][1] < TARGET + 2:
            for d in (0.15, 0.30):
                E2 = prev_prime(int((E / N - d) * N))
                pts.append((E2 / N, z_at(E2, spec)))
        print(f"// {shape}: " + ", ".join(f"bpw {b:.3f} Z {z:.2f}" for b, z in pts), file=sys.stderr, flush=True)
        if pts[0][1] >= TARGET:
            continue
        for (b1, z1), (b2, z2) in zip(pts, pts[1:]):
            if z1 < TARGET <= z2:
                cap = b1 + (b2 - b1) * (TARGET - z1) / (z2 - z1)
                w, m, h = (parse_num(x) for x in shape.split(":"))
                rows.append(f"{{{w}, {m}, {h}, {cap:.3f}}},")
                break
        else:
            raise SystemExit(f"{shape}: Z = {TARGET} not bracketed by {pts}")
    print("// This engine's max bits/word (WIDTH, MIDDLE, HEIGHT, bpw) where it is below the reference's")
    print("// table: fitted at Gumbel Z = 30 on this engine's kernels by gen_engine_bpw.py (MI355X).")
    print("\n".join(rows))


if __name__ == "__main__":
    main()
