#!/usr/bin/env python3
"""Checks that the oracle's pins catch one-constant slips (CPU only).

TEST INFRASTRUCTURE ONLY. Plants one mutation at a time in oracle.cpp -- a
float literal of the chosen functions scaled by 1.1 -- builds the mutant into
a temporary library, runs the oracle pin tests against it (ORACLE_LIB) and
reports the mutants no test fails ("survivors"). oracle/liboracle.so is not
touched.

    python oracle/mutation_sweep.py --seed 7 --count 120 [--set step|render|extensions]

`step` mutates advect-react, force, light, metrics, beds, MacCormack, the
room colour and its indirect light; `render` the raymarch, embers, tonemap,
reset, velocity advection, projection and the source-model overlay;
`extensions` the curl-noise potential, the directional-light transmittance
and its use in the render, and the gizmos.
"""
import argparse
import os
import random
import re
import shutil
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
SRC = os.path.join(HERE, "oracle.cpp")

SETS = {
    "step": [
        ("advect_react", r"^static void advect_react", r"^static void force_boundary_divergence"),
        ("force", r"^static void force_boundary_divergence", r"^static void sor_half"),
        ("light", r"^static void light_build", r"^static void light_pass"),
        ("room", r"^static V3 room_background", r"^static bool intersect_box"),
        ("camera", r"^static V3 camera_response", r"^static void tonemap"),
        ("metrics", r"^static void metrics", r"^static void run_stage"),
        ("beds", r"^static inline R capsule_mask", r"^static void reset"),
        ("maccormack", r"^static inline R maccormack", r"^static inline V4 sample_light"),
        ("indirect", r"^static V3 gather_irradiance", r"^static inline R rect_mask"),
    ],
    "render": [
        ("render", r"^static V4 render_pixel", r"^static void splat_embers"),
        ("embers", r"^static void splat_embers", r"^static void render_hdr"),
        ("tonemap", r"^static void tonemap", r"^static void metrics"),
        ("reset", r"^static void reset", r"^static void advect_velocity"),
        ("advect_velocity", r"^static void advect_velocity", r"^static void advect_react"),
        ("project", r"^static void project", r"^static V3 blackbody"),
        ("overlay", r"^static void source_model_overlay", r"^static V3 room_background"),
    ],
    "extensions": [
        ("curl", r"^static R curl_edge_potential", r"^// -+ force \+ boundary"),
        ("dir_trans", r"^static const R kDirKappa", r"^// Build, then two propagations"),
        ("gizmos", r"^static inline R segment_alpha", r"^static void tonemap"),
    ],
}
TESTS = ["tests/test_oracle_pins.py", "tests/test_oracle.py", "tests/test_oracle_extensions.py"]


def candidates(lines, spans):
    lit = re.compile(r"(?<![\w.])(\d+\.\d+)f")
    out = []
    for name, start, end in spans:
        s = next(i for i, line in enumerate(lines) if re.search(start, line))
        e = next(i for i in range(s + 1, len(lines)) if re.search(end, lines[i]))
        for i in range(s, e):
            if lines[i].strip().startswith("//"):
                continue
            out += [(name, i, m.start(1), m.end(1), m.group(1)) for m in lit.finditer(lines[i])]
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--count", type=int, default=60)
    ap.add_argument("--set", choices=sorted(SETS), default="step")
    args = ap.parse_args()
    lines = open(SRC).read().split("\n")
    cands = candidates(lines, SETS[args.set])
    random.seed(args.seed)
    pick = random.sample(cands, min(args.count, len(cands)))
    tmp = tempfile.mkdtemp(prefix="oracle_mutants_")
    lib = os.path.join(tmp, "liboracle_mutant.so")
    env = dict(os.environ, ORACLE_LIB=lib)
    survivors = []
    try:
        for k, (name, i, a, b, val) in enumerate(pick):
            new = repr(float(val) * 1.1) if float(val) != 0 else "0.5"
            mutated = lines[:]
            mutated[i] = mutated[i][:a] + new + mutated[i][b:]
            src = os.path.join(tmp, "oracle_mutant.cpp")
            text = "\n".join(mutated).replace('#include "../include/firesim.h"', f'#include "{os.path.join(REPO, "include", "firesim.h")}"')
            open(src, "w").write(text)
            r = subprocess.run(["g++", "-O1", "-march=x86-64-v3", "-fopenmp", "-fPIC", "-std=c++17", "-shared", "-o", lib, src],
                               capture_output=True)
            if r.returncode:
                print(f"{k:3d} {name} line {i + 1}: build failed", flush=True)
                continue
            t = subprocess.run([sys.executable, "-m", "pytest", "-x", "-q", "-p", "no:cacheprovider", *TESTS], cwd=REPO, env=env,
                               capture_output=True, text=True, timeout=900)
            killed = t.returncode != 0
            print(f"{k:3d} {name:16s} line {i + 1:5d} {val:>10s} -> {new:>22s}  {'killed' if killed else 'SURVIVED'}", flush=True)
            if not killed:
                survivors.append((name, i + 1, val, lines[i].strip()))
    finally:
        shutil.rmtree(tmp, ignore_errors=True)
    print(f"survivors: {len(survivors)} of {len(pick)}")
    for s in survivors:
        print(s)
    return 0


if __name__ == "__main__":
    sys.exit(main())
