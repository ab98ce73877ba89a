#!/usr/bin/env python3
"""Writes the golden fixtures of this directory.  Run in the build container
(it reads the reference tree); the fixtures are committed and the tests never
read the reference.

fortran_bm.json
    Data of the reference's own benchmark outputs
    TeaLeaf/Benchmarks/tea_bm_{1..6}.out (Fortran TeaLeaf_ref): grid, steps,
    per-step CG iteration counts, and the step-0 and final field summaries
    (volume, mass, internal energy, temperature).

refhost_oracle.json
    Output of the reference's own host program (TeaLeaf/*.c and
    TeaLeaf/drivers/*.c compiled unchanged, see oracle/Makefile) linked to the
    CPU oracle backend, on a few decks for each solver (CG, Jacobi, Chebyshev,
    PPCG): per-step "CG: N" / "Jacobi: N" iteration counts, the Chebyshev and
    PPCG counts ("Cheby: N iterations (E estimated)", "PPCG: N iterations"),
    the eigenvalue bounds as printed (%.12e) and the final "Actual"
    temperature as printed (%.15e).  This pins the host control flow of this
    project's drivers against the reference's.

config2_4096_oracle.json (`make_golden.py config2`, no reference needed)
    One timestep of the benchmark workload at 4096^2 on this project's host
    and the CPU oracle: iteration count, the four field summaries as exact
    doubles and the SHA-256 of the final interior u, which the GPU must
    reproduce bit for bit.
"""

import json
import os
import re
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, REPO)

REF = "/root/reference/TeaLeaf"
REFHOST = os.path.join(REPO, "oracle", "_ref", "tealeaf_refhost_oracle")


def fortran_fixtures():
    out = {}
    for b in range(1, 7):
        path = os.path.join(REF, "Benchmarks", f"tea_bm_{b}.out")
        text = open(path).read()
        cells = int(re.search(r"x_cells\s*=\s*(\d+)", text).group(1))
        steps = int(re.search(r"end_step\s*=\s*(\d+)", text).group(1))
        iters = [int(v) for v in re.findall(r"Iteration count\s+(\d+)", text)]
        rows = re.findall(r"step:\s+(\d+)\s+(\S+)\s+(\S+)\s+(\S+)\s+(\S+)\s+(\S+)", text)
        summaries = []
        for r in rows:
            step = int(r[0])
            vol, mass, dens, ie, temp = (float(v) for v in r[1:])
            summaries.append(dict(step=step, vol=vol, mass=mass, ie=ie, temp=temp))
        passed = "PASSED" in text and "NOT PASSED" not in text
        out[f"bm{b}"] = dict(cells=cells, steps=steps, iterations=iters,
                             summaries=summaries, passed=passed,
                             source=f"TeaLeaf/Benchmarks/tea_bm_{b}.out")
    return out


def refhost_decks():
    from exploringsycl_amd.deck import Deck, problem_seed
    return {
        "bm1_10x10x10": Deck(10, 10, end_step=10),
        "tea_64x64x1": Deck(64, 64, end_step=1),
        "rect_120x90x3_recip": Deck.with_random_rectangles(
            120, 90, 8, problem_seed(9, 1), end_step=3, coefficient="inverse"),
        "bm2_250x250x10": Deck(250, 250, end_step=10),
        "jacobi_64x64x1": Deck(64, 64, end_step=1, extra=["use_jacobi"]),
        "jacobi_rect_40x30x2": Deck.with_random_rectangles(
            40, 30, 4, problem_seed(9, 2), end_step=2,
            extra=["use_jacobi", "eps 1e-12"]),
        "cheby_250x250x2": Deck(250, 250, end_step=2, extra=["use_chebyshev"]),
        "cheby_errswitch_64x64x1": Deck(64, 64, end_step=1, extra=[
            "use_chebyshev", "errswitch", "epslim 1e-4"]),
        "cheby_rect_recip_120x90x3": Deck.with_random_rectangles(
            120, 90, 8, problem_seed(9, 1), end_step=3, coefficient="inverse",
            extra=["use_chebyshev", "presteps 40"]),
        "ppcg_250x250x2": Deck(250, 250, end_step=2, extra=["use_ppcg"]),
        "ppcg_errswitch_64x64x1": Deck(64, 64, end_step=1, extra=[
            "use_ppcg", "errswitch"]),
        "ppcg_rect_recip_120x90x3": Deck.with_random_rectangles(
            120, 90, 8, problem_seed(9, 1), end_step=3, coefficient="inverse",
            extra=["use_ppcg", "ppcg_inner_steps 6"]),
    }


def run_refhost(deck):
    with tempfile.TemporaryDirectory() as wd:
        deck.write(os.path.join(wd, "tea.in"))
        with open(os.path.join(REPO, "exploringsycl_amd", "data", "tea.problems")) as f:
            open(os.path.join(wd, "tea.problems"), "w").write(f.read())
        p = subprocess.run([REFHOST], cwd=wd, capture_output=True, text=True,
                           env=dict(os.environ, OMP_NUM_THREADS="8"), check=True)
    lines = p.stdout.splitlines()
    iters = [int(l.split()[1]) for l in lines
             if l.startswith("CG:") or l.startswith("Jacobi:")]
    after = []
    for l in lines:
        m = re.match(r"(Cheby|PPCG):\s+(\d+) iterations(?: \((\d+) estimated\))?", l)
        if m:
            after.append([int(m.group(2)), int(m.group(3)) if m.group(3) else -1])
    eig = [l.split()[2] for l in lines if l.startswith("Min. eigenvalue")]
    eig = list(zip(eig, [l.split()[2] for l in lines if l.startswith("Max. eigenvalue")]))
    actual = [l.split()[1] for l in lines if l.startswith("Actual")]
    return dict(iterations=iters, after_switch=after, eigenvalues=[list(e) for e in eig],
                final_temp_printed=actual[-1] if actual else None)


ORACLE_HOST = os.path.join(REPO, "oracle", "build", "tealeaf_oracle")


def config2_fixture(cells=4096, steps=1):
    """The bench workload (4096^2, K = 1/density, tea.in + 16 seeded
    rectangles, eps 1e-15, max 10000 iterations), first timestep, on this
    project's host driving the CPU oracle: iteration counts, the four field
    summaries as exact doubles (hex) and the SHA-256 of the final interior u
    (float64, row-major)."""
    import hashlib
    from exploringsycl_amd.deck import config_deck
    deck = config_deck(2, x_cells=cells, end_step=steps)
    with tempfile.TemporaryDirectory() as wd:
        deck.write(os.path.join(wd, "tea.in"))
        subprocess.run([ORACLE_HOST, "--quiet", "--json", "run.json", "--field-out", "u.bin"],
                       cwd=wd, check=True, env=dict(os.environ, OMP_NUM_THREADS="8"))
        rec = json.load(open(os.path.join(wd, "run.json")))
        with open(os.path.join(wd, "u.bin"), "rb") as f:
            digest = hashlib.sha256(f.read()).hexdigest()
    return dict(cells=cells, steps=steps, deck=deck.render(),
                iterations=[s["iterations"] for s in rec["steps"]],
                executed=[s["executed"] for s in rec["steps"]],
                summaries=[{k: (float(v).hex() if k in ("vol", "mass", "ie", "temp") else v)
                            for k, v in sm.items()} for sm in rec["summaries"]],
                u_sha256=digest)


def main():
    if sys.argv[1:] == ["config2"]:
        fx = config2_fixture()
        with open(os.path.join(HERE, "config2_4096_oracle.json"), "w") as f:
            json.dump(fx, f, indent=1)
        print("wrote config2_4096_oracle.json", fx["iterations"])
        return
    fx = fortran_fixtures()
    with open(os.path.join(HERE, "fortran_bm.json"), "w") as f:
        json.dump(fx, f, indent=1)
    if not os.path.exists(REFHOST):
        sys.exit(f"{REFHOST} missing: make -C oracle ref")
    rh = {}
    for name, deck in refhost_decks().items():
        rh[name] = dict(run_refhost(deck), deck=deck.render())
    with open(os.path.join(HERE, "refhost_oracle.json"), "w") as f:
        json.dump(rh, f, indent=1)
    print("wrote", sorted(fx), sorted(rh))


if __name__ == "__main__":
    main()
