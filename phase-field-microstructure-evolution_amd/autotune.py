#!/usr/bin/env python3
"""Tuner of the fused step kernels' launch shapes (pf_set_tuning).

For each requested (grid, var, form) it times, on the GPU it runs on, every
candidate of the x-line's shape (lines per block x grid) by the x-line stage's
own HIP-event time, and every candidate of the z-line's shape (grid x XCD run
length) by the update stage's time, over a few steps of the synthetic
polycrystal (pf_put_voronoi).  The winners are then checked on whole steps
(graph replay, no events) against the library's defaults and kept only when
the step is faster.  With --write the table the library compiles in
(csrc/pf_tuned.inc) is rewritten, so every later run of that build uses the
same shapes: the choice is made once, here, not at pf_create.

The launch shapes decide which blocks do what, not what each point or column
computes: a tuned run's state equals the default run's, and its summary
energies (sums of per-block partials) can differ in the last bits.  The z
transforms' radix sequence (zline_plan, FP32 only) is a candidate too; it
changes the rounding of the state within FP32's own accuracy.

    python phase-field-microstructure-evolution_amd/autotune.py --shapes 256:16,512:36 [--write]
    python phase-field-microstructure-evolution_amd/autotune.py --from-json tuning.json
"""
from __future__ import annotations

import argparse
import importlib.util
import itertools
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
INC = os.path.join(HERE, "csrc", "pf_tuned.inc")
FORMS = {"fp32_mixed": 1, "fp64_reduced": 0}

XLINE = dict(xline_lines=(0, 1, 2, 4), xline_grid=(4096, 8192, 16384))
ZLINE = dict(zline_grid=(16384, 32768, 65536), zline_run_rows=(0, 1, 2, 4), zline_plan=(0, 1, 2))


def load_pf():
    if "pf_amd" in sys.modules:
        return sys.modules["pf_amd"]
    spec = importlib.util.spec_from_file_location("pf_amd", os.path.join(HERE, "__init__.py"),
                                                  submodule_search_locations=[HERE])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["pf_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


def log(msg):
    print(f"[autotune] {msg}", file=sys.stderr, flush=True)


def stage_ms(sim, stage, steps, reps):
    """Median over reps of the mean per-launch time of `stage` over `steps` steps."""
    vals = []
    for _ in range(reps):
        sim.set_timing(False)
        sim.set_timing(True)
        sim.step(steps, diag_every=2)
        st = sim.stage_times()[stage]
        vals.append(st["ms"] / max(st["count"], 1))
    sim.set_timing(False)
    vals.sort()
    return vals[len(vals) // 2]


def step_ms(sim, steps, reps):
    vals = []
    for _ in range(reps):
        t0 = time.perf_counter()
        sim.step(steps, diag_every=2)
        vals.append((time.perf_counter() - t0) / steps * 1e3)
    vals.sort()
    return vals[len(vals) // 2]


def tune_one(pf, n, var, form, seeds, steps, reps):
    prm = pf.make_params(t_step=0.005, grad_coeff_phi=1.0, grad_coeff_c=1.0)
    with pf.Simulation(prm, n, n, n, var, form) as sim:
        sim.put_voronoi(seeds)
        sim.set_tuning()  # defaults (the compiled table is ignored while tuning)
        sim.step(4, diag_every=2)
        base_step = step_ms(sim, steps, reps)
        base_x = stage_ms(sim, "nonlinear", steps, reps)
        base_z = stage_ms(sim, "kspace_update", steps, reps)
        log(f"{n}^3 x {var + 1} {form}: default step {base_step:.4f} ms, x-line {base_x:.4f}, z-line {base_z:.4f}")
        best_x, best_x_ms = {}, base_x
        for vals in itertools.product(*XLINE.values()):
            kw = dict(zip(XLINE, vals))
            try:
                sim.set_tuning(**kw)
            except pf.PfError:
                continue
            ms = stage_ms(sim, "nonlinear", steps, reps)
            log(f"  x-line {kw}: {ms:.4f} ms")
            if ms < best_x_ms:
                best_x, best_x_ms = kw, ms
        best_z, best_z_ms = {}, base_z
        for vals in itertools.product(*ZLINE.values()):
            kw = dict(zip(ZLINE, vals))
            try:
                sim.set_tuning(**kw)
            except pf.PfError:
                continue
            ms = stage_ms(sim, "kspace_update", steps, reps)
            log(f"  z-line {kw}: {ms:.4f} ms")
            if ms < best_z_ms:
                best_z, best_z_ms = kw, ms
        chosen = dict(best_x, **best_z)
        sim.set_tuning(**chosen)
        sim.step(2, diag_every=2)
        tuned_step = step_ms(sim, steps, reps)
        sim.set_tuning()
        sim.step(2, diag_every=2)
        base_step2 = step_ms(sim, steps, reps)
        base = min(base_step, base_step2)
        keep = tuned_step < base
        log(f"{n}^3 x {var + 1} {form}: tuned {chosen} step {tuned_step:.4f} ms vs default {base:.4f} ms"
            f" -> {'kept' if keep else 'defaults kept'}")
        t = dict(pf.TUNING_DEFAULTS, **(chosen if keep else {}))
        return dict(grid=n, var=var, form=form, prec=FORMS[form], tuning=t, kept=keep,
                    default_step_ms=base, tuned_step_ms=tuned_step,
                    default_stage_ms={"x": base_x, "z": base_z}, tuned_stage_ms={"x": best_x_ms, "z": best_z_ms})


def write_inc(results):
    lines = ["// Tuned launch shapes, one entry per (nx, ny, nz, var, precision):",
             "//   {nx, ny, nz, var, prec, {xline_lines, xline_grid, zline_grid, zline_run_rows, zline_plan}}",
             "// Written by autotune.py from timings on an MI355X; an empty table means the",
             "// library's defaults everywhere."]
    for r in sorted(results, key=lambda r: (r["grid"], r["var"], r["prec"])):
        if not r["kept"]:
            continue
        t = r["tuning"]
        n = r["grid"]
        lines.append(f"{{{n}, {n}, {n}, {r['var']}, {r['prec']}, {{{t['xline_lines']}, {t['xline_grid']}, "
                     f"{t['zline_grid']}, {t['zline_run_rows']}, {t.get('zline_plan', 0)}}}}},  "
                     f"// {r['form']}: {r['default_step_ms']:.4f} -> {r['tuned_step_ms']:.4f} ms/step")
    with open(INC, "w") as f:
        f.write("\n".join(lines) + "\n")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--shapes", default="256:16", help="comma list of grid:var (cubic grids)")
    ap.add_argument("--forms", default="fp32_mixed", help="comma list of fp32_mixed, fp64_reduced")
    ap.add_argument("--seeds", type=int, default=64)
    ap.add_argument("--steps", type=int, default=8)
    ap.add_argument("--reps", type=int, default=3)
    ap.add_argument("--write", action="store_true", help="rewrite csrc/pf_tuned.inc with the winners")
    ap.add_argument("--from-json", metavar="FILE",
                    help="no timing: rewrite csrc/pf_tuned.inc from the JSON line a tuning run printed")
    args = ap.parse_args()
    if args.from_json:
        with open(args.from_json) as f:
            results = [r for line in f if line.startswith("{") for r in json.loads(line).get("autotune", [])]
        write_inc(results)
        log(f"wrote {INC} from {args.from_json} ({len(results)} shapes)")
        return
    pf = load_pf()
    pf.lib()
    results = []
    for form in args.forms.split(","):
        for sh in args.shapes.split(","):
            n, var = (int(x) for x in sh.split(":"))
            results.append(tune_one(pf, n, var, form, args.seeds, args.steps, args.reps))
    print(json.dumps({"autotune": results}), flush=True)
    if args.write:
        write_inc(results)
        log(f"wrote {INC}")


if __name__ == "__main__":
    main()
