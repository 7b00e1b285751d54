"""Write the committed golden fixtures under tests/golden/.

1. ``filter_masks.json``: the literal expected boolean masks held by the
   reference's own tests (exponax tests/test_filter_masks.py), extracted by
   parsing that file as text with ``ast`` (nothing is imported or run). Each
   record is the called function, its arguments and the expected array.
2. ``oracle_*.npz``: seeded inputs and the oracle's FP64 and FP32 outputs at
   1 and 10 steps for reduced batches of the benchmark workloads. They pin
   the oracle against accidental change (tests/test_oracle_regression.py)
   and give the GPU tests a stored reference.

Usage (from the repository root; step 1 needs the reference checkout):
    python tests/golden/make_golden.py [--reference /root/reference]
"""
from __future__ import annotations

import argparse
import ast
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)


def extract_masks(reference: str) -> list:
    path = os.path.join(reference, "tests", "test_filter_masks.py")
    tree = ast.parse(open(path).read())
    out = []
    for node in ast.walk(tree):
        if not (isinstance(node, ast.Call) and getattr(node.func, "attr", "") == "assert_equal"):
            continue
        call, expected = node.args[0], node.args[1]
        fname = call.func.attr
        args = [ast.literal_eval(a) for a in call.args]
        kwargs = {k.arg: ast.literal_eval(k.value) for k in call.keywords}
        if expected.func.attr == "ones":
            # np.ones(<shape of integer constants>, dtype=bool)
            shape_expr = ast.Expression(expected.args[0])
            for n in ast.walk(shape_expr):
                if not isinstance(n, (ast.Expression, ast.Tuple, ast.Constant, ast.BinOp, ast.FloorDiv, ast.Add,
                                      ast.Sub, ast.Mult, ast.Load)):
                    raise ValueError(f"unexpected node in shape: {ast.dump(n)}")
            shape = eval(compile(shape_expr, "<shape>", "eval"), {"__builtins__": {}})
            arr = np.ones(shape, dtype=bool).tolist()
        else:
            arr = ast.literal_eval(expected.args[0])
        out.append({"function": fname, "args": args, "kwargs": kwargs, "expected": arr})
    return out


def oracle_fixtures():
    from oracle import ic as oic
    from oracle import stepper as ost

    cases = {
        # KS conservative 1-D, L=100, N=256, dt=0.1 (the single-trajectory workload)
        "ks_cons_1d": dict(cls="KuramotoSivashinskyConservative", args=(1, 100.0, 256, 0.1), kw={},
                           batch=2, cutoff=5, seed=1),
        # Burgers 1-D, N=1024, L=1, dt=1e-3, per-trajectory viscosity (the batch workload)
        "burgers_1d": dict(cls="Burgers", args=(1, 1.0, 1024, 1e-3), kw={}, batch=8, cutoff=8, seed=2,
                           nu=True),
        # KS (gradient norm) 1-D
        "ks_1d": dict(cls="KuramotoSivashinsky", args=(1, 60.0, 128, 0.05), kw={}, batch=2, cutoff=5, seed=4),
    }
    written = []
    for name, c in cases.items():
        u0, rng = oic.tfs_batch(c["batch"], c["args"][0], c["args"][1], c["args"][2], c["cutoff"], c["seed"])
        res = {"u0": u0}
        for dt_name, dtype in (("f64", np.float64), ("f32", np.float32)):
            cls = getattr(ost, c["cls"])
            if c.get("nu"):
                nus = 10 ** np.random.default_rng(c["seed"] + 1000).uniform(-3.0, -2.0, c["batch"])
                res["nu"] = nus
                st = ost.Ensemble([cls(*c["args"], diffusivity=float(n), dtype=dtype) for n in nus], dtype=dtype)
            else:
                st = cls(*c["args"], dtype=dtype, **c["kw"])
            u = u0.astype(dtype)
            for s in range(1, 11):
                u = st.step_batch(u)
                if s in (1, 10):
                    res[f"{dt_name}_step{s}"] = u.copy()
        path = os.path.join(HERE, f"oracle_{name}.npz")
        np.savez_compressed(path, **res)
        written.append(path)
    return written


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--reference", default="/root/reference")
    args = ap.parse_args()
    if os.path.isdir(args.reference):
        masks = extract_masks(args.reference)
        with open(os.path.join(HERE, "filter_masks.json"), "w") as f:
            json.dump({"source": "exponax tests/test_filter_masks.py (literal expected arrays)",
                       "cases": masks}, f)
        print(f"filter_masks.json: {len(masks)} cases")
    else:
        print("reference checkout not found; filter_masks.json left unchanged")
    for p in oracle_fixtures():
        print("wrote", os.path.relpath(p, ROOT))


if __name__ == "__main__":
    main()
