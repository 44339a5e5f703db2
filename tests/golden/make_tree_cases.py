"""Writes tests/golden/tree_cases.json: seeded inputs with Morton keys, the stable sort
permutation and the Karras children from the independent constructions in independent.py."""
import json
from pathlib import Path

import numpy as np

from independent import tree_case

rng = np.random.default_rng(20240603)
cases = []
c = np.array([[i >> 2 & 1, i >> 1 & 1, i & 1] for i in range(8)], np.float32)
cases.append(tree_case("n8_cube_corners", c[:, 0], c[:, 1], c[:, 2], np.ones(8)))
p = rng.normal(0, 1, (64, 3)).astype(np.float32)
p[:16] = p[0]
cases.append(tree_case("n64_sixteen_duplicates", p[:, 0], p[:, 1], p[:, 2], rng.uniform(0.5, 2, 64)))
t = rng.random(33).astype(np.float32)
cases.append(tree_case("n33_line_zero_span", t, np.full(33, 2.0), np.full(33, -1.0), np.ones(33)))
cases.append(tree_case("n2", [0.0, 1.0], [0.0, 0.5], [0.0, 0.25], [1.0, 2.0]))
cases.append(tree_case("n1", [0.3], [0.2], [0.1], [1.0]))
p = rng.normal(0, 5, (100, 3)).astype(np.float32)
cases.append(tree_case("n100_random", p[:, 0], p[:, 1], p[:, 2], rng.uniform(0.5, 2, 100)))
p = rng.integers(0, 4, (200, 3)).astype(np.float32)  # many exact duplicates on a coarse grid
cases.append(tree_case("n200_coarse_grid", p[:, 0], p[:, 1], p[:, 2], np.ones(200)))
out = Path(__file__).resolve().parent / "tree_cases.json"
out.write_text(json.dumps({"cases": cases}))
print(f"wrote {out} ({len(cases)} cases)")
