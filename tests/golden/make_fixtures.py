"""Writes tests/golden/generators.npz: the first elements of every oracle
input generator, so a change to a generator is caught.  Run from the repo
root: python tests/golden/make_fixtures.py"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
import oracle  # noqa: E402

out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "generators.npz")
np.savez(out,
         uniform_i32_seed0=oracle.gen_uniform_i32(0, -1000, 1000, 64),
         uniform_f32_seed0=oracle.gen_uniform_f32(0, 64),
         normal_f64_seed1=oracle.gen_normal(1, 64),
         runs_geometric_seed0=oracle.gen_runs_geometric(0, 256)[0],
         ref_hash_i32=oracle.ref_hash_i32(64),
         ref_hash_mod10_i32=oracle.ref_hash_i32(64, mod10=True))
print("wrote", out)
