"""Copy the reference's KKT parity fixtures into tests/golden/parity_kkt/.

Source: tests/data/parity/ of the reference (Matrix Market matrices from
interior-point iterations of CUTEst problems, each with a sidecar .json
holding the right-hand side and the KKT inertia (n, m, 0), and the inertia
and relative residual that MUMPS 5.8.2 and SPRAL SSIDS recorded for it).
These are data files (inputs and recorded outputs), copied unchanged so the
tests can run where the reference tree is absent.

Usage: python tests/golden/copy_parity_kkt.py [reference_root]
"""
import os
import shutil
import sys

ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
src = os.path.join(ref, "tests", "data", "parity")
dst = os.path.join(os.path.dirname(os.path.abspath(__file__)), "parity_kkt")
os.makedirs(dst, exist_ok=True)
n = 0
for d in sorted(os.listdir(src)):
    p = os.path.join(src, d)
    if os.path.isdir(p):
        for f in sorted(os.listdir(p)):
            if f.endswith((".mtx", ".json")):
                shutil.copyfile(os.path.join(p, f), os.path.join(dst, f))
                n += 1
    elif d == "manifest.json":
        shutil.copyfile(p, os.path.join(dst, d))
        n += 1
print(f"copied {n} files to {dst}")
