"""Plain form through worker processes (MoleculePool): the same energies and
iteration counts as one-at-a-time runs of the same molecules.

The pool's workers are started by fork + exec, which must come before the
starting process uses the GPU; this test's process already has, so the check
runs in a fresh interpreter started as a child process."""
import json
import os
import subprocess
import sys

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

SCRIPT = r"""
import json, sys
sys.path.insert(0, {root!r})
from pyscf_ipu_amd import molgen
from pyscf_ipu_amd.gto import METHANE, build_mol
from pyscf_ipu_amd.nanoDFT.pool import MoleculePool
mols = [build_mol(s) for s in [METHANE] + molgen.plain_set(3, 3, seed=11)]
with MoleculePool(workers=2) as pool:
    got = pool.run(mols, its=30, conv_tol=1e-9)
from pyscf_ipu_amd.nanoDFT.nanoDFT import RKS
ref = [RKS(m, level=0).run(its=30, conv_tol=1e-9) for m in mols]
print(json.dumps({{"got": [[r.energy, r.iterations, r.converged] for r in got],
                  "ref": [[r.energy, r.iterations, r.converged] for r in ref]}}))
"""


def test_pool_equals_serial(hip):
    print("pool: 2 workers, 4 molecules, in a child interpreter", flush=True)
    out = subprocess.run([sys.executable, "-c", SCRIPT.format(root=ROOT)], capture_output=True, text=True,
                         timeout=300)
    assert out.returncode == 0, out.stderr[-3000:]
    d = json.loads(out.stdout.strip().splitlines()[-1])
    for (eg, ig, cg), (er, ir, cr) in zip(d["got"], d["ref"]):
        assert cg and cr
        assert abs(eg - er) < 1e-9  # FP64 J/K sums in LDS-atomic order
        assert ig == ir
