"""The entry points the round-end checks call, each in a fresh process: smoke() loads the HIP
library through the geometry generator before anything else touches the GPU."""
import os
import subprocess
import sys

import pytest

from tests.gpu_util import say

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_smoke_in_fresh_process():
    say("smoke() in a fresh process")
    r = subprocess.run([sys.executable, "-c", "import __graft_entry__ as g; g.smoke()"], cwd=ROOT,
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    assert "smoke ok" in r.stdout, r.stdout


def test_library_loaded_before_torch():
    # the library loaded by a CPU entry point first, torch afterwards, then a GPU run
    say("library before torch, then a fused run")
    code = ("from fluidmaxxer_amd import geometry\n"
            "p = geometry.sphere_pack(24, 12, 10, 3, 2.0, 3.0, 0.6, overlap=False)\n"
            "import numpy as np\n"
            "from fluidmaxxer_amd.coupled import CoupledConfig, CoupledSim\n"
            "cfg = CoupledConfig(24, 12, 10, solid=p.solid, region=p.sphere)\n"
            "sim = CoupledSim(cfg, prec='f32')\n"
            "c = sim.run(20).cpu().numpy()\n"
            "assert np.all(np.isfinite(c))\n"
            "print('ok')\n")
    r = subprocess.run([sys.executable, "-c", code], cwd=ROOT, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0 and "ok" in r.stdout, r.stdout[-2000:] + r.stderr[-2000:]
