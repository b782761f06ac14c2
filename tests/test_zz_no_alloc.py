"""fmx_reduce_mass_static, the no-allocation variant of fmx_reduce_mass (partials in the library's
per-GPU static array instead of a caller workspace), against fmx_reduce_mass on the same inputs: the
same kernels in the same order, so the same bits.  Two calls in a row, the second on a smaller input
(the array keeps the first call's partials; the second must not see them), in every storage
precision; and the size limit is an FMX_ERR_ARG."""
import numpy as np
import pytest

from fluidmaxxer_amd import abi
from fluidmaxxer_amd.coupled import CoupledSim
from tests.gpu_util import host, say
from tests.test_gpu_coupled import small_config

pytestmark = pytest.mark.gpu


def _both(sim, n):
    import torch

    ws = sim._mass_workspace()
    a = torch.full((abi.MASS_COLS,), -1.0, dtype=torch.float64, device=sim.device)
    b = torch.full((abi.MASS_COLS,), -2.0, dtype=torch.float64, device=sim.device)
    args = (sim.prec, sim.phi.data_ptr(), sim.q.data_ptr(), sim.f[sim.cur].data_ptr(), n, sim.f_ref)
    abi.call("fmx_reduce_mass", *args, a.data_ptr(), ws.data_ptr(), ws.numel(), sim._s())
    abi.call("fmx_reduce_mass_static", *args, b.data_ptr(), sim._s())
    return host(a), host(b)


@pytest.mark.parametrize("prec", ["f32", "f64", "f16", "bf16"])
def test_static_mass_matches_workspace_mass(prec):
    cfg = small_config("langmuir_series", dims=(96, 40, 36))  # 138,240 cells: 34 partial blocks
    say(f"fmx_reduce_mass_static vs fmx_reduce_mass prec={prec} n={cfg.n}")
    sim = CoupledSim(cfg, prec=prec)
    sim.run(20)
    for n in (cfg.n, cfg.n // 7):  # the second call smaller: fewer partials than the first left behind
        a, b = _both(sim, n)
        assert np.array_equal(a, b), (n, a, b)
        assert a[3] == 0 and a[2] > 0
    big = (1 << 35) + 1
    rc = abi.lib().fmx_reduce_mass_static(sim.prec, None, None, None, big, 1.0, sim.q.data_ptr(), sim._s())
    assert rc == abi.ERR_ARG and b"static scratch" in abi.lib().fmx_last_error()
