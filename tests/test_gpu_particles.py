"""GPU tracer particles against the CPU restatement: lowbias32 seeding is
integer + single IEEE multiplies, advection is compiled without contraction,
so both must be bit-identical to the oracle."""
import sys

import numpy as np
import pytest

import foil_cfd_amd as fc
from foil_cfd_amd import cases
from oracle import oracle as orc

pytestmark = pytest.mark.gpu


def test_seed_bit_exact_full_pool(gpu):
    print("case: particle seed 1M on 512x192x240", flush=True)
    case = cases.config3_wing()
    s = cases.make_solver(case, "f16")
    s.particles_seed(1_000_000, 0xF01DCAFE)
    got = s.particles_read()
    want = orc.particles_seed(1_000_000, case.dims, 0xF01DCAFE)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
    s.close()


def test_advect_bit_exact(gpu):
    print("case: particle advection tunnel 192x64x16", flush=True)
    case = cases.naca_tunnel(192, 64, 16, 48, aoa=8.0, cs=0.1, ramp=False)
    s = cases.make_solver(case, "f32")
    s.step_n(200)
    s.synchronize()
    rho, u, v, w = s.read_macro()
    s.particles_seed(200_000, 0xF01DCAFE)
    pos = s.particles_read()
    for frame in range(1, 6):
        seed = (frame * 2654435761) & 0xFFFFFFFF
        s.particles_advect(8.0, seed)
        orc.particles_advect(pos, u, v, w, case.flags, case.dims, 8.0, seed)
    got = s.particles_read()
    mismatch = (got.view(np.uint32) != pos.view(np.uint32)).any(axis=1)
    print(f"particle mismatches: {mismatch.sum()} of {len(pos)}", flush=True)
    assert mismatch.sum() == 0
    s.close()
