"""Host-side mirror of the reference interface (src/particles.jl,
src/update_particles_field.jl), checked without a GPU."""
import numpy as np
import pytest

import fmm4rb_hip as F
from fmm4rb_hip import workloads as W


def test_particles_construction():
    # test/test_particles.jl:2-18
    npar = 10
    rng = np.random.default_rng(0)
    pos, mom = rng.random((3, npar)), rng.random((3, npar))
    p = F.Particles(pos=pos, mom=mom)
    assert p.npar == npar
    assert np.array_equal(p.positions, pos.T) and np.array_equal(p.momenta, mom.T)
    assert np.array_equal(p.efields, np.zeros((npar, 3)))
    assert np.array_equal(p.bfields, np.zeros((npar, 3)))
    assert p.charge == -1.0 and p.mass == 1.0


def test_beam_alias():
    # test/test_particles.jl:20-28
    assert F.Beam is F.Particles
    b = F.Beam(pos=np.zeros((3, 10)), mom=np.zeros((3, 10)))
    assert isinstance(b, F.Particles)


def test_particles_asserts():
    # src/particles.jl:12-13
    with pytest.raises(AssertionError):
        F.Particles(pos=np.zeros((3, 4)), mom=np.zeros((3, 5)))
    with pytest.raises(AssertionError):
        F.Particles(pos=np.zeros((2, 4)), mom=np.zeros((2, 4)))


def test_algorithm_type_fields():
    a = F.FMMHIP(eta=0.5, N0=125, n=4)
    assert (a.n, a.N0, a.eta, a.precision) == (4, 125, 0.5, "fp64")
    assert F.FMMHIP(n=6, N0=125, eta=0.5, precision="mixed").precision == "mixed"


def test_unknown_algorithm_rejected_before_device_use():
    b = F.Beam(pos=np.zeros((3, 4)), mom=np.zeros((3, 4)))
    with pytest.raises(TypeError):
        F.update_particles_field(b, object(), lambda_=1.0)


def test_maxlevel_matches_oracle(orc):
    for N, N0 in [(8192, 125), (1 << 20, 125), (1 << 22, 125), (25600000, 125), (251, 125),
                  (4, 2), (11, 3), (1, 1), (1 << 13, 1 << 6)]:
        assert F.maxlevel(N, N0) == orc.maxlevel(N, N0)


def test_workloads_shapes():
    pos, mom = W.bunch(1000, seed=1)
    assert pos.shape == mom.shape == (1000, 3)
    assert abs(mom[:, 2].mean() - 100) < 0.5
    pos, mom = W.ellipsoid(2000, seed=1)
    r = (pos / np.array([1e-3, 1e-3, 1e-4])) ** 2
    assert np.all(r.sum(axis=1) <= 1.0 + 1e-12)
    assert W.rel_l2([1.0, 2.0], [1.0, 2.0]) == 0.0
