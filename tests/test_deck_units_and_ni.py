"""Operator decks with physical units and the input / check operators of the reference's sample
decks: quantities convert to the internal units (angstrom, picosecond, Dalton, Da A^2/ps^2);
a deck written like contribs/microStamp/samples/benchmark_lj_snap/input_lj_Ni.msp (lattice,
deterministic gaussian_noise_r, LJ in eV, forces divided by the mass, check_values) run
through `default_simulation` passes check_values against the reference's golden file."""
import os

import numpy as np
import pytest

from exanbody_amd import XnbError
from exanbody_amd.deck import Deck

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden", "check_values_lj_Ni.dat")

NI_DECK = """
global:
  dt: 2e-3 ps
  rcut_inc: 2.0 ang
  rcut_max: 4.1 ang
  simulation_end_iteration: 100
  deterministic_noise: true

simulation: default_simulation

+first_iteration:
  - grid_stats
  - chunk_neighbors_stats

input_data:
  - particle_types:
      particle_type_map: { Ni: 0 }
  - particle_type_add_properties:
      Ni: { mass: 58.693 Da , z: 28 }
  - domain:
      cell_size: 13.92 ang
      grid_dims: [ 4 , 4 , 4 ]
      bounds: [ [ 0.0 um , 0.0 um , 0.0 um ] , [ 55.68 ang , 55.68 ang , 55.68 ang ] ]
      periodic: [ true , true , true ]
      expandable: false
      xform: [ [ 1 , 0 , 0 ] , [ 0 , 1 , 0 ] , [ 0 , 0 , 1 ] ]
  - init_rcb_grid
  - lattice:
      structure: FCC
      types: [ Ni , Ni , Ni , Ni ]
      size: [ 3.48 ang , 3.48 ang , 3.48 ang ]
  - gaussian_noise_r:
      sigma: 0.05 ang
  - replicate_domain

replicate_domain:
  repeat: [ 1 , 1 , 1 ]

lennard_jones_force:
  config: { epsilon: %(eps)s , sigma: 2.2808 ang }
  rcut: 4.1 ang

compute_all_forces_energy:
  - zero_particle_force: { ghost: true }
  - lennard_jones_force
  - update_force_from_ghost
  - divide_force_by_type_scalar: mass

simulation_epilog:
  - check_values:
      file: %(file)s
      samples: 100
      pos_threshold: 1.e-5
      vel_threshold: 1.e-5
      acc_threshold: 1.e-5
"""


@pytest.mark.parametrize("dt,expect", [("2e-3 ps", 0.002), ("2 fs", 0.002), ("1 ns", 1000.0), ("0.5", 0.5),
                                       ("1e-15 s", 1e-3)])
def test_time_quantities_convert_to_picoseconds(dt, expect):
    d = Deck("global:\n  dt: %s\n  physical_time: 0.0\n" % dt, files=())
    d.run("next_time_step")
    assert float(d.slot("physical_time")) == pytest.approx(expect, rel=1e-15)


def test_unknown_unit_is_reported():
    d = Deck("global:\n  dt: 2 parsec\n  physical_time: 0.0\n", files=())
    with pytest.raises(XnbError, match="unknown unit"):
        d.run("next_time_step")


@pytest.mark.gpu
def test_ni_sample_deck_passes_reference_check_values():
    d = Deck(NI_DECK % dict(eps="0.3729 eV", file=GOLDEN))
    log = d.run()
    print(log[-400:])
    assert "lattice FCC: 16384 particles" in log
    assert d.slot("timestep") == "101"
    for k in ("check_values_max_dr", "check_values_max_da", "check_values_max_dv"):
        assert float(d.slot(k)) < 1e-5, (k, d.slot(k))
    # the library's operators reproduce the golden as closely as the oracle does
    assert float(d.slot("check_values_max_da")) < 1e-7


@pytest.mark.gpu
def test_ni_sample_deck_check_values_fails_on_a_wrong_parameter():
    d = Deck(NI_DECK % dict(eps="0.3730 eV", file=GOLDEN))
    with pytest.raises(XnbError, match="check_values"):
        d.run()


@pytest.mark.gpu
def test_divide_force_by_type_scalar_bitwise():
    from exanbody_amd import Box, cubic_domain, generate_fcc
    r, v, ids, L = generate_fcc(6, 0.8442, jitter=0.05, temperature=0.5, seed=3)
    types = (np.arange(r.shape[0]) % 3).astype(np.uint8)
    dom = cubic_domain(L, 2.8)
    b = Box(dom["bounds_min"], dom["bounds_max"], dom["grid_dims"], dom["cell_size"])
    b.nbh_dist(2.5, 0.3)
    b.set_particles(r, v, ids, types)
    b.chunk_neighbors(2.8)
    b.lennard_jones_force(1.0, 1.0, 2.5)
    f0 = b.export()
    mass = np.array([1.5, 58.693, 0.3])
    b.divide_force_by_type_scalar(mass)
    f1 = b.export()
    want = f0["f"] / mass[f0["type"].astype(int)][:, None]
    np.testing.assert_array_equal(f1["f"].view(np.uint64), want.view(np.uint64))
    with pytest.raises(XnbError):
        b.divide_force_by_type_scalar(mass[:2])  # type 2 has no value
    b.close()


@pytest.mark.gpu
def test_check_values_creates_then_checks_its_own_file(tmp_path):
    """A missing reference file is written (hex floats); the same deck run again matches it
    exactly (deterministic noise, same operators)."""
    f = str(tmp_path / "ni_10_steps.dat")
    text = NI_DECK % dict(eps="0.3729 eV", file=f) + "\nglobal:\n  simulation_end_iteration: 10\n"
    d = Deck(text)
    assert "created" in d.run()
    assert sum(1 for line in open(f) if line.startswith("- [")) == 100
    d2 = Deck(text)
    d2.run()
    for k in ("check_values_max_dr", "check_values_max_da", "check_values_max_dv"):
        assert float(d2.slot(k)) == 0.0, (k, d2.slot(k))
