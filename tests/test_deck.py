"""Operator decks: YAML-subset reader and slot store on the CPU; deck-driven runs on the GPU
equal the oracle's operator sequence bit for bit."""
import numpy as np
import pytest

from exanbody_amd import XnbError
from exanbody_amd.deck import Deck, domain_block


def test_deck_parses_reference_style_structure():
    d = Deck("""
global:
  dt: 0.002            # comment
  rcut_max: 4.1
  name: "a b # c"
+first_iteration:
  - message: { text: extra }
lennard_jones_force:
  config: { epsilon: 0.5 , sigma: [1, 2] }
""")
    assert d.slot("dt") == "0.002"
    assert d.slot("rcut_max") == "4.1"
    assert d.slot("name") == "a b # c"
    assert d.slot("simulation_end_iteration") == "100"  # from the bundled step-loop deck


def test_deck_reports_unknown_operator_without_gpu():
    d = Deck("simulation:\n  - no_such_operator\n", files=())
    with pytest.raises(XnbError, match="unknown operator"):
        d.run()


def test_deck_bad_yaml():
    with pytest.raises(XnbError):
        Deck("global:\n  x: {a: 1\n", files=())


@pytest.mark.gpu
@pytest.mark.parametrize("trigger,rho", [("displacement", 0.8442), ("interval", 0.8442), ("interval", 1.6)])
def test_deck_run_matches_oracle(trigger, rho):
    """rho 1.6: about 62 atoms per cell, so the deck's rebuild_amr step re-sorts every cell."""
    from exanbody_amd import cubic_domain, generate_fcc
    from oracle.oracle import Oracle
    if rho > 1.0:
        r, v, ids, L = generate_fcc(10, rho, jitter=0.05, temperature=0.72, seed=19)
    else:
        r, v, ids, L = generate_fcc(7, rho, jitter=0.05, temperature=1.3, seed=17)
    dom = cubic_domain(L, 2.8)
    extra = ""
    if trigger == "interval":
        extra = """
trigger_move_particles:
  rebind: { freq: rebuild_freq, result: trigger_move_particles }
  body:
    - nth_timestep
"""
    d = Deck(domain_block(dom["bounds_min"], dom["bounds_max"], dom["grid_dims"], dom["cell_size"]),
             "global:\n  simulation_end_iteration: 40\n  rebuild_freq: 10\n" + extra)
    d.set_particles(r, v, ids)
    d.run()
    assert d.slot("timestep") == "41"  # advanced after each of the 40 steps, from 1
    o = Oracle(dom["bounds_min"], dom["bounds_max"], dom["grid_dims"], dom["cell_size"], ghost_dist=2.8)
    o.set_particles(r, v, ids)
    o.init(1.0, 1.0, 2.5, 0.3)
    o.run(40, 0.005, 1.0, 1.0, 2.5, 0.3, rebuild_every=10 if trigger == "interval" else 0)
    oe, ge = o.export(), d.box().export()
    np.testing.assert_array_equal(oe["id"], ge["id"])
    for k in ("r", "v", "f"):
        np.testing.assert_array_equal(oe[k].view(np.uint64), ge[k].view(np.uint64), err_msg=k)
    ke, pe = o.energies()
    assert float(d.slot("kinetic_energy")) == pytest.approx(ke, rel=1e-12)
    assert float(d.slot("potential_energy")) == pytest.approx(pe, rel=1e-12)
