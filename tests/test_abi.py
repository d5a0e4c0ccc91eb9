"""C ABI surface of libtau_leaping_amd.so on the CPU (no GPU compute calls):
every entry point declared in include/tl_api.h is exported, models build and
validate with the documented status codes, and the StochKit reader agrees
with the channel-major assembly of the same network."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import tau_leaping_amd as T
import tl_oracle as O
from helpers import oracle_of
from tau_leaping_amd import models
from tau_leaping_amd.models import Network, Reaction

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_EXAMPLES = "/root/reference/examples"


def _declared():
    txt = open(os.path.join(ROOT, "include", "tl_api.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(tl_[a-z0-9_]+)\s*\(", txt)))


def test_every_declared_symbol_is_exported():
    names = _declared()
    assert len(names) >= 25
    lib = C.CDLL(T.LIB_PATH)
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_version_and_errors():
    assert b"gfx950" in T.lib().tl_version()
    with pytest.raises(T.TLError) as e:
        T.Model.from_stochkit("/definitely/not/here.xml")
    assert e.value.status == 1


def test_model_structure_matches_oracle():
    for net in (models.dimer_decay(), models.toy_network(40, 90, seed=1), models.colloidal(30, seed=2),
                models.cyclic_chain(50, seed=3)):
        m = T.Model.from_network(net)
        om = oracle_of(net)
        order, s0, s1, rp, csr = m.structure()
        assert np.array_equal(order, om.order)
        assert np.array_equal(s0, om.s0) and np.array_equal(s1, om.s1)
        assert np.array_equal(rp, om.row_ptr) and np.array_equal(csr, om.csr_packed)
        assert m.n_species == net.n_species and m.n_channels == net.n_channels
        np.testing.assert_array_equal(m.rates, [r.rate for r in net.reactions])
        np.testing.assert_array_equal(m.x0, net.x0)


def test_net_stoichiometry_layout_matches_oracle():
    # the engine merges per column, the oracle per species row; both must
    # give the same packed rows, and switching back restores the entries as read
    for net in (models.gene_expression(), models.dimer_decay(), models.toy_network(40, 90, seed=1)):
        m = T.Model.from_network(net)
        lit = m.structure()
        assert not m.net_stoichiometry and m.layout_nnz == m.nnz
        m.net_stoichiometry = True
        assert m.net_stoichiometry
        om = O.OracleModel(net.n_species, net.n_channels, *net.to_csc(), net_stoichiometry=True)
        order, s0, s1, rp, csr = m.structure()
        # reaction orders come from the entries as read
        assert np.array_equal(order, lit[0]) and np.array_equal(s0, lit[1]) and np.array_equal(s1, lit[2])
        assert np.array_equal(rp, om.row_ptr) and np.array_equal(csr, om.csr_packed)
        assert m.layout_nnz == rp[-1] and m.nnz == lit[3][-1]
        m.net_stoichiometry = False
        back = m.structure()
        assert all(np.array_equal(u, v) for u, v in zip(back, lit))
    ge = T.Model.from_network(models.gene_expression())
    ge.net_stoichiometry = True
    # G (catalyst of transcription) and the dimer D (catalyst of repression)
    # lose those entries; 2P -> P keeps one entry of -1
    assert ge.nnz - ge.layout_nnz == 7


def test_net_stoichiometry_range_error():
    # two product entries of one species: 7 + 7 leaves the 4-bit range [-3, 12] once merged
    m = T.Model.from_csc(2, [0, 3], [0, 1, 1], [-1, 7, 7], [1.0], [10, 0])
    with pytest.raises(T.TLError) as e:
        m.net_stoichiometry = True
    assert e.value.status == 4
    assert not m.net_stoichiometry and m.layout_nnz == 3


def test_colloid_size():
    net = models.colloidal(250, seed=44)
    m = T.Model.from_network(net)
    assert (m.n_species, m.n_channels, m.nnz) == (250, 31250, 93500)


def test_validation_status_codes():
    def mk(reacs, x0=(1, 2, 3)):
        return T.Model.from_network(Network(3, reacs, np.array(x0, np.int32)))

    with pytest.raises(T.TLError) as e:
        mk([Reaction({0: 2, 1: 2}, {}, 1.0)])  # four reactant molecules
    assert e.value.status == 3
    order, s0, s1, _, _ = mk([Reaction({}, {0: 1}, 1.0), Reaction({0: 1}, {}, 1.0)]).structure()
    assert list(order) == [0, 1]  # a source is order 0
    assert list(mk([Reaction({0: 3}, {1: 1}, 1.0)]).structure()[0]) == [6]  # 3S
    with pytest.raises(T.TLError) as e:
        mk([Reaction({0: 1}, {1: 13}, 1.0)])  # stoichiometry 13 does not pack in 4 bits
    assert e.value.status == 4
    with pytest.raises(T.TLError) as e:
        mk([Reaction({0: 4}, {1: 1}, 1.0)])  # four molecules of one species
    assert e.value.status in (3, 4)
    with pytest.raises(T.TLError) as e:
        mk([Reaction({0: 1}, {1: 1}, -1.0)])
    assert e.value.status == 9
    with pytest.raises(T.TLError) as e:
        mk([Reaction({0: 1}, {1: 1}, 1.0)], x0=(1, -2, 3))
    assert e.value.status == 9


def test_setters_and_names():
    net = models.dimer_decay()
    m = T.Model.from_network(net)
    assert m.species_names == ["S1", "S2", "S3"]
    assert m.species_index("S2") == 1
    m.rates = [2.0, 0.001, 0.25, 0.08]
    np.testing.assert_array_equal(m.rates, [2.0, 0.001, 0.25, 0.08])
    m.x0 = [5, 6, 7]
    np.testing.assert_array_equal(m.x0, [5, 6, 7])
    with pytest.raises(T.TLError):
        m.species_index("nope")


def test_stochkit_roundtrip(tmp_path):
    for net in (models.dimer_decay(12345), models.toy_network(20, 45, seed=6), models.colloidal(12, seed=1)):
        p = tmp_path / f"{net.name}.xml"
        p.write_text(net.to_stochkit_xml())
        m = T.Model.from_stochkit(str(p))
        ref = T.Model.from_network(net)
        for a, b in zip(m.structure(), ref.structure()):
            assert np.array_equal(a, b)
        np.testing.assert_array_equal(m.rates, ref.rates)
        np.testing.assert_array_equal(m.x0, ref.x0)


@pytest.mark.parametrize("text,status", [
    ("<Model></Model>", 2),
    ("<Model><NumberOfReactions>1</NumberOfReactions><NumberOfSpecies>1</NumberOfSpecies>"
     "<ParametersList><Parameter><Id>c1</Id><Expression>1</Expression></Parameter></ParametersList>"
     "<ReactionsList><Reaction><Rate>c9</Rate><Reactants/></Reaction></ReactionsList>"
     "<SpeciesList><Species><Id>S1</Id><InitialPopulation>3</InitialPopulation></Species></SpeciesList></Model>", 2),
    ("<Model><NumberOfReactions>2</NumberOfReactions><NumberOfSpecies>1</NumberOfSpecies>"
     "<ParametersList><Parameter><Id>c1</Id><Expression>1</Expression></Parameter></ParametersList>"
     "<ReactionsList><Reaction><Rate>c1</Rate><Reactants><SpeciesReference id=\"S1\" stoichiometry=\"1\"/>"
     "</Reactants></Reaction></ReactionsList>"
     "<SpeciesList><Species><Id>S1</Id><InitialPopulation>3</InitialPopulation></Species></SpeciesList></Model>", 2),
    ("<Model><NumberOfReactions>1</NumberOfReactions>", 2),
])
def test_reader_errors(tmp_path, text, status):
    p = tmp_path / "bad.xml"
    p.write_text(text)
    with pytest.raises(T.TLError) as e:
        T.Model.from_stochkit(str(p))
    assert e.value.status == status


@pytest.mark.skipif(not os.path.isdir(REF_EXAMPLES), reason="reference examples not mounted")
def test_reference_examples_parse():
    """The reference's three example models (examples/*.xml)."""
    expect = {"dimer_decay.xml": (3, 4, 7), "test_128.xml": (128, 128, 301), "test_1024.xml": (1024, 1024, 2325)}
    for f, (n, mch, nnz) in expect.items():
        m = T.Model.from_stochkit(os.path.join(REF_EXAMPLES, f))
        assert (m.n_species, m.n_channels, m.nnz) == (n, mch, nnz)
    d = T.Model.from_stochkit(os.path.join(REF_EXAMPLES, "dimer_decay.xml"))
    np.testing.assert_array_equal(d.rates, [1.0, 0.002, 0.5, 0.04])
    np.testing.assert_array_equal(d.x0, [10000, 0, 0])
    ref = T.Model.from_network(models.dimer_decay(10000))
    for a, b in zip(d.structure(), ref.structure()):
        assert np.array_equal(a, b)


def test_simulate_without_gpu_fails_loudly():
    import torch

    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    m = T.Model.from_network(models.dimer_decay())
    with pytest.raises(T.TLError) as e:
        m.simulate(1.0, 10)
    assert e.value.status == 10
