"""The streaming StochKit reader (tau_leaping_amd/csrc/tl_stochkit.cpp) against
an independent parse of the same file.

The expected model is built from Python's own XML parser (xml.etree) with the
reference's rules, StochKitReader.cpp:35-179 and CSCHostMatrixData.cpp:19-107:
species indexed in SpeciesList order, every reaction a CSC column with its
entries sorted by species, a species that is both reactant and product giving
a reactant entry (-v) then a product entry (+v).  That expected model goes
through tl_model_from_csc, whose layouts the oracle tests pin, and the two
models must agree in every array.  The files exercise what a DOM reader never
sees but a streaming one must get right: text, comments and CDATA longer than
the read window, markup split across refills, entities, declarations,
self-closing tags, and the selection rules for repeated elements."""
import os
import subprocess
import time
import xml.etree.ElementTree as ET

import numpy as np
import pytest

import tau_leaping_amd as T

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GEN = os.path.join(ROOT, "tau_leaping_amd", "bin", "toy_system_gen")


def expected_model(path):
    """Model built from xml.etree's parse of a StochKit file."""
    model = ET.parse(path).getroot()
    assert model.tag == "Model"
    params = {}
    for p in model.find("ParametersList").findall("Parameter"):
        params[p.find("Id").text.strip()] = float(p.find("Expression").text.strip())
    names, pops, idx = [], [], {}
    for s in model.find("SpeciesList").findall("Species"):
        sid = s.find("Id").text.strip()
        pop = int(s.find("InitialPopulation").text.strip())
        if sid in idx:  # a repeated Id keeps its first index and takes the last population
            pops[idx[sid]] = pop
            continue
        idx[sid] = len(names)
        names.append(sid)
        pops.append(pop)
    col_ptr, rows, vals, rates = [0], [], [], []
    for r in model.find("ReactionsList").findall("Reaction"):
        sides = []
        for tag in ("Reactants", "Products"):
            lst = r.find(tag)
            d = {}
            if lst is not None:
                for sr in lst.findall("SpeciesReference"):
                    d[sr.get("id")] = int(sr.get("stoichiometry"))  # a repeat keeps the last value
            sides.append(d)
        ent = sorted([(idx[k] * 2, -v) for k, v in sides[0].items()] + [(idx[k] * 2 + 1, v) for k, v in sides[1].items()])
        rows += [e[0] // 2 for e in ent]
        vals += [e[1] for e in ent]
        col_ptr.append(len(rows))
        rates.append(params[r.find("Rate").text.strip()])
    m = T.Model.from_csc(len(names), np.array(col_ptr, np.int32), np.array(rows, np.int32), np.array(vals, np.int8),
                         np.array(rates), np.array(pops, np.int32))
    return m, names


def assert_same(path):
    got = T.Model.from_stochkit(str(path))
    ref, names = expected_model(str(path))
    assert (got.n_species, got.n_channels, got.nnz) == (ref.n_species, ref.n_channels, ref.nnz)
    for a, b in zip(got.structure(), ref.structure()):
        np.testing.assert_array_equal(a, b)
    np.testing.assert_array_equal(got.rates, ref.rates)
    np.testing.assert_array_equal(got.x0, ref.x0)
    assert got.species_names == names
    return got


def stochkit(reactions, species, params=None, head="", between="", n_reactions=None):
    """A small StochKit document; reactions are (rate_id, reactants, products)
    with [(species, stoichiometry)] lists, species are (id, population)."""
    params = params if params is not None else sorted({r[0] for r in reactions})
    out = [head, "<Model>\n<Description>t</Description>\n"]
    out.append(f"<NumberOfReactions>{len(reactions) if n_reactions is None else n_reactions}</NumberOfReactions>\n")
    out.append(f"<NumberOfSpecies>{len({s[0] for s in species})}</NumberOfSpecies>\n<ParametersList>\n")
    for i, p in enumerate(params):
        pid, val = p if isinstance(p, tuple) else (p, 0.5 + i)
        out.append(f"<Parameter><Id>{pid}</Id><Expression>{val!r}</Expression></Parameter>\n")
    out.append("</ParametersList>\n<ReactionsList>\n")
    for j, (rate, reac, prod) in enumerate(reactions):
        out.append(f"<Reaction><Id>R{j}</Id><Type>mass-action</Type><Rate>{rate}</Rate><Reactants>")
        out += [f'<SpeciesReference id="{s}" stoichiometry="{v}"/>' for s, v in reac]
        out.append("</Reactants><Products>")
        out += [f'<SpeciesReference id="{s}" stoichiometry="{v}"/>' for s, v in prod]
        out.append("</Products></Reaction>\n" + between)
    out.append("</ReactionsList>\n<SpeciesList>\n")
    for sid, pop in species:
        out.append(f"<Species><Id>{sid}</Id><InitialPopulation>{pop}</InitialPopulation></Species>\n")
    out.append("</SpeciesList>\n</Model>\n")
    return "".join(out)


DIMER = ([("c1", [("S1", 1)], []), ("c2", [("S1", 2)], [("S2", 1)]), ("c3", [("S2", 1)], [("S1", 2)]),
          ("c4", [("S2", 1)], [("S3", 1)])],
         [("S1", 100000), ("S2", 0), ("S3", 0)])


def test_generated_networks(tmp_path):
    # the reference generator's normal family (types 1-3, product stoichiometry 2), cyclic chain, colloid
    for i, args in enumerate((["-r", "300", "-c", "900", "--seed", "11"], ["-r", "40", "-w", "--seed", "2"],
                              ["-r", "30", "-s", "--seed", "5", "--log-rates", "1e-4", "0.1"])):
        p = tmp_path / f"g{i}.xml"
        with open(p, "w") as f:
            subprocess.run([GEN] + args, stdout=f, stderr=subprocess.DEVNULL, check=True)
        assert_same(p)


def test_markup_variants(tmp_path):
    """Comments inside elements, CDATA, entities, single-quoted attributes,
    declarations, processing instructions, spacing in end tags, a species on
    both sides (catalyst), species declared after use."""
    txt = ('<?xml version="1.0"?>\n<!DOCTYPE Model>\n<!-- lead -->\n<Model>'
           "<NumberOfReactions> 3 </NumberOfReactions ><NumberOfSpecies>3</NumberOfSpecies>"
           "<ParametersList><Parameter><Id>k&amp;1</Id><!-- c --><Expression><![CDATA[ 0.25 ]]></Expression>"
           "</Parameter><Parameter><Id>k2</Id><Expression>1.5e-3</Expression></Parameter></ParametersList>"
           "<ReactionsList><?pi x?>"
           "<Reaction><Rate>k&amp;1</Rate><Reactants><SpeciesReference id='A' stoichiometry='1'/>"
           "<SpeciesReference id=\"B\" stoichiometry=\"1\"></SpeciesReference></Reactants>"
           "<Products><SpeciesReference id='A' stoichiometry='1'/><SpeciesReference id='C' stoichiometry='1'/>"
           "</Products></Reaction>"
           "<Reaction><Rate>k2</Rate><Reactants><SpeciesReference id='C' stoichiometry='2'/></Reactants></Reaction>"
           "<Reaction><Rate>k2</Rate><Reactants><SpeciesReference id='B' stoichiometry='1'/></Reactants>"
           "<Products><SpeciesReference id='A' stoichiometry='2'/></Products></Reaction>"
           "</ReactionsList><SpeciesList>"
           "<Species><Id>C</Id><InitialPopulation>7</InitialPopulation></Species>"
           "<Species><Id> A </Id><InitialPopulation>&#53;</InitialPopulation></Species>"
           "<Species><Id>B</Id><InitialPopulation>9</InitialPopulation></Species>"
           "</SpeciesList></Model>\n<!-- trailing -->\n")
    p = tmp_path / "v.xml"
    p.write_text(txt)
    m = assert_same(p)
    assert m.species_names == ["C", "A", "B"]
    np.testing.assert_array_equal(m.x0, [7, 5, 9])
    np.testing.assert_array_equal(m.rates, [0.25, 1.5e-3, 1.5e-3])


def test_longer_than_the_read_window(tmp_path):
    """Character data, a comment and CDATA of several MiB (the reader's window
    is 1 MiB), and a tag straddling a refill boundary at every offset."""
    big = "x" * (3 << 20)
    reac, spec = DIMER
    txt = stochkit(reac, spec, head=f"<!-- {big} -->\n", between=f"<!--{big[:1 << 20]}-->")
    txt = txt.replace("<Description>t</Description>", f"<Description>{big}</Description>")
    txt = txt.replace("<Expression>0.5</Expression>", f"<Expression><![CDATA[{' ' * (2 << 20)}0.5]]></Expression>")
    p = tmp_path / "w.xml"
    p.write_text(txt)
    m = assert_same(p)
    assert m.rates[0] == 0.5
    # shift the whole document so markup crosses the first refill at many offsets
    base = stochkit(reac, spec)
    for pad in range((1 << 20) - len(base) - 8, (1 << 20) + 8, 3):
        if pad < 0:
            continue
        q = tmp_path / "s.xml"
        q.write_text(" " * pad + base)
        got = T.Model.from_stochkit(str(q))
        assert (got.n_species, got.n_channels, got.nnz) == (3, 4, 7)


def test_repeated_elements(tmp_path):
    """Selection rules: a repeated parameter Id takes the last value; a repeated
    species Id keeps its first index and takes the last population; a repeated
    SpeciesReference keeps the last stoichiometry; only the first Rate /
    ParametersList counts; later roots are ignored."""
    reac = [("c1", [("S1", 1), ("S1", 2)], [("S2", 1)]), ("c2", [("S2", 1)], [])]
    spec = [("S1", 10), ("S2", 20), ("S1", 30)]
    txt = stochkit(reac, spec, params=[("c1", 1.0), ("c2", 2.0), ("c1", 3.0)])
    txt = txt.replace("<Rate>c2</Rate>", "<Rate>c2</Rate><Rate>c1</Rate>")
    txt = txt.replace("</ParametersList>", "</ParametersList><ParametersList><Parameter><Id>c2</Id>"
                      "<Expression>9</Expression></Parameter></ParametersList>")
    txt += "<Model><NumberOfReactions>0</NumberOfReactions></Model>\n"
    p = tmp_path / "r.xml"
    p.write_text(txt)
    m = T.Model.from_stochkit(str(p))
    assert m.species_names == ["S1", "S2"]
    np.testing.assert_array_equal(m.x0, [30, 20])
    np.testing.assert_array_equal(m.rates, [3.0, 2.0])
    order, s0, s1, _, _ = m.structure()
    assert order[0] == 3 and s0[0] == 0  # 2 S1 -> S2: the second reference wins
    ET.fromstring(txt[:txt.index("</Model>") + 8])  # the first document is well formed


@pytest.mark.parametrize("mutate", [
    lambda t: t.replace("<!-- lead -->", "<!-- lead "),
    lambda t: t.replace("</Reaction>", "</Reactions>", 1),
    lambda t: t.replace("<Rate>c1</Rate>", "", 1),
    lambda t: t.replace("<Reactants>", "<Reactant>", 1).replace("</Reactants>", "</Reactant>", 1),
    lambda t: t.replace('id="S3"', 'id="S9"', 1),
    lambda t: t.replace('stoichiometry="1"', "stoichiometry=1", 1),
    lambda t: t.replace("</Model>", ""),
    lambda t: t + "junk",
    lambda t: t.replace("<Expression>0.5</Expression>", "<Expression>fast</Expression>"),
    lambda t: t.replace("<NumberOfReactions>4", "<NumberOfReactions>5"),
])
def test_malformed(tmp_path, mutate):
    reac, spec = DIMER
    good = stochkit(reac, spec, head="<!-- lead -->\n")
    p = tmp_path / "ok.xml"
    p.write_text(good)
    T.Model.from_stochkit(str(p))
    p = tmp_path / "bad.xml"
    p.write_text(mutate(good))
    with pytest.raises(T.TLError) as e:
        T.Model.from_stochkit(str(p))
    assert e.value.status == 2


def test_large_generated_model(tmp_path):
    """1e5 channels over 2e4 species from the generator: the reader keeps up
    (seconds) and agrees with the independent parse."""
    p = tmp_path / "big.xml"
    t0 = time.time()
    with open(p, "w") as f:
        subprocess.run([GEN, "-r", "20000", "-c", "100000", "--toy", "--seed", "3", "--log-rates", "0.001", "1"],
                       stdout=f, stderr=subprocess.DEVNULL, check=True)
    t_gen = time.time() - t0
    t0 = time.time()
    m = T.Model.from_stochkit(str(p))
    t_read = time.time() - t0
    assert (m.n_species, m.n_channels) == (20000, 100000)
    assert t_gen < 10 and t_read < 10, (t_gen, t_read)
    assert_same(p)


def test_extended_forms_round_trip(tmp_path):
    """Sources (an empty Reactants list) and channels with three reactant
    molecules read back as orders 0, 4, 5 and 6."""
    from helpers import trimolecular_net

    net = trimolecular_net()
    p = tmp_path / "tri.xml"
    p.write_text(net.to_stochkit_xml())
    m = assert_same(p)
    order = list(m.structure()[0])
    assert order[:7] == [4, 1, 5, 1, 4, 1, 0] and order[-2:] == [5, 6]
