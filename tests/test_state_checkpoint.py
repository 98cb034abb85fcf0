"""State checkpoints (OpenMM State XML, version 1) with the ATM global
parameters: the reference's equilibrated checkpoint format
(python/tests/temoa-g1-equil.xml, read by Simulation.loadState in
example/abfe/abfe.py:121) parsed and written back.

Fixture: tests/golden/temoa-g1-equil-excerpt.xml, the reference checkpoint cut
to 16 particles by tests/golden/make_state_excerpt.py (every kept line
unchanged)."""
import os

import numpy as np
import pytest

from atmmetaforce import openmm as mm

HERE = os.path.dirname(os.path.abspath(__file__))
EXCERPT = os.path.join(HERE, "golden", "temoa-g1-equil-excerpt.xml")


def load():
    return mm.XmlSerializer.deserialize(open(EXCERPT).read())


def test_reference_checkpoint_values():
    s = load()
    assert isinstance(s, mm.State)
    assert s.getTime() == 399.99999999745313
    a, b, c = s.getPeriodicBoxVectors()
    assert tuple(a) == (4.217498273277886, 0.0, 0.0)
    assert tuple(b) == (0.0, 4.659005774446288, 0.0)
    assert tuple(c) == (0.0, 0.0, 4.3635687055440755)
    p = s.getParameters()
    # the ATM globals of the checkpoint, including OpenMM's leading-dot number format
    assert p == {"ATMAcore": 0.0625, "ATMAlpha": 0.0, "ATMLambda1": 0.5000000000000003,
                 "ATMLambda2": 0.5000000000000003, "ATMU0": 0.0, "ATMUbcore": 2092.0, "ATMUmax": 4184.0,
                 "ATMW0": 0.0, "MonteCarloPressure": 1.0, "MonteCarloTemperature": 300.0}
    x = s.getPositions(asNumpy=True)
    v = s.getVelocities(asNumpy=True)
    assert x.shape == (16, 3) and v.shape == (16, 3)
    assert tuple(x[0]) == (1.9480619430541992, 2.921112298965454, 2.4037277698516846)
    assert tuple(x[15]) == tuple(float(t) for t in _last("Position"))
    assert tuple(v[15]) == tuple(float(t) for t in _last("Velocity"))


def _last(tag):
    import re
    m = re.findall(r'<%s x="([^"]+)" y="([^"]+)" z="([^"]+)"/>' % tag, open(EXCERPT).read())
    return m[-1]


def test_checkpoint_round_trip_is_exact():
    s = load()
    xml = mm.XmlSerializer.serialize(s)
    assert xml.startswith('<?xml version="1.0" ?>\n<State ')
    assert 'type="State"' in xml and 'version="1"' in xml
    t = mm.XmlSerializer.deserialize(xml)
    assert t.getTime() == s.getTime()
    assert t.getParameters() == s.getParameters()
    assert np.array_equal(t.getPositions(asNumpy=True), s.getPositions(asNumpy=True))
    assert np.array_equal(t.getVelocities(asNumpy=True), s.getVelocities(asNumpy=True))
    assert [tuple(v) for v in t.getPeriodicBoxVectors()] == [tuple(v) for v in s.getPeriodicBoxVectors()]
    with pytest.raises(Exception, match="does not contain forces|Forces"):
        t.getForces()


def test_bad_checkpoints():
    xml = open(EXCERPT).read()
    with pytest.raises(Exception, match="Unsupported version"):
        mm.XmlSerializer.deserialize(xml.replace('version="1"', 'version="2"', 1))
    with pytest.raises(Exception, match="not a number"):
        mm.XmlSerializer.deserialize(xml.replace('ATMLambda1=".5000000000000003"', 'ATMLambda1="half"'))
    with pytest.raises(Exception):
        mm.XmlSerializer.deserialize(xml[: len(xml) // 2])


def test_barostat_parameter_holder():
    b = mm.MonteCarloBarostat(1.0, 300.0, 10000000)
    assert b.Pressure() == "MonteCarloPressure" and mm.MonteCarloBarostat.Temperature() == "MonteCarloTemperature"
    assert (b.getDefaultPressure(), b.getDefaultTemperature(), b.getFrequency()) == (1.0, 300.0, 10000000)
    with pytest.raises(Exception, match="negative"):
        b.setFrequency(-1)
