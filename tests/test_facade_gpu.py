"""The reference's Calculations and Coordinates suites, re-expressed on the
C++ Lattice facade (tests/cpp), run on the GPU."""
import pytest

from test_facade_cpu import run_suite

pytestmark = pytest.mark.gpu


def test_calculations_suite():
    # reference test/CalculationsTest.cpp: streaming, collision, full step,
    # reflections (with and without force), point/plane flow, compression
    out = run_suite("CalculationsTest")
    assert "0 failed" in out


def test_coordinates_suite():
    # reference test/CoordinatesTest.cpp: load/retrieve, 100^3 index
    # mapping host- and device-side, model flags, distributeVector,
    # reflection tables
    out = run_suite("CoordinatesTest")
    assert "0 failed" in out
